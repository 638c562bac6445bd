"""The drop-in boundary, driven the way the reference drives its own backends.

- The factories are wired as ``run_hysteresis_loop`` wires the JAX ones
  (``src/hysteresis_loop.py:309-372``): ``make_energy_kernels``, ``make_solve_U`` with the
  Dirichlet mask, ``compute_exchange_diagonal``, ``make_minimizer(solve_U=...)``; the geometry is
  the reference's default ``TetGeom`` (``src/loop.py:899-908``): oriented connectivity, volumes and
  J^-T, no node coordinates. ``minimize`` gets the sweep's full keyword set
  (``src/hysteresis_loop.py:504-537``). Every result is checked against the oracle.
- ``make_poisson_ops`` and ``make_pcg_solve`` (``src/poisson_solve.py:124-480``) against the
  oracle's assembled operators and a direct solve.
- ``mm_minimize`` called through the plain ctypes stub of INTEGRATION.md (no package code), as
  ``cpp_minimize`` is called (``src/cpp_minimizer.py:78-193``), against the oracle's field step.
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from conftest import ndfeb_materials, oracle_levels, tilted_m0

pytestmark = pytest.mark.gpu

H_DIR = np.array([0.017452, 0.0, 0.999848])


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-300))


def sweep_kwargs(method, L=None):
    """The keyword set run_hysteresis_loop passes to minimize (LoopParams defaults, tau_f and
    cg_tol of the cube workload)."""
    return dict(gamma=0.0, max_iter=2000, tau_f=1e-10, eps_a=1e-12, tau0=1e-2, tau_min=1e-12, tau_max=1e3,
                ls_eta1=0.1, ls_eta2=0.1, ls_C=2.0, ls_c=0.5, ls_s0=1.0, ls_max_evals=15, h=H_DIR, mfinal=None,
                Js_ref=1.61, verbose=False, pc_iters=10, pc_auto=True, pc_force_eta=0.5, pc_force_alpha=0.5,
                pc_stagnation_nu=0.01, memory=5, tn_iters=10, lr=1e-3, mu=0.9, pc_reg=0.0, phi_extrapolate=False,
                L=L, sparse_ops={})


@pytest.fixture(scope="module")
def wired(cube_c0):
    """The factories on the reference's default geometry, plus the oracle on the same mesh."""
    from mumag_matrixfree2_amd.energy_kernels import compute_exchange_diagonal, make_energy_kernels
    from mumag_matrixfree2_amd.fem_utils import TetGeom, clear_context_cache
    from mumag_matrixfree2_amd.minimizers import make_minimizer
    from mumag_matrixfree2_amd.poisson_solve import make_solve_U
    from oracle.fem import Problem, compute_volume_JinvT

    mesh = cube_c0
    knt, ijk, bnd = mesh["knt"], mesh["ijk"], mesh["boundary_mask"]
    A, K1, Js, k = ndfeb_materials()
    conn32, volume, JinvT = compute_volume_JinvT(knt, ijk[:, :4])
    mat_id = ijk[:, 4].astype(np.int32)
    geom = TetGeom(conn=conn32, volume=volume, mat_id=mat_id, JinvT=JinvT, x_nodes=None)
    prob = Problem(knt, ijk, A, K1, Js, k, boundary_mask=bnd)
    V_mag = float(volume[np.isin(mat_id, np.where(Js > 0)[0] + 1)].sum())
    M_nodal = prob.M_nodal
    clear_context_cache()
    eg, e_only, g_only, loc = make_energy_kernels(geom, A_lookup=A, K1_lookup=K1, Js_lookup=Js, k_easy_lookup=k,
                                                  V_mag=V_mag, M_nodal=M_nodal, chunk_elems=200_000,
                                                  assembly="scatter", grad_backend="stored_JinvT", mode="assembled")
    solve_U = make_solve_U(geom, Js, precond_type="amgcl", order=1, chunk_elems=200_000, cg_maxiter=2000,
                           cg_tol=1e-10, poisson_reg=1e-12, grad_backend="stored_JinvT", boundary_mask=bnd,
                           mode="assembled", A_sparse=None, cpu_spmv_backend=None, poisson_solver="pcg")
    d_diag = compute_exchange_diagonal(geom, A, V_mag, chunk_elems=200_000, assembly="scatter",
                                       grad_backend="stored_JinvT")
    mins = {meth: make_minimizer(geom, A_lookup=A, K1_lookup=K1, Js_lookup=Js, k_easy_lookup=k, V_mag=V_mag,
                                 node_volumes=None, M_nodal=M_nodal, solve_U=solve_U, cg_tol=1e-10, method=meth,
                                 B_bias=None, chunk_elems=200_000, energy_assembly="scatter",
                                 grad_backend="stored_JinvT", mode="assembled")
            for meth in ("pcohen_hs", "lbfgs")}
    print(f"[wired] N={prob.N} E={len(conn32)} V_mag={V_mag:.6e}", flush=True)
    yield dict(geom=geom, prob=prob, eg=(eg, e_only, g_only, loc), solve_U=solve_U, d_diag=d_diag, mins=mins,
               A=A, K1=K1, Js=Js, k=k, bnd=bnd, V_mag=V_mag)
    clear_context_cache()


def test_factories_share_one_solver_context(wired):
    """The minimisers run on the solver's context, which carries the full material tables taken
    from the energy kernels built before it (the energy kernels' own context has no Dirichlet
    mask, which energies do not need)."""
    ctx = wired["solve_U"]._ctx
    assert wired["mins"]["pcohen_hs"]._ctx is ctx and wired["mins"]["lbfgs"]._ctx is ctx
    assert np.array_equal(ctx.exchange_diagonal(), wired["eg"][0]._ctx.exchange_diagonal())


def test_solve_U_matches_direct_solve(wired):
    from oracle.solvers import solve_U as oracle_solve_U

    p = wired["prob"]
    m = tilted_m0(p.N, seed=40, sigma=0.3)
    U, it, rr = wired["solve_U"](m, np.zeros(p.N), return_info=True)
    U_ref, _, _ = oracle_solve_U(p, m, np.zeros(p.N), 1e-10, precond="direct")
    b = p.rhs(m)
    print(f"[solve_U] it={it} rel_res={rr:.3e} err={rel(U, U_ref):.3e}", flush=True)
    assert it > 0 and rr <= 1e-10
    assert np.linalg.norm(b - p.A @ U) <= 1.001e-10 * np.linalg.norm(b)
    assert rel(U, U_ref) <= 1e-8
    U2 = wired["solve_U"](m, U, tol=1e-6)  # warm start at a looser tolerance: no iteration needed
    assert isinstance(U2, np.ndarray) and rel(U2, U) <= 1e-12


def test_energy_kernels_match_oracle(wired):
    eg, e_only, g_only, loc = wired["eg"]
    p = wired["prob"]
    m = tilted_m0(p.N, seed=41, sigma=0.3)
    U = np.random.default_rng(41).standard_normal(p.N) * wired["bnd"]
    B = 0.4 * H_DIR
    E, g = eg(m, U, B, sparse_ops={})
    E_ref, g_ref = p.energy_and_grad(m, U, B)
    scale = sum(abs(v) for v in p.energy_terms(m, U, B).values())
    assert abs(E - E_ref) <= 1e-12 * scale and rel(g, g_ref) <= 1e-12
    # shared nodes add block sums with FP64 atomics in arrival order: equal to the last bits
    assert abs(e_only(m, U, B) - E) <= 1e-14 * scale and rel(g_only(m, U, B), g) <= 1e-14
    v = np.random.default_rng(42).standard_normal((p.N, 3))
    assert rel(loc(v), p.local_grad(v)) <= 1e-12


def test_exchange_diagonal_matches_oracle(wired):
    from oracle.fem import exchange_diagonal, grad_phi_from_JinvT

    geom, p = wired["geom"], wired["prob"]
    d_ref = exchange_diagonal(geom.conn, geom.volume, grad_phi_from_JinvT(geom.JinvT), wired["A"], geom.mat_id,
                              wired["V_mag"], N=p.N)
    assert rel(wired["d_diag"], d_ref) <= 1e-12


def test_energy_kernels_reject_a_wrong_magnetic_volume(wired):
    from mumag_matrixfree2_amd.energy_kernels import make_energy_kernels

    with pytest.raises(ValueError, match="V_mag"):
        make_energy_kernels(wired["geom"], wired["A"], wired["K1"], wired["Js"], wired["k"], 2.0 * wired["V_mag"],
                            None)


@pytest.mark.parametrize("method", ["pcohen_hs", "lbfgs"])
def test_minimize_with_the_sweep_keywords_matches_oracle(wired, method):
    """minimize(m0, B_ext, U0=U, **sweep keywords): the field step equals the oracle's field step
    run with the solver's own AMG hierarchy: the same minimiser iterations and energy
    evaluations, energies within 1e-9 of the energy scale."""
    from oracle.minimize import Minimizer
    from oracle.solvers import solve_U as oracle_solve_U

    p = wired["prob"]
    ctx = wired["solve_U"]._ctx
    lv = oracle_levels(ctx.amg_levels, "amgcl")

    def sU(mm, x0, tol):
        return oracle_solve_U(p, mm, x0, tol, precond="amg", levels=lv, kind="amgcl")

    m0 = tilted_m0(p.N, seed=43, sigma=0.2)
    B = -0.3 * H_DIR
    m, U, info = wired["mins"][method](m0, B, U0=np.zeros(p.N), **sweep_kwargs(method))
    m_ref, U_ref, info_ref = Minimizer(p, sU, method=method, cg_tol=1e-10).minimize(m0, B, tau_f=1e-10, eps_a=1e-12)
    scale = sum(abs(v) for v in p.energy_terms(m_ref, U_ref, B).values())
    print(f"[{method}] gpu it={info['iters']} nf={info['evals']} E={info['E']:.15e} | ref it={info_ref['iters']} "
          f"nf={info_ref['evals']} E={info_ref['E']:.15e}", flush=True)
    assert set(info) >= {"iters", "time", "E", "gnorm", "preco_iters", "evals", "demag_iters"}
    assert info["iters"] == info_ref["iters"] and info["evals"] == info_ref["evals"]
    assert abs(info["E"] - info_ref["E"]) <= 1e-9 * scale
    mag = p.M_nodal > 0
    assert np.max(np.abs(m[mag] - m_ref[mag])) < 1e-6


# --------------------------------------------------------------------------- the ctypes stub

f64p, i32p = C.POINTER(C.c_double), C.POINTER(C.c_int32)


class MMParams(C.Structure):
    _fields_ = [("method", C.c_int), ("max_iter", C.c_int), ("tau_f", C.c_double), ("eps_a", C.c_double),
                ("tau0", C.c_double), ("ls_eta1", C.c_double), ("ls_eta2", C.c_double), ("ls_C", C.c_double),
                ("ls_c", C.c_double), ("ls_max_evals", C.c_int), ("pc_iters", C.c_int), ("pc_auto", C.c_int),
                ("pc_force_eta", C.c_double), ("pc_force_alpha", C.c_double), ("pc_stagnation_nu", C.c_double),
                ("pc_reg", C.c_double), ("L", C.c_int), ("cg_tol", C.c_double), ("phi_tol", C.c_double),
                ("cg_maxiter", C.c_int), ("memory", C.c_int)]


class MMInfo(C.Structure):
    _fields_ = [("iters", C.c_int), ("evals", C.c_int), ("demag_iters", C.c_int), ("preco_iters", C.c_int),
                ("converged", C.c_int), ("E", C.c_double), ("gnorm", C.c_double), ("time", C.c_double)]


def _p(a, t=np.float64):
    return np.ascontiguousarray(a, dtype=t)


def test_mm_minimize_through_the_ctypes_stub(cube_c0):
    """The INTEGRATION.md stub: mm_ctx_create + AMG upload + mm_minimize with caller-owned host
    arrays, against the oracle's pcohen_hs field step with the same hierarchy."""
    from mumag_matrixfree2_amd import _hip
    from mumag_matrixfree2_amd.amg import build_hierarchy, fine_matrix, upload_hierarchy
    from oracle.fem import Problem
    from oracle.minimize import Minimizer
    from oracle.solvers import solve_U as oracle_solve_U

    lib = C.CDLL(str(_hip.LIB_PATH))
    lib.mm_ctx_create.argtypes = [C.c_int, C.c_int, f64p, i32p, i32p, C.c_int, f64p, f64p, f64p, f64p, f64p,
                                  C.c_double, C.c_int, C.POINTER(C.c_int64)]
    lib.mm_minimize.argtypes = [C.c_int64, C.POINTER(MMParams), f64p, f64p, f64p, C.POINTER(MMInfo)]
    lib.mm_ctx_destroy.argtypes = [C.c_int64]
    mesh = cube_c0
    knt, ijk, bnd = mesh["knt"], mesh["ijk"], mesh["boundary_mask"]
    A, K1, Js, k = ndfeb_materials()
    h = C.c_int64()
    a = [_p(knt), _p(ijk[:, :4], np.int32), _p(ijk[:, 4], np.int32), _p(A), _p(K1), _p(Js), _p(k), _p(bnd)]
    rc = lib.mm_ctx_create(len(knt), len(ijk), a[0].ctypes.data_as(f64p), a[1].ctypes.data_as(i32p),
                           a[2].ctypes.data_as(i32p), len(A), *(x.ctypes.data_as(f64p) for x in a[3:8]), 1e-12, 0,
                           C.byref(h))
    assert rc == 0

    class _Handle:  # the package's upload helper takes an object with the handle and sizes
        handle, N = h.value, len(knt)

    p = Problem(knt, ijk, A, K1, Js, k, boundary_mask=bnd)
    levels = build_hierarchy(p.A, "amgcl")
    upload_hierarchy(_Handle, levels, "amgcl")
    par = MMParams(method=0, max_iter=2000, tau_f=1e-10, eps_a=1e-12, tau0=1e-2, ls_eta1=0.1, ls_eta2=0.1, ls_C=2.0,
                   ls_c=0.5, ls_max_evals=15, pc_iters=10, pc_auto=1, pc_force_eta=0.5, pc_force_alpha=0.5,
                   pc_stagnation_nu=0.01, pc_reg=0.0, L=0, cg_tol=1e-10, phi_tol=min(1e-10, 0.1 * 1e-10),
                   cg_maxiter=2000, memory=10)
    m0 = tilted_m0(len(knt), seed=44, sigma=0.2)
    m, U, B = _p(m0).copy(), np.zeros(len(knt)), _p(0.5 * H_DIR)
    info = MMInfo()
    rc = lib.mm_minimize(h.value, C.byref(par), m.ctypes.data_as(f64p), U.ctypes.data_as(f64p),
                         B.ctypes.data_as(f64p), C.byref(info))
    lib.mm_ctx_destroy(h.value)
    assert rc == 0
    lv = oracle_levels(levels, "amgcl")

    def sU(mm, x0, tol):
        return oracle_solve_U(p, mm, x0, tol, precond="amg", levels=lv, kind="amgcl")

    m_ref, U_ref, info_ref = Minimizer(p, sU, "pcohen_hs", cg_tol=1e-10).minimize(m0, B, tau_f=1e-10)
    scale = sum(abs(v) for v in p.energy_terms(m_ref, U_ref, B).values())
    print(f"[stub] it={info.iters} nf={info.evals} E={info.E:.15e} | ref it={info_ref['iters']} "
          f"nf={info_ref['evals']} E={info_ref['E']:.15e}", flush=True)
    assert info.iters == info_ref["iters"] and info.evals == info_ref["evals"]
    assert abs(info.E - info_ref["E"]) <= 1e-9 * scale
    mag = p.M_nodal > 0
    assert np.max(np.abs(m.reshape(-1, 3)[mag] - m_ref[mag])) < 1e-6
    assert rel(U, U_ref) < 1e-6


@pytest.fixture(scope="module")
def poisson_factories(cube_c0):
    from mumag_matrixfree2_amd.fem_utils import TetGeom, clear_context_cache
    from mumag_matrixfree2_amd.poisson_solve import make_pcg_solve, make_poisson_ops
    from oracle.fem import Problem

    mesh = cube_c0
    A, K1, Js, k = ndfeb_materials()
    geom = TetGeom(conn=mesh["ijk"][:, :4].astype(np.int32), mat_id=mesh["ijk"][:, 4].astype(np.int32),
                   x_nodes=mesh["knt"])
    apply_A, rhs_from_m, assemble_diag = make_poisson_ops(geom, Js, chunk_elems=200_000, reg=1e-12,
                                                          grad_backend="on_the_fly", boundary_mask=mesh["boundary_mask"],
                                                          mode="matrix_free")
    solve = make_pcg_solve(apply_A, None, precond_type="amgcl", order=1, maxiter=2000, tol=1e-10,
                           boundary_mask=None)
    prob = Problem(mesh["knt"], mesh["ijk"], A, K1, Js, k, boundary_mask=mesh["boundary_mask"])
    yield apply_A, rhs_from_m, assemble_diag, solve, prob
    clear_context_cache()


def test_poisson_ops_match_assembled_operators(poisson_factories):
    apply_A, rhs_from_m, assemble_diag, _, p = poisson_factories
    x = np.random.default_rng(45).standard_normal(p.N)
    m = tilted_m0(p.N, seed=45, sigma=0.4)
    assert rel(apply_A({}, x), p.A @ x) <= 1e-12
    assert rel(rhs_from_m({}, m), p.rhs(m)) <= 1e-12
    assert rel(assemble_diag({}, p.N), p.A.diagonal()) <= 1e-12


def test_pcg_solve_meets_its_tolerance(poisson_factories):
    import scipy.sparse.linalg as spla

    _, rhs_from_m, _, solve, p = poisson_factories
    b = rhs_from_m({}, tilted_m0(p.N, seed=46, sigma=0.4))
    x, it, r2 = solve({}, b, np.zeros(p.N))
    x_ref = spla.spsolve(p.A.tocsc(), b)
    print(f"[pcg] it={it} r2={r2:.3e}", flush=True)
    assert it > 0 and r2 <= 1e-20 * (b @ b)
    assert np.linalg.norm(b - p.A @ x) <= 1.001e-10 * np.linalg.norm(b)
    assert rel(x, x_ref) <= 1e-8
    x2, it2, _ = solve({}, b, np.zeros(p.N), tol=1e-4)
    assert it2 < it and np.linalg.norm(b - p.A @ x2) <= 1e-4 * np.linalg.norm(b)
