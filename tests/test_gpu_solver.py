"""GPU parity of the preconditioner and solve_U (PCG + AMG V-cycle) against the oracle (FP64)."""

from __future__ import annotations

import numpy as np
import pytest

from conftest import ndfeb_materials, oracle_levels, tilted_m0

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-300))


@pytest.fixture(scope="module", params=["amgcl", "amg"])
def solver_ctx(request, cube_c0):
    from mumag_matrixfree2_amd.context import MeshContext

    A, K1, Js, k = ndfeb_materials()
    m = cube_c0
    ctx = MeshContext(m["knt"], m["ijk"][:, :4], m["ijk"][:, 4], A, K1, Js, k, boundary_mask=m["boundary_mask"])
    ctx.setup_precond(request.param)
    print(f"[solver_ctx] kind={request.param} levels={[l['A'].shape[0] for l in ctx.amg_levels]}", flush=True)
    return request.param, ctx


@pytest.mark.parametrize("nb", [256, 512, 1024, 2048])
def test_vcycle_matches_oracle(solver_ctx, oracle_c0, nb):
    """The FP64 V-cycle on every cluster block size (its first apply stages x = M r from the
    residual, the second the corrected x)."""
    from oracle.solvers import vcycle

    kind, ctx = solver_ctx
    r = np.random.default_rng(11).standard_normal(ctx.N)
    ctx.set_cluster_block_nodes(nb)
    try:
        z = ctx.precond_apply(r)
    finally:
        ctx.set_cluster_block_nodes(256)
    z_ref = vcycle(oracle_levels(ctx.amg_levels, kind), oracle_c0.apply_A, r, kind)
    assert rel(z, z_ref) <= 1e-11
    assert float(r @ z) > 0.0


@pytest.mark.parametrize("lanes", [1, 2, 4, 8, 16, 32, 64])
def test_vcycle_matches_oracle_at_every_lane_count(solver_ctx, oracle_c0, lanes):
    """The CSR products of the coarse levels give the same V-cycle with any lanes per row (the
    values the SpMV tuner chooses between)."""
    from oracle.solvers import vcycle

    kind, ctx = solver_ctx
    saved = ctx.amg_lanes()
    r = np.random.default_rng(13).standard_normal(ctx.N)
    try:
        for lev, L in enumerate(saved):
            for which in ("A", "P", "R"):
                ctx.set_amg_lanes(lev, which, lanes)
        assert all(v == lanes for d in ctx.amg_lanes() for v in d.values())
        z = ctx.precond_apply(r)
    finally:
        for lev, L in enumerate(saved):
            for which, v in L.items():
                ctx.set_amg_lanes(lev, which, v)
    z_ref = vcycle(oracle_levels(ctx.amg_levels, kind), oracle_c0.apply_A, r, kind)
    assert rel(z, z_ref) <= 1e-11


def test_spmv_lane_tuner_keeps_the_vcycle(solver_ctx, oracle_c0):
    from mumag_matrixfree2_amd.autotune import tune_spmv_lanes
    from oracle.solvers import vcycle

    kind, ctx = solver_ctx
    saved = ctx.amg_lanes()
    try:
        res = tune_spmv_lanes(ctx, reps=3)
        print(f"[{kind}] spmv lanes {res}", flush=True)
        assert len(res["lanes"]) == len(ctx.amg_levels)
        r = np.random.default_rng(14).standard_normal(ctx.N)
        z = ctx.precond_apply(r)
    finally:
        for lev, L in enumerate(saved):
            for which, v in L.items():
                ctx.set_amg_lanes(lev, which, v)
    z_ref = vcycle(oracle_levels(ctx.amg_levels, kind), oracle_c0.apply_A, r, kind)
    assert rel(z, z_ref) <= 1e-11


@pytest.mark.parametrize("nb", [256, 1024, 2048])
def test_colour_sequenced_scatter_matches_oracle(solver_ctx, oracle_c0, nb):
    """The colour-sequenced scatter of the cluster apply (no two blocks of a launch share a node;
    one launch per colour; shared-node sums added with plain loads and stores): the oracle's
    A x and V-cycle, and FP64 and mixed solves to the solver tolerance against the oracle's
    direct solution. The block sums themselves still come from LDS atomics in arrival order, so
    repeated applies agree to a few ulp (the bound of the atomic scatter), not bit for bit."""
    from oracle.solvers import solve_U as oracle_solve_U
    from oracle.solvers import vcycle

    kind, ctx = solver_ctx
    rng = np.random.default_rng(21)
    x = rng.standard_normal(ctx.N)
    m = tilted_m0(ctx.N, seed=22)
    U_ref, _, _ = oracle_solve_U(oracle_c0, m, np.zeros(ctx.N), 1e-12, precond="direct")
    ctx.set_cluster_block_nodes(nb)
    ctx.set_apply_scatter("colour")
    try:
        assert ctx.n_block_colours >= 2 or ctx.n_blocks_cl == 1
        ys = [ctx.apply_A(x) for _ in range(10)]
        assert rel(ys[0], oracle_c0.A @ x) <= 1e-12
        assert max(float(np.max(np.abs(y - ys[0]))) for y in ys) <= 4 * np.spacing(np.abs(ys[0]).max())
        z = ctx.precond_apply(x)
        z_ref = vcycle(oracle_levels(ctx.amg_levels, kind), oracle_c0.apply_A, x, kind)
        assert rel(z, z_ref) <= 1e-11
        U1, it1, rr1 = ctx.solve_U(m, np.zeros(ctx.N), 1e-10)
        U2, it2, rr2 = ctx.solve_U(m, np.zeros(ctx.N), 1e-10)
        assert rr1 <= 1e-10 and rel(U1, U_ref) <= 1e-7
        assert abs(it1 - it2) <= 1 and rel(U1, U2) <= 1e-13
        ctx.set_precision("mixed")
        try:
            Um, _, rrm = ctx.solve_U(m, np.zeros(ctx.N), 1e-10)
        finally:
            ctx.set_precision("fp64")
        assert rrm <= 1e-10 and rel(Um, U_ref) <= 1e-7
    finally:
        ctx.set_apply_scatter("atomic")
        ctx.set_cluster_block_nodes(256)
    assert ctx.n_block_colours == 0


def test_solve_U_matches_oracle_pcg(solver_ctx, oracle_c0):
    from oracle.solvers import solve_U

    kind, ctx = solver_ctx
    m = tilted_m0(ctx.N, seed=12, sigma=0.4)
    x0 = np.zeros(ctx.N)
    U, it, rr = ctx.solve_U(m, x0, 1e-10)
    U_ref, it_ref, rr_ref = solve_U(oracle_c0, m, x0, 1e-10, precond="amg",
                                    levels=oracle_levels(ctx.amg_levels, kind), kind=kind)
    print(f"[{kind}] it={it} it_ref={it_ref} rr={rr:.3e} rr_ref={rr_ref:.3e}", flush=True)
    assert abs(it - it_ref) <= 1
    assert rr <= 1e-10
    assert rel(U, U_ref) <= 1e-9
    # true residual of the GPU solution, computed by the oracle
    b = oracle_c0.rhs(m)
    assert np.linalg.norm(b - oracle_c0.A @ U) <= 1.0001e-10 * np.linalg.norm(b)


def test_solve_U_matches_direct(solver_ctx, oracle_c0):
    from oracle.solvers import solve_U

    kind, ctx = solver_ctx
    m = np.zeros((ctx.N, 3))
    m[:, 0] = 1.0
    U, it, rr = ctx.solve_U(m, np.zeros(ctx.N), 1e-12)
    U_ref, _, _ = solve_U(oracle_c0, m, np.zeros(ctx.N), 1e-12, precond="direct")
    assert rel(U, U_ref) <= 1e-9


def test_solve_U_warm_start_zero_iterations(solver_ctx, oracle_c0):
    """A converged initial guess returns after 0 iterations with U unchanged."""
    from oracle.solvers import solve_U

    kind, ctx = solver_ctx
    m = tilted_m0(ctx.N, seed=13, sigma=0.2)
    U_ref, _, _ = solve_U(oracle_c0, m, np.zeros(ctx.N), 1e-13, precond="direct")
    U, it, rr = ctx.solve_U(m, U_ref, 1e-8)
    assert it == 0
    assert np.array_equal(U, U_ref)


def test_jacobi_pcg(cube_c0, oracle_c0):
    from mumag_matrixfree2_amd.context import MeshContext
    from oracle.solvers import pcg

    A, K1, Js, k = ndfeb_materials()
    m = cube_c0
    ctx = MeshContext(m["knt"], m["ijk"][:, :4], m["ijk"][:, 4], A, K1, Js, k, boundary_mask=m["boundary_mask"])
    ctx.setup_precond("jacobi")
    mm = tilted_m0(ctx.N, seed=14, sigma=0.4)
    U, it, rr = ctx.solve_U(mm, np.zeros(ctx.N), 1e-9, maxiter=5000)
    d = oracle_c0.A.diagonal()
    b = oracle_c0.rhs(mm)
    U_ref, it_ref, _ = pcg(oracle_c0.apply_A, lambda r: r / (d + 1e-30), b, np.zeros(ctx.N), 1e-9, 5000)
    assert abs(it - it_ref) <= 2
    assert rel(U, U_ref) <= 1e-7
    ctx.close()


def test_amg_upload_rejects_malformed_csr(cube_c0):
    """A column index outside its matrix, or decreasing row offsets, in a level's A, P or R is an
    error at upload, before any kernel could read through it; the good hierarchy still works."""
    import scipy.sparse as sp

    from mumag_matrixfree2_amd._hip import call, f64, i32, ptr
    from mumag_matrixfree2_amd.amg import build_hierarchy, fine_matrix, upload_hierarchy
    from mumag_matrixfree2_amd.context import MeshContext

    A, K1, Js, k = ndfeb_materials()
    m = cube_c0
    ctx = MeshContext(m["knt"], m["ijk"][:, :4], m["ijk"][:, 4], A, K1, Js, k, boundary_mask=m["boundary_mask"])
    levels = build_hierarchy(fine_matrix(ctx), smoother="amgcl")
    assert len(levels) >= 2
    lv0, lv1 = levels[0], levels[1]
    n0, n1 = lv0["A"].shape[0], lv0["P"].shape[1]

    def set0(P, R):
        call("mm_amg_set_level", ctx.handle, 0, n0, n1, None, None, None, ptr(f64(lv0["M"])),
             ptr(i32(P.indptr)), ptr(i32(P.indices)), ptr(f64(P.data)),
             ptr(i32(R.indptr)), ptr(i32(R.indices)), ptr(f64(R.data)), None)

    P, R = sp.csr_matrix(lv0["P"]), sp.csr_matrix(lv0["R"])
    bad_P = P.copy()
    bad_P.indices = bad_P.indices.copy()
    bad_P.indices[len(bad_P.indices) // 2] = n1  # one past the last coarse unknown
    with pytest.raises(RuntimeError, match="prolongation P: malformed"):
        set0(bad_P, R)
    bad_R = R.copy()
    bad_R.indices = bad_R.indices.copy()
    bad_R.indices[0] = -1
    with pytest.raises(RuntimeError, match="restriction R: malformed"):
        set0(P, bad_R)
    A1 = sp.csr_matrix(lv1["A"])
    ptr_bad = A1.indptr.copy()
    ptr_bad[1], ptr_bad[2] = ptr_bad[2], ptr_bad[1]  # decreasing offsets
    n_next1 = 0 if len(levels) == 2 else levels[1]["P"].shape[1]
    with pytest.raises(RuntimeError, match="level matrix A: malformed"):
        call("mm_amg_set_level", ctx.handle, 1, n1, n_next1, ptr(i32(ptr_bad)), ptr(i32(A1.indices)),
             ptr(f64(A1.data)), ptr(f64(lv1["M"])), None, None, None, None, None, None,
             ptr(f64(lv1["A_dense_inv"])) if len(levels) == 2 else None)
    upload_hierarchy(ctx, levels, "amgcl")
    r = np.random.default_rng(15).standard_normal(ctx.N)
    assert float(r @ ctx.precond_apply(r)) > 0.0
    ctx.close()


def test_partials_in_several_batches_match_oracle():
    """16-node element blocks on an 83k-tet cube: more than 4,096 cluster blocks, so the p.Ap
    partials of the PCG apply are summed in several load batches by their consumer, and the
    per-tet kernels see blocks with fewer elements than threads (their prefetched element words
    are clamped). The apply, rhs, energy and gradient, local Hessian and solve_U match the oracle."""
    from mumag_matrixfree2_amd.context import MeshContext
    from mumag_matrixfree2_amd.mesh import config_cube_mesh
    from oracle.fem import Problem
    from oracle.solvers import solve_U

    mesh = config_cube_mesh(20.0, 18, 3)
    A, K1, Js, k = ndfeb_materials()
    ctx = MeshContext(mesh["knt"], mesh["ijk"][:, :4], mesh["ijk"][:, 4], A, K1, Js, k,
                      boundary_mask=mesh["boundary_mask"], block_nodes=16)
    ctx.setup_precond("amgcl")
    print(f"[many blocks] N={ctx.N} E={ctx.E} cluster blocks={ctx.n_blocks_cl} "
          f"magnet blocks={ctx.n_blocks_mag}", flush=True)
    assert ctx.n_blocks_cl > 4096
    ref = Problem(mesh["knt"], mesh["ijk"], A, K1, Js, k, boundary_mask=mesh["boundary_mask"])
    rng = np.random.default_rng(21)
    x = rng.standard_normal(ctx.N)
    assert rel(ctx.apply_A(x), ref.A @ x) <= 1e-12
    m = tilted_m0(ctx.N, seed=22, sigma=0.3)
    assert rel(ctx.rhs(m), ref.rhs(m)) <= 1e-12
    U = rng.standard_normal(ctx.N) * 0.1
    B = np.array([0.01, -0.02, -0.3])
    E, g = ctx.energy_grad(m, U, B)
    E_ref, g_ref = ref.energy_and_grad(m, U, B)
    scale = sum(abs(v) for v in ref.energy_terms(m, U, B).values())
    assert abs(E - E_ref) <= 1e-12 * scale
    assert rel(g, g_ref) <= 1e-12
    v = rng.standard_normal((ctx.N, 3))
    assert rel(ctx.local_grad(v), ref.local_grad(v)) <= 1e-12
    x0 = np.zeros(ctx.N)
    Us, it, rr = ctx.solve_U(m, x0, 1e-10)
    U_ref, it_ref, _ = solve_U(ref, m, x0, 1e-10, precond="amg", levels=oracle_levels(ctx.amg_levels, "amgcl"),
                               kind="amgcl")
    print(f"[many blocks] it={it} it_ref={it_ref} rr={rr:.3e}", flush=True)
    assert abs(it - it_ref) <= 1
    assert rel(Us, U_ref) <= 1e-9
    b = ref.rhs(m)
    assert np.linalg.norm(b - ref.A @ Us) <= 1.0001e-10 * np.linalg.norm(b)
    ctx.close()
