"""The benchmark configurations checked against the oracle.

- The plain FP64 configuration at its full size (80 nm cube, 60^3 hexes x 6 tets + 5-layer
  airbox, 2,058,000 tets): the first field steps of the +2 T -> -8 T sweep on the GPU against
  the oracle's compiled pcohen_hs restatement (``oracle/cpu_minimizer.cpp``) with the same AMG
  hierarchy: identical minimiser iterations and energy evaluations per step, energies within
  1e-9 of the energy scale.
- The polycrystal configuration's form at reduced size (8 Voronoi grains with a grain-boundary
  phase and per-grain easy axes, 30 nm cube + airbox) in the mixed-precision form: a sweep
  against the FP64 GPU sweep and the oracle's sweep, same switching step, energies within 1e-8.
- Scatter determinism: the FP64 atomic scatter of the apply and a full solve repeated in one
  process differ by no more than a stated bound.
- The 24.6M-tet polycrystal at its real size through size-independent properties (symmetry,
  linear exactness, the solver tolerance in all three precision forms).
"""

from __future__ import annotations

import numpy as np
import pytest

from conftest import MU0, ndfeb_materials, tilted_m0

pytestmark = pytest.mark.gpu

H_DIR = np.array([np.sin(np.deg2rad(1.0)), 0.0, np.cos(np.deg2rad(1.0))])
JS = 1.61


def _energy_scale(prob, m, U, B):
    return sum(abs(v) for v in prob.energy_terms(m, U, B).values())


def test_full_size_cube_first_steps_match_the_compiled_oracle():
    from mumag_matrixfree2_amd.context import MeshContext, make_params
    from mumag_matrixfree2_amd.hysteresis_loop import field_values
    from mumag_matrixfree2_amd.mesh import config_cube_mesh
    from oracle import cpu
    from oracle.fem import Problem

    if not cpu.LIB.exists():
        cpu.build()
    mesh = config_cube_mesh(80.0, 60, 5)
    A, K1, Js, k = ndfeb_materials()
    ctx = MeshContext(mesh["knt"], mesh["ijk"][:, :4], mesh["ijk"][:, 4], A, K1, Js, k,
                      boundary_mask=mesh["boundary_mask"])
    ctx.setup_precond("amgcl")
    prob = Problem(mesh["knt"], mesh["ijk"], A, K1, Js, k, boundary_mask=mesh["boundary_mask"])
    print(f"[cube2M] N={ctx.N} E={ctx.E} levels={[l['A'].shape[0] for l in ctx.amg_levels]}", flush=True)
    cm = cpu.CpuMinimizer(prob, ctx.amg_levels)
    par = make_params("pcohen_hs", tau_f=1e-10, eps_a=1e-12, cg_tol=1e-10)
    B_vals = field_values(2.0 / JS, -8.0 / JS, -0.25 / JS, False)
    m0 = tilted_m0(ctx.N, seed=0, sigma=0.01)
    ctx.state_set(m0, np.zeros(ctx.N))
    m_c, U_c = m0.copy(), np.zeros(ctx.N)
    for kstep in range(3):
        B = B_vals[kstep] * H_DIR
        info = ctx.minimize_resident(par, B)
        m_c, U_c, info_c = cm.minimize(m_c, B, U0=U_c, tau_f=1e-10, eps_a=1e-12, cg_tol=1e-10)
        m_g, U_g = ctx.state_get()
        scale = _energy_scale(prob, m_c, U_c, B)
        print(f"[cube2M step {kstep}] gpu it={info['iters']} nf={info['evals']} E={info['E']:.15e} | cpu "
              f"it={info_c['iters']} nf={info_c['evals']} E={info_c['E']:.15e} | max dm "
              f"{np.max(np.abs(m_g - m_c)):.2e}", flush=True)
        assert info["iters"] == info_c["iters"] and info["evals"] == info_c["evals"]
        assert abs(info["E"] - info_c["E"]) <= 1e-9 * scale
    cm.close()
    ctx.close()


def _poly_materials(n_grains=8, seed=5):
    from mumag_matrixfree2_amd.mesh import random_easy_axes

    Kd = JS**2 / (2 * MU0)
    G = n_grains + 2
    A, K1, Js, k = np.zeros(G), np.zeros(G), np.zeros(G), np.tile([0.0, 0.0, 1.0], (G, 1))
    A[:n_grains], K1[:n_grains], Js[:n_grains] = 7.7e-12 * 1e18 / Kd, 4.3e6 / Kd, 1.0
    k[:n_grains] = random_easy_axes(n_grains, seed)
    A[n_grains], Js[n_grains] = 1.5e-12 * 1e18 / Kd, 1.0
    return A, K1, Js, k


def test_polycrystal_mixed_sweep_matches_fp64_and_the_oracle():
    """+2 T -> -8 T in 0.5 T steps on an 8-grain polycrystal with a grain-boundary phase: the
    mixed form and the FP64 form on the GPU and the oracle's compiled pcohen_hs sweep switch at
    the same step; per-step energies agree to 1e-8 of the energy scale (mixed vs FP64) and to
    1e-9 (FP64 vs oracle)."""
    from mumag_matrixfree2_amd.context import MeshContext
    from mumag_matrixfree2_amd.hysteresis_loop import LoopParams, field_values, run_hysteresis_loop, switching_step
    from mumag_matrixfree2_amd.mesh import polycrystal_cube
    from oracle import cpu
    from oracle.fem import Problem

    if not cpu.LIB.exists():
        cpu.build()
    mesh = polycrystal_cube(30.0, 12, 3, n_grains=8, gb_width=1.5, seed=4)
    mats = _poly_materials()
    ctx = MeshContext(mesh["knt"], mesh["ijk"][:, :4], mesh["ijk"][:, 4], *mats, boundary_mask=mesh["boundary_mask"])
    ctx.setup_precond("amgcl", max_coarse=200)
    prob = Problem(mesh["knt"], mesh["ijk"], *mats, boundary_mask=mesh["boundary_mask"])
    params = LoopParams(h_dir=H_DIR, B_start=2.0 / JS, B_end=-8.0 / JS, dB=-0.5 / JS, tau_f=1e-10, cg_tol=1e-10,
                        Js_ref=JS)
    m0 = tilted_m0(ctx.N, seed=0, sigma=0.01)
    r64 = run_hysteresis_loop(ctx, m0, params)
    ctx.set_precision("mixed")
    rmx = run_hysteresis_loop(ctx, m0, params)
    ctx.set_precision("fp64")
    cm = cpu.CpuMinimizer(prob, ctx.amg_levels)
    m, U = m0.copy(), np.zeros(ctx.N)
    J_ref, E_ref, scales = [], [], []
    for Bmag in field_values(params.B_start, params.B_end, params.dB, False):
        B = Bmag * H_DIR
        m, U, info = cm.minimize(m, B, U0=U, tau_f=1e-10, eps_a=1e-12, cg_tol=1e-10)
        J_ref.append(prob.J_parallel(m, H_DIR))
        E_ref.append(info["E"])
        scales.append(_energy_scale(prob, m, U, B))
    cm.close()
    ctx.close()
    E64, Emx, scales = r64["history"][:, 2], rmx["history"][:, 2], np.array(scales)
    sw = (switching_step(r64["history"][:, 1]), switching_step(rmx["history"][:, 1]), switching_step(np.array(J_ref)))
    print(f"[poly sweep] switching steps fp64/mixed/oracle {sw}; max |dE|/scale mixed-fp64 "
          f"{np.max(np.abs(Emx - E64) / scales):.2e}, fp64-oracle {np.max(np.abs(E64 - E_ref) / scales):.2e}",
          flush=True)
    assert sw[0] == sw[1] == sw[2] > 0
    assert np.max(np.abs(Emx - E64) / scales) <= 1e-8
    assert np.max(np.abs(E64 - np.array(E_ref)) / scales) <= 1e-9


def test_scatter_spread_is_bounded(cube_c0):
    """The apply's FP64 atomics add shared-node block sums in arrival order, so repeated applies
    may differ in the last bits: the spread of 20 applies is at most 4 ulp of the largest entry,
    and two full solves of the same system differ by less than 1e-13 relative."""
    from mumag_matrixfree2_amd.context import MeshContext

    A, K1, Js, k = ndfeb_materials()
    m = cube_c0
    ctx = MeshContext(m["knt"], m["ijk"][:, :4], m["ijk"][:, 4], A, K1, Js, k, boundary_mask=m["boundary_mask"])
    ctx.setup_precond("amgcl")
    x = np.random.default_rng(50).standard_normal(ctx.N)
    ys = np.array([ctx.apply_A(x) for _ in range(20)])
    spread = float(np.max(np.abs(ys - ys[0])))
    mm = tilted_m0(ctx.N, seed=51, sigma=0.3)
    U1 = ctx.solve_U(mm, np.zeros(ctx.N), 1e-10)[0]
    U2 = ctx.solve_U(mm, np.zeros(ctx.N), 1e-10)[0]
    ctx.close()
    du = float(np.linalg.norm(U1 - U2) / np.linalg.norm(U1))
    print(f"[scatter] apply spread {spread:.3e} (|y|max {np.abs(ys).max():.3e}), solve spread {du:.3e}", flush=True)
    assert spread <= 4 * np.spacing(np.abs(ys).max())
    assert du <= 1e-13


def test_full_size_polycrystal_size_independent_properties():
    """The 24.6M-tet polycrystal at its real size, checked through properties that hold at any
    size: A is symmetric (x.Ay = y.Ax), A (a.x + c) = reg (a.x + c) on nodes whose neighbours are all
    interior (the stiffness annihilates linear functions), and the FP64, mixed and FP32-PCG solves
    all meet the 1e-10 tolerance on the FP64 true residual and agree with each other to it."""
    from mumag_matrixfree2_amd.autotune import tune_apply
    from mumag_matrixfree2_amd.context import MeshContext
    from mumag_matrixfree2_amd.mesh import polycrystal_cube, random_easy_axes

    mesh = polycrystal_cube(200.0, 150, 5, n_grains=27, gb_width=2.0, seed=4)
    Kd = JS**2 / (2 * MU0)
    G = 29
    A, K1, Js, k = np.zeros(G), np.zeros(G), np.zeros(G), np.tile([0.0, 0.0, 1.0], (G, 1))
    A[:27], K1[:27], Js[:27] = 7.7e-12 * 1e18 / Kd, 4.3e6 / Kd, 1.0
    k[:27] = random_easy_axes(27, 40)
    A[27], Js[27] = 1.5e-12 * 1e18 / Kd, 1.0
    ctx = MeshContext(mesh["knt"], mesh["ijk"][:, :4], mesh["ijk"][:, 4], A, K1, Js, k,
                      boundary_mask=mesh["boundary_mask"])
    print(f"[poly25M] N={ctx.N} E={ctx.E}", flush=True)
    assert ctx.E == 24_576_000
    # the geometry the context computed on the device and copied back (the AMG set-up assembles
    # its matrix from it) holds every element: the same volumes as the host's own determinants
    _, Ve_ctx, _ = ctx.geometry()
    _, Ve_host, _ = ctx.element_table()
    assert np.array_equal(np.sort(Ve_ctx) > 0, np.ones(ctx.E, dtype=bool))
    assert np.max(np.abs(np.sort(Ve_ctx) - np.sort(Ve_host))) <= 1e-12 * Ve_host.max()
    tuned = tune_apply(ctx, reps=5)
    print(f"[poly25M] apply shape {tuned['unit']}:{tuned['param']}:{tuned['cluster_block_nodes']}", flush=True)
    rng = np.random.default_rng(80)
    x, y = rng.standard_normal(ctx.N), rng.standard_normal(ctx.N)
    Ax, Ay = ctx.apply_A(x), ctx.apply_A(y)
    assert abs(x @ Ay - y @ Ax) <= 1e-12 * np.linalg.norm(x) * np.linalg.norm(Ay)
    # linear functions: rows whose whole node star is interior
    knt = mesh["knt"]
    lin = knt @ np.array([0.3, -0.2, 0.5]) + 1.7
    conn = mesh["ijk"][:, :4]
    bad = np.zeros(ctx.N, dtype=bool)
    on_bnd = mesh["boundary_mask"] == 0
    touches_bnd = on_bnd[conn].any(axis=1)
    bad[conn[touches_bnd].ravel()] = True
    inner = ~bad
    Al = ctx.apply_A(lin)
    assert np.max(np.abs(Al[inner] - 1e-12 * lin[inner])) <= 1e-12 * np.max(np.abs(Al))
    # solves in the three precision forms
    ctx.setup_precond("amgcl")
    m = np.zeros((ctx.N, 3))
    m[:, 2] = 1.0
    m += 0.3 * rng.standard_normal((ctx.N, 3))
    m /= np.linalg.norm(m, axis=1, keepdims=True)
    Us = {}
    for prec in ("fp64", "mixed", "fp32pcg"):
        ctx.set_precision(prec)
        U, it, rr = ctx.solve_U(m, np.zeros(ctx.N), 1e-10)
        print(f"[poly25M solve {prec}] it={it} rel_res={rr:.2e}", flush=True)
        assert rr <= 1e-10
        Us[prec] = U
    ctx.set_precision("fp64")
    b = ctx.rhs(m)
    for prec, U in Us.items():
        assert np.linalg.norm(b - ctx.apply_A(U)) <= 1.01e-10 * np.linalg.norm(b), prec
        assert np.linalg.norm(U - Us["fp64"]) <= 1e-8 * np.linalg.norm(Us["fp64"]), prec
    ctx.close()
