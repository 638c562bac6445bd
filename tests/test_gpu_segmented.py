"""Segmented Poisson solves: a small grain set concatenated into one context, one PCG per grain
in shared launches (``mm_ctx_create_segmented`` / ``mm_solve_U_segmented``).

Bar: the union operators equal each grain's assembled oracle operators to 1e-12 relative; every
grain's solve meets its own true-residual tolerance against its oracle matrix and agrees with
the oracle's direct solve to 1e-7 at tol 1e-10; a grain that starts converged takes no
iteration and keeps its potential bit for bit; per-grain iteration counts follow the per-grain
tolerances.
"""

from __future__ import annotations

import numpy as np
import pytest

from conftest import ndfeb_materials, tilted_m0

pytestmark = pytest.mark.gpu

TOLS = np.array([1e-6, 1e-8, 1e-10, 1e-10, 1e-9])


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-300))


@pytest.fixture(scope="module")
def grainset():
    from mumag_matrixfree2_amd.grainset import grain_set_context
    from mumag_matrixfree2_amd.mesh import grain_set
    from oracle.fem import Problem

    gs = grain_set(5, seed=7, t_min=3e3, t_max=3e4)
    tables = [ndfeb_materials(axis=g["easy_axis"]) for g in gs]
    ctx = grain_set_context(gs, tables)
    ctx.setup_precond("amgcl", max_coarse=300)
    probs = [Problem(g["knt"], g["ijk"], *t, boundary_mask=g["boundary_mask"]) for g, t in zip(gs, tables)]
    print(f"[grainset] N={ctx.N} E={ctx.E} segments={ctx.segments.tolist()}", flush=True)
    yield ctx, probs
    ctx.close()


def union_m(ctx, seed):
    return tilted_m0(ctx.N, seed=seed, sigma=0.3)


def test_union_operators_match_each_grain(grainset):
    from mumag_matrixfree2_amd.grainset import split

    ctx, probs = grainset
    x = np.random.default_rng(1).standard_normal(ctx.N)
    m = union_m(ctx, 2)
    y, b = split(ctx, ctx.apply_A(x)), split(ctx, ctx.rhs(m))
    for i, (p, xg, mg) in enumerate(zip(probs, split(ctx, x), split(ctx, m))):
        assert rel(y[i], p.A @ xg) <= 1e-12
        assert rel(b[i], p.D @ mg.ravel()) <= 1e-12


@pytest.mark.parametrize("shape", ["cluster:256", "persistent:2", "element:128"])
def test_segmented_solve_per_grain_tolerance(grainset, shape):
    from mumag_matrixfree2_amd.grainset import split
    from oracle.solvers import solve_U as oracle_solve_U

    ctx, probs = grainset
    unit, param = shape.split(":")
    ctx.set_apply_shape(unit, int(param))
    try:
        m = union_m(ctx, 3)
        U, its, rrs = ctx.solve_U_segmented(m, np.zeros(ctx.N), TOLS)
    finally:
        ctx.set_apply_shape("cluster", 256)
    print(f"[segmented {shape}] its={its.tolist()} rr={[f'{r:.1e}' for r in rrs]}", flush=True)
    for i, (p, Ug, mg) in enumerate(zip(probs, split(ctx, U), split(ctx, m))):
        b = p.rhs(mg)
        true_rr = np.linalg.norm(b - p.A @ Ug) / np.linalg.norm(b)
        assert rrs[i] <= TOLS[i] and true_rr <= TOLS[i] * 1.001
        if TOLS[i] <= 1e-10:
            U_ref, _, _ = oracle_solve_U(p, mg, np.zeros(p.N), 1e-12, precond="direct")
            assert rel(Ug, U_ref) <= 1e-7
    assert its[0] < its[2] and its[1] <= its[2]


def test_segmented_solve_mixed_precision(grainset):
    """FP32 V-cycle inside the FP64 segmented PCG: every grain still meets its FP64 tolerance."""
    from mumag_matrixfree2_amd.grainset import split

    ctx, probs = grainset
    m = union_m(ctx, 4)
    ctx.set_precision("mixed")
    try:
        U, its, rrs = ctx.solve_U_segmented(m, np.zeros(ctx.N), 1e-10)
    finally:
        ctx.set_precision("fp64")
    U64, its64, _ = ctx.solve_U_segmented(m, np.zeros(ctx.N), 1e-10)
    print(f"[segmented mixed] its={its.tolist()} fp64 its={its64.tolist()}", flush=True)
    for p, Ug, U64g, mg in zip(probs, split(ctx, U), split(ctx, U64), split(ctx, m)):
        b = p.rhs(mg)
        assert np.linalg.norm(b - p.A @ Ug) / np.linalg.norm(b) <= 1e-10 * 1.001
        assert rel(Ug, U64g) <= 1e-7
    assert np.all(np.abs(its - its64) <= 2)


def test_converged_segment_takes_no_iteration(grainset):
    ctx, probs = grainset
    m = union_m(ctx, 5)
    U1, _, _ = ctx.solve_U_segmented(m, np.zeros(ctx.N), 1e-11)
    x0 = np.zeros(ctx.N)
    s = ctx.segments
    x0[s[1]:s[2]] = U1[s[1]:s[2]]
    U2, its, rrs = ctx.solve_U_segmented(m, x0, 1e-10)
    assert its[1] == 0 and all(its[i] > 0 for i in range(len(its)) if i != 1)
    assert np.array_equal(U2[s[1]:s[2]], U1[s[1]:s[2]])
    assert np.all(rrs <= 1e-10)
    _, its0, _ = ctx.solve_U_segmented(m, np.zeros(ctx.N), 1e-10, maxiter=0)
    assert np.all(its0 == 0)


def test_segmented_rejects_bad_input(grainset):
    from mumag_matrixfree2_amd.context import MeshContext
    from mumag_matrixfree2_amd.mesh import grid_box

    ctx, _ = grainset
    with pytest.raises(ValueError):
        ctx.solve_U_segmented(union_m(ctx, 6), np.zeros(ctx.N), np.ones(2))  # wrong tolerance count
    knt, tets = grid_box(2.0, 2.0, 2.0, 1.0)
    A, K1, Js, k = ndfeb_materials()
    mat = np.ones(len(tets), dtype=np.int32)
    with pytest.raises(RuntimeError, match="couples two segments"):
        MeshContext(knt, tets, mat, A, K1, Js, k, segments=[0, knt.shape[0] // 2, knt.shape[0]])
    plain = MeshContext(knt, tets, mat, A, K1, Js, k)
    try:
        with pytest.raises(RuntimeError, match="segments"):
            plain.solve_U_segmented(np.zeros((plain.N, 3)), np.zeros(plain.N), 1e-8)
    finally:
        plain.close()


def test_minimize_segmented_equals_per_grain(grainset):
    """One field step of every grain in shared launches against each grain minimised in a
    context of its own. The two differ only through the AMG hierarchies (one on the union
    matrix, one per grain) and the order of the FP64 atomic adds, so the solves agree to their
    tolerance and the minimisers to well below the energy tolerance."""
    from mumag_matrixfree2_amd.context import MeshContext, make_params
    from mumag_matrixfree2_amd.grainset import split
    from mumag_matrixfree2_amd.mesh import grain_set

    ctx, probs = grainset
    gs = grain_set(5, seed=7, t_min=3e3, t_max=3e4)
    tables = [ndfeb_materials(axis=g["easy_axis"]) for g in gs]
    p = make_params("pcohen_hs", tau_f=1e-10, cg_tol=1e-10)
    B = np.array([0.017452, 0.0, 0.999848]) * -0.6
    m0 = [tilted_m0(g["knt"].shape[0], seed=60 + i, sigma=0.05) for i, g in enumerate(gs)]
    singles = []
    for g, t, m in zip(gs, tables, m0):
        c = MeshContext(g["knt"], g["ijk"][:, :4], g["ijk"][:, 4], *t, boundary_mask=g["boundary_mask"])
        c.setup_precond("amgcl", max_coarse=300)
        c.state_set(m, np.zeros(c.N))
        info = c.minimize_resident(p, B)
        singles.append((info, c.state_get()[0], c.state_j_parallel(B), c.V_mag))
        c.close()
    ctx.state_set(np.vstack(m0), np.zeros(ctx.N))
    infos = ctx.minimize_segmented(p, B)
    m_seg = split(ctx, ctx.state_get()[0])
    J_seg = ctx.state_j_parallel_segmented(B)
    V_seg = ctx.segment_volumes()
    for i, ((a, m_a, J_a, V_a), b) in enumerate(zip(singles, infos)):
        print(f"[grain {i}] single E={a['E']:.15e} it={a['iters']} nf={a['evals']} | segmented E={b['E']:.15e} "
              f"it={b['iters']} nf={b['evals']} dm={np.max(np.abs(m_seg[i] - m_a)):.1e}", flush=True)
        assert a["converged"] and b["converged"]
        assert abs(a["E"] - b["E"]) <= 1e-12 * abs(a["E"])
        assert abs(a["iters"] - b["iters"]) <= 1
        assert np.max(np.abs(m_seg[i] - m_a)) < 1e-9
        assert abs(J_seg[i] - J_a) <= 1e-8 and abs(V_seg[i] - V_a) <= 1e-12 * V_a
    with pytest.raises(RuntimeError, match="segmented"):
        ctx.minimize_resident(p, B)
    # the subset launches are undone: whole-context operators cover every grain again
    x = np.random.default_rng(8).standard_normal(ctx.N)
    for pr, yg, xg in zip(probs, split(ctx, ctx.apply_A(x)), split(ctx, x)):
        assert rel(yg, pr.A @ xg) <= 1e-12


def test_minimize_segmented_mixed_precision(grainset):
    """The mixed form (FP32 V-cycle inside the FP64 PCG) of the segmented field step: every
    grain's energy within 1e-8 of the FP64 segmented step, the same iterations within 2."""
    from mumag_matrixfree2_amd.context import make_params
    from mumag_matrixfree2_amd.mesh import grain_set

    ctx, _ = grainset
    gs = grain_set(5, seed=7, t_min=3e3, t_max=3e4)
    p = make_params("pcohen_hs", tau_f=1e-10, cg_tol=1e-10)
    B = np.array([0.017452, 0.0, 0.999848]) * -0.4
    m0 = np.vstack([tilted_m0(g["knt"].shape[0], seed=70 + i, sigma=0.05) for i, g in enumerate(gs)])
    out = {}
    for prec in ("fp64", "mixed"):
        ctx.set_precision(prec)
        ctx.state_set(m0, np.zeros(ctx.N))
        out[prec] = ctx.minimize_segmented(p, B)
    ctx.set_precision("fp64")
    for a, b in zip(out["fp64"], out["mixed"]):
        assert a["converged"] and b["converged"]
        assert abs(a["E"] - b["E"]) <= 1e-8 * max(1.0, abs(a["E"]))
        assert abs(a["iters"] - b["iters"]) <= 2


def test_hysteresis_sweep_of_a_grain_set(grainset, tmp_path):
    """run_hysteresis_loop on the segmented context: per-grain J_par equal to each grain's own
    sweep in a context of its own, the volume-averaged J_par of the rows, and the result files."""
    from mumag_matrixfree2_amd.context import MeshContext
    from mumag_matrixfree2_amd.grainset import split
    from mumag_matrixfree2_amd.hysteresis_loop import LoopParams, run_hysteresis_loop
    from mumag_matrixfree2_amd.mesh import grain_set

    ctx, _ = grainset
    gs = grain_set(5, seed=7, t_min=3e3, t_max=3e4)
    h = np.array([0.017452, 0.0, 0.999848])
    params = LoopParams(h_dir=h, B_start=1.0, B_end=-1.0, dB=-1.0, tau_f=1e-10, cg_tol=1e-10)
    m0 = np.vstack([tilted_m0(g["knt"].shape[0], seed=80 + i, sigma=0.05) for i, g in enumerate(gs)])
    res = run_hysteresis_loop(ctx, m0, params, out_dir=tmp_path, name="grains")
    gJ = res["grain_J"]
    assert gJ.shape == (3, len(gs))
    vol = ctx.segment_volumes()
    assert np.allclose(res["history"][:, 1], gJ @ vol / vol.sum(), rtol=0, atol=1e-14)
    for i, (g, m_i) in enumerate(zip(gs, split(ctx, m0))):
        c = MeshContext(g["knt"], g["ijk"][:, :4], g["ijk"][:, 4], *ndfeb_materials(axis=g["easy_axis"]),
                        boundary_mask=g["boundary_mask"])
        c.setup_precond("amgcl", max_coarse=300)
        r = run_hysteresis_loop(c, m_i, params)
        c.close()
        assert np.max(np.abs(r["history"][:, 1] - gJ[:, i])) < 1e-9
    assert (tmp_path / "hysteresis.csv").exists() and (tmp_path / "grains.mh").exists()
