"""Mixed-precision form: FP32 V-cycle under the FP64 PCG.

Bar: the FP32 preconditioner is the FP64 one to FP32 accuracy (<= 1e-5 relative); the solve
still meets the FP64 tolerance (true residual computed by the oracle in FP64); energies of a
field step agree with the FP64 GPU path to 1e-10 of the energy scale; a sweep switches at the
same field step.
"""

from __future__ import annotations

import numpy as np
import pytest

from conftest import ndfeb_materials, oracle_levels, tilted_m0

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-300))


@pytest.fixture(scope="module")
def mixed_ctx(cube_c0):
    from mumag_matrixfree2_amd.context import MeshContext

    A, K1, Js, k = ndfeb_materials()
    m = cube_c0
    ctx = MeshContext(m["knt"], m["ijk"][:, :4], m["ijk"][:, 4], A, K1, Js, k, boundary_mask=m["boundary_mask"])
    ctx.setup_precond("amgcl")
    print(f"[mixed_ctx] levels={[l['A'].shape[0] for l in ctx.amg_levels]}", flush=True)
    yield ctx
    ctx.set_precision("fp64")


def test_fp32_vcycle_is_fp64_vcycle_to_fp32_accuracy(mixed_ctx, oracle_c0):
    from oracle.solvers import vcycle

    ctx = mixed_ctx
    r = np.random.default_rng(31).standard_normal(ctx.N)
    ctx.set_precision("mixed")
    z = ctx.precond_apply(r)
    ctx.set_precision("fp64")
    z_ref = vcycle(oracle_levels(ctx.amg_levels, "amgcl"), oracle_c0.apply_A, r, "amgcl")
    assert rel(z, z_ref) <= 1e-5
    assert float(r @ z) > 0


@pytest.mark.parametrize("nb", [512, 1024, 2048])
def test_fp32_vcycle_same_under_both_stagings(mixed_ctx, oracle_c0, nb):
    """The FP32 applies of the V-cycle stage the same values from entry records as from gathers:
    the two V-cycles agree to the rounding of the atomic summation order, and both are the FP64
    oracle's V-cycle to FP32 accuracy."""
    from oracle.solvers import vcycle

    ctx = mixed_ctx
    r = np.random.default_rng(32).standard_normal(ctx.N)
    z_ref = vcycle(oracle_levels(ctx.amg_levels, "amgcl"), oracle_c0.apply_A, r, "amgcl")
    ctx.set_cluster_block_nodes(nb)
    ctx.set_precision("mixed")
    try:
        z = {}
        for staging in ("gather", "records"):
            ctx.set_cluster_staging(staging)
            z[staging] = ctx.precond_apply(r)
    finally:
        ctx.set_precision("fp64")
        ctx.set_cluster_staging("records")
        ctx.set_cluster_block_nodes(256)
    assert rel(z["records"], z["gather"]) <= 1e-6
    assert rel(z["records"], z_ref) <= 1e-5 and rel(z["gather"], z_ref) <= 1e-5


def test_mixed_solve_meets_fp64_tolerance(mixed_ctx, oracle_c0):
    ctx = mixed_ctx
    m = tilted_m0(ctx.N, seed=32, sigma=0.4)
    U64, it64, _ = ctx.solve_U(m, np.zeros(ctx.N), 1e-10)
    ctx.set_precision("mixed")
    U, it, rr = ctx.solve_U(m, np.zeros(ctx.N), 1e-10)
    ctx.set_precision("fp64")
    b = oracle_c0.rhs(m)
    true_rr = np.linalg.norm(b - oracle_c0.A @ U) / np.linalg.norm(b)
    print(f"[mixed solve] it={it} (fp64 {it64}) rr={rr:.3e} true={true_rr:.3e}", flush=True)
    assert true_rr <= 1.001e-10
    assert it <= it64 + 2
    assert rel(U, U64) <= 1e-8


@pytest.mark.parametrize("method", ["pcohen_hs", "lbfgs"])
def test_mixed_field_step_energy(mixed_ctx, oracle_c0, method):
    from mumag_matrixfree2_amd.context import make_params

    ctx = mixed_ctx
    m0 = tilted_m0(ctx.N, seed=33, sigma=0.1)
    B = np.array([0.017452, 0.0, 0.999848]) * (-0.4)
    p = make_params(method, tau_f=1e-10, cg_tol=1e-10)
    out = {}
    for prec in ("fp64", "mixed"):
        ctx.set_precision(prec)
        ctx.state_set(m0, np.zeros(ctx.N))
        out[prec] = ctx.minimize_resident(p, B)
    ctx.set_precision("fp64")
    m, U = ctx.state_get()
    terms = oracle_c0.energy_terms(m, U, B)
    scale = sum(abs(v) for v in terms.values())
    print(f"[mixed {method}] E64={out['fp64']['E']:.15e} Emix={out['mixed']['E']:.15e} "
          f"it={out['fp64']['iters']}/{out['mixed']['iters']}", flush=True)
    assert abs(out["mixed"]["E"] - out["fp64"]["E"]) <= 1e-10 * scale


def test_mixed_sweep_same_switching_step(small_cube):
    from mumag_matrixfree2_amd.context import MeshContext
    from mumag_matrixfree2_amd.hysteresis_loop import LoopParams, run_hysteresis_loop

    A, K1, Js, k = ndfeb_materials()
    m = small_cube
    ctx = MeshContext(m["knt"], m["ijk"][:, :4], m["ijk"][:, 4], A, K1, Js, k, boundary_mask=m["boundary_mask"])
    ctx.setup_precond("amgcl", max_coarse=50)
    h = np.array([0.017452, 0.0, 0.999848])
    params = LoopParams(h_dir=h, B_start=2.0 / 1.61, B_end=-8.0 / 1.61, dB=-0.5 / 1.61, tau_f=1e-10, cg_tol=1e-10)
    m0 = tilted_m0(ctx.N, seed=0, sigma=0.01)
    r64 = run_hysteresis_loop(ctx, m0, params)
    ctx.set_precision("mixed")
    rmx = run_hysteresis_loop(ctx, m0, params)
    ctx.close()
    E64, Emx = r64["history"][:, 2], rmx["history"][:, 2]
    assert r64["switching_step"] == rmx["switching_step"] > 0
    assert np.max(np.abs(Emx - E64) / np.maximum(np.abs(E64), 1.0)) <= 1e-8


@pytest.mark.parametrize("blocks", ["256", "1024:records", "1024:gather", "2048:records"])
def test_fp32_pcg_meets_the_fp64_tolerance(mixed_ctx, oracle_c0, blocks):
    """The FP32 PCG with FP64 refinement stops on the FP64 true residual: r.r <= tol^2 b.b
    checked by the oracle in FP64, and the solution agrees with the FP64 solve to the tolerance
    (from zero and from a warm start close to the solution); on 256-node cluster blocks and on
    1024-node blocks with either staging of the FP32 kernels."""
    ctx = mixed_ctx
    nb, *staging = blocks.split(":")
    m = tilted_m0(ctx.N, seed=34, sigma=0.4)
    U64, it64, _ = ctx.solve_U(m, np.zeros(ctx.N), 1e-10)
    b = oracle_c0.rhs(m)
    ctx.set_cluster_block_nodes(int(nb))
    ctx.set_cluster_staging(staging[0] if staging else "records")
    ctx.set_precision("fp32pcg")
    try:
        U, it, rr = ctx.solve_U(m, np.zeros(ctx.N), 1e-10)
        U2, it2, rr2 = ctx.solve_U(m, U64 * (1 + 1e-4), 1e-12)
    finally:
        ctx.set_precision("fp64")
        ctx.set_cluster_staging("records")
        ctx.set_cluster_block_nodes(256)
    true_rr = np.linalg.norm(b - oracle_c0.A @ U) / np.linalg.norm(b)
    true_rr2 = np.linalg.norm(b - oracle_c0.A @ U2) / np.linalg.norm(b)
    print(f"[fp32pcg solve] it={it} (fp64 {it64}) rr={rr:.3e} true={true_rr:.3e}; warm it={it2} true={true_rr2:.3e}",
          flush=True)
    assert rr <= 1e-10 and true_rr <= 1.001e-10
    assert rr2 <= 1e-12 and true_rr2 <= 1.001e-12
    assert rel(U, U64) <= 1e-8
    assert it <= 2 * it64 + 5


@pytest.mark.parametrize("eta", [1e-2, 1e-3])
def test_fp32_pcg_residual_replacement_keeps_the_krylov_iteration(mixed_ctx, oracle_c0, eta):
    """The FP32 PCG replaces its recursive residual by the FP64 true residual after each eta
    reduction and keeps its search direction: it stops on the FP64 true residual, reaches the FP64
    solution to the tolerance, and needs no more than two iterations beyond the FP64 PCG (a
    restarted refinement loses the Krylov space at every cycle). An eta outside (0, 1) is an error."""
    ctx = mixed_ctx
    m = tilted_m0(ctx.N, seed=35, sigma=0.4)
    U64, it64, _ = ctx.solve_U(m, np.zeros(ctx.N), 1e-10)
    b = oracle_c0.rhs(m)
    ctx.set_precision("fp32pcg")
    ctx.set_refinement(eta)
    try:
        U, it, rr = ctx.solve_U(m, np.zeros(ctx.N), 1e-10)
        with pytest.raises(RuntimeError, match="replacement factor"):
            ctx.set_refinement(1.0)
    finally:
        ctx.set_refinement(1e-3)  # the default
        ctx.set_precision("fp64")
    true_rr = np.linalg.norm(b - oracle_c0.A @ U) / np.linalg.norm(b)
    print(f"[fp32pcg eta={eta}] it={it} (fp64 {it64}) rr={rr:.3e} true={true_rr:.3e}", flush=True)
    assert rr <= 1e-10 and true_rr <= 1.001e-10
    assert rel(U, U64) <= 1e-8
    assert it <= it64 + 2


@pytest.mark.parametrize("h", [1e-3, 1e3])
def test_fp32_pcg_vcycle_independent_of_the_length_unit(cube_c0, h):
    """The FP32 PCG's V-cycle keeps FP16 copies of its level matrices, each scaled by a power of
    two: the same mesh in a length unit h times the nanometre (the stiffness scales as h, the
    right-hand side as h^2, so the coarse matrices at h = 1e3 exceed the FP16 maximum and at
    h = 1e-3 sit near its subnormals unless scaled) takes the iterations it takes in nanometres
    on a hierarchy of three or more levels, and its solution is the nanometre one times h."""
    from mumag_matrixfree2_amd.context import MeshContext

    A, K1, Js, k = ndfeb_materials()
    msh = cube_c0
    out = {}
    for unit in (1.0, h):
        ctx = MeshContext(msh["knt"] * unit, msh["ijk"][:, :4], msh["ijk"][:, 4], A, K1, Js, k,
                          boundary_mask=msh["boundary_mask"])
        try:
            ctx.setup_precond("amgcl", max_coarse=60)
            levels = [lv["A"].shape[0] for lv in ctx.amg_levels]
            m = tilted_m0(ctx.N, seed=36, sigma=0.4)
            U64, it64, _ = ctx.solve_U(m, np.zeros(ctx.N), 1e-10)
            ctx.set_precision("fp32pcg")
            U, it, rr = ctx.solve_U(m, np.zeros(ctx.N), 1e-10)
        finally:
            ctx.close()
        print(f"[length unit {unit:g}] levels={levels} it={it} (fp64 {it64}) rr={rr:.3e}", flush=True)
        assert len(levels) >= 3 and rr <= 1e-10 and rel(U, U64) <= 1e-8
        out[unit] = (U, it)
    assert abs(out[h][1] - out[1.0][1]) <= 1
    assert rel(out[h][0], h * out[1.0][0]) <= 1e-8


@pytest.mark.parametrize("method", ["pcohen_hs", "lbfgs"])
def test_fp32_pcg_field_step_energy(mixed_ctx, oracle_c0, method):
    from mumag_matrixfree2_amd.context import make_params

    ctx = mixed_ctx
    m0 = tilted_m0(ctx.N, seed=35, sigma=0.1)
    B = np.array([0.017452, 0.0, 0.999848]) * (-0.4)
    p = make_params(method, tau_f=1e-10, cg_tol=1e-10)
    out = {}
    for prec in ("fp64", "fp32pcg"):
        ctx.set_precision(prec)
        ctx.state_set(m0, np.zeros(ctx.N))
        out[prec] = ctx.minimize_resident(p, B)
    ctx.set_precision("fp64")
    m, U = ctx.state_get()
    scale = sum(abs(v) for v in oracle_c0.energy_terms(m, U, B).values())
    print(f"[fp32pcg {method}] E64={out['fp64']['E']:.15e} E32={out['fp32pcg']['E']:.15e} "
          f"it={out['fp64']['iters']}/{out['fp32pcg']['iters']} pcg={out['fp64']['demag_iters']}/"
          f"{out['fp32pcg']['demag_iters']}", flush=True)
    assert out["fp32pcg"]["converged"]
    assert abs(out["fp32pcg"]["E"] - out["fp64"]["E"]) <= 1e-10 * scale


def test_fp32_energy_densities_to_fp32_accuracy(mixed_ctx, oracle_c0):
    """Form 'fp32all' computes the element energy densities and gradient terms in FP32 (FP64
    accumulation and compensated sum): energy and gradient agree with the oracle to FP32
    accuracy (1e-6 of the energy scale, 1e-5 relative gradient), not to the FP64 bar."""
    ctx = mixed_ctx
    m = tilted_m0(ctx.N, seed=36, sigma=0.3)
    U = np.random.default_rng(36).standard_normal(ctx.N)
    B = np.array([0.01, 0.0, -0.3])
    ctx.set_precision("fp32all")
    try:
        E, g = ctx.energy_grad(m, U, B)
    finally:
        ctx.set_precision("fp64")
    E_ref, g_ref = oracle_c0.energy_and_grad(m, U, B)
    scale = sum(abs(v) for v in oracle_c0.energy_terms(m, U, B).values())
    print(f"[fp32 energy] |dE|/scale={abs(E - E_ref) / scale:.2e} |dg|/|g|={rel(g, g_ref):.2e}", flush=True)
    assert abs(E - E_ref) <= 1e-6 * scale and rel(g, g_ref) <= 1e-5
