"""GPU parity of the element kernels against the oracle, through the C-ABI (FP64).

Bar: relative L2 difference <= 1e-12 (the FP64 path's parity target); the only
differences allowed are summation order (LDS / global FP64 atomics) and the on-the-fly
gradient arithmetic.
"""

from __future__ import annotations

import numpy as np
import pytest

from conftest import ndfeb_materials, tilted_m0

pytestmark = pytest.mark.gpu

TOL = 1e-12


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-300))


@pytest.fixture(scope="module")
def ctx_c0(cube_c0):
    from mumag_matrixfree2_amd.context import MeshContext

    A, K1, Js, k = ndfeb_materials()
    m = cube_c0
    print(f"[ctx_c0] N={m['knt'].shape[0]} E={m['ijk'].shape[0]}", flush=True)
    return MeshContext(m["knt"], m["ijk"][:, :4], m["ijk"][:, 4], A, K1, Js, k, boundary_mask=m["boundary_mask"])


def test_context_setup_matches_oracle(ctx_c0, oracle_c0):
    assert abs(ctx_c0.V_mag - oracle_c0.V_mag) <= 1e-12 * oracle_c0.V_mag
    M, invMr, invMp = ctx_c0.nodal()
    assert rel(M, oracle_c0.M_nodal) <= TOL
    assert rel(invMr, oracle_c0.inv_M_rel) <= TOL
    assert rel(invMp, oracle_c0.inv_M_prec) <= TOL
    assert ctx_c0.max_block_nodes <= 256


@pytest.mark.parametrize("geom", ["on_the_fly", "stored_f64"])
def test_poisson_apply(ctx_c0, oracle_c0, geom):
    rng = np.random.default_rng(1)
    x = rng.standard_normal(ctx_c0.N)
    ctx_c0.set_geometry_mode(geom)
    try:
        y = ctx_c0.apply_A(x)
    finally:
        ctx_c0.set_geometry_mode("on_the_fly")
    assert rel(y, oracle_c0.A @ x) <= TOL


@pytest.mark.parametrize("shape", ["persistent:1", "persistent:3", "cluster:256", "cluster:64", "element:128",
                                   "element:64", "cluster:256:512", "cluster:256:1024", "cluster:256:2048",
                                   "cluster:256:512:gather", "cluster:256:1024:gather",
                                   "cluster:256:2048:gather"])
def test_poisson_apply_every_work_shape(ctx_c0, oracle_c0, shape):
    """Every apply shape the autotuner can pick gives the oracle's A x, in one apply and inside
    a PCG solve (the persistent grid writes its x.Ax partials per block); the third field is the
    cluster block size (slot indices above 255 carry their high bits in the cluster mask), the
    fourth the staging of those blocks in the FP32 kernels (entry records instead of gathers; the
    FP32 PCG's kernels with it are checked in the mixed-precision tests)."""
    from oracle.solvers import solve_U as oracle_solve_U

    unit, param, *rest = shape.split(":")
    nb = int(rest[0]) if rest else 256
    staging = rest[1] if len(rest) > 1 else "records"
    rng = np.random.default_rng(5)
    x = rng.standard_normal(ctx_c0.N)
    ctx_c0.set_cluster_block_nodes(nb)
    ctx_c0.set_apply_shape(unit, int(param))
    ctx_c0.set_cluster_staging(staging)
    try:
        assert rel(ctx_c0.apply_A(x), oracle_c0.A @ x) <= TOL
        if ctx_c0.precond_type is None:
            ctx_c0.setup_precond("jacobi")
        m = tilted_m0(ctx_c0.N, seed=3)
        U, it, rr = ctx_c0.solve_U(m, np.zeros(ctx_c0.N), 1e-10)
        U_ref, _, _ = oracle_solve_U(oracle_c0, m, np.zeros(ctx_c0.N), 1e-12, precond="direct")
        assert rr <= 1e-10 and rel(U, U_ref) <= 1e-7
        if rest:
            assert ctx_c0.cluster_block_nodes == nb and ctx_c0.n_blocks_cl >= 1
            ctx_c0.set_precision("mixed")
            try:
                Um, _, rrm = ctx_c0.solve_U(m, np.zeros(ctx_c0.N), 1e-10)
            finally:
                ctx_c0.set_precision("fp64")
            assert rrm <= 1e-10 and rel(Um, U_ref) <= 1e-7
    finally:
        ctx_c0.set_cluster_staging("records")
        ctx_c0.set_cluster_block_nodes(256)
        ctx_c0.set_apply_shape("cluster", 256)


def test_autotuner_keeps_the_fastest_shape(ctx_c0):
    from mumag_matrixfree2_amd.autotune import tune_apply

    res = tune_apply(ctx_c0, reps=5)
    times = res["us_per_apply"]
    from mumag_matrixfree2_amd.autotune import CLUSTER_BLOCK_NODES, SHAPES

    shapes = {k: v for k, v in times.items() if not k.endswith(":colour")}
    assert len(shapes) == len(SHAPES) + len(CLUSTER_BLOCK_NODES) and all(t > 0 for t in times.values())
    key = f"{res['unit']}:{res['param']}" + ("" if res["cluster_block_nodes"] == 256 else
                                            f":{res['cluster_block_nodes']}")
    assert shapes[key] == min(shapes.values())
    if res["unit"] == "cluster":  # the colour-sequenced scatter was timed on the chosen shape
        assert f"{key}:colour" in times
        assert res["scatter"] == ("colour" if times[f"{key}:colour"] < times[key] else "atomic")
    assert ctx_c0.apply_scatter == res["scatter"]
    assert ctx_c0.apply_shape == {"unit": res["unit"], "param": res["param"]}
    assert ctx_c0.cluster_block_nodes == res["cluster_block_nodes"]
    ctx_c0.set_apply_scatter("atomic")
    ctx_c0.set_cluster_block_nodes(256)
    ctx_c0.set_apply_shape("cluster", 256)


def test_poisson_apply_fp32_geometry(ctx_c0, oracle_c0):
    rng = np.random.default_rng(2)
    x = rng.standard_normal(ctx_c0.N)
    ctx_c0.set_geometry_mode("stored_f32")
    try:
        y = ctx_c0.apply_A(x)
    finally:
        ctx_c0.set_geometry_mode("on_the_fly")
    assert rel(y, oracle_c0.A @ x) <= 1e-6


def test_poisson_diag(ctx_c0, oracle_c0):
    assert rel(ctx_c0.diag(), oracle_c0.A.diagonal()) <= TOL


def test_rhs(ctx_c0, oracle_c0):
    m = tilted_m0(ctx_c0.N, seed=3, sigma=0.3)
    assert rel(ctx_c0.rhs(m), oracle_c0.rhs(m)) <= TOL


def test_energy_and_grad(ctx_c0, oracle_c0):
    m = tilted_m0(ctx_c0.N, seed=4, sigma=0.3)
    U = np.random.default_rng(5).standard_normal(ctx_c0.N) * 0.1
    B = np.array([0.01, -0.02, -0.3])
    E, g = ctx_c0.energy_grad(m, U, B)
    E_ref, g_ref = oracle_c0.energy_and_grad(m, U, B)
    terms = oracle_c0.energy_terms(m, U, B)
    scale = sum(abs(v) for v in terms.values())
    assert abs(E - E_ref) <= TOL * scale
    assert rel(g, g_ref) <= TOL


def test_local_grad(ctx_c0, oracle_c0):
    v = np.random.default_rng(6).standard_normal((ctx_c0.N, 3))
    assert rel(ctx_c0.local_grad(v), oracle_c0.local_grad(v)) <= TOL


def test_local_grad_per_tet_materials():
    """The local Hessian action on an 8-grain polycrystal with a grain-boundary phase (per-grain
    easy axes, A and K1 changing across the boundaries, an airbox) against the oracle: in FP64 to
    the parity bar, and in the FP32 element arithmetic of the mixed forms to FP32 accuracy."""
    from mumag_matrixfree2_amd.context import MeshContext
    from mumag_matrixfree2_amd.mesh import polycrystal_cube, random_easy_axes
    from oracle.fem import Problem

    from conftest import MU0

    mesh = polycrystal_cube(30.0, 12, 3, n_grains=8, gb_width=1.5, seed=4)
    Kd = 1.61**2 / (2 * MU0)
    G = 10
    A, K1, Js, k = np.zeros(G), np.zeros(G), np.zeros(G), np.tile([0.0, 0.0, 1.0], (G, 1))
    A[:8], K1[:8], Js[:8] = 7.7e-12 * 1e18 / Kd, 4.3e6 / Kd, 1.0
    k[:8] = random_easy_axes(8, 5)
    A[8], Js[8] = 1.5e-12 * 1e18 / Kd, 1.0
    ctx = MeshContext(mesh["knt"], mesh["ijk"][:, :4], mesh["ijk"][:, 4], A, K1, Js, k,
                      boundary_mask=mesh["boundary_mask"])
    prob = Problem(mesh["knt"], mesh["ijk"], A, K1, Js, k, boundary_mask=mesh["boundary_mask"])
    try:
        v = np.random.default_rng(16).standard_normal((ctx.N, 3))
        ref = prob.local_grad(v)
        out = {}
        for prec in ("fp64", "mixed"):
            ctx.set_precision(prec)
            out[prec] = ctx.local_grad(v)
        ctx.set_precision("fp64")
        assert rel(out["fp64"], ref) <= TOL
        assert rel(out["mixed"], ref) <= 1e-5
    finally:
        ctx.close()


def test_j_parallel(ctx_c0, oracle_c0):
    m = tilted_m0(ctx_c0.N, seed=7, sigma=0.5)
    h = np.array([0.017452, 0.0, 0.999848])
    assert abs(ctx_c0.j_parallel(m, h) - oracle_c0.J_parallel(m, h)) <= 1e-13


@pytest.mark.parametrize("k", [1, 2, 4, 8])
def test_multi_vector_apply(ctx_c0, oracle_c0, k):
    """k vectors at once: each column equals the oracle's A x (and the single-vector apply)."""
    X = np.random.default_rng(70 + k).standard_normal((k, ctx_c0.N))
    Y = ctx_c0.apply_A_multi(X)
    for r in range(k):
        assert rel(Y[r], oracle_c0.A @ X[r]) <= TOL
        assert rel(Y[r], ctx_c0.apply_A(X[r])) <= 1e-14


def test_rhs_width_tuner_reports_per_vector_cost(ctx_c0):
    from mumag_matrixfree2_amd.autotune import RHS_WIDTHS, tune_rhs_width

    res = tune_rhs_width(ctx_c0, reps=3)
    assert set(res["us_per_vector"]) == set(RHS_WIDTHS) and res["single_us"] > 0
    assert res["width"] in RHS_WIDTHS
    if res["width"] > 1:
        assert res["us_per_vector"][res["width"]] < res["single_us"]
    assert ctx_c0.cluster_block_nodes == 256 and ctx_c0.apply_shape == {"unit": "cluster", "param": 256}
