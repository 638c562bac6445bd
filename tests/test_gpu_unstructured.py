"""GPU parity on an unstructured tet mesh (a Delaunay tetrahedralisation of scattered points).

The other GPU tests use six-tet hex splits, where every edge-star cluster is a whole hex. Here
the longest-edge groups are partial stars, so the apply runs clusters of every length and
shape (chains, packed chains, rings closed through a duplicated slot) and the per-element path
runs ordinary tets; both are checked against the oracle's assembled operators, and the full
solve and energy kernels run on the same mesh.
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
def delaunay_pair():
    from scipy.spatial import Delaunay

    from mumag_matrixfree2_amd.context import MeshContext
    from mumag_matrixfree2_amd.mesh import morton_reorder, outer_boundary_mask
    from oracle.fem import Problem

    rng = np.random.default_rng(11)
    inner = rng.uniform(-9.0, 9.0, size=(2500, 3))
    g = np.linspace(-10.0, 10.0, 9)
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    grid = np.column_stack([X.ravel(), Y.ravel(), Z.ravel()])
    shell = grid[np.max(np.abs(grid), axis=1) >= 10.0]
    pts = np.vstack([inner, shell])
    tri = Delaunay(pts)
    tets = tri.simplices.astype(np.int32)
    x0 = pts[tets]
    vol = np.abs(np.linalg.det(x0[:, 1:] - x0[:, :1])) / 6.0
    tets = tets[vol > 1e-6]
    cen = pts[tets].mean(axis=1)
    mat = np.where(np.linalg.norm(cen, axis=1) < 6.0, 1, 2).astype(np.int32)
    mesh = morton_reorder(pts, np.column_stack([tets, mat]))
    knt, ijk = mesh["knt"], mesh["ijk"]
    bnd = outer_boundary_mask(ijk[:, :4], knt.shape[0])
    A, K1, Js, k = ndfeb_materials(axis=(0.3, 0.1, 1.0))
    ctx = MeshContext(knt, ijk[:, :4], ijk[:, 4], A, K1, Js, k, boundary_mask=bnd)
    prob = Problem(knt, ijk, A, K1, Js, k, boundary_mask=bnd)
    print(f"[delaunay] N={ctx.N} E={ctx.E} clusters={ctx.n_clusters}", flush=True)
    return ctx, prob


@pytest.mark.parametrize("shape", ["cluster:256", "cluster:64", "persistent:2", "element:256", "cluster:256:512",
                                   "cluster:256:1024"])
def test_apply_on_unstructured_mesh(delaunay_pair, shape):
    ctx, prob = delaunay_pair
    assert ctx.n_clusters > ctx.E / 6  # partial stars: more clusters than on a hex split
    unit, param, *nb = shape.split(":")
    ctx.set_cluster_block_nodes(int(nb[0]) if nb else 256)
    ctx.set_apply_shape(unit, int(param))
    try:
        x = np.random.default_rng(2).standard_normal(ctx.N)
        assert rel(ctx.apply_A(x), prob.A @ x) <= TOL
    finally:
        ctx.set_cluster_block_nodes(256)
        ctx.set_apply_shape("cluster", 256)


def test_solve_and_energy_on_unstructured_mesh(delaunay_pair):
    from oracle.solvers import solve_U as oracle_solve_U

    ctx, prob = delaunay_pair
    ctx.setup_precond("amgcl")
    m = tilted_m0(ctx.N, seed=4, sigma=0.2)
    assert rel(ctx.rhs(m), prob.D @ m.ravel()) <= TOL
    U, it, rr = ctx.solve_U(m, np.zeros(ctx.N), 1e-10)
    U_ref, _, _ = oracle_solve_U(prob, m, np.zeros(ctx.N), 1e-12, precond="direct")
    assert rr <= 1e-10 and rel(U, U_ref) <= 1e-7
    B = np.array([0.1, 0.0, -0.4])
    E, g = ctx.energy_grad(m, U_ref, B)
    E_ref, g_ref = prob.energy_and_grad(m, U_ref, B)
    assert abs(E - E_ref) <= 1e-12 * max(1.0, abs(E_ref))
    assert rel(g, g_ref) <= TOL
    v = np.random.default_rng(3).standard_normal((ctx.N, 3))
    assert rel(ctx.local_grad(v), prob.local_grad(v)) <= TOL
