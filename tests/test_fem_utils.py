"""Host side of the geometry boundary: the reference's default TetGeom carries J^-T and volumes
but no node coordinates (``src/loop.py:899-908``); the coordinates are recovered from the
Jacobians up to a translation per connected component."""

from __future__ import annotations

import numpy as np
import pytest


def _delaunay(seed=11, n=800):
    from scipy.spatial import Delaunay

    rng = np.random.default_rng(seed)
    pts = rng.uniform(-5.0, 5.0, size=(n, 3))
    tets = Delaunay(pts).simplices.astype(np.int32)
    x = pts[tets]
    keep = np.abs(np.linalg.det(x[:, 1:] - x[:, :1])) / 6.0 > 1e-6
    return pts, tets[keep]


@pytest.mark.parametrize("field", ["JinvT", "grad_phi"])
def test_coordinates_recovered_from_reference_geometry(field):
    """On an unstructured mesh (random corner order, so about half the tets are flipped by the
    orientation fix) the recovered coordinates equal the originals shifted by one vector, to
    rounding; the recomputed J^-T equals the given one."""
    from mumag_matrixfree2_amd.fem_utils import TetGeom, node_coordinates
    from oracle.fem import compute_volume_JinvT, grad_phi_from_JinvT

    pts, tets = _delaunay()
    conn, vol, JinvT = compute_volume_JinvT(pts, tets)
    kw = {"JinvT": JinvT} if field == "JinvT" else {"grad_phi": grad_phi_from_JinvT(JinvT)}
    geom = TetGeom(conn=conn, volume=vol, mat_id=np.ones(len(conn), np.int32), x_nodes=None, **kw)
    x = node_coordinates(geom)
    shift = x - pts
    assert np.abs(shift - shift[conn[0, 0]]).max() <= 1e-12 * np.abs(pts).max()
    _, vol2, JinvT2 = compute_volume_JinvT(x, conn)
    np.testing.assert_allclose(vol2, vol, rtol=1e-11, atol=1e-12 * vol.max())
    np.testing.assert_allclose(JinvT2, JinvT, rtol=1e-9, atol=1e-12 * np.abs(JinvT).max())


def test_coordinates_of_disconnected_meshes_and_unused_nodes():
    """Two disjoint meshes (one translation each) and an unused node (left at the origin)."""
    from mumag_matrixfree2_amd.fem_utils import coordinates_from_jacobians
    from oracle.fem import compute_volume_JinvT

    p1, t1 = _delaunay(1, 200)
    p2, t2 = _delaunay(2, 150)
    pts = np.vstack([p1, p2 + 100.0, [[7.0, 7.0, 7.0]]])
    tets = np.vstack([t1, t2 + len(p1)])
    conn, vol, JinvT = compute_volume_JinvT(pts, tets)
    x = coordinates_from_jacobians(conn, JinvT, len(pts))
    for sl in (slice(0, len(p1)), slice(len(p1), len(p1) + len(p2))):
        d = x[sl] - pts[sl]
        assert np.abs(d - d[0]).max() <= 1e-11 * np.abs(pts).max()
    assert np.array_equal(x[-1], np.zeros(3))


def test_geometry_without_coordinates_or_jacobians_is_rejected():
    from mumag_matrixfree2_amd.fem_utils import TetGeom, node_coordinates

    with pytest.raises(ValueError, match="x_nodes"):
        node_coordinates(TetGeom(conn=np.zeros((1, 4), np.int32), volume=np.ones(1)))
