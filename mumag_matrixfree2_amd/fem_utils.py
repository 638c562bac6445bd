"""Geometry container of the reference API (``src/fem_utils.py:22-40``) for the HIP backend.

``TetGeom`` keeps the reference's field names. The HIP backend needs the node
coordinates (``x_nodes``): shape-function gradients and volumes are computed on the GPU
from them, so ``volume`` / ``grad_phi`` / ``JinvT`` may be left as None.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass(frozen=True)
class TetGeom:
    conn: np.ndarray
    volume: np.ndarray | None = None
    mat_id: np.ndarray | None = None
    grad_phi: np.ndarray | None = None
    JinvT: np.ndarray | None = None
    x_nodes: np.ndarray | None = None

    @property
    def n_nodes(self) -> int:
        if self.x_nodes is not None:
            return int(np.asarray(self.x_nodes).shape[0])
        return int(np.max(self.conn)) + 1

    def material_ids(self) -> np.ndarray:
        if self.mat_id is None:
            return np.ones(np.asarray(self.conn).shape[0], dtype=np.int32)
        return np.asarray(self.mat_id, dtype=np.int32)


def geom_from_mesh(knt: np.ndarray, ijk: np.ndarray) -> TetGeom:
    """TetGeom from a mesh ``(knt (N,3), ijk (E,4|5))``; material ids default to 1, and
    0-based ids are shifted to 1-based as ``src/loop.py:791-801`` does."""
    ijk = np.asarray(ijk)
    conn = ijk[:, :4].astype(np.int32)
    if ijk.shape[1] == 5:
        mat = ijk[:, 4].astype(np.int32)
        if mat.min() == 0:
            mat = mat + 1
    else:
        mat = np.ones(conn.shape[0], dtype=np.int32)
    return TetGeom(conn=conn, mat_id=mat, x_nodes=np.asarray(knt, dtype=np.float64))


def context_for(geom: TetGeom, A_lookup, K1_lookup, Js_lookup, k_easy_lookup, boundary_mask=None, reg=1e-12,
                geom_mode="on_the_fly"):
    """Create the device context of a TetGeom (raises when the GPU library is unavailable)."""
    from .context import MeshContext

    if geom.x_nodes is None:
        raise ValueError("the HIP backend needs geom.x_nodes (node coordinates)")
    G = int(np.asarray(A_lookup).size)
    k = np.asarray(k_easy_lookup, dtype=np.float64).reshape(-1, 3)
    if k.shape[0] < G:
        k = np.vstack([k, np.tile([0.0, 0.0, 1.0], (G - k.shape[0], 1))])
    return MeshContext(geom.x_nodes, geom.conn, geom.material_ids(), A_lookup, K1_lookup, Js_lookup, k,
                       boundary_mask=boundary_mask, reg=reg, geom=geom_mode)
