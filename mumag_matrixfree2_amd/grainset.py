"""A set of independent grains as one segmented context.

The grains of a grain set (each its own mesh, airbox and materials) are concatenated into one
mesh with disjoint node ranges, one segment per grain. One context then holds them all, so every
kernel launch covers every grain: the element kernels see a block-diagonal operator, the host
AMG set-up on the union matrix produces a block-diagonal hierarchy (aggregates never cross a
grain), and ``solve_U_segmented`` runs one PCG per grain with per-grain scalars and stopping
tests in shared launches. The reference has no batched form: it solves one system per call
(``src/poisson_solve.py:748-771``).
"""

from __future__ import annotations

import numpy as np

from .context import MeshContext


def union_mesh(grains: list[dict], tables: list[tuple]) -> dict:
    """Concatenate grain meshes. ``grains[i]``: ``knt`` (N_i,3), ``ijk`` (E_i,5) with 1-based
    material ids into ``tables[i] = (A, K1, Js, k)``, ``boundary_mask`` (N_i,). Returns the union
    ``knt``, ``conn``, ``mat_id``, material tables, ``boundary_mask`` and node offsets ``segments``."""
    knt, conn, mat, bnd = [], [], [], []
    A, K1, Js, k = [], [], [], []
    off, moff = [0], 0
    for g, (A_g, K1_g, Js_g, k_g) in zip(grains, tables):
        n = g["knt"].shape[0]
        knt.append(np.asarray(g["knt"], dtype=np.float64))
        conn.append(np.asarray(g["ijk"])[:, :4].astype(np.int64) + off[-1])
        mat.append(np.asarray(g["ijk"])[:, 4].astype(np.int64) + moff)
        bnd.append(np.asarray(g["boundary_mask"], dtype=np.float64))
        A.append(np.asarray(A_g, dtype=np.float64))
        K1.append(np.asarray(K1_g, dtype=np.float64))
        Js.append(np.asarray(Js_g, dtype=np.float64))
        k.append(np.asarray(k_g, dtype=np.float64).reshape(-1, 3))
        moff += len(A_g)
        off.append(off[-1] + n)
    return {"knt": np.vstack(knt), "conn": np.vstack(conn).astype(np.int32),
            "mat_id": np.concatenate(mat).astype(np.int32), "A": np.concatenate(A), "K1": np.concatenate(K1),
            "Js": np.concatenate(Js), "k": np.vstack(k), "boundary_mask": np.concatenate(bnd),
            "segments": np.asarray(off, dtype=np.int32)}


def grain_set_context(grains: list[dict], tables: list[tuple], **ctx_kwargs) -> MeshContext:
    """One segmented ``MeshContext`` over all grains (``union_mesh``); at most 255 materials."""
    u = union_mesh(grains, tables)
    return MeshContext(u["knt"], u["conn"], u["mat_id"], u["A"], u["K1"], u["Js"], u["k"],
                       boundary_mask=u["boundary_mask"], segments=u["segments"], **ctx_kwargs)


def split(ctx: MeshContext, v: np.ndarray) -> list[np.ndarray]:
    """Per-grain views of a nodal array of the union (first axis = nodes)."""
    s = ctx.segments
    return [v[s[i]:s[i + 1]] for i in range(s.size - 1)]
