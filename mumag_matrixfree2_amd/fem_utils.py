"""Geometry container of the reference API (``src/fem_utils.py:22-40``) for the HIP backend.

``TetGeom`` keeps the reference's field names. The HIP backend computes shape-function
gradients and volumes on the GPU from node coordinates. The reference's default geometry
(``--geom-backend stored_JinvT``, ``src/loop.py:899-908``) carries ``JinvT`` and ``volume`` but
no ``x_nodes``; for it the node coordinates are recovered from the Jacobians
(``coordinates_from_jacobians``): J = (J^-T)^-T holds the edge vectors x_a - x_0 of each tet, so
walking a spanning tree of the mesh gives every node's position up to one translation per
connected component, which no operator depends on. ``grad_phi`` (the ``stored_grad_phi``
backend) carries J^-T in its rows 1..3.
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


def coordinates_from_jacobians(conn: np.ndarray, JinvT: np.ndarray, n_nodes: int | None = None) -> np.ndarray:
    """Node coordinates (N, 3) whose tets have the given inverse-transposed Jacobians.

    J_e = inv(JinvT_e^T) has columns x_a - x_0 (a = 1..3) (``compute_volume_JinvT``,
    ``src/loop.py:56-96``). The edges (corner 0 -> corner a) of all tets form a graph; a
    breadth-first spanning tree of each connected component fixes every node relative to the
    component's root (placed at the origin), and the positions are summed along the tree by
    pointer doubling (log2 depth vectorised passes). Nodes in no tet stay at the origin.
    """
    import scipy.sparse as sp
    from scipy.sparse.csgraph import breadth_first_order, connected_components

    conn = np.asarray(conn, dtype=np.int64)[:, :4]
    E = conn.shape[0]
    N = int(conn.max()) + 1 if n_nodes is None else int(n_nodes)
    J = np.linalg.inv(np.swapaxes(np.asarray(JinvT, dtype=np.float64), 1, 2))  # (E, 3, 3), columns = edges
    src = np.repeat(conn[:, 0], 3)
    dst = conn[:, 1:].reshape(-1)
    vec = np.swapaxes(J, 1, 2).reshape(-1, 3)  # row 3e + (a-1) = x_a - x_0
    # directed edges in both directions, first occurrence of each pair
    u = np.concatenate([src, dst])
    v = np.concatenate([dst, src])
    w = np.concatenate([vec, -vec])
    key = u * N + v
    key_u, first = np.unique(key, return_index=True)
    w = w[first]
    G = sp.csr_matrix((np.ones(key_u.size), (key_u // N, key_u % N)), shape=(N, N))
    ncomp, label = connected_components(G, directed=False)
    used = np.zeros(N, dtype=bool)
    used[conn.reshape(-1)] = True
    pred = np.arange(N)
    step = np.zeros((N, 3))
    for c in range(ncomp):
        nodes = np.where(label == c)[0]
        if not used[nodes[0]]:
            continue
        order, pr = breadth_first_order(G, int(nodes[0]), directed=False, return_predecessors=True)
        child = order[1:]
        par = pr[child]
        idx = np.searchsorted(key_u, par * N + child)
        pred[child] = par
        step[child] = w[idx]
    # pointer doubling: pos[v] = step[v] + pos[pred[v]], roots have pred = self and step = 0
    pos = step.copy()  # invariant: pos[v] = x_v - x_anc[v]
    anc = pred.copy()
    while np.any(anc[anc] != anc):
        pos = pos + pos[anc]
        anc = anc[anc]
    return pos


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


def _norm_tables(geom, A_lookup, K1_lookup, Js_lookup, k_easy_lookup):
    A = np.ascontiguousarray(A_lookup, dtype=np.float64).reshape(-1)
    G = A.size
    K1 = np.ascontiguousarray(K1_lookup, dtype=np.float64).reshape(-1)
    Js = np.ascontiguousarray(Js_lookup, dtype=np.float64).reshape(-1)
    k = np.asarray(k_easy_lookup, dtype=np.float64).reshape(-1, 3)
    if k.shape[0] < G:
        k = np.vstack([k, np.tile([0.0, 0.0, 1.0], (G - k.shape[0], 1))])
    return A, K1, Js, np.ascontiguousarray(k[:G])


def _same(a, b) -> bool:
    if a is None or b is None:
        return a is None and b is None
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and np.array_equal(a, b)


# Contexts built for one TetGeom by the factories, so that make_energy_kernels, make_solve_U and
# make_minimizer called on the same geometry share device state where their inputs agree.
# Entries hold a reference to the geometry, so its id stays unique while cached.
_CACHE: list[dict] = []


def clear_context_cache() -> None:
    """Drop the factories' shared contexts (their device memory is freed with the last reference)."""
    _CACHE.clear()


def _find(geom, **want):
    for e in _CACHE:
        if e["geom"] is not geom:
            continue
        if all(_same(e[k], v) if k in ("A", "K1", "Js", "k", "mask") else e[k] == v for k, v in want.items()):
            return e
    return None


def context_for(geom: TetGeom, A_lookup, K1_lookup, Js_lookup, k_easy_lookup, boundary_mask=None, reg=1e-12,
                geom_mode="on_the_fly", shared: bool = True):
    """The device context of a TetGeom with these material tables, Dirichlet mask and
    regularisation (raises when the GPU library is unavailable). With ``shared`` an existing
    context of the same geometry and inputs is reused."""
    A, K1, Js, k = _norm_tables(geom, A_lookup, K1_lookup, Js_lookup, k_easy_lookup)
    mask = None if boundary_mask is None else np.asarray(boundary_mask, dtype=np.float64)
    if shared:
        e = _find(geom, A=A, K1=K1, Js=Js, k=k, mask=mask, reg=float(reg), geom_mode=geom_mode)
        if e is not None:
            return e["ctx"]
    from .context import MeshContext

    knt = node_coordinates(geom)
    ctx = MeshContext(knt, geom.conn, geom.material_ids(), A, K1, Js, k, boundary_mask=mask, reg=reg,
                      geom=geom_mode)
    if geom.volume is not None:
        _, Ve, _ = ctx.element_table()
        vol = np.asarray(geom.volume, dtype=np.float64)[: Ve.size]
        if np.max(np.abs(Ve - vol)) > 1e-9 * max(float(np.max(np.abs(vol))), 1e-300):
            raise ValueError("geom.volume does not match the volumes of the mesh geometry")
    if shared:
        _CACHE.append({"geom": geom, "A": A, "K1": K1, "Js": Js, "k": k, "mask": mask, "reg": float(reg),
                       "geom_mode": geom_mode, "ctx": ctx})
    return ctx


def context_with_tables(geom: TetGeom, A_lookup, K1_lookup, Js_lookup, k_easy_lookup):
    """A shared context of this geometry built with these material tables (any Dirichlet mask),
    or None. Energies and gradients do not depend on the Poisson mask."""
    A, K1, Js, k = _norm_tables(geom, A_lookup, K1_lookup, Js_lookup, k_easy_lookup)
    e = _find(geom, A=A, K1=K1, Js=Js, k=k)
    return None if e is None else e["ctx"]


def material_tables_of(geom: TetGeom, Js_lookup):
    """Full material tables (A, K1, Js, k) of a context already built for this geometry whose
    Js table equals ``Js_lookup`` (e.g. by make_energy_kernels), or None. The Poisson operators
    depend on Js only; a solver context built with the full tables also serves the minimiser."""
    Js = np.ascontiguousarray(Js_lookup, dtype=np.float64).reshape(-1)
    for e in _CACHE:
        if e["geom"] is geom and _same(e["Js"], Js) and (np.any(e["A"] != 0) or np.any(e["K1"] != 0)):
            return e["A"], e["K1"], e["Js"], e["k"]
    return None


def tables_match(ctx, A_lookup, K1_lookup, Js_lookup, k_easy_lookup) -> bool:
    """Whether a context was built with these material tables."""
    for e in _CACHE:
        if e["ctx"] is ctx:
            A, K1, Js, k = _norm_tables(e["geom"], A_lookup, K1_lookup, Js_lookup, k_easy_lookup)
            return _same(e["A"], A) and _same(e["K1"], K1) and _same(e["Js"], Js) and _same(e["k"], k)
    return False


def cached_inputs(ctx):
    """(boundary_mask, reg) a shared context was built with, or None for a context not built here."""
    for e in _CACHE:
        if e["ctx"] is ctx:
            return e["mask"], e["reg"]
    return None


def node_coordinates(geom: TetGeom) -> np.ndarray:
    """Node coordinates of a TetGeom: ``x_nodes``, else recovered from ``JinvT`` or from
    ``grad_phi`` (rows 1..3 of grad_phi are the columns of J^-T). Raises if none is given."""
    if geom.x_nodes is not None:
        return np.asarray(geom.x_nodes, dtype=np.float64)
    conn = np.asarray(geom.conn)
    JinvT = geom.JinvT
    if JinvT is None and geom.grad_phi is not None:
        JinvT = np.swapaxes(np.asarray(geom.grad_phi, dtype=np.float64)[:, 1:4, :], 1, 2)
    if JinvT is None:
        raise ValueError("the HIP backend needs geom.x_nodes, geom.JinvT or geom.grad_phi")
    E = conn.shape[0]
    return coordinates_from_jacobians(conn, np.asarray(JinvT)[:E])
