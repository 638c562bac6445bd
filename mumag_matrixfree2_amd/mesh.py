"""Structured tetrahedral meshes for the hysteresis workloads, and Morton reordering.

Three pieces, all host-side NumPy, vectorised (no per-node Python loops):

* ``grid_box`` -- the reference's regular-grid box mesher
  (``src/mesh.py:1202-1272``, ``mesh_backend_grid_box``): nodes on
  ``linspace(-L/2, L/2, n+1)`` per axis, node index ``i + (nx+1)*(j + (ny+1)*k)``,
  and every hex split into the same six tets in the same order
  ``(A,B,D,H) (A,B,F,H) (A,C,D,H) (A,C,G,H) (A,E,F,H) (A,E,G,H)``.
  The 20 nm / h=2 case reproduces the reference fixture ``tests/single_solid.npz``.
* ``graded_airbox_cube`` -- a magnet cube embedded in a structured, geometrically
  graded air shell. It takes the role of ``add_shell.run_add_shell_pipeline``
  (``src/add_shell.py:691``), which needs TetGen; here each air layer is one
  hex cell whose thickness grows by ``grading`` per layer (``h*grading**k``).
  Air tets get material id 2 (Js=A=K1=0); the outer surface nodes are the
  Dirichlet set (``boundary_mask``, semantics of ``src/add_shell.py:105-136``).
* ``morton_reorder`` -- the reference's GPU reordering (``src/reorder_mesh.py:19-30,86-101``):
  10-bit-per-axis quantisation over the bounding box, interleaved x<<2|y<<1|z,
  nodes sorted by their key and elements by the key of their centroid.
"""

from __future__ import annotations

import numpy as np

# Six-tet split of one hex; corner letters follow src/mesh.py:1244-1265.
# Corners: A=(0,0,0) B=(1,0,0) C=(0,1,0) D=(1,1,0) E=(0,0,1) F=(1,0,1) G=(0,1,1) H=(1,1,1)
_CORNER_OFFSETS = {
    "A": (0, 0, 0), "B": (1, 0, 0), "C": (0, 1, 0), "D": (1, 1, 0),
    "E": (0, 0, 1), "F": (1, 0, 1), "G": (0, 1, 1), "H": (1, 1, 1),
}
_SIX_TETS = ("ABDH", "ABFH", "ACDH", "ACGH", "AEFH", "AEGH")


def tensor_grid_tets(xs: np.ndarray, ys: np.ndarray, zs: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """Tetrahedralise the tensor grid ``xs x ys x zs`` (six tets per hex).

    Node and tet numbering match the reference grid mesher for any coordinate
    arrays: node ``i + nxp*(j + nyp*k)``; hexes visited k-outer, j, i-inner.

    Returns:
        (knt (N,3) float64, tets (E,4) int32).
    """
    xs = np.asarray(xs, dtype=np.float64)
    ys = np.asarray(ys, dtype=np.float64)
    zs = np.asarray(zs, dtype=np.float64)
    nxp, nyp, nzp = xs.size, ys.size, zs.size
    kk, jj, ii = np.meshgrid(np.arange(nzp), np.arange(nyp), np.arange(nxp), indexing="ij")
    knt = np.column_stack([xs[ii.ravel()], ys[jj.ravel()], zs[kk.ravel()]])

    hk, hj, hi = np.meshgrid(np.arange(nzp - 1), np.arange(nyp - 1), np.arange(nxp - 1), indexing="ij")
    hi, hj, hk = hi.ravel(), hj.ravel(), hk.ravel()

    def corner(letter: str) -> np.ndarray:
        di, dj, dk = _CORNER_OFFSETS[letter]
        return (hi + di) + nxp * ((hj + dj) + nyp * (hk + dk))

    cols = {c: corner(c) for c in "ABCDEFGH"}
    tets = np.empty((hi.size, 6, 4), dtype=np.int64)
    for t, word in enumerate(_SIX_TETS):
        for a, c in enumerate(word):
            tets[:, t, a] = cols[c]
    return knt, tets.reshape(-1, 4).astype(np.int32)


def grid_box(Lx: float, Ly: float, Lz: float, h: float) -> tuple[np.ndarray, np.ndarray]:
    """Regular-grid box centred at the origin, as ``src/mesh.py:1202-1272`` (axis-aligned frame).

    Returns:
        (knt (N,3), ijk (E,5) int32 with material id 1 in the last column).
    """
    nx = max(1, int(np.ceil(Lx / h)))
    ny = max(1, int(np.ceil(Ly / h)))
    nz = max(1, int(np.ceil(Lz / h)))
    xs = np.linspace(-Lx / 2, Lx / 2, nx + 1)
    ys = np.linspace(-Ly / 2, Ly / 2, ny + 1)
    zs = np.linspace(-Lz / 2, Lz / 2, nz + 1)
    knt, tets = tensor_grid_tets(xs, ys, zs)
    ijk = np.hstack([tets, np.ones((tets.shape[0], 1), dtype=np.int32)])
    return knt, ijk


def graded_axis(L: float, n: int, layers: int, grading: float) -> np.ndarray:
    """Axis coordinates: ``n`` uniform cells over [-L/2, L/2] plus ``layers`` graded air cells per side.

    Air cell k (k = 1..layers) has thickness ``h*grading**k`` with ``h = L/n``.
    """
    inner = np.linspace(-L / 2, L / 2, n + 1)
    h = L / n
    steps = h * grading ** np.arange(1, layers + 1)
    outer = L / 2 + np.cumsum(steps)
    return np.concatenate([-outer[::-1], inner, outer])


def graded_airbox_cube(L: float, n: int, layers: int, grading: float = 2.0) -> dict:
    """Magnet cube of edge ``L`` (``n`` cells per edge) inside a graded structured air shell.

    Returns a dict with ``knt`` (N,3), ``ijk`` (E,5) (material 1 = magnet, 2 = air),
    and ``boundary_mask`` (N,) float64 (0.0 on the outer surface, 1.0 elsewhere).
    With ``layers == 0`` the mesh is the plain magnet cube (single material).
    """
    xs = graded_axis(L, n, layers, grading)
    knt, tets = tensor_grid_tets(xs, xs, xs)
    cen = knt[tets].mean(axis=1)
    inside = np.all(np.abs(cen) < L / 2, axis=1)
    mat = np.where(inside, 1, 2).astype(np.int32)
    ijk = np.hstack([tets, mat[:, None]])
    bnd = box_boundary_mask(knt)
    return {"knt": knt, "ijk": ijk, "boundary_mask": bnd}


def outer_boundary_mask(tets: np.ndarray, num_nodes: int) -> np.ndarray:
    """Nodes on faces that belong to exactly one tet get 0.0, all others 1.0.

    Same semantics as ``find_outer_surface_tris`` + ``find_outer_boundary_mask``
    (``src/add_shell.py:105-136``).
    """
    t4 = np.asarray(tets)[:, :4].astype(np.int64)
    faces = t4[:, [[0, 1, 2], [0, 1, 3], [0, 2, 3], [1, 2, 3]]].reshape(-1, 3)
    keys = np.sort(faces, axis=1)
    if num_nodes < (1 << 21):
        # three node ids (< 2^21 each) packed into one int64 key
        packed = (keys[:, 0] << 42) | (keys[:, 1] << 21) | keys[:, 2]
        _, inv, cnt = np.unique(packed, return_inverse=True, return_counts=True)
        single = cnt[inv.ravel()] == 1
    else:
        # (k0, k1) packed, k2 as the minor sort key; a face is single if it equals neither neighbour
        hi = keys[:, 0] * num_nodes + keys[:, 1]
        order = np.lexsort((keys[:, 2], hi))
        hs, ls = hi[order], keys[order, 2]
        same = (hs[1:] == hs[:-1]) & (ls[1:] == ls[:-1])
        dup = np.zeros(len(order), dtype=bool)
        dup[1:] |= same
        dup[:-1] |= same
        single = np.empty(len(order), dtype=bool)
        single[order] = ~dup
    surf = faces[single]
    mask = np.ones(num_nodes, dtype=np.float64)
    mask[np.unique(surf)] = 0.0
    return mask


def box_boundary_mask(knt: np.ndarray) -> np.ndarray:
    """0.0 on nodes that lie on the faces of the axis-aligned bounding box, 1.0 elsewhere.

    For a conforming tet mesh that fills its bounding box (the structured generators here) the
    faces that belong to exactly one tet are exactly those on the box surface, so this equals
    ``outer_boundary_mask`` without the face sort.
    """
    knt = np.asarray(knt, dtype=np.float64)
    on = np.any((knt == knt.min(axis=0)) | (knt == knt.max(axis=0)), axis=1)
    return np.where(on, 0.0, 1.0)


def morton_encode(q: np.ndarray) -> np.ndarray:
    """30-bit Morton key of 10-bit integer coordinates (x bit at 3i+2, y at 3i+1, z at 3i)."""
    x = q[:, 0].astype(np.int64)
    y = q[:, 1].astype(np.int64)
    z = q[:, 2].astype(np.int64)
    code = np.zeros(q.shape[0], dtype=np.int64)
    for i in range(10):
        code |= ((x >> i) & 1) << (3 * i + 2)
        code |= ((y >> i) & 1) << (3 * i + 1)
        code |= ((z >> i) & 1) << (3 * i)
    return code


def morton_reorder(knt: np.ndarray, ijk: np.ndarray, extra_node_arrays: tuple = ()) -> dict:
    """Morton-sort nodes and elements (``src/reorder_mesh.py:86-101``, target 'gpu').

    Both sorts use NumPy's default argsort kind, as the reference does, so keys that tie (the
    centroids of a hex's tets can share a quantised key) are ordered the same way.

    Returns a dict with the permuted ``knt``, ``ijk``, ``node_perm`` (new -> old),
    ``inv_node_perm`` (old -> new), ``elem_perm`` and the permuted ``extra`` node arrays.
    """
    knt = np.asarray(knt, dtype=np.float64)
    ijk = np.asarray(ijk)
    lo = knt.min(axis=0)
    denom = knt.max(axis=0) - lo
    denom = np.where(denom == 0, 1.0, denom)
    qn = ((knt - lo) / denom * 1023).astype(np.int32)
    node_perm = np.argsort(morton_encode(qn))
    inv = np.empty(knt.shape[0], dtype=np.int32)
    inv[node_perm] = np.arange(knt.shape[0], dtype=np.int32)
    cen = knt[ijk[:, :4].astype(np.int64)].mean(axis=1)
    qc = ((cen - lo) / denom * 1023).astype(np.int32)
    elem_perm = np.argsort(morton_encode(qc))
    ijk_new = ijk[elem_perm].copy()
    ijk_new[:, :4] = inv[ijk_new[:, :4].astype(np.int64)]
    return {
        "knt": knt[node_perm],
        "ijk": ijk_new,
        "node_perm": node_perm,
        "inv_node_perm": inv,
        "elem_perm": elem_perm,
        "extra": tuple(np.asarray(a)[node_perm] for a in extra_node_arrays),
    }


def config_cube_mesh(L: float, n: int, layers: int, grading: float = 2.0, morton: bool = True) -> dict:
    """Cube magnet + graded airbox, optionally Morton-ordered: the mesh of the cube workloads.

    The 6,000-tet magnet with 3 air layers is ``config_cube_mesh(20, 10, 3)``;
    the 1.3M-tet magnet with 5 air layers (~2.06M tets) is ``config_cube_mesh(80, 60, 5)``.
    """
    mesh = graded_airbox_cube(L, n, layers, grading)
    if morton:
        r = morton_reorder(mesh["knt"], mesh["ijk"], (mesh["boundary_mask"],))
        mesh = {"knt": r["knt"], "ijk": r["ijk"], "boundary_mask": r["extra"][0]}
    return mesh


def polycrystal_cube(L: float = 200.0, n: int = 150, layers: int = 5, n_grains: int = 27, gb_width: float = 2.0,
                     seed: int = 4, grading: float = 2.0, morton: bool = True) -> dict:
    """Cube polycrystal with Voronoi grains and a grain-boundary phase, inside a graded airbox.

    The role of the reference's neper-based polycrystal meshers (``src/mesh.py:1868-2044``),
    on the structured grid: ``n_grains`` centres are drawn uniformly in the cube with
    ``numpy.random.default_rng(seed)``; each magnet tet belongs to the grain of the centre nearest
    to its centroid; tets whose centroid lies within ``gb_width`` of the bisector plane between its
    nearest and second-nearest centre form the grain-boundary phase.

    Material ids: 1..n_grains grains, n_grains + 1 grain boundary, n_grains + 2 air. Returns a
    dict with ``knt``, ``ijk``, ``boundary_mask`` and ``centres``.
    """
    rng = np.random.default_rng(seed)
    centres = rng.uniform(-L / 2, L / 2, size=(n_grains, 3))
    xs = graded_axis(L, n, layers, grading)
    knt, tets = tensor_grid_tets(xs, xs, xs)
    E = tets.shape[0]
    mat = np.full(E, n_grains + 2, dtype=np.int32)
    chunk = 2_000_000
    for s in range(0, E, chunk):
        t = tets[s:s + chunk].astype(np.int64)
        cen = knt[t].mean(axis=1)
        inside = np.all(np.abs(cen) < L / 2, axis=1)
        # nearest and second-nearest centre, one pass per centre
        d1 = np.full(len(t), np.inf)
        d2 = np.full(len(t), np.inf)
        i1 = np.zeros(len(t), dtype=np.int64)
        i2 = np.zeros(len(t), dtype=np.int64)
        for g in range(n_grains):
            d = ((cen - centres[g]) ** 2).sum(axis=1)
            closer = d < d1
            second = ~closer & (d < d2)
            d2 = np.where(closer, d1, np.where(second, d, d2))
            i2 = np.where(closer, i1, np.where(second, g, i2))
            d1 = np.where(closer, d, d1)
            i1 = np.where(closer, g, i1)
        # distance from the centroid to the bisector plane of the two nearest centres
        dist = (d2 - d1) / (2.0 * np.linalg.norm(centres[i2] - centres[i1], axis=1))
        m = np.where(dist < gb_width, n_grains + 1, i1 + 1).astype(np.int32)
        mat[s:s + chunk] = np.where(inside, m, n_grains + 2)
    ijk = np.hstack([tets, mat[:, None]])
    bnd = box_boundary_mask(knt)
    out = {"knt": knt, "ijk": ijk, "boundary_mask": bnd, "centres": centres}
    if morton:
        r = morton_reorder(knt, ijk, (bnd,))
        out.update(knt=r["knt"], ijk=r["ijk"], boundary_mask=r["extra"][0])
    return out


def random_easy_axes(n: int, seed: int) -> np.ndarray:
    """Easy axes uniform on the unit sphere (normalised Gaussian vectors)."""
    v = np.random.default_rng(seed).standard_normal((n, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def grain_set(n_grains: int = 64, seed: int = 3, h: float = 2.0, layers: int = 3, t_min: float = 2e4,
              t_max: float = 4e5, grading: float = 2.0, morton: bool = True) -> list[dict]:
    """A set of independent grains, each a cube or a sphere inside its own graded airbox.

    Drawn from ``numpy.random.default_rng(seed)`` in this order per grain: the total tet count
    T (log-uniform in [t_min, t_max]), the shape (cube or sphere, a fair coin), the easy axis
    (normalised Gaussian vector). The structured grid has ``n = round((T/6)^(1/3)) - 2*layers``
    magnet cells per edge (at least 2) of size ``h``; a sphere grain keeps the tets whose
    centroid lies within L/2 of the centre as magnet (material 1) and the rest as air
    (material 2). Each dict holds ``knt``, ``ijk``, ``boundary_mask``, ``shape``, ``easy_axis``.
    """
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_grains):
        T = float(np.exp(rng.uniform(np.log(t_min), np.log(t_max))))
        shape = "sphere" if rng.uniform() < 0.5 else "cube"
        k = rng.standard_normal(3)
        k /= np.linalg.norm(k)
        n = max(2, int(round((T / 6.0) ** (1.0 / 3.0))) - 2 * layers)
        L = n * h
        xs = graded_axis(L, n, layers, grading)
        knt, tets = tensor_grid_tets(xs, xs, xs)
        cen = knt[tets].mean(axis=1)
        if shape == "cube":
            inside = np.all(np.abs(cen) < L / 2, axis=1)
        else:
            inside = np.linalg.norm(cen, axis=1) < L / 2
        ijk = np.hstack([tets, np.where(inside, 1, 2).astype(np.int32)[:, None]])
        bnd = box_boundary_mask(knt)
        g = {"knt": knt, "ijk": ijk, "boundary_mask": bnd, "shape": shape, "easy_axis": k}
        if morton:
            r = morton_reorder(knt, ijk, (bnd,))
            g.update(knt=r["knt"], ijk=r["ijk"], boundary_mask=r["extra"][0])
        out.append(g)
    return out
