"""Host mesh generators: boundary masks, Morton reordering, the Voronoi polycrystal."""

from __future__ import annotations

import numpy as np


def test_boundary_masks_agree():
    """Packed-key and lexsort branches of the face count, and the bounding-box shortcut."""
    from mumag_matrixfree2_amd.mesh import box_boundary_mask, graded_axis, outer_boundary_mask, tensor_grid_tets

    xs = graded_axis(20.0, 6, 2, 2.0)
    ys = graded_axis(10.0, 3, 1, 1.5)
    knt, tets = tensor_grid_tets(xs, ys, xs)
    N = knt.shape[0]
    packed = outer_boundary_mask(tets, N)
    lex = outer_boundary_mask(tets, 1 << 21)[:N]  # num_nodes >= 2^21 takes the lexsort branch
    box = box_boundary_mask(knt)
    assert np.array_equal(packed, lex)
    assert np.array_equal(packed, box)
    nb = (xs.size * ys.size * xs.size) - (xs.size - 2) * (ys.size - 2) * (xs.size - 2)
    assert int((packed == 0).sum()) == nb


def test_morton_reorder_is_a_relabelling():
    from mumag_matrixfree2_amd.mesh import config_cube_mesh

    a = config_cube_mesh(20.0, 5, 2, morton=False)
    b = config_cube_mesh(20.0, 5, 2, morton=True)
    assert a["knt"].shape == b["knt"].shape and a["ijk"].shape == b["ijk"].shape
    # the same set of elements as coordinate quadruples, in some order
    ka = np.sort(np.round(a["knt"][a["ijk"][:, :4]], 9).reshape(len(a["ijk"]), -1), axis=0)
    kb = np.sort(np.round(b["knt"][b["ijk"][:, :4]], 9).reshape(len(b["ijk"]), -1), axis=0)
    assert np.array_equal(ka, kb)
    assert np.array_equal(np.sort(a["boundary_mask"]), np.sort(b["boundary_mask"]))


def test_morton_key_bit_layout():
    """x bit i at 3i+2, y at 3i+1, z at 3i (src/reorder_mesh.py:19-30), checked on single bits
    and against a bit-string interleave."""
    from mumag_matrixfree2_amd.mesh import morton_encode

    q = np.array([[1, 0, 0], [0, 1, 0], [0, 0, 1], [1023, 1023, 1023], [512, 0, 0], [0, 0, 512]])
    assert morton_encode(q).tolist() == [4, 2, 1, 2**30 - 1, 1 << 29, 1 << 27]
    rng = np.random.default_rng(5)
    q = rng.integers(0, 1024, (200, 3))
    ref = [int("".join(f"{x:010b}"[i] + f"{y:010b}"[i] + f"{z:010b}"[i] for i in range(10)), 2) for x, y, z in q]
    assert morton_encode(q).tolist() == ref


def test_morton_order_follows_the_reference_sort_with_ties():
    """Nodes and elements are sorted by Morton key with the reference's argsort (default kind);
    keys are non-decreasing along the new order. A far-away unused node stretches the bounding
    box so that the tets of a hex share quantised centroids: the keys tie, and the tie order is
    the reference's."""
    from mumag_matrixfree2_amd.mesh import graded_airbox_cube, morton_encode, morton_reorder

    mesh = graded_airbox_cube(8.0, 4, 2, 2.0)
    knt = np.vstack([mesh["knt"], [[4000.0, 4000.0, 4000.0]]])
    ijk = mesh["ijk"]
    r = morton_reorder(knt, ijk)
    lo, hi = knt.min(axis=0), knt.max(axis=0)
    qn = ((knt - lo) / (hi - lo) * 1023).astype(np.int32)
    kn = morton_encode(qn)
    qc = ((knt[ijk[:, :4]].mean(axis=1) - lo) / (hi - lo) * 1023).astype(np.int32)
    kc = morton_encode(qc)
    assert np.unique(kc).size < kc.size  # ties exist
    assert np.array_equal(r["node_perm"], np.argsort(kn))
    assert np.array_equal(r["elem_perm"], np.argsort(kc))
    assert np.all(np.diff(kn[r["node_perm"]]) >= 0) and np.all(np.diff(kc[r["elem_perm"]]) >= 0)
    assert np.array_equal(r["knt"], knt[r["node_perm"]])
    assert np.array_equal(r["node_perm"][r["ijk"][:, :4]], ijk[r["elem_perm"], :4])


def test_polycrystal_phases_by_brute_force():
    """Grain = nearest centre, grain boundary = within gb_width of the bisector plane of the
    two nearest centres, air = outside the cube; checked tet by tet with a full sort."""
    from mumag_matrixfree2_amd.mesh import polycrystal_cube

    L, ng, w = 30.0, 6, 1.5
    m = polycrystal_cube(L, 8, 2, n_grains=ng, gb_width=w, seed=9, morton=False)
    knt, ijk, C = m["knt"], m["ijk"], m["centres"]
    cen = knt[ijk[:, :4]].mean(axis=1)
    expect = np.empty(len(ijk), dtype=np.int64)
    for e, c in enumerate(cen):
        if np.any(np.abs(c) >= L / 2):
            expect[e] = ng + 2
            continue
        d = np.linalg.norm(C - c, axis=1)
        i1, i2 = np.argsort(d)[:2]
        n = (C[i2] - C[i1]) / np.linalg.norm(C[i2] - C[i1])
        dist = np.dot(c - 0.5 * (C[i1] + C[i2]), n)  # signed distance to the bisector, <= 0 on i1's side
        expect[e] = ng + 1 if -dist < w else i1 + 1
    assert np.array_equal(ijk[:, 4], expect)
    assert set(np.unique(ijk[:, 4])) <= set(range(1, ng + 3))
    assert np.all(m["boundary_mask"][ijk[ijk[:, 4] <= ng, :4]] == 1.0)  # magnet never touches the Dirichlet set


def test_grain_set_draws():
    from mumag_matrixfree2_amd.mesh import grain_set

    a = grain_set(6, seed=3, t_min=2e3, t_max=2e4, morton=False)
    b = grain_set(6, seed=3, t_min=2e3, t_max=2e4, morton=False)
    for g, h in zip(a, b):
        assert np.array_equal(g["ijk"], h["ijk"]) and g["shape"] == h["shape"]
        assert abs(np.linalg.norm(g["easy_axis"]) - 1.0) < 1e-14
        mag = g["ijk"][:, 4] == 1
        cen = g["knt"][g["ijk"][:, :4]].mean(axis=1)
        if g["shape"] == "sphere":
            r = np.linalg.norm(cen, axis=1)
            assert r[mag].max() < r[~mag].min()
        assert 0 < mag.sum() < len(mag)
    assert {g["shape"] for g in a} <= {"cube", "sphere"}


def test_grain_set_union_is_block_diagonal():
    """The union of a grain set: node ranges per grain, connectivity shifted into its range,
    material ids shifted to the grain's own table rows, the assembled operators block-diagonal
    with each block equal to the grain's own, and the aggregates of the host AMG set-up numbered
    in node order so every grain's coarse rows are contiguous on every level."""
    import scipy.sparse as sp

    from conftest import ndfeb_materials
    from mumag_matrixfree2_amd.amg import build_hierarchy
    from mumag_matrixfree2_amd.grainset import union_mesh
    from mumag_matrixfree2_amd.mesh import grain_set
    from oracle.fem import Problem

    gs = grain_set(3, seed=7, t_min=2e3, t_max=6e3)
    tables = [ndfeb_materials(axis=g["easy_axis"]) for g in gs]
    u = union_mesh(gs, tables)
    off = u["segments"]
    assert off[0] == 0 and off[-1] == u["knt"].shape[0] and len(off) == len(gs) + 1
    ijk = np.column_stack([u["conn"], u["mat_id"]])
    pu = Problem(u["knt"], ijk, u["A"], u["K1"], u["Js"], u["k"], boundary_mask=u["boundary_mask"])
    blocks = []
    for i, (g, t) in enumerate(zip(gs, tables)):
        conn = u["conn"][u["conn"][:, 0] >= off[i]]
        conn = conn[conn[:, 0] < off[i + 1]]
        assert conn.min() >= off[i] and conn.max() < off[i + 1]
        p = Problem(g["knt"], g["ijk"], *t, boundary_mask=g["boundary_mask"])
        blocks.append(p.A)
        s = slice(off[i], off[i + 1])
        assert abs(pu.A[s, s] - p.A).max() <= 1e-13 * abs(p.A).max()
    assert abs(pu.A - sp.block_diag(blocks)).max() <= 1e-13 * abs(pu.A).max()
    seg = off.copy()
    for lv in build_hierarchy(pu.A.tocsr(), max_coarse=50)[:-1]:
        P = lv["P"].tocsr()
        nxt, expect = [], 0
        for i in range(len(gs)):
            cols = P.indices[P.indptr[seg[i]]:P.indptr[seg[i + 1]]]
            assert cols.size > 0 and cols.min() == expect
            nxt.append(cols.min())
            expect = cols.max() + 1
        assert expect == P.shape[1]
        seg = np.array(nxt + [P.shape[1]])
