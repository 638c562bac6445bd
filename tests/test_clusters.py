"""Edge-star clusters (the work unit of the Poisson apply), checked on the host.

A cluster is up to 6 tets sharing one edge on 8 node slots [E0, E1, R0..R5]; tet j is
(E0, E1, Rj, R(j+1) mod 6). The checks: the clusters partition the mesh's tets exactly
(as node sets), every listed tet shares the cluster's edge, and on the six-tet hex split
every cluster is a whole hex (6 tets on 8 distinct nodes).
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import pytest


def clusters(knt, tets):
    from mumag_matrixfree2_amd._hip import call, f64, i32, ptr

    E = tets.shape[0]
    nodes8 = np.full((E, 8), -2, dtype=np.int32)
    masks = np.zeros(E, dtype=np.uint32)
    n = C.c_int32(0)
    call("mm_edge_clusters", knt.shape[0], E, ptr(f64(knt)), ptr(i32(tets)), ptr(nodes8),
         masks.ctypes.data_as(C.POINTER(C.c_uint32)), C.byref(n))
    return nodes8[: n.value], masks[: n.value]


def expand(nodes8, masks):
    out = []
    for nd, m in zip(nodes8, masks):
        for j in range(6):
            if m >> j & 1:
                a, b = nd[2 + j], nd[2 + (j + 1) % 6]
                assert a >= 0 and b >= 0, "tet uses an unused slot"
                out.append(tuple(sorted((nd[0], nd[1], a, b))))
    return out


def check_partition(knt, tets):
    nodes8, masks = clusters(knt, tets)
    got = expand(nodes8, masks)
    want = [tuple(sorted(t)) for t in tets.tolist()]
    assert sorted(got) == sorted(want)
    assert len(set(want)) == len(want)
    return nodes8, masks


def test_hex_split_gives_whole_hexes():
    from mumag_matrixfree2_amd.mesh import graded_axis, tensor_grid_tets

    xs = graded_axis(6.0, 3, 2, 2.0)
    knt, tets = tensor_grid_tets(xs, xs, xs)
    nodes8, masks = check_partition(knt, tets)
    assert len(masks) * 6 == tets.shape[0]
    assert np.all(masks == 63)
    assert all(len(set(nd.tolist())) == 8 for nd in nodes8)


def test_morton_reordered_cube_mesh():
    from mumag_matrixfree2_amd.mesh import config_cube_mesh

    mesh = config_cube_mesh(8.0, 4, 2)
    nodes8, masks = check_partition(mesh["knt"], mesh["ijk"][:, :4])
    assert np.all(masks == 63)


@pytest.mark.parametrize("seed", [0, 1])
def test_unstructured_subsets_are_partitioned(seed):
    """Random subsets of a jittered hex split: broken stars become chains of face-sharing
    tets, split and packed into clusters; every tet still appears exactly once."""
    from mumag_matrixfree2_amd.mesh import tensor_grid_tets

    rng = np.random.default_rng(seed)
    xs = np.linspace(0.0, 5.0, 6)
    knt, tets = tensor_grid_tets(xs, xs, xs)
    knt = knt + 0.05 * rng.standard_normal(knt.shape)
    keep = rng.random(tets.shape[0]) < 0.7
    sub = tets[keep]
    nodes8, masks = check_partition(knt, sub)
    assert len(masks) < sub.shape[0]
    for nd, m in zip(nodes8, masks):
        used = {q for j in range(6) if m >> j & 1 for q in (2 + j, 2 + (j + 1) % 6)}
        assert all(nd[q] == -1 for q in range(2, 8) if q not in used and nd[q] < 0)
