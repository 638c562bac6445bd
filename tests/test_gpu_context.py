"""Context creation rejects malformed meshes with a clear error (C-ABI status, ``mm_last_error``)
before any kernel runs, and a rejected creation leaves no context behind."""

from __future__ import annotations

import numpy as np
import pytest

from conftest import ndfeb_materials

pytestmark = pytest.mark.gpu


def _create(knt, ijk, bnd):
    """mm_ctx_create through the C-ABI itself (the Python wrapper checks some of this earlier)."""
    import ctypes as C

    from mumag_matrixfree2_amd._hip import call, f64, i32, ptr

    A, K1, Js, k = ndfeb_materials()
    conn, mat = i32(np.ascontiguousarray(ijk[:, :4])), i32(np.ascontiguousarray(ijk[:, 4]))
    knt = f64(np.ascontiguousarray(knt))
    h = C.c_int64(0)
    call("mm_ctx_create", len(knt), len(conn), ptr(knt), ptr(conn), ptr(mat), len(A), ptr(f64(A)), ptr(f64(K1)),
         ptr(f64(Js)), ptr(f64(np.asarray(k).reshape(-1))), ptr(f64(bnd)), 1e-12, 0, C.byref(h))
    return h.value


def test_valid_small_cube_builds(small_cube):
    from mumag_matrixfree2_amd._hip import call

    h = _create(small_cube["knt"], small_cube["ijk"], small_cube["boundary_mask"])
    assert h > 0
    call("mm_ctx_destroy", h)


@pytest.mark.parametrize("case", ["node_out_of_range", "negative_node", "material_zero", "material_above_G",
                                  "flat_element", "nan_coordinate"])
def test_malformed_mesh_is_rejected(small_cube, case):
    knt = np.array(small_cube["knt"], dtype=np.float64, copy=True)
    ijk = np.array(small_cube["ijk"], dtype=np.int32, copy=True)
    msg = {"node_out_of_range": "connectivity index out of range", "negative_node": "connectivity index out of range",
           "material_zero": "material id out of range", "material_above_G": "material id out of range",
           "flat_element": "degenerate element", "nan_coordinate": "degenerate element"}[case]
    if case == "node_out_of_range":
        ijk[7, 2] = len(knt)
    elif case == "negative_node":
        ijk[3, 0] = -1
    elif case == "material_zero":
        ijk[5, 4] = 0
    elif case == "material_above_G":
        ijk[5, 4] = 3  # two materials (magnet, air)
    elif case == "flat_element":
        ijk[11, 3] = ijk[11, 2]  # two equal corners: zero volume
    else:
        knt[ijk[13, 1]] = np.nan
    with pytest.raises(RuntimeError, match=msg):
        _create(knt, ijk, small_cube["boundary_mask"])


def test_geometry_copy_holds_every_element_on_recycled_memory():
    """Contexts created one after another, each on device memory the previous one freed: the
    element volumes each context computes on the device and copies back (the AMG set-up assembles
    its matrix from them) equal the host's own determinants, element for element."""
    from mumag_matrixfree2_amd.context import MeshContext
    from mumag_matrixfree2_amd.mesh import config_cube_mesh

    mesh = config_cube_mesh(40.0, 20, 3)
    A, K1, Js, k = ndfeb_materials()
    for rep in range(4):
        ctx = MeshContext(mesh["knt"], mesh["ijk"][:, :4], mesh["ijk"][:, 4], A, K1, Js, k,
                          boundary_mask=mesh["boundary_mask"])
        _, Ve_ctx, _ = ctx.geometry()
        _, Ve_host, _ = ctx.element_table()
        ctx.close()
        assert np.all(Ve_ctx > 0), rep
        assert np.max(np.abs(np.sort(Ve_ctx) - np.sort(Ve_host))) <= 1e-12 * Ve_host.max(), rep
