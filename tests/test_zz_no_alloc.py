"""The allocation-free energy entry point (mm_energy_grad_noalloc) against mm_energy_grad.

Device buffers are allocated by the test with the HIP runtime (the caller's memory); the library call allocates nothing and
keeps its block partials in the library's fixed device scratch array. Two calls in a row, the
second on a smaller mesh, each equal to the allocating entry point; a mesh with more blocks
than the scratch holds is rejected with an error."""

from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from conftest import ndfeb_materials, tilted_m0

pytestmark = pytest.mark.gpu


def _hip_runtime():
    """The HIP runtime the library itself links (the caller allocates its own buffers)."""
    from mumag_matrixfree2_amd import _hip

    _hip.load()
    rt = C.CDLL("libamdhip64.so", mode=C.RTLD_GLOBAL)
    rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    rt.hipFree.argtypes = [C.c_void_p]
    rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return rt


def _noalloc(ctx, m, U, B):
    """Upload m and U into caller-allocated device buffers, call the entry point on the
    context's stream (NULL), wait for the device, download E and g."""
    from mumag_matrixfree2_amd import _hip

    rt = _hip_runtime()
    host = {"m": np.ascontiguousarray(m, dtype=np.float64).reshape(-1),
            "U": np.ascontiguousarray(U, dtype=np.float64), "E": np.full(1, np.nan), "g": np.full(3 * ctx.N, np.nan)}
    dev = {}
    try:
        for k, a in host.items():
            p = C.c_void_p()
            assert rt.hipMalloc(C.byref(p), a.nbytes) == 0
            dev[k] = p
            assert rt.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        Bh = np.ascontiguousarray(B, dtype=np.float64)
        rc = _hip.load().mm_energy_grad_noalloc(ctx.handle, dev["m"], dev["U"], _hip.ptr(Bh), dev["E"], dev["g"],
                                                None)
        assert rt.hipDeviceSynchronize() == 0
        if rc != 0:
            raise RuntimeError(_hip.last_error())
        for k in ("E", "g"):
            assert rt.hipMemcpy(host[k].ctypes.data, dev[k], host[k].nbytes, 2) == 0  # device to host
    finally:
        for p in dev.values():
            rt.hipFree(p)
    return float(host["E"][0]), host["g"].reshape(-1, 3)


def _ctx(mesh, **kw):
    from mumag_matrixfree2_amd.context import MeshContext

    A, K1, Js, k = ndfeb_materials()
    return MeshContext(mesh["knt"], mesh["ijk"][:, :4], mesh["ijk"][:, 4], A, K1, Js, k,
                       boundary_mask=mesh["boundary_mask"], **kw)


def test_noalloc_energy_matches_the_allocating_entry(cube_c0, small_cube):
    big, small = _ctx(cube_c0), _ctx(small_cube)
    print(f"[noalloc] magnetic blocks {big.n_blocks_mag} then {small.n_blocks_mag}", flush=True)
    assert small.n_blocks_mag < big.n_blocks_mag
    rng = np.random.default_rng(60)
    for ctx, B in ((big, np.array([0.01, 0.0, -0.4])), (small, np.array([0.0, 0.2, 0.3]))):
        m = tilted_m0(ctx.N, seed=61, sigma=0.3)
        U = rng.standard_normal(ctx.N) * ctx.boundary_mask
        E0, g0 = ctx.energy_grad(m, U, B)
        E1, g1 = _noalloc(ctx, m, U, B)
        print(f"[noalloc N={ctx.N}] E={E1:.17e} ref={E0:.17e} |dg|/|g|="
              f"{np.linalg.norm(g1 - g0) / np.linalg.norm(g0):.2e}", flush=True)
        assert abs(E1 - E0) <= 1e-14 * max(1.0, abs(E0))
        assert np.linalg.norm(g1 - g0) <= 1e-14 * np.linalg.norm(g0)
    big.close()
    small.close()


def test_noalloc_energy_rejects_meshes_beyond_the_scratch():
    """80 nm cube (1.3M magnet tets) cut into 16-node blocks: more than 32768 magnetic blocks."""
    from mumag_matrixfree2_amd.mesh import config_cube_mesh

    ctx = _ctx(config_cube_mesh(80.0, 60, 5), block_nodes=16)
    print(f"[noalloc] 16-node blocks: {ctx.n_blocks_mag} magnetic blocks", flush=True)
    assert 2 * ctx.n_blocks_mag > 65536
    with pytest.raises(RuntimeError, match="scratch"):
        _noalloc(ctx, tilted_m0(ctx.N), np.zeros(ctx.N), np.zeros(3))
    ctx.close()
