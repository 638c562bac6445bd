"""ctypes binding of ``libmumag_hip.so`` (the C-ABI declared in ``include/mumag_hip.h``).

There is no CPU fallback: if the library is missing or the GPU is absent, every entry
point raises.
"""

from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libmumag_hip.so"

i32p = C.POINTER(C.c_int32)
i64p = C.POINTER(C.c_int64)
f64p = C.POINTER(C.c_double)


class MMParams(C.Structure):
    _fields_ = [
        ("method", C.c_int),
        ("max_iter", C.c_int),
        ("tau_f", C.c_double),
        ("eps_a", C.c_double),
        ("tau0", C.c_double),
        ("ls_eta1", C.c_double),
        ("ls_eta2", C.c_double),
        ("ls_C", C.c_double),
        ("ls_c", C.c_double),
        ("ls_max_evals", C.c_int),
        ("pc_iters", C.c_int),
        ("pc_auto", C.c_int),
        ("pc_force_eta", C.c_double),
        ("pc_force_alpha", C.c_double),
        ("pc_stagnation_nu", C.c_double),
        ("pc_reg", C.c_double),
        ("L", C.c_int),
        ("cg_tol", C.c_double),
        ("phi_tol", C.c_double),
        ("cg_maxiter", C.c_int),
        ("memory", C.c_int),
    ]


class MMInfo(C.Structure):
    _fields_ = [
        ("iters", C.c_int),
        ("evals", C.c_int),
        ("demag_iters", C.c_int),
        ("preco_iters", C.c_int),
        ("converged", C.c_int),
        ("E", C.c_double),
        ("gnorm", C.c_double),
        ("time", C.c_double),
    ]


# name -> argtypes (restype is always c_int)
SIGNATURES = {
    "mm_last_error": [C.c_char_p, C.c_int],
    "mm_device_info": [C.c_char_p, C.c_int],
    "mm_ctx_create": [C.c_int, C.c_int, f64p, i32p, i32p, C.c_int, f64p, f64p, f64p, f64p, f64p, C.c_double,
                      C.c_int, i64p],
    "mm_ctx_destroy": [C.c_int64],
    "mm_ctx_create_segmented": [C.c_int, C.c_int, f64p, i32p, i32p, C.c_int, f64p, f64p, f64p, f64p, f64p, C.c_double,
                                C.c_int, C.c_int, i32p, i64p],
    "mm_ctx_info": [C.c_int64, f64p],
    "mm_ctx_get_geometry": [C.c_int64, i32p, f64p, f64p],
    "mm_ctx_get_nodal": [C.c_int64, f64p, f64p, f64p],
    "mm_set_geometry_mode": [C.c_int64, C.c_int],
    "mm_set_precision": [C.c_int64, C.c_int],
    "mm_set_apply_shape": [C.c_int64, C.c_int, C.c_int],
    "mm_set_cluster_block_nodes": [C.c_int64, C.c_int],
    "mm_set_cluster_staging": [C.c_int64, C.c_int],
    "mm_set_apply_scatter": [C.c_int64, C.c_int],
    "mm_set_refinement": [C.c_int64, C.c_double],
    "mm_poisson_apply": [C.c_int64, f64p, f64p],
    "mm_poisson_rhs": [C.c_int64, f64p, f64p],
    "mm_poisson_diag": [C.c_int64, f64p],
    "mm_amg_set_level": [C.c_int64, C.c_int, C.c_int, C.c_int, i32p, i32p, f64p, f64p, i32p, i32p, f64p, i32p, i32p,
                         f64p, f64p],
    "mm_amg_commit": [C.c_int64, C.c_int, C.c_int],
    "mm_amg_set_lanes": [C.c_int64, C.c_int, C.c_int, C.c_int],
    "mm_amg_get_lanes": [C.c_int64, C.c_int, i32p],
    "mm_precond_jacobi": [C.c_int64],
    "mm_precond_apply": [C.c_int64, f64p, f64p],
    "mm_ctx_get_exchange_diag": [C.c_int64, f64p],
    "mm_pcg_solve": [C.c_int64, f64p, f64p, C.c_double, C.c_int, i32p, f64p],
    "mm_solve_U": [C.c_int64, f64p, f64p, C.c_double, C.c_int, i32p, f64p],
    "mm_solve_U_segmented": [C.c_int64, f64p, f64p, f64p, C.c_int, i32p, f64p],
    "mm_energy_grad": [C.c_int64, f64p, f64p, f64p, f64p, f64p],
    "mm_local_grad": [C.c_int64, f64p, f64p],
    "mm_energy_grad_noalloc": [C.c_int64, C.c_void_p, C.c_void_p, f64p, C.c_void_p, C.c_void_p, C.c_void_p],
    "mm_j_parallel": [C.c_int64, f64p, f64p, f64p],
    "mm_state_set": [C.c_int64, f64p, f64p],
    "mm_state_get": [C.c_int64, f64p, f64p],
    "mm_params_check": [C.POINTER(MMParams)],
    "mm_minimize_resident": [C.c_int64, C.POINTER(MMParams), f64p, C.POINTER(MMInfo)],
    "mm_minimize": [C.c_int64, C.POINTER(MMParams), f64p, f64p, f64p, C.POINTER(MMInfo)],
    "mm_state_j_parallel": [C.c_int64, f64p, f64p],
    "mm_minimize_segmented": [C.c_int64, C.POINTER(MMParams), f64p, C.POINTER(MMInfo)],
    "mm_state_j_parallel_segmented": [C.c_int64, f64p, f64p],
    "mm_seg_volumes": [C.c_int64, f64p],
    "mm_set_solve_guess": [C.c_int64, C.c_int],
    "mm_minimize_batched": [C.c_int, C.POINTER(C.c_int64), C.POINTER(MMParams), f64p, C.c_int, C.POINTER(MMInfo)],
    "mm_solve_U_batched": [C.c_int, C.POINTER(C.c_int64), C.POINTER(f64p), C.POINTER(f64p), f64p, C.c_int, C.c_int,
                           i32p, f64p],
    "mm_sa_aggregate": [C.c_int, i32p, i32p, i32p, i32p],
    "mm_edge_clusters": [C.c_int, C.c_int, f64p, i32p, i32p, C.POINTER(C.c_uint32), i32p],
    "mm_apply_timing": [C.c_int64, C.c_int, f64p, i64p],
    "mm_counters": [C.c_int64, i64p],
    "mm_bench_apply": [C.c_int64, C.c_int, C.c_int, f64p],
    "mm_bench_copy": [C.c_int64, C.c_int, f64p],
    "mm_bench_apply_multi": [C.c_int64, C.c_int, C.c_int, f64p],
    "mm_poisson_apply_multi": [C.c_int64, C.c_int, f64p, f64p],
    "mm_bench_precond": [C.c_int64, C.c_int, C.c_int, f64p],
}

_lib = None


def load() -> C.CDLL:
    """Load the library once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"HIP library not built: {LIB_PATH} is missing (run __graft_entry__.build() or "
            "python -m mumag_matrixfree2_amd._build)"
        )
    lib = C.CDLL(str(LIB_PATH))
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = C.c_int
    _lib = lib
    return lib


def last_error() -> str:
    buf = C.create_string_buffer(4096)
    load().mm_last_error(buf, 4096)
    return buf.value.decode(errors="replace")


def check(ret: int, what: str = "") -> None:
    if ret != 0:
        raise RuntimeError(f"{what or 'mumag_hip'} failed ({ret}): {last_error()}")


def call(name: str, *args) -> None:
    check(getattr(load(), name)(*args), name)


def f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def i32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)


def ptr(a: np.ndarray | None):
    if a is None:
        return None
    if a.dtype == np.float64:
        return a.ctypes.data_as(f64p)
    if a.dtype == np.int32:
        return a.ctypes.data_as(i32p)
    if a.dtype == np.int64:
        return a.ctypes.data_as(i64p)
    raise TypeError(a.dtype)


def device_info() -> str:
    buf = C.create_string_buffer(512)
    check(load().mm_device_info(buf, 512), "mm_device_info")
    return buf.value.decode()
