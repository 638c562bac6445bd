"""Multi-core CPU restatement of the pcohen_hs field step -- TEST INFRASTRUCTURE ONLY.

``cpu_minimizer.cpp`` restates ``oracle/minimize.py`` and ``oracle/solvers.py`` (the Armijo
line search, Steihaug-PCG, pcohen_hs and the AMG-preconditioned PCG) in C++ with OpenMP over the
assembled operators of ``oracle.fem.Problem``, in the role of the reference's MKL C++ path
(``src/cpp/cpp_mkl_minimizer.cpp:542-763``). It serves as the CPU baseline of ``bench.py`` (all
host cores) and is checked against the NumPy oracle in ``tests/test_oracle_cpu.py``.
"""

from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np
import scipy.sparse as sp

HERE = Path(__file__).resolve().parent
SRC = HERE / "cpu_minimizer.cpp"
LIB = HERE / "libocpu.so"

_lib = None


def build() -> Path:
    """Compile ``cpu_minimizer.cpp`` into ``oracle/libocpu.so`` (g++, OpenMP)."""
    tmp = LIB.with_suffix(".so.tmp")
    cmd = ["g++", "-O3", "-std=c++17", "-fopenmp", "-fPIC", "-shared", str(SRC), "-o", str(tmp)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"g++ failed ({res.returncode}):\n{res.stdout}\n{res.stderr}")
    tmp.replace(LIB)
    return LIB


def load() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise RuntimeError(f"{LIB} is missing (oracle.cpu.build())")
        lib = C.CDLL(str(LIB))
        i32p, f64p = C.POINTER(C.c_int32), C.POINTER(C.c_double)
        lib.ocpu_threads.restype = C.c_int
        lib.ocpu_create.restype = C.c_int64
        lib.ocpu_create.argtypes = [C.c_int] + [i32p, i32p, f64p] * 4 + [f64p, f64p, f64p, C.c_double]
        lib.ocpu_add_level.restype = C.c_int
        lib.ocpu_add_level.argtypes = [C.c_int64, C.c_int, i32p, i32p, f64p, f64p, C.c_int, i32p, i32p, f64p,
                                       i32p, i32p, f64p, f64p]
        lib.ocpu_minimize.restype = C.c_int
        lib.ocpu_minimize.argtypes = [C.c_int64, f64p, i32p, f64p, f64p, f64p, i32p, f64p]
        lib.ocpu_destroy.restype = C.c_int
        lib.ocpu_destroy.argtypes = [C.c_int64]
        _lib = lib
    return _lib


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def _csr(M):
    M = sp.csr_matrix(M)
    M.sort_indices()
    return (np.ascontiguousarray(M.indptr, dtype=np.int32), np.ascontiguousarray(M.indices, dtype=np.int32),
            np.ascontiguousarray(M.data, dtype=np.float64))


def threads() -> int:
    return int(load().ocpu_threads())


class CpuMinimizer:
    """pcohen_hs field steps on the CPU (all OpenMP threads) for ``problem`` (``oracle.fem.Problem``)
    with the V-cycle on ``levels`` (dicts with ``A``, ``P``, ``R``; the last level is solved with
    its dense inverse). The smoother is SPAI0 (``oracle.fem.spai0_diagonal``), as the oracle's."""

    def __init__(self, problem, levels):
        from .fem import spai0_diagonal

        lib = load()
        self._keep = []
        ops = []
        for M in (problem.A, problem.D, problem.G, problem.K_eff):
            arrs = _csr(M)
            self._keep.append(arrs)
            ops += [_p(arrs[0], C.c_int32), _p(arrs[1], C.c_int32), _p(arrs[2], C.c_double)]
        vecs = [np.ascontiguousarray(v, dtype=np.float64) for v in (problem.M_nodal, problem.inv_M_rel,
                                                                     problem.inv_M_prec)]
        self._keep.append(vecs)
        self.h = lib.ocpu_create(problem.N, *ops, *[_p(v, C.c_double) for v in vecs], float(problem.V_mag))
        self.N = problem.N
        L = len(levels)
        for i, lv in enumerate(levels):
            A = sp.csr_matrix(lv["A"])
            n = A.shape[0]
            Mv = np.ascontiguousarray(spai0_diagonal(A), dtype=np.float64)
            a = _csr(A) if i > 0 else None
            last = i == L - 1
            pp = _csr(lv["P"]) if not last else None
            rr = _csr(lv["R"]) if not last else None
            inv = np.ascontiguousarray(np.linalg.inv(A.toarray())) if last else None
            self._keep.append((Mv, a, pp, rr, inv))
            null_i, null_d = C.POINTER(C.c_int32)(), C.POINTER(C.c_double)()
            ret = lib.ocpu_add_level(
                self.h, n,
                _p(a[0], C.c_int32) if a else null_i, _p(a[1], C.c_int32) if a else null_i,
                _p(a[2], C.c_double) if a else null_d, _p(Mv, C.c_double),
                0 if last else lv["P"].shape[1],
                _p(pp[0], C.c_int32) if pp else null_i, _p(pp[1], C.c_int32) if pp else null_i,
                _p(pp[2], C.c_double) if pp else null_d,
                _p(rr[0], C.c_int32) if rr else null_i, _p(rr[1], C.c_int32) if rr else null_i,
                _p(rr[2], C.c_double) if rr else null_d,
                _p(inv, C.c_double) if inv is not None else null_d)
            if ret != 0:
                raise RuntimeError("ocpu_add_level failed")

    def minimize(self, m0, B_ext, U0=None, tau_f=1e-10, eps_a=1e-12, max_iter=2000, cg_tol=1e-10,
                 ls_eta1=0.1, ls_eta2=0.1, ls_C=2.0, ls_c=0.5, pc_iters=10, pc_auto=True, pc_force_eta=0.5,
                 pc_force_alpha=0.5, pc_stagnation_nu=0.01, pc_reg=0.0, cg_maxiter=2000):
        """(m, U, info) as ``oracle.minimize.Minimizer.minimize`` for 'pcohen_hs'."""
        m = np.ascontiguousarray(np.array(m0, dtype=np.float64).reshape(-1, 3))
        U = np.ascontiguousarray(np.zeros(self.N) if U0 is None else np.array(U0, dtype=np.float64))
        par = np.array([tau_f, eps_a, ls_eta1, ls_eta2, ls_C, ls_c, pc_force_eta, pc_force_alpha,
                        pc_stagnation_nu, pc_reg, cg_tol], dtype=np.float64)
        ipar = np.array([max_iter, 15, pc_iters, int(bool(pc_auto)), cg_maxiter], dtype=np.int32)
        B = np.ascontiguousarray(B_ext, dtype=np.float64).reshape(3)
        ii = np.zeros(5, dtype=np.int32)
        dd = np.zeros(2)
        ret = load().ocpu_minimize(self.h, _p(par, C.c_double), _p(ipar, C.c_int32), _p(m, C.c_double),
                                   _p(U, C.c_double), _p(B, C.c_double), _p(ii, C.c_int32), _p(dd, C.c_double))
        if ret != 0:
            raise RuntimeError("ocpu_minimize failed")
        info = dict(iters=int(ii[0]), evals=int(ii[1]), demag_iters=int(ii[2]), preco_iters=int(ii[3]),
                    converged=bool(ii[4]), E=float(dd[0]), gnorm=float(dd[1]))
        return m, U, info

    def close(self):
        if self.h:
            load().ocpu_destroy(self.h)
            self.h = 0
