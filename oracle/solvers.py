"""Oracle: PCG, AMG V-cycles and solve_U (test infrastructure only)."""

from __future__ import annotations

import numpy as np
import scipy.sparse as sp


def pcg(apply_A, apply_Minv, b, x0, tol, maxiter, boundary_mask=None):
    """Standard PCG exactly as ``make_pcg_solve.solve`` (``src/poisson_solve.py:422-478``).

    b and x0 are masked when a boundary mask is given; the loop runs while
    ``it < maxiter and r.r > tol^2 b.b``; alpha and beta carry the 1e-30 guards.
    Returns (x, it, r2).
    """
    eps = 1e-30
    if boundary_mask is not None:
        b = b * boundary_mask
        x = x0 * boundary_mask
    else:
        x = x0.copy()
    r = b - apply_A(x)
    if boundary_mask is not None:
        r = r * boundary_mask
    z = apply_Minv(r)
    p = z.copy()
    rz = r @ z
    r2 = r @ r
    tol2 = tol * tol * (b @ b)
    it = 0
    while it < maxiter and r2 > tol2:
        Ap = apply_A(p)
        if boundary_mask is not None:
            Ap = Ap * boundary_mask
        alpha = rz / (p @ Ap + eps)
        x = x + alpha * p
        r = r - alpha * Ap
        z = apply_Minv(r)
        rz_new = r @ z
        beta = rz_new / (rz + eps)
        p = z + beta * p
        r2 = r @ r
        rz = rz_new
        it += 1
    return x, it, r2


def vcycle(levels, apply_A0, r, kind="amgcl", omega=0.6667):
    """One V-cycle, following ``make_jax_amgcl_vcycle`` / ``make_jax_amg_vcycle`` (``src/amg_utils.py:491-623``).

    ``levels[l]`` holds ``A`` (CSR; unused on level 0 where ``apply_A0`` is used),
    ``M`` (SPAI0 diagonal for 'amgcl', the matrix diagonal for 'amg'), ``P`` and ``R``
    (all but the last level) and ``A_dense`` on the last level. One pre- and one
    post-smoothing sweep per level; x starts at zero; the coarsest level is a dense solve.
    """
    L = len(levels)

    def smooth(apply_A, b, x, M):
        res = b - apply_A(x)
        if kind == "amgcl":
            return x + M * res
        return x + omega * (res / (M + 1e-30))

    def rec(l, b):
        lv = levels[l]
        if l == L - 1:
            return np.linalg.solve(lv["A_dense"], b)
        apply_A = apply_A0 if l == 0 else (lambda v, A=lv["A"]: A @ v)
        x = smooth(apply_A, b, np.zeros_like(b), lv["M"])
        res = b - apply_A(x)
        e = rec(l + 1, lv["R"] @ res)
        x = x + lv["P"] @ e
        return smooth(apply_A, b, x, lv["M"])

    return rec(0, r)


def solve_U(problem, m, x0, tol, maxiter=2000, precond="direct", levels=None, kind="amgcl"):
    """``solve_U`` (``src/poisson_solve.py:748-771``), returning (U, it, rel_res).

    precond 'direct' solves the masked system with a sparse direct factorisation
    (the reference CPU path uses PARDISO, ``src/amg_utils.py:792-882``);
    'amg' runs the PCG above with the given V-cycle hierarchy.
    """
    b = problem.rhs(m)
    bnorm2 = b @ b
    if precond == "direct":
        import scipy.sparse.linalg as spla

        U = spla.spsolve(sp.csc_matrix(problem.A), b)
        r = b - problem.apply_A(U)
        return U, 0, float(np.sqrt((r @ r) / (bnorm2 + 1e-30)))
    # the driver passes no boundary mask through sparse_ops: b and x0 are not masked
    U, it, r2 = pcg(problem.apply_A, lambda r: vcycle(levels, problem.apply_A, r, kind), b, x0, tol, maxiter)
    return U, it, float(np.sqrt(r2 / (bnorm2 + 1e-30)))
