"""Demagnetising-potential operators and solver with the reference's factory signatures
(``src/poisson_solve.py:124-776``), running on the HIP library.

``sparse_ops`` is accepted for signature compatibility; the assembled operators it carries
in the reference are replaced by the matrix-free GPU kernels of the context.
Semantics are those of the reference's assembled (production) mode: A has identity rows
and zero columns at Dirichlet nodes and ``reg`` on the diagonal; b = D m; PCG stops at
``it >= maxiter`` or ``r.r <= tol^2 b.b``.
"""

from __future__ import annotations

from collections.abc import Callable

import numpy as np

from .fem_utils import TetGeom, context_for, material_tables_of


def _poisson_context(geom, Js_lookup, boundary_mask, reg, context):
    """The Poisson operators depend on the geometry, Js, the Dirichlet mask and reg. When the
    energy kernels of the same geometry were built first (the order of
    src/hysteresis_loop.py:309-338), their full material tables are used, so the minimiser built
    on this solver shares its context."""
    if context is not None:
        return context
    full = material_tables_of(geom, Js_lookup)
    if full is not None:
        return context_for(geom, *full, boundary_mask, reg)
    Js = np.asarray(Js_lookup, dtype=np.float64)
    z = np.zeros_like(Js)
    return context_for(geom, z, z, Js, np.tile([0.0, 0.0, 1.0], (Js.size, 1)), boundary_mask, reg)


def make_poisson_ops(geom: TetGeom, Js_lookup, *, reg: float = 1e-12, boundary_mask=None, context=None, **_ignored):
    """(apply_A(sparse_ops, U), rhs_from_m(sparse_ops, m), assemble_diag(sparse_ops, N)) — ``src/poisson_solve.py:124-305``."""
    ctx = _poisson_context(geom, Js_lookup, boundary_mask, reg, context)

    def apply_A(sparse_ops, U):
        return ctx.apply_A(U)

    def rhs_from_m(sparse_ops, m):
        return ctx.rhs(m)

    def assemble_diag(sparse_ops, N):
        return ctx.diag()

    apply_A._ctx = ctx
    return apply_A, rhs_from_m, assemble_diag


def make_pcg_solve(apply_A: Callable, Mdiag=None, *, precond_type: str = "amgcl", maxiter: int = 500,
                   tol: float = 1e-8, **_ignored):
    """solve(sparse_ops, b, x0, tol=None, hierarchy=None) -> (x, it, r2) — ``src/poisson_solve.py:343-480``.

    ``apply_A`` must come from ``make_poisson_ops`` (it carries the device context). The PCG
    runs on the GPU with the context's preconditioner ('jacobi', 'amgcl' or 'amg').
    """
    ctx = getattr(apply_A, "_ctx", None)
    if ctx is None:
        raise TypeError("make_pcg_solve needs an apply_A from mumag_matrixfree2_amd.make_poisson_ops")
    if ctx.precond_type != precond_type:
        ctx.setup_precond(precond_type)
    default_tol = float(tol)

    def solve(sparse_ops, b, x0, tol=None, hierarchy=None):
        t = default_tol if tol is None else float(tol)
        return ctx.pcg(b, x0, t, maxiter)

    return solve


def make_solve_U(geom: TetGeom, Js_lookup, *, precond_type: str = "amgcl", cg_maxiter: int = 2000,
                 cg_tol: float = 1e-8, poisson_reg: float = 1e-12, enforce_zero_mean=None, boundary_mask=None,
                 context=None, **_ignored):
    """solve_U(m, x0, tol=None, return_info=False, sparse_ops=None) — ``src/poisson_solve.py:483-776``.

    Returns U, or (U, iterations, relative residual) with ``return_info``.
    """
    ctx = _poisson_context(geom, Js_lookup, boundary_mask, poisson_reg, context)
    if ctx.precond_type != precond_type:
        ctx.setup_precond(precond_type)
    if enforce_zero_mean is None:
        enforce_zero_mean = boundary_mask is None

    def solve_U(m, x0, tol=None, return_info=False, sparse_ops=None):
        t = cg_tol if tol is None else float(tol)
        x0 = np.asarray(x0, dtype=np.float64)
        if enforce_zero_mean:
            x0 = x0 - x0.mean()
        U, it, rr = ctx.solve_U(m, x0, t, cg_maxiter)
        if enforce_zero_mean:
            U = U - U.mean()
        return (U, it, rr) if return_info else U

    solve_U._ctx = ctx
    return solve_U
