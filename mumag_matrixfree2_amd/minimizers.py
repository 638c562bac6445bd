"""Field-step minimiser with the reference's factory signature (``src/minimizers.py:2927-4098``),
running device-resident on the HIP library.

Methods: 'pcohen_hs' (``src/minimizers.py:950-1073``, the reference default and its only C++
method) and 'lbfgs' (``src/minimizers.py:1216-1381``, memory 10 as on the reference's
``loop.py`` path). The whole field step (initial solve, iterations, line searches, inner
Steihaug-PCG) runs in one library call; only scalars cross PCIe during it.
"""

from __future__ import annotations

import numpy as np

from .context import make_params
from .fem_utils import TetGeom, cached_inputs, context_for, tables_match

SUPPORTED = ("pcohen_hs", "lbfgs")


def make_minimizer(geom: TetGeom, A_lookup, K1_lookup, Js_lookup, k_easy_lookup, V_mag, node_volumes, M_nodal,
                   solve_U, cg_tol, method: str = "pcohen_hs", *, context=None, boundary_mask=None,
                   **kwargs):
    """minimize(m0, B_ext, **params) -> (m, U, info).

    ``solve_U`` must be the one returned by ``make_solve_U`` (its device context and
    preconditioner are reused). ``params`` takes the reference's keys (tau_f, eps_a, max_iter,
    tau0, ls_*, pc_*, L, U0, phi_tol, cg_maxiter).
    """
    if method not in SUPPORTED:
        raise NotImplementedError(f"method {method!r} is not implemented by the HIP backend ({SUPPORTED})")
    ctx = context
    if ctx is None:
        sctx = getattr(solve_U, "_ctx", None)
        if sctx is not None and tables_match(sctx, A_lookup, K1_lookup, Js_lookup, k_easy_lookup):
            ctx = sctx
        else:
            # the solver's context lacks the exchange/anisotropy tables: a context with the full
            # tables and the solver's Dirichlet mask, regularisation and preconditioner
            mask, reg = (boundary_mask, 1e-12) if sctx is None else (cached_inputs(sctx) or (boundary_mask, sctx.reg))
            ctx = context_for(geom, A_lookup, K1_lookup, Js_lookup, k_easy_lookup, mask, reg)
            if sctx is not None and sctx.precond_type is not None and ctx.precond_type != sctx.precond_type:
                ctx.setup_precond(sctx.precond_type)
            if sctx is not None:
                ctx.set_precision(sctx.precision)
    if ctx.precond_type is None:
        ctx.setup_precond("amgcl")
    memory = int(kwargs.get("memory", 10))  # src/minimizers.py:3099
    cg_maxiter = int(kwargs.get("cg_maxiter", 2000))

    def minimize(m0, B_ext, **params):
        U0 = params.pop("U0", None)
        params.pop("sparse_ops", None)
        # the sweep passes memory=params.memory with every call (src/hysteresis_loop.py:504-531), but
        # the L-BFGS history length is the factory's (src/minimizers.py:3099): the call's value is
        # dropped, as the reference drops it
        params.pop("memory", None)
        params.pop("cg_tol", None)
        # both methods call the line search with 15 evaluations per phase whatever the caller's
        # ls_max_evals (src/minimizers.py:1035,1335)
        params.pop("ls_max_evals", None)
        p = make_params(method, cg_tol=cg_tol, memory=memory, cg_maxiter=params.pop("cg_maxiter", cg_maxiter),
                        **params)
        m0 = np.asarray(m0, dtype=np.float64)
        ctx.state_set(m0, None if U0 is None else np.asarray(U0, dtype=np.float64))
        info = ctx.minimize_resident(p, np.asarray(B_ext, dtype=np.float64))
        m, U = ctx.state_get()
        return m, U, info

    minimize._ctx = ctx
    return minimize
