"""Energy and gradient kernels with the reference's factory signatures
(``src/energy_kernels.py:82-566``), running on the HIP library.

E = 1/2 sum m.(g + g_z) / V_mag with g = exchange + anisotropy + demag + Zeeman, all per
magnetic element on the GPU (matrix-free; the reference's assembled K_eff/G products are
the same sums).
"""

from __future__ import annotations

import numpy as np

from .fem_utils import TetGeom, context_for, context_with_tables


def make_energy_kernels(geom: TetGeom, A_lookup, K1_lookup, Js_lookup, k_easy_lookup, V_mag, M_nodal, *,
                        B_bias=None, boundary_mask=None, reg: float = 1e-12, context=None, **_ignored):
    """(energy_and_grad, energy_only, grad_only, local_grad_only) — ``src/energy_kernels.py:82-489``.

    ``V_mag`` and ``M_nodal`` are recomputed by the context from the mesh; a mismatch with the
    caller's values raises, since it would mean a different unit normalisation.
    """
    if B_bias is not None and np.any(np.asarray(B_bias) != 0):
        raise NotImplementedError("per-node bias fields are not supported by the HIP backend")
    ctx = (context or context_with_tables(geom, A_lookup, K1_lookup, Js_lookup, k_easy_lookup)
           or context_for(geom, A_lookup, K1_lookup, Js_lookup, k_easy_lookup, boundary_mask, reg))
    if abs(ctx.V_mag - float(V_mag)) > 1e-9 * max(1.0, abs(float(V_mag))):
        raise ValueError(f"V_mag {V_mag} differs from the mesh's magnetic volume {ctx.V_mag}")
    if M_nodal is not None:
        M_dev = ctx.nodal()[0]
        if np.max(np.abs(M_dev - np.asarray(M_nodal))) > 1e-9 * max(1.0, float(np.max(np.abs(M_dev)))):
            raise ValueError("M_nodal differs from the mesh's lumped magnetic moments")

    def energy_and_grad(m, U, B_ext, sparse_ops=None):
        return ctx.energy_grad(m, U, B_ext)

    def energy_only(m, U, B_ext, sparse_ops=None):
        return ctx.energy_grad(m, U, B_ext)[0]

    def grad_only(m, U, B_ext, sparse_ops=None):
        return ctx.energy_grad(m, U, B_ext)[1]

    def local_grad_only(v, sparse_ops=None):
        return ctx.local_grad(v)

    energy_and_grad._ctx = ctx
    return energy_and_grad, energy_only, grad_only, local_grad_only


def compute_exchange_diagonal(geom: TetGeom, A_lookup, V_mag, *, context=None, **_ignored):
    """diag(K_ex)/V_mag (``src/energy_kernels.py:492-566``).

    Computed at context creation; a standalone call builds a context with Js = 1 on every
    material with A != 0 so that V_mag is the magnetic volume of the exchange materials.
    """
    if context is None:
        A = np.asarray(A_lookup, dtype=np.float64)
        from .fem_utils import _CACHE

        for e in _CACHE:  # any context of this geometry with the same exchange constants
            if e["geom"] is geom and e["A"].shape == A.shape and np.array_equal(e["A"], A):
                context = e["ctx"]
                break
        else:
            context = context_for(geom, A, np.zeros_like(A), (A != 0).astype(np.float64),
                                  np.tile([0, 0, 1.0], (A.size, 1)))
    d = context.exchange_diagonal()
    return d * context.V_mag / float(V_mag)
