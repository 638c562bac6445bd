"""Host set-up of the smoothed-aggregation AMG hierarchy used as the PCG preconditioner.

The reference builds it with ``pyamg.smoothed_aggregation_solver(A, strength='evolution',
smooth='energy', max_levels=10, max_coarse=1000)`` on the masked, regularised Poisson
matrix (``src/poisson_solve.py:570-639``). pyamg is not available here, so this module
builds a classical SA hierarchy itself:

* symmetric strength of connection |a_ij| >= theta sqrt(|a_ii a_jj|) with theta = 0.1 on the
  fine level and 0.05 below: the graded air cells are strongly anisotropic, and keeping their
  weak couplings out of the aggregates cuts the PCG iterations of the 2.06M-tet cube from
  173 (theta = 0) to 22; Dirichlet identity rows are isolated and never aggregated;
* greedy three-pass aggregation (``mm_sa_aggregate`` in the library, host code);
* tentative prolongator with normalised constant columns, smoothed once with damped
  Jacobi, omega = 1.5 / rho(D^-1 A) (the textbook 4/3 gave 11% more PCG iterations on the
  2.06M-tet cube sweep); on the fine level with the filtered matrix (weak couplings lumped into
  the diagonal): on the 24.6M-tet polycrystal the fine prolongator keeps 3.3 instead of 5.7
  entries per row and the level-1 operator 44 instead of 67, for 20 instead of 19 PCG
  iterations from a zero guess (filtering the coarser levels too costs a third more iterations);
* R = P^T, A_c = R A P, down to at most ``max_coarse`` rows (N/64 clamped to [1000, 6000]) or
  ``max_levels`` levels;
* per-level smoother multiplier: SPAI0 diagonal a_ii / sum_j a_ij^2 for 'amgcl'
  (``src/amg_utils.py:87-111``), 0.6667 / a_ii for 'amg' (``src/amg_utils.py:377-404``);
* the coarsest level is inverted once (dense), instead of the reference's per-call
  ``jnp.linalg.solve`` (``src/amg_utils.py:577-578``): the same solution.

Level 0 is applied matrix-free on the GPU; only its P, R and smoother diagonal are uploaded.
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from ._hip import call, f64, i32, ptr


def fine_matrix(ctx) -> sp.csr_matrix:
    """Assembled Poisson matrix of a context: Ve grad phi_a . grad phi_b, Dirichlet rows and
    columns removed with unit diagonal, reg on the interior diagonal (``src/amg_utils.py:21-84``)."""
    conn, Ve, grad = ctx.geometry()
    N = ctx.N
    g = np.empty((conn.shape[0], 4, 3))
    g[:, 1:, :] = grad
    g[:, 0, :] = -grad.sum(axis=1)
    Ke = Ve[:, None, None] * np.einsum("eai,ebi->eab", g, g)
    c64 = conn.astype(np.int64)
    rows = np.repeat(c64, 4, axis=1).ravel()
    cols = np.tile(c64, (1, 4)).ravel()
    A = sp.csr_matrix((Ke.ravel(), (rows, cols)), shape=(N, N))
    A.sum_duplicates()
    diag_extra = np.full(N, ctx.reg)
    if ctx.boundary_mask is not None:
        b = ctx.boundary_mask == 0
        coo = A.tocoo()
        keep = ~(b[coo.row] | b[coo.col])
        A = sp.csr_matrix((coo.data[keep], (coo.row[keep], coo.col[keep])), shape=(N, N))
        diag_extra[b] = 1.0
    A = (A + sp.diags(diag_extra)).tocsr()
    A.sort_indices()
    return A


def strength(A: sp.csr_matrix, theta: float = 0.0) -> sp.csr_matrix:
    """Symmetric strength graph without the diagonal."""
    C_ = A.tocoo()
    d = np.abs(A.diagonal())
    off = C_.row != C_.col
    keep = off & (C_.data != 0)
    if theta > 0:
        keep &= np.abs(C_.data) >= theta * np.sqrt(d[C_.row] * d[C_.col])
    S = sp.csr_matrix((np.ones(int(keep.sum())), (C_.row[keep], C_.col[keep])), shape=A.shape)
    S.sort_indices()
    return S


def aggregate(S: sp.csr_matrix) -> tuple[np.ndarray, int]:
    """Aggregate ids per node (-1: unaggregated), numbered in the order of each aggregate's
    first node, so the coarse unknowns follow the node order: the coarse rows of a set of
    independent systems on contiguous node ranges (a segmented context) are contiguous too."""
    Sp = i32(S.indptr)
    Sj = i32(S.indices)
    agg = np.empty(S.shape[0], dtype=np.int32)
    na = C.c_int32(0)
    call("mm_sa_aggregate", S.shape[0], ptr(Sp), ptr(Sj), ptr(agg), C.byref(na))
    na = int(na.value)
    rows = np.flatnonzero(agg >= 0)
    first = np.full(na, S.shape[0], dtype=np.int64)
    np.minimum.at(first, agg[rows], rows)
    new_id = np.empty(na, dtype=np.int32)
    new_id[np.argsort(first, kind="stable")] = np.arange(na, dtype=np.int32)
    agg[rows] = new_id[agg[rows]]
    return agg, na


def spectral_radius_DinvA(A: sp.csr_matrix) -> float:
    d = A.diagonal()
    s = 1.0 / np.sqrt(np.where(d > 0, d, 1.0))
    B = sp.diags(s) @ A @ sp.diags(s)
    n = A.shape[0]
    if n < 50:
        return float(np.max(np.abs(np.linalg.eigvalsh(B.toarray()))))
    v0 = np.random.default_rng(0).standard_normal(n)
    lam = spla.eigsh(B, k=1, which="LA", tol=1e-3, maxiter=200, v0=v0, return_eigenvectors=False)
    return float(lam[0])


def filtered(A: sp.csr_matrix, theta: float) -> sp.csr_matrix:
    """A with its weak off-diagonal entries (|a_ij| < theta sqrt(|a_ii a_jj|), the strength test)
    added to the diagonal instead: the same row sums, only strong couplings off the diagonal."""
    C_ = A.tocoo()
    d = np.abs(A.diagonal())
    off = C_.row != C_.col
    weak = off & (np.abs(C_.data) < theta * np.sqrt(d[C_.row] * d[C_.col]))
    lump = np.bincount(C_.row[weak], weights=C_.data[weak], minlength=A.shape[0])
    keep = ~weak
    Af = sp.csr_matrix((C_.data[keep], (C_.row[keep], C_.col[keep])), shape=A.shape) + sp.diags(lump)
    Af = Af.tocsr()
    Af.sort_indices()
    return Af


def build_hierarchy(A: sp.csr_matrix, smoother: str = "amgcl", max_levels: int = 10, max_coarse: int | None = None,
                    theta=(0.1, 0.05), omega: float = 1.5, filter_levels=(0,), log=None) -> list[dict]:
    """List of level dicts: ``A`` (CSR), ``M`` (smoother multiplier), ``P``, ``R`` (all but last),
    ``A_dense_inv`` (last). ``theta`` is one strength threshold or one per level (the last
    entry repeats). ``max_coarse`` defaults to N/64 clamped to [1000, 6000]: on the GPU a
    dense inverse of up to ~6000 rows (one symmetric GEMV streaming its lower triangle) is
    cheaper than the several small kernels of one or two more levels, and it is the exact coarse
    solve (24.6M-tet polycrystal: a 5,742-row coarsest level instead of a fifth level of 278 rows,
    PCG iterations -2%, field-steps/s +2.6% in the FP32-PCG form and +4% in FP64). On the levels in
    ``filter_levels`` the prolongator is smoothed with the filtered matrix (weak couplings
    lumped into the diagonal), so it spreads only along strong couplings."""
    if max_coarse is None:
        max_coarse = int(min(6000, max(1000, A.shape[0] // 64)))
    thetas = [float(theta)] if np.isscalar(theta) else [float(t) for t in theta]
    levels = []
    A = A.tocsr()
    while A.shape[0] > max_coarse and len(levels) < max_levels - 1:
        S = strength(A, thetas[min(len(levels), len(thetas) - 1)])
        agg, na = aggregate(S)
        if na == 0 or na >= A.shape[0]:
            break
        rows = np.where(agg >= 0)[0]
        cnt = np.bincount(agg[rows], minlength=na).astype(np.float64)
        T = sp.csr_matrix((1.0 / np.sqrt(cnt[agg[rows]]), (rows, agg[rows])), shape=(A.shape[0], na))
        As = filtered(A, thetas[min(len(levels), len(thetas) - 1)]) if len(levels) in tuple(filter_levels) else A
        d = As.diagonal()
        Dinv = 1.0 / np.where(d != 0, d, 1.0)
        rho = spectral_radius_DinvA(As)
        P = (T - (omega / rho) * (sp.diags(Dinv) @ (As @ T))).tocsr()
        P.eliminate_zeros()
        P.sort_indices()
        R = P.T.tocsr()
        R.sort_indices()
        levels.append({"A": A, "P": P, "R": R})
        A = (R @ A @ P).tocsr()
        A.sort_indices()
        if log is not None:
            log(f"[amg] level {len(levels)}: {A.shape[0]} rows, {A.nnz} nonzeros")
    levels.append({"A": A})
    for lv in levels:
        Al = lv["A"]
        if smoother == "amgcl":
            A2 = sp.csr_matrix((Al.data**2, Al.indices, Al.indptr), shape=Al.shape)
            lv["M"] = Al.diagonal() / (np.asarray(A2.sum(axis=1)).ravel() + 1e-30)
        else:
            lv["M"] = 0.6667 / (Al.diagonal() + 1e-30)
    levels[-1]["A_dense_inv"] = np.linalg.inv(levels[-1]["A"].toarray())
    return levels


def upload_hierarchy(ctx, levels: list[dict], kind: str) -> None:
    """Copy the hierarchy to the device (level 0 keeps its matrix-free operator)."""
    n_lev = len(levels)
    for l, lv in enumerate(levels):
        A = lv["A"]
        n = A.shape[0]
        last = l == n_lev - 1
        Ap = Ai = Ax = None
        if l > 0:
            Ap, Ai, Ax = i32(A.indptr), i32(A.indices), f64(A.data)
        P = lv.get("P")
        R = lv.get("R")
        n_next = 0 if last else P.shape[1]
        Pp = Pi = Px = Rp = Ri = Rx = None
        if not last:
            Pp, Pi, Px = i32(P.indptr), i32(P.indices), f64(P.data)
            Rp, Ri, Rx = i32(R.indptr), i32(R.indices), f64(R.data)
        dinv = f64(lv["A_dense_inv"]) if last else None
        M = f64(lv["M"])
        call("mm_amg_set_level", ctx.handle, l, n, n_next, ptr(Ap), ptr(Ai), ptr(Ax), ptr(M), ptr(Pp), ptr(Pi),
             ptr(Px), ptr(Rp), ptr(Ri), ptr(Rx), ptr(dinv))
    call("mm_amg_commit", ctx.handle, 1 if kind == "amgcl" else 2, n_lev)
