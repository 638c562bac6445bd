"""Oracle: manifold tools, Armijo line search, Steihaug-PCG preconditioner, pcohen_hs and L-BFGS
(test infrastructure only). Semantics follow the reference's GPU (JAX) path in ``src/minimizers.py``.
"""

from __future__ import annotations

import numpy as np


def cayley_update(m, H, tau):
    """Unit-length-preserving Cayley rotation of m about k = tau H / 2 (``src/minimizers.py:71-79``)."""
    k = 0.5 * tau * H
    k2 = np.sum(k * k, axis=1, keepdims=True)
    km = np.cross(k, m)
    kdm = np.sum(k * m, axis=1, keepdims=True)
    mn = ((1.0 - k2) * m + 2.0 * km + 2.0 * kdm * k) / (1.0 + k2)
    return mn / np.linalg.norm(mn, axis=1, keepdims=True)


def cayley_transport(v, H, tau):
    """The same rotation applied to a tangent vector, without renormalisation (``src/minimizers.py:82-91``)."""
    k = 0.5 * tau * H
    k2 = np.sum(k * k, axis=1, keepdims=True)
    return ((1.0 - k2) * v + 2.0 * np.cross(k, v) + 2.0 * np.sum(k * v, axis=1, keepdims=True) * k) / (1.0 + k2)


def tangent_grad(m, g):
    """g - (m.g) m (``src/minimizers.py:94-96``)."""
    return g - np.sum(m * g, axis=1, keepdims=True) * m


def check_convergence(it, E, E_prev, m, m_new, gnorm_inf, tau_f, eps_a):
    """Gill-Murray U1..U4 (``src/minimizers.py:160-172``)."""
    if it <= 0:
        return False
    dm = np.max(np.abs(m_new - m))
    u1 = (E_prev - E) < tau_f * (1.0 + abs(E))
    u2 = dm < np.sqrt(tau_f) * 2.0
    u3 = gnorm_inf <= tau_f ** (1.0 / 3.0) * (1.0 + abs(E))
    u4 = gnorm_inf < eps_a
    return bool((u1 and u2 and u3) or u4)


def armijo_ls(energy_and_grad, solve_U, m, pg, H, E0, U_base, g_raw_init, B_ext, phi_tol,
              eta1=0.1, eta2=0.1, C=2.0, c=0.5, s0=1.0, max_evals=15):
    """Curvilinear Armijo line search ``make_armijo_ls_v2`` (``src/minimizers.py:265-385``).

    Each trial: m_t = cayley(m, H, s); U_t = solve_U(m_t, U of the previous trial, phi_tol);
    E_t, g_t = energy_and_grad; d = (E_t - E0)/(s pg + 1e-30). Expansion while
    |1-d| < eta2 and d >= 0 (s *= C), then contraction s = s_min + c (s - s_min) until
    eta1 <= d < 1e10. Returns (tau, E, g_raw, U, m, evals, demag_iters).
    """
    if not pg < 0:
        return 0.0, E0, g_raw_init, U_base, m, 0, 0

    def D(s, U_guess):
        mt = cayley_update(m, H, s)
        Ut, it, _ = solve_U(mt, U_guess, phi_tol)
        Et, gt = energy_and_grad(mt, Ut, B_ext)
        if not np.isfinite(Et):
            Et = 1e20
        return (Et - E0) / (s * pg + 1e-30), Et, gt, Ut, mt, it

    s, s_min, it, done = float(s0), 0.0, 0, False
    E_v, g_v, U_v, m_v, d_v, demag = E0, g_raw_init, U_base, m, 0.0, 0
    while it < max_evals and not done:
        d_n, E_v, g_v, U_v, m_v, di = D(s, U_v)
        stop = abs(1.0 - d_n) >= eta2 or d_n < 0
        s_next = s if stop else C * s
        s_min = s_min if stop else s
        s = s_next
        it += 1
        done = stop
        d_v = d_n
        demag += di
    it_exp, s_min_exp = it, s_min
    done = (d_v >= eta1) and (d_v < 1e10)
    it_con = 0
    while it_con < max_evals and not done:
        s = s_min_exp + c * (s - s_min_exp)
        d_v, E_v, g_v, U_v, m_v, di = D(s, U_v)
        done = (d_v >= eta1) and (d_v < 1e10)
        it_con += 1
        demag += di
    if d_v >= 0:
        return s, E_v, g_v, U_v, m_v, it_exp + it_con, demag
    return 0.0, E0, g_raw_init, U_base, m, it_exp + it_con, demag


def solve_Py_g(local_grad, m, g_ext, g_tan_ext, inv_M_prec, max_iter=10, tol=0.0, reg=0.0, stagnation_nu=1e-3):
    """Steihaug-PCG for P y = g_tan with Jacobi scaling (``src/minimizers.py:393-492``).

    P v = C v - (m.Cv) m - (m.g) v + reg v with C the local (exchange+anisotropy)
    Hessian. Exits on negative curvature or stagnation (dq <= nu Q); the result is
    clipped to 10 |z| and replaced by z = M g_tan if it is not a descent direction.
    Returns (y, iterations) with the JAX iteration count.
    """
    Mp = np.asarray(inv_M_prec).reshape(-1, 1)
    mg = np.sum(m * g_ext, axis=1, keepdims=True)

    def P(v):
        Cv = local_grad(v)
        return Cv - np.sum(m * Cv, axis=1, keepdims=True) * m - mg * v + reg * v

    y = np.zeros_like(g_tan_ext)
    r = g_tan_ext.copy()
    z = r * Mp
    p = z.copy()
    rho = np.vdot(r, z)
    target = tol**2 * rho
    Q, it, done = 0.0, 0, False
    while it < max_iter and rho > target and rho > 1e-25 and rho < 1e20 and not done:
        Ap = P(p)
        pAp = np.vdot(p, Ap)
        neg = pAp <= 0.0
        alpha = rho / (pAp + 1e-30)
        dq = 0.5 * alpha * rho
        stag = it > 0 and dq <= stagnation_nu * Q
        done = neg or stag
        if not done:
            y = y + alpha * p
            r = r - alpha * Ap
        z = r * Mp
        rho_n = np.vdot(r, z)
        if not done:
            p = z + (rho_n / (rho + 1e-30)) * p
        rho = rho_n
        Q = Q + dq
        it += 1
    zf = g_tan_ext * Mp
    yn, zn = np.linalg.norm(y), np.linalg.norm(zf)
    if yn > 10.0 * zn:
        y = y * (10.0 * zn / (yn + 1e-30))
    if not np.vdot(y, g_tan_ext) > 1e-12:
        y = zf
    return y, it


def lbfgs_two_loop(q, S, Y, rho, hist, invMr):
    """H q for the L-BFGS inverse Hessian (two-loop recursion, src/minimizers.py:1284-1312).

    S, Y (M, N, 3) are a ring of M pairs, rho[i] = 1 / (s_i . y_i); the newest of the ``hist``
    pairs stored so far sits at (hist - 1) % M. H0 = gam * diag(invMr) with
    gam = s.y / (y . invMr y) of the newest pair, clipped to [1e-3, 1e3] (1 with no pairs).
    """
    M = len(rho)
    nh = min(hist, M)
    q = np.array(q, dtype=np.float64, copy=True)
    alphas = np.zeros(M)
    for i in range(nh):
        idx = (hist - 1 - i) % M
        alphas[idx] = rho[idx] * np.vdot(S[idx], q)
        q = q - alphas[idx] * Y[idx]
    last = (hist - 1) % M
    if hist > 0:
        gam = np.vdot(S[last], Y[last]) / (np.vdot(Y[last], Y[last] * invMr) + 1e-30)
    else:
        gam = 1.0
    gam = float(np.clip(gam, 1e-3, 1e3))
    r = gam * q * invMr
    for i in range(nh):
        idx = (hist - nh + i) % M
        b = rho[idx] * np.vdot(Y[idx], r)
        r = r + S[idx] * (alphas[idx] - b)
    return r


class Minimizer:
    """Field-step minimiser: ``make_minimizer`` + ``solve_and_minimize`` + ``kernel`` (``src/minimizers.py:3402-3452``).

    ``method`` is 'pcohen_hs' (``src/minimizers.py:950-1073``, beta_type 'hs') or
    'lbfgs' (``src/minimizers.py:1216-1381``, memory 10).
    """

    def __init__(self, problem, solve_U, method="pcohen_hs", cg_tol=1e-10, memory=10):
        self.p = problem
        self.solve_U = solve_U
        self.method = method
        self.cg_tol = cg_tol
        self.memory = memory

    def minimize(self, m0, B_ext, U0=None, tau_f=1e-10, eps_a=1e-12, max_iter=2000, tau0=1e-2,
                 ls_eta1=0.1, ls_eta2=0.1, ls_C=2.0, ls_c=0.5, pc_iters=10, pc_auto=True,
                 pc_force_eta=0.5, pc_force_alpha=0.5, pc_stagnation_nu=0.01, pc_reg=0.0, L=None,
                 trace=None, trace_vectors=False, beta_type="hs"):
        """One field step. ``beta_type`` 'hs' (Hestenes-Stiefel, the reference's pcohen_hs) or 'pr'
        (Polak-Ribiere), the two branches of ``make_pcohen_minimizer`` (``src/minimizers.py:1000-1010``).
        ``trace`` (a list) receives one dict per iteration; with ``trace_vectors`` the pcohen_hs
        entries also hold m, y, the gradient difference, d_prev and d."""
        p = self.p
        N = p.N
        phi_tol = float(min(self.cg_tol, tau_f * 0.1))
        invMr = p.inv_M_rel[:, None]
        m = np.asarray(m0, dtype=np.float64)
        m = m / np.linalg.norm(m, axis=1, keepdims=True)
        U0 = np.zeros(N) if U0 is None else U0
        U, demag, _ = self.solve_U(m, U0, self.cg_tol)
        eg = p.energy_and_grad
        E, g_raw = eg(m, U, B_ext)
        g_tan = tangent_grad(m, g_raw * invMr)
        g_tan_ext = tangent_grad(m, g_raw)
        gnorm = float(np.max(np.abs(g_tan)))
        evals, preco = 1, 0
        it, conv = 0, False
        Lp = L or N
        ls_args = dict(eta1=ls_eta1, eta2=ls_eta2, C=ls_C, c=ls_c, max_evals=15)
        if self.method == "pcohen_hs":
            g_prev, y_prev, d_prev = g_tan_ext, g_tan_ext, -g_tan_ext
            while not conv and it < max_iter:
                g_tan = tangent_grad(m, g_raw * invMr)
                g_tan_ext = tangent_grad(m, g_raw)
                gi = float(np.max(np.abs(g_tan)))
                pc_tol = min(pc_force_eta, gi**pc_force_alpha) if pc_auto else 0.0
                y, pit = solve_Py_g(p.local_grad, m, g_raw, g_tan_ext, p.inv_M_prec, pc_iters, pc_tol,
                                    pc_reg, pc_stagnation_nu)
                gs = float(np.max(np.abs(y)))
                dg = g_tan_ext - g_prev
                if beta_type == "pr":
                    den = np.vdot(y_prev, g_prev) + 1e-30
                else:
                    den = np.vdot(d_prev, dg) + 1e-30
                beta = 0.0 if it % Lp == 0 else max(0.0, np.vdot(y, dg) / den)
                d = -y + beta * tangent_grad(m, d_prev)
                fallback = bool(np.vdot(d, g_tan_ext) > 0)
                if fallback:
                    d = -y
                H = np.cross(m, d)
                pg = float(np.vdot(g_raw, d))
                tau, E_new, g_new, U_new, m_new, le, ld = armijo_ls(
                    eg, self.solve_U, m, pg, H, E, U, g_raw, B_ext, phi_tol, s0=1.0, **ls_args)
                conv = check_convergence(it, E_new, E, m, m_new, gs, tau_f, eps_a)
                if trace is not None:
                    t = dict(it=it, E=E_new, tau=tau, pg=pg, evals=le, preco=pit, gnorm=gs, beta=beta,
                             fallback=fallback)
                    if trace_vectors:
                        t.update(m=m, y=y, dg=dg, d_prev=d_prev, d=d)
                    trace.append(t)
                g_prev, y_prev, d_prev = g_tan_ext, y, d
                m, U, g_raw, E, gnorm = m_new, U_new, g_new, E_new, gs
                it += 1
                evals += le
                preco += pit
                demag += ld
        elif self.method == "lbfgs":
            M = self.memory
            S = np.zeros((M, N, 3))
            Y = np.zeros((M, N, 3))
            rho = np.zeros(M)
            hist = 0
            while not conv and it < max_iter:
                g_tan = tangent_grad(m, g_raw * invMr)
                g_tan_ext = tangent_grad(m, g_raw)
                gi = float(np.max(np.abs(g_tan)))
                d = -lbfgs_two_loop(g_tan_ext, S, Y, rho, hist, invMr)
                if it == 0 or np.vdot(d, g_tan_ext) > 0:
                    d = -g_tan
                H = np.cross(m, d)
                pg = float(np.vdot(g_raw, d))
                s0 = tau0 if it == 0 else 1.0
                tau, E_new, g_new, U_new, m_new, le, ld = armijo_ls(
                    eg, self.solve_U, m, pg, H, E, U, g_raw, B_ext, phi_tol, s0=s0, **ls_args)
                conv = check_convergence(it, E_new, E, m, m_new, gi, tau_f, eps_a)
                s_new = tangent_grad(m_new, d * tau)
                y_new = tangent_grad(m_new, g_new) - tangent_grad(m_new, g_tan_ext)
                curv = np.vdot(y_new, s_new)
                ok = curv > 1e-12 * np.vdot(s_new, s_new)
                idx = hist % M
                S = np.stack([tangent_grad(m_new, h) for h in S])
                Y = np.stack([tangent_grad(m_new, h) for h in Y])
                if ok:
                    S[idx], Y[idx], rho[idx] = s_new, y_new, 1.0 / (curv + 1e-30)
                    hist += 1
                if trace is not None:
                    trace.append(dict(it=it, E=E_new, tau=tau, pg=pg, evals=le, curv=curv, ok=ok, gnorm=gi))
                m, U, g_raw, E, gnorm = m_new, U_new, g_new, E_new, gi
                it += 1
                evals += le
                demag += ld
        else:
            raise NotImplementedError(self.method)
        return m, U, dict(iters=it, E=float(E), gnorm=float(gnorm), evals=evals, preco_iters=preco,
                          demag_iters=int(demag), converged=bool(conv))


def field_values(H_start, H_end, dH, loop=False):
    """Field sweep values (``src/hysteresis_loop.py:128-150``)."""
    if dH == 0:
        raise ValueError("dH must be nonzero")
    if (H_end - H_start) * dH < 0:
        dH = -dH
    n = int(np.floor((H_end - H_start) / dH + 1e-12))
    vals = H_start + dH * np.arange(n + 1)
    if abs(vals[-1] - H_end) > 1e-12 * max(1.0, abs(H_end)):
        vals = np.append(vals, H_end)
    if not loop:
        return vals
    return np.concatenate([vals, vals[-2::-1]])
