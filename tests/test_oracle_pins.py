"""Checks fixed by the mathematics that pin the oracle's V-cycle, Armijo line search, pcohen_hs
step and Steihaug exits, each with a planted mistake that the check rejects.

- V-cycle (``src/amg_utils.py:557-623``): with one level it is the exact inverse; on two and
  three levels its error propagation is (I - MA)(I - P B_c R A)(I - MA) (pre-smoothing, coarse
  correction with the coarse V-cycle B_c, post-smoothing); it is symmetric and positive definite.
- Armijo (``src/minimizers.py:265-385``): a synthetic energy whose ratio d(s) = (E(s) - E0)/(s pg)
  is prescribed along a Cayley rotation; the accepted step, the evaluation count and the warm-start
  chain of the Poisson solves are derived by hand from the expansion/contraction rule.
- pcohen_hs (``src/minimizers.py:950-1073``): with restart period L = 1 it is preconditioned
  steepest descent, iterate for iterate; the Hestenes-Stiefel beta makes the new direction
  conjugate to the gradient difference up to the projection defect (an identity the
  Polak-Ribiere beta does not satisfy).
- Steihaug (``src/minimizers.py:393-492``): the stagnation exit after one step returns the Cauchy
  point; negative curvature and zero iterations fall back to z = M g_tan; the 10x clip.
"""

from __future__ import annotations

import numpy as np
import pytest

from conftest import ndfeb_materials, oracle_levels, tilted_m0

# ----------------------------------------------------------------------------- V-cycle


def _dense_levels(sizes, kind, seed):
    """A random SPD Galerkin hierarchy: A_0 SPD, P_l full rank, A_{l+1} = P^T A_l P, R = P^T."""
    from oracle.fem import spai0_diagonal
    import scipy.sparse as sp

    rng = np.random.default_rng(seed)
    n0 = sizes[0]
    Q = rng.standard_normal((n0, n0))
    A = Q @ Q.T + n0 * np.eye(n0)
    levels = []
    for i, n in enumerate(sizes):
        As = sp.csr_matrix(A)
        M = spai0_diagonal(As) if kind == "amgcl" else A.diagonal().copy()
        lv = {"A": As, "M": M}
        if i + 1 < len(sizes):
            P = rng.standard_normal((n, sizes[i + 1]))
            lv["P"], lv["R"] = sp.csr_matrix(P), sp.csr_matrix(P.T)
            A = P.T @ A @ P
        else:
            lv["A_dense"] = A
        levels.append(lv)
    return levels


def _as_matrix(op, n):
    return np.column_stack([op(e) for e in np.eye(n)])


def _smoother(lv, kind, omega=0.6667):
    A = lv["A"].toarray()
    if kind == "amgcl":
        return np.diag(lv["M"])
    return np.diag(omega / (lv["M"] + 1e-30)) @ np.eye(A.shape[0])


def _error_propagation_expected(levels, kind):
    """I - B_0 A_0 built recursively from the two-grid formula with the coarse V-cycle as the
    coarse solver: B_L = A_L^-1 on the last level."""
    L = len(levels)
    B = np.linalg.inv(levels[-1]["A_dense"])
    for l in range(L - 2, -1, -1):
        lv = levels[l]
        A = lv["A"].toarray()
        n = A.shape[0]
        M = _smoother(lv, kind)
        P, R = lv["P"].toarray(), lv["R"].toarray()
        I = np.eye(n)
        E = (I - M @ A) @ (I - P @ B @ R @ A) @ (I - M @ A)
        B = (I - E) @ np.linalg.inv(A)  # the V-cycle operator of level l
    return I - B @ levels[0]["A"].toarray(), B


def _vcycle_matrix(levels, kind, vcycle=None):
    from oracle.solvers import vcycle as vc

    vcycle = vcycle or vc
    A0 = levels[0]["A"]
    return _as_matrix(lambda r: vcycle(levels, lambda v: A0 @ v, r, kind), A0.shape[0])


@pytest.mark.parametrize("kind", ["amgcl", "amg"])
def test_vcycle_one_level_is_the_exact_inverse(kind):
    levels = _dense_levels([12], kind, 1)
    B = _vcycle_matrix(levels, kind)
    np.testing.assert_allclose(B @ levels[0]["A"].toarray(), np.eye(12), atol=1e-12)


@pytest.mark.parametrize("kind", ["amgcl", "amg"])
@pytest.mark.parametrize("sizes", [[24, 9], [30, 12, 5]])
def test_vcycle_error_propagation(kind, sizes):
    levels = _dense_levels(sizes, kind, 2 + len(sizes))
    B = _vcycle_matrix(levels, kind)
    Eexp, Bexp = _error_propagation_expected(levels, kind)
    A = levels[0]["A"].toarray()
    np.testing.assert_allclose(np.eye(A.shape[0]) - B @ A, Eexp, atol=1e-10)
    np.testing.assert_allclose(B, Bexp, atol=1e-12 * np.abs(Bexp).max() * 1e3)
    # symmetric pre/post smoothing and R = P^T make B symmetric
    np.testing.assert_allclose(B, B.T, atol=1e-12 * np.abs(B).max())


def test_vcycle_planted_mistakes_fail_the_error_propagation_check():
    """A V-cycle without post-smoothing, or with the coarse correction subtracted, does not give
    the (I - MA)(I - P B_c R A)(I - MA) error propagation."""
    from oracle import solvers

    levels = _dense_levels([24, 9], "amgcl", 4)
    Eexp, _ = _error_propagation_expected(levels, "amgcl")
    A = levels[0]["A"].toarray()

    def no_post(levels, apply_A0, r, kind="amgcl", omega=0.6667):
        lv0, lv1 = levels
        x = lv0["M"] * r
        res = r - apply_A0(x)
        e = np.linalg.solve(lv1["A_dense"], lv0["R"] @ res)
        return x + lv0["P"] @ e

    def minus_coarse(levels, apply_A0, r, kind="amgcl", omega=0.6667):
        lv0, lv1 = levels
        x = lv0["M"] * r
        res = r - apply_A0(x)
        x = x - lv0["P"] @ np.linalg.solve(lv1["A_dense"], lv0["R"] @ res)
        return x + lv0["M"] * (r - apply_A0(x))

    planted = _mutated_oracle_function(solvers, "vcycle", "        return smooth(apply_A, b, x, lv[\"M\"])",
                                       "        return x")
    for bad in (no_post, minus_coarse, planted):
        B = _vcycle_matrix(levels, "amgcl", bad)
        assert np.abs((np.eye(24) - B @ A) - Eexp).max() > 1e-3
    B = _vcycle_matrix(levels, "amgcl", solvers.vcycle)
    assert np.abs((np.eye(24) - B @ A) - Eexp).max() < 1e-10


def test_vcycle_on_the_mesh_hierarchy_is_spd(small_cube):
    """On the small cube's SA hierarchy (matrix-free level 0, CSR coarse levels, dense last
    level) the V-cycle is a symmetric positive definite operator."""
    from mumag_matrixfree2_amd.amg import build_hierarchy
    from oracle.fem import Problem

    A, K1, Js, k = ndfeb_materials()
    p = Problem(small_cube["knt"], small_cube["ijk"], A, K1, Js, k, boundary_mask=small_cube["boundary_mask"])
    lv = oracle_levels(build_hierarchy(p.A, "amgcl", max_coarse=40), "amgcl")
    assert len(lv) >= 3
    from oracle.solvers import vcycle

    n = p.N
    B = _as_matrix(lambda r: vcycle(lv, p.apply_A, r, "amgcl"), n)
    np.testing.assert_allclose(B, B.T, atol=1e-12 * np.abs(B).max())
    assert np.linalg.eigvalsh(0.5 * (B + B.T)).min() > 0
    Eexp, _ = _error_propagation_expected([{**l, "A": l["A"]} for l in lv], "amgcl")
    np.testing.assert_allclose(np.eye(n) - B @ p.A.toarray(), Eexp, atol=1e-9)


# ----------------------------------------------------------------------------- Armijo


class _Path:
    """A one-node system on the Cayley path m(s) = cayley(m0, H, s) about H = x: the rotation
    angle is 2 atan(s/2), so s is read back from m; the energy is E0 + pg s d(s) for a
    prescribed d. solve_U returns U_guess + 1 (one 'iteration'), so the warm-start chain is
    visible in the returned U."""

    def __init__(self, dfun, pg=-1.0, E0=0.5):
        self.m0 = np.array([[0.0, 0.0, 1.0]])
        self.H = np.array([[1.0, 0.0, 0.0]])
        self.dfun, self.pg, self.E0 = dfun, pg, E0
        self.seen = []

    def s_of(self, m):
        ang = np.arctan2(np.linalg.norm(np.cross(self.m0[0], m[0])), self.m0[0] @ m[0])
        return 2.0 * np.tan(0.5 * ang)

    def energy_and_grad(self, m, U, B):
        s = self.s_of(m)
        self.seen.append(s)
        return self.E0 + self.pg * s * self.dfun(s), np.full((1, 3), s)

    def solve_U(self, m, x0, tol):
        return x0 + 1.0, 1, 0.0

    def run(self, s0=1.0):
        from oracle.minimize import armijo_ls

        return armijo_ls(self.energy_and_grad, self.solve_U, self.m0, self.pg, self.H, self.E0, np.zeros(1),
                         np.zeros((1, 3)), np.zeros(3), 1e-11, eta1=0.1, eta2=0.1, C=2.0, c=0.5, s0=s0, max_evals=15)


def _close(a, b):
    return abs(a - b) <= 1e-9 * max(1.0, abs(b))


def test_armijo_expansion_then_accept():
    """d(s) = 1 - 0.005 s (a quadratic with curvature 0.01): expansion continues while
    |1 - d| = 0.005 s < 0.1, i.e. at s = 1, 2, 4, 8, 16, and stops at s = 32 (|1 - d| = 0.16);
    d(32) = 0.84 >= 0.1 accepts without contraction: tau = 32 after 6 evaluations."""
    path = _Path(lambda s: 1.0 - 0.005 * s)
    tau, E, g, U, m, evals, demag = path.run()
    assert _close(tau, 32.0) and evals == 6 and demag == 6
    assert [round(s, 6) for s in path.seen] == [1, 2, 4, 8, 16, 32]
    assert _close(E, 0.5 - 32.0 * 0.84) and _close(path.s_of(m), 32.0)
    assert U[0] == 6.0  # each trial's solve starts from the previous trial's U


def test_armijo_contraction_to_the_midpoint():
    """d(s) = 1 - s: the first trial gives d = 0 (stop expanding, s_min = 0, d < eta1), the
    contraction s = 0 + 0.5 (1 - 0) = 0.5 gives d = 0.5: tau = 0.5 after 2 evaluations."""
    path = _Path(lambda s: 1.0 - s)
    tau, E, g, U, m, evals, demag = path.run()
    assert _close(tau, 0.5) and evals == 2 and U[0] == 2.0
    assert _close(E, 0.5 - 0.5 * 0.5)


def test_armijo_expansion_then_contraction_towards_s_min():
    """d(s) = 1 - 0.01 s below s = 6 and -1 (energy increase) from 6 on: expansion at 1, 2, 4
    (|1 - d| < 0.1, s_min becomes 4) stops at 8 (d = -1); contraction from s_min = 4:
    s = 4 + 0.5 (8 - 4) = 6 (d = -1), then s = 4 + 0.5 (6 - 4) = 5 (d = 0.95, accepted):
    tau = 5 after 6 evaluations. Contracting towards 0 instead of s_min would accept s = 4."""
    path = _Path(lambda s: 1.0 - 0.01 * s if s < 6.0 - 1e-9 else -1.0)
    tau, E, g, U, m, evals, demag = path.run()
    assert [round(s, 6) for s in path.seen] == [1, 2, 4, 8, 6, 5]
    assert _close(tau, 5.0) and evals == 6
    assert _close(E, 0.5 - 5.0 * 0.95)


def test_armijo_failure_returns_the_start():
    """d < 0 everywhere: one expansion trial, then 15 contractions (the cap), all rejected; the
    safety rule returns tau = 0, E0, the input U and m, with 16 evaluations counted."""
    path = _Path(lambda s: -1.0)
    tau, E, g, U, m, evals, demag = path.run()
    assert tau == 0.0 and E == 0.5 and evals == 16 and U[0] == 0.0
    assert np.array_equal(m, path.m0)
    assert _close(path.seen[-1], 2.0 ** -15)


def test_armijo_expansion_cap_returns_the_next_step():
    """d = 1 (a linear energy): the expansion never stops and ends at the cap of 15 evaluations
    (s = 1 .. 2^14). The carried step is then C s = 2^15 while the returned state is that of the
    last trial, 2^14 (the reference returns s_final of its loop state)."""
    path = _Path(lambda s: 1.0)
    tau, E, g, U, m, evals, demag = path.run()
    assert evals == 15 and _close(tau, 2.0**15)
    assert abs(path.s_of(m) - 2.0**14) <= 1e-6 * 2.0**14
    assert _close(E, 0.5 - 2.0**14)


def test_armijo_non_finite_energy_is_an_increase():
    """A non-finite trial energy counts as 1e20 (src/minimizers.py:297): at s = 1 the energy is
    NaN, the contraction to 0.5 is a plain decrease (d = 1 - 0.1 s): tau = 0.5, 2 evaluations."""
    path = _Path(lambda s: float("nan") if s > 0.75 else 1.0 - 0.1 * s)
    tau, E, g, U, m, evals, demag = path.run()
    assert _close(tau, 0.5) and evals == 2


def _mutated_oracle_function(module, name, old, new):
    """A copy of an oracle function with one line changed (a planted mistake), compiled in the
    oracle module's namespace."""
    import inspect
    import textwrap

    src = textwrap.dedent(inspect.getsource(getattr(module, name)))
    assert old in src, old
    ns = dict(vars(module))
    exec(compile(src.replace(old, new), f"<mutated {name}>", "exec"), ns)
    return ns[name]


@pytest.mark.parametrize("old,new", [
    ("s = s_min_exp + c * (s - s_min_exp)", "s = c * s"),            # contraction towards 0
    ("s_min = s_min if stop else s", "s_min = s_min"),                # s_min never advances
    ("stop = abs(1.0 - d_n) >= eta2 or d_n < 0", "stop = d_n < 0"),   # no eta2 test
])
def test_armijo_planted_mistakes_are_caught(monkeypatch, old, new):
    """Each planted change of the expansion/contraction rule in a copy of the oracle's
    armijo_ls changes the hand-derived step or evaluation count of the cases above."""
    from oracle import minimize

    bad = _mutated_oracle_function(minimize, "armijo_ls", old, new)
    monkeypatch.setattr(minimize, "armijo_ls", bad)
    results = []
    for dfun, want in ((lambda s: 1.0 - 0.005 * s, (32.0, 6)),
                       (lambda s: 1.0 - 0.01 * s if s < 6.0 - 1e-9 else -1.0, (5.0, 6))):
        tau, *_, evals, _ = _Path(dfun).run()
        results.append(_close(tau, want[0]) and evals == want[1])
    assert not all(results)


# ----------------------------------------------------------------------------- pcohen_hs


@pytest.fixture(scope="module")
def small_problem(small_cube):
    from mumag_matrixfree2_amd.amg import build_hierarchy
    from oracle.fem import Problem
    from oracle.solvers import solve_U

    A, K1, Js, k = ndfeb_materials()
    p = Problem(small_cube["knt"], small_cube["ijk"], A, K1, Js, k, boundary_mask=small_cube["boundary_mask"])
    lv = oracle_levels(build_hierarchy(p.A, "amgcl"), "amgcl")

    def sU(m, x0, tol):
        return solve_U(p, m, x0, tol, precond="amg", levels=lv, kind="amgcl")

    return p, sU


def _steepest_descent(p, sU, m0, B, tau_f, max_iter):
    """Preconditioned steepest descent on the sphere written from the pieces: d = -P^-1 g_tan
    (Steihaug), Armijo on the Cayley path, the Gill-Murray test on max |y|."""
    from oracle.minimize import armijo_ls, check_convergence, solve_Py_g, tangent_grad

    cg_tol = 1e-10
    phi_tol = min(cg_tol, 0.1 * tau_f)
    m = m0 / np.linalg.norm(m0, axis=1, keepdims=True)
    U, _, _ = sU(m, np.zeros(p.N), cg_tol)
    E, g = p.energy_and_grad(m, U, B)
    out = []
    for it in range(max_iter):
        g_tan = tangent_grad(m, g * p.inv_M_rel[:, None])
        gte = tangent_grad(m, g)
        gi = float(np.max(np.abs(g_tan)))
        y, _ = solve_Py_g(p.local_grad, m, g, gte, p.inv_M_prec, 10, min(0.5, gi**0.5), 0.0, 0.01)
        d = -y
        pg = float(np.vdot(g, d))
        tau, E1, g1, U1, m1, _, _ = armijo_ls(p.energy_and_grad, sU, m, pg, np.cross(m, d), E, U, g, B, phi_tol)
        conv = check_convergence(it, E1, E, m, m1, float(np.max(np.abs(y))), tau_f, 1e-12)
        m, U, g, E = m1, U1, g1, E1
        out.append(E)
        if conv:
            break
    return m, out


def test_pcohen_with_restart_period_one_is_steepest_descent(small_problem):
    from oracle.minimize import Minimizer

    p, sU = small_problem
    m0 = tilted_m0(p.N, seed=7, sigma=0.3)
    B = np.array([0.0, 0.0, -0.2])
    trace = []
    m, U, info = Minimizer(p, sU, "pcohen_hs", cg_tol=1e-10).minimize(m0, B, tau_f=1e-10, L=1, max_iter=60,
                                                                      trace=trace)
    m_sd, E_sd = _steepest_descent(p, sU, m0, B, 1e-10, 60)
    assert all(t["beta"] == 0.0 for t in trace)
    assert [t["E"] for t in trace] == E_sd
    assert np.array_equal(m, m_sd)


def _hs_defect(t):
    """d . dg + beta sum_i (m_i . d_prev_i)(m_i . dg_i): zero for the HS beta, since
    d = -y + beta (d_prev - (m.d_prev) m) and beta d_prev.dg = y.dg."""
    m, dg, dp = t["m"], t["dg"], t["d_prev"]
    proj = np.sum(np.sum(m * dp, axis=1) * np.sum(m * dg, axis=1))
    return np.vdot(t["d"], dg) + t["beta"] * proj, np.linalg.norm(t["y"]) * np.linalg.norm(dg)


@pytest.mark.parametrize("beta_type", ["hs", "pr"])
def test_pcohen_hs_direction_is_conjugate_to_the_gradient_difference(small_problem, beta_type):
    """For every iteration with beta > 0 (no restart, no descent fallback), the HS direction
    satisfies d . dg = -beta sum_i (m_i . d_prev_i)(m_i . dg_i) to rounding; the Polak-Ribiere
    beta (planted through beta_type='pr') violates it."""
    from oracle.minimize import Minimizer

    p, sU = small_problem
    m0 = tilted_m0(p.N, seed=8, sigma=0.3)
    trace = []
    Minimizer(p, sU, "pcohen_hs", cg_tol=1e-10).minimize(m0, np.array([0.0, 0.0, -0.3]), tau_f=1e-10, max_iter=40,
                                                         trace=trace, trace_vectors=True, beta_type=beta_type)
    active = [t for t in trace if t["beta"] > 0 and not t["fallback"]]
    assert len(active) >= 3
    rel = [abs(a) / b for a, b in map(_hs_defect, active)]
    if beta_type == "hs":
        assert max(rel) < 1e-10
    else:
        assert max(rel) > 1e-4


def test_pcohen_restart_period(small_problem):
    """With L = 3, beta is 0 at iterations 0, 3, 6, ... and there d = -y exactly."""
    from oracle.minimize import Minimizer

    p, sU = small_problem
    trace = []
    Minimizer(p, sU, "pcohen_hs", cg_tol=1e-10).minimize(tilted_m0(p.N, seed=9, sigma=0.3), np.array([0, 0, -0.3]),
                                                         tau_f=1e-10, max_iter=12, L=3, trace=trace,
                                                         trace_vectors=True)
    for t in trace:
        if t["it"] % 3 == 0:
            assert t["beta"] == 0.0 and np.array_equal(t["d"], -t["y"])
    assert any(t["beta"] > 0 for t in trace if t["it"] % 3)


# ----------------------------------------------------------------------------- Steihaug exits


def _synthetic(n=5, seed=0, C=None, gscale=1.0):
    rng = np.random.default_rng(seed)
    m = rng.standard_normal((n, 3))
    m /= np.linalg.norm(m, axis=1, keepdims=True)
    if C is None:
        Q = rng.standard_normal((3 * n, 3 * n))
        C = Q @ Q.T + 3 * n * np.eye(3 * n)
    g = gscale * rng.standard_normal((n, 3))
    Mp = rng.uniform(0.5, 2.0, n)

    def local_grad(v):
        return (C @ v.ravel()).reshape(-1, 3)

    def P(v):
        Cv = local_grad(v)
        return Cv - np.sum(m * Cv, axis=1, keepdims=True) * m - np.sum(m * g, axis=1, keepdims=True) * v

    gt = g - np.sum(m * g, axis=1, keepdims=True) * m
    return m, g, gt, Mp, local_grad, P


def test_steihaug_stagnation_exit_returns_the_cauchy_point():
    """With nu = 1e6 the second iteration sees dq <= nu Q and stops before updating: y is the
    first CG step alpha_0 z_0 with alpha_0 = rho_0 / (z_0 . P z_0), and 2 iterations are counted.
    With nu = 0 the loop runs to max_iter."""
    from oracle.minimize import solve_Py_g

    m, g, gt, Mp, lg, P = _synthetic()
    z0 = gt * Mp[:, None]
    rho0 = np.vdot(gt, z0)
    cauchy = rho0 / np.vdot(z0, P(z0)) * z0
    y, it = solve_Py_g(lg, m, g, gt, Mp, max_iter=10, tol=0.0, stagnation_nu=1e6)
    assert it == 2
    np.testing.assert_allclose(y, cauchy, rtol=1e-13, atol=1e-15)
    y0, it0 = solve_Py_g(lg, m, g, gt, Mp, max_iter=10, tol=0.0, stagnation_nu=0.0)
    assert it0 == 10 and np.linalg.norm(y0 - cauchy) > 1e-6 * np.linalg.norm(cauchy)


def test_steihaug_negative_curvature_falls_back_to_z():
    """C = -I with m.g >= 0 makes P negative on tangent vectors: the first iteration exits
    without a step (y = 0), which is no descent direction, so y = z = M g_tan; 1 iteration."""
    from oracle.minimize import solve_Py_g

    n = 4
    m, g, gt, Mp, lg, P = _synthetic(n=n, seed=3, C=-np.eye(3 * n))
    g = g + (np.abs(np.sum(m * g, axis=1)) + 1.0)[:, None] * m  # m.g > 0 on every node
    gt = g - np.sum(m * g, axis=1, keepdims=True) * m
    y, it = solve_Py_g(lg, m, g, gt, Mp, max_iter=10, tol=0.0)
    assert it == 1
    np.testing.assert_array_equal(y, gt * Mp[:, None])


def test_steihaug_zero_iterations_fall_back_to_z():
    from oracle.minimize import solve_Py_g

    m, g, gt, Mp, lg, P = _synthetic(seed=4)
    y, it = solve_Py_g(lg, m, g, gt, Mp, max_iter=0)
    assert it == 0
    np.testing.assert_array_equal(y, gt * Mp[:, None])


def test_steihaug_clips_to_ten_times_z():
    """A nearly singular positive P (C = 1e-6 I, m.g < 0 tiny) gives a first step alpha_0 z_0
    with alpha_0 >> 10: the result is clipped to length 10 |z| along z, i.e. y = 10 z."""
    from oracle.minimize import solve_Py_g

    n = 4
    m, g, gt, Mp, lg, P = _synthetic(n=n, seed=5, C=1e-6 * np.eye(3 * n))
    g = gt - 1e-6 * m  # m.g = -1e-6: P = (1e-6 + 1e-6) I on tangent vectors
    y, it = solve_Py_g(lg, m, g, gt, Mp, max_iter=1, tol=0.0)
    z = gt * Mp[:, None]
    np.testing.assert_allclose(y, 10.0 * z, rtol=1e-12)
