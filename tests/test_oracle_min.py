"""CPU checks that pin the oracle's manifold tools, line search, preconditioner solve and the
pcohen_hs / L-BFGS field-step minimisers against closed forms and invariants."""

from __future__ import annotations

import numpy as np
import pytest

from conftest import ndfeb_materials, oracle_levels, tilted_m0


def test_cayley_is_rotation_by_2atan_k():
    from oracle.minimize import cayley_transport, cayley_update

    rng = np.random.default_rng(0)
    m = rng.standard_normal((50, 3))
    m /= np.linalg.norm(m, axis=1, keepdims=True)
    H = rng.standard_normal((50, 3))
    tau = 0.7
    mn = cayley_update(m, H, tau)
    assert np.allclose(np.linalg.norm(mn, axis=1), 1.0, atol=1e-15)
    # rotation about k = tau H / 2 by the angle 2 atan|k|: the component along k is preserved
    k = 0.5 * tau * H
    kh = k / np.linalg.norm(k, axis=1, keepdims=True)
    assert np.allclose(np.sum(mn * kh, axis=1), np.sum(m * kh, axis=1), atol=1e-14)
    mp = m - np.sum(m * kh, axis=1, keepdims=True) * kh
    mnp = mn - np.sum(mn * kh, axis=1, keepdims=True) * kh
    cosang = np.sum(mp * mnp, axis=1) / (np.linalg.norm(mp, axis=1) * np.linalg.norm(mnp, axis=1))
    assert np.allclose(np.arccos(np.clip(cosang, -1, 1)), 2 * np.arctan(np.linalg.norm(k, axis=1)), atol=1e-7)
    # transport is the same orthogonal map: preserves inner products
    v = rng.standard_normal((50, 3))
    w = rng.standard_normal((50, 3))
    assert np.allclose(np.sum(cayley_transport(v, H, tau) * cayley_transport(w, H, tau), axis=1),
                       np.sum(v * w, axis=1), atol=1e-13)


def test_tangent_projection():
    from oracle.minimize import tangent_grad

    rng = np.random.default_rng(1)
    m = rng.standard_normal((20, 3))
    m /= np.linalg.norm(m, axis=1, keepdims=True)
    g = rng.standard_normal((20, 3))
    t = tangent_grad(m, g)
    assert np.allclose(np.sum(m * t, axis=1), 0.0, atol=1e-15)
    assert np.allclose(tangent_grad(m, t), t, atol=1e-15)


def test_check_convergence_cases():
    from oracle.minimize import check_convergence

    m = np.zeros((2, 3))
    assert not check_convergence(0, 1.0, 1.0, m, m, 0.0, 1e-8, 1e-12)  # never at it == 0
    assert check_convergence(1, 1.0, 1.0, m, m, 0.0, 1e-8, 1e-12)       # U4: gnorm < eps_a
    assert check_convergence(1, 1.0, 1.0, m, m + 1e-5, 1e-3, 1e-8, 1e-12)  # U1-U3
    assert not check_convergence(1, 1.0, 1.0, m, m + 1e-3, 1e-3, 1e-8, 1e-12)  # U2 fails


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


def test_steihaug_one_step_is_cauchy_point(small_problem):
    from oracle.minimize import solve_Py_g, tangent_grad

    p, _ = small_problem
    m = tilted_m0(p.N, seed=2, sigma=0.3)
    g = p.energy_and_grad(m, np.zeros(p.N), np.array([0, 0, -0.5]))[1]
    gt = tangent_grad(m, g)
    Mp = p.inv_M_prec[:, None]
    y, it = solve_Py_g(p.local_grad, m, g, gt, p.inv_M_prec, max_iter=1, tol=0.0, stagnation_nu=0.0)
    z = gt * Mp
    Cz = p.local_grad(z)
    Pz = Cz - np.sum(m * Cz, axis=1, keepdims=True) * m - np.sum(m * g, axis=1, keepdims=True) * z
    cauchy = (np.vdot(gt, z) / np.vdot(z, Pz)) * z
    assert it == 1
    if np.vdot(z, Pz) > 0:
        assert np.linalg.norm(y - cauchy) <= 1e-12 * np.linalg.norm(cauchy) or np.allclose(y, z)


def test_steihaug_converges_on_spd_tangent_system(small_problem):
    """With P SPD on the tangent space, the inner PCG solves P y = g_tan (residual -> 0)."""
    from oracle.minimize import solve_Py_g, tangent_grad

    p, _ = small_problem
    m = tilted_m0(p.N, seed=3, sigma=0.05)
    # a field along m keeps -(m.g) >= 0 so P stays positive definite on the tangent space
    g = p.energy_and_grad(m, np.zeros(p.N), np.array([0.0, 0.0, 2.0]))[1]
    gt = tangent_grad(m, g)
    y, it = solve_Py_g(p.local_grad, m, g, gt, p.inv_M_prec, max_iter=500, tol=1e-10, reg=0.0, stagnation_nu=0.0)
    Cy = p.local_grad(y)
    Py = Cy - np.sum(m * Cy, axis=1, keepdims=True) * m - np.sum(m * g, axis=1, keepdims=True) * y
    mag = p.M_nodal > 0
    assert np.linalg.norm((Py - gt)[mag]) <= 1e-6 * np.linalg.norm(gt[mag])
    assert np.vdot(y, gt) > 0


def test_armijo_accepts_sufficient_decrease(small_problem):
    from oracle.minimize import armijo_ls, tangent_grad

    p, sU = small_problem
    m = tilted_m0(p.N, seed=4, sigma=0.2)
    B = np.array([0.0, 0.0, 0.5])
    U, _, _ = sU(m, np.zeros(p.N), 1e-10)
    E0, g = p.energy_and_grad(m, U, B)
    d = -tangent_grad(m, g * p.inv_M_rel[:, None])
    H = np.cross(m, d)
    pg = float(np.vdot(g, d))
    tau, E, g1, U1, m1, ev, dem = armijo_ls(p.energy_and_grad, sU, m, pg, H, E0, U, g, B, 1e-11, s0=1e-2)
    assert pg < 0 and tau > 0 and E < E0
    assert (E - E0) / (tau * pg) >= 0.1
    assert np.allclose(np.linalg.norm(m1, axis=1), 1.0)
    # an ascent direction is rejected without evaluation
    t2, E2, *_ , ev2, _ = armijo_ls(p.energy_and_grad, sU, m, -pg, -H, E0, U, g, B, 1e-11)
    assert t2 == 0.0 and E2 == E0 and ev2 == 0


@pytest.mark.parametrize("method", ["pcohen_hs", "lbfgs"])
def test_minimizer_invariants(small_problem, method):
    from oracle.minimize import Minimizer

    p, sU = small_problem
    mn = Minimizer(p, sU, method=method, cg_tol=1e-10)
    trace = []
    m0 = tilted_m0(p.N, seed=5, sigma=0.2)
    m, U, info = mn.minimize(m0, np.array([0.0, 0.0, 0.5]), tau_f=1e-10, trace=trace)
    Es = [t["E"] for t in trace]
    assert info["converged"]
    assert np.all(np.diff(Es) <= 1e-14 * np.abs(Es[:-1]).max())  # energy never increases
    assert np.allclose(np.linalg.norm(m, axis=1), 1.0, atol=1e-14)
    if method == "lbfgs":
        assert all(t["curv"] > 0 for t in trace if t["ok"])


def test_minimizers_agree_on_the_minimum(small_problem):
    from oracle.minimize import Minimizer

    p, sU = small_problem
    m0 = tilted_m0(p.N, seed=6, sigma=0.2)
    B = np.array([0.01, 0.0, -0.2])
    res = {}
    for method in ("pcohen_hs", "lbfgs"):
        m, U, info = Minimizer(p, sU, method=method, cg_tol=1e-10).minimize(m0, B, tau_f=1e-12)
        res[method] = (m, info)
    Ea, Eb = res["pcohen_hs"][1]["E"], res["lbfgs"][1]["E"]
    assert abs(Ea - Eb) <= 1e-8 * max(1.0, abs(Ea))
    mag = p.M_nodal > 0
    assert np.max(np.abs(res["pcohen_hs"][0][mag] - res["lbfgs"][0][mag])) < 1e-3


def test_field_values():
    from oracle.minimize import field_values

    v = field_values(2.0, -8.0, -0.25)
    assert v.size == 41 and v[0] == 2.0 and abs(v[-1] + 8.0) < 1e-12
    v = field_values(2.0, -8.0, 0.25)  # sign of dH follows the sweep direction
    assert v.size == 41
    v = field_values(0.0, 1.0, 0.3)
    assert np.allclose(v, [0.0, 0.3, 0.6, 0.9, 1.0])
    assert field_values(0.0, 1.0, 0.5, loop=True).tolist() == [0.0, 0.5, 1.0, 0.5, 0.0]


def _dense_inverse_bfgs(S, Y, rho, hist, invMr):
    """The L-BFGS inverse Hessian as an explicit matrix: H0 = gam D, then the BFGS inverse update
    H <- (I - rho s y^T)^T H (I - rho y s^T) + rho s s^T for each stored pair, oldest first
    (Nocedal & Wright, Numerical Optimization, eqs. 6.17 and 7.19)."""
    M = len(rho)
    nh = min(hist, M)
    n = S[0].size
    d = np.repeat(invMr.ravel(), 3)
    if hist > 0:
        s, y = S[(hist - 1) % M].ravel(), Y[(hist - 1) % M].ravel()
        gam = float(np.clip(s @ y / (y @ (d * y)), 1e-3, 1e3))
    else:
        gam = 1.0
    H = gam * np.diag(d)
    for i in range(nh):
        idx = (hist - nh + i) % M
        s, y = S[idx].ravel(), Y[idx].ravel()
        V = np.eye(n) - rho[idx] * np.outer(y, s)
        H = V.T @ H @ V + rho[idx] * np.outer(s, s)
    return H


@pytest.mark.parametrize("hist", [0, 1, 3, 4, 7])
def test_lbfgs_two_loop_is_the_dense_inverse_bfgs_update(hist):
    """The two-loop recursion equals the explicit product of BFGS inverse updates on a ring of
    4 pairs, also after the ring has wrapped (hist 7): a wrong loop order, ring index or sign of
    a correction term fails it."""
    from oracle.minimize import lbfgs_two_loop

    rng = np.random.default_rng(11 + hist)
    N, M = 6, 4
    S = rng.standard_normal((M, N, 3))
    Y = 0.7 * S + 0.2 * rng.standard_normal((M, N, 3))
    rho = np.array([1.0 / np.vdot(S[i], Y[i]) for i in range(M)])
    assert np.all(rho > 0)
    invMr = rng.uniform(0.5, 2.0, (N, 1))
    q = rng.standard_normal((N, 3))
    r = lbfgs_two_loop(q, S, Y, rho, hist, invMr)
    H = _dense_inverse_bfgs(S, Y, rho, hist, invMr)
    np.testing.assert_allclose(r.ravel(), H @ q.ravel(), rtol=1e-12, atol=1e-12)
    if hist > 0:
        # secant condition of the newest pair: H y = s
        last = (hist - 1) % M
        np.testing.assert_allclose(lbfgs_two_loop(Y[last], S, Y, rho, hist, invMr), S[last],
                                   rtol=1e-11, atol=1e-11)
