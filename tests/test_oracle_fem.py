"""CPU checks that pin the oracle's geometry, operators and energy against the mathematics and
the reference's own analytic tests (tests/test_energy.py, tests/test_gradients.py)."""

from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

from conftest import MU0, ndfeb_materials, tilted_m0

GOLDEN = Path(__file__).resolve().parent / "golden"


def test_grid_mesher_reproduces_reference_fixture():
    from mumag_matrixfree2_amd.mesh import grid_box

    d = np.load(GOLDEN / "single_solid.npz")
    knt, ijk = grid_box(20.0, 20.0, 20.0, 2.0)
    assert np.array_equal(knt, d["knt"])
    assert np.array_equal(ijk, d["ijk"])


def test_geometry_basis_property(cube_c0):
    """grad phi_a . (x_b - x_c) = delta_ab - delta_ac: the defining property of P1 gradients."""
    from oracle.fem import compute_volume_JinvT, grad_phi_from_JinvT

    knt, ijk = cube_c0["knt"], cube_c0["ijk"]
    conn, vol, JinvT = compute_volume_JinvT(knt, ijk[:, :4])
    g = grad_phi_from_JinvT(JinvT)
    X = knt[conn]  # (E,4,3)
    M = np.einsum("eai,ebi->eab", g, X - X[:, :1, :])  # grad phi_a . (x_b - x_0)
    target = np.eye(4)[None] - np.eye(4)[:, :1][None]
    assert np.max(np.abs(M - target)) < 1e-12
    # volumes tile the box exactly; orientation fix leaves all dets positive
    L = knt.max() - knt.min()
    assert abs(vol.sum() - L**3) < 1e-9 * L**3
    x0 = knt[conn[:, 0]]
    J = np.stack([knt[conn[:, 1]] - x0, knt[conn[:, 2]] - x0, knt[conn[:, 3]] - x0], axis=2)
    assert np.all(np.linalg.det(J) > 0)


def test_orientation_fix_swaps_nodes_1_2():
    from oracle.fem import compute_volume_JinvT

    knt = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], dtype=float)
    conn = np.array([[0, 2, 1, 3]])  # negative orientation
    c, vol, _ = compute_volume_JinvT(knt, conn)
    assert c.tolist() == [[0, 1, 2, 3]]
    assert abs(vol[0] - 1.0 / 6.0) < 1e-15


def test_poisson_matrix_structure(oracle_c0):
    p = oracle_c0
    A = p.A
    assert abs(A - A.T).max() < 1e-15 * abs(A).max()
    bnd = p.boundary_mask == 0
    # Dirichlet rows are identity rows
    Ab = A[np.where(bnd)[0]]
    assert Ab.nnz == int(bnd.sum()) and np.all(Ab.data == 1.0)
    # interior rows away from the boundary: A.1 = reg and A.(linear) = reg * linear
    touches_bnd = bnd[p.conn].any(axis=1)
    adj = np.zeros(p.N, dtype=bool)
    adj[p.conn[touches_bnd].ravel()] = True
    inner = ~adj
    one = np.ones(p.N)
    assert np.max(np.abs((A @ one)[inner] - p.reg)) < 1e-13
    x = p.knt @ np.array([0.3, -1.1, 0.7])
    assert np.max(np.abs((A @ x)[inner] - p.reg * x[inner])) < 1e-11


def test_rhs_assembled_equals_element_loop(oracle_c0):
    m = tilted_m0(oracle_c0.N, seed=1, sigma=0.5)
    b1 = oracle_c0.rhs(m)
    b2 = oracle_c0.rhs_elementwise(m)
    assert np.linalg.norm(b1 - b2) <= 1e-13 * np.linalg.norm(b1)


def test_energy_assembled_equals_element_loop(oracle_c0):
    p = oracle_c0
    m = tilted_m0(p.N, seed=2, sigma=0.5)
    U = np.random.default_rng(3).standard_normal(p.N)
    B = np.array([0.1, 0.2, -0.4])
    E1, g1 = p.energy_and_grad(m, U, B)
    E2, g2 = p.energy_and_grad_elementwise(m, U, B)
    assert abs(E1 - E2) <= 1e-13 * (abs(E1) + 1)
    assert np.linalg.norm(g1 - g2) <= 1e-13 * np.linalg.norm(g1)


def _cube_problem(Js=1.6, K1=4.3e6, A_si=7.7e-12, layers=3):
    from mumag_matrixfree2_amd.mesh import config_cube_mesh
    from oracle.fem import Problem

    mesh = config_cube_mesh(20.0, 10, layers)
    A, K1r, Jsr, k = ndfeb_materials(Js=Js, K1=K1, A_si=A_si)
    p = Problem(mesh["knt"], mesh["ijk"], A, K1r, Jsr, k, boundary_mask=mesh["boundary_mask"])
    return p, Js**2 / (2 * MU0)


def test_energy_kats():
    """Analytic energies of tests/test_energy.py:77-145 (same tolerances)."""
    from oracle.solvers import solve_U

    Js, K1, A_si, L = 1.6, 4.3e6, 7.7e-12, 20.0
    p, Kd = _cube_problem(Js, K1, A_si)
    V_si = p.V_mag * 1e-27
    SI = Kd * V_si
    N = p.N
    x = p.knt[:, 0]
    kw = np.pi / L
    zero = np.zeros(N)
    # exchange of a helix
    m = np.column_stack([np.cos(kw * x), np.sin(kw * x), np.zeros(N)])
    E, _ = p.energy_and_grad(m, zero, np.zeros(3))
    # the helix includes the exchange-free air region; only the magnet carries A
    E_ana = A_si * (kw * 1e9) ** 2 * V_si
    assert abs(E * SI - E_ana) / E_ana < 0.02
    # Zeeman of uniform m_x in 0.1 T
    mx = np.zeros((N, 3))
    mx[:, 0] = 1.0
    E, _ = p.energy_and_grad(mx, zero, np.array([0.1 / Js, 0, 0]))
    # anisotropy of this material (K1 != 0) adds nothing for m _|_ k; the helix term is absent
    E_ana = -(1.0 / MU0) * Js * V_si * 0.1
    assert abs(E * SI - E_ana) / abs(E_ana) < 1e-6
    # anisotropy at 45 degrees
    m45 = np.zeros((N, 3))
    m45[:, 0] = m45[:, 2] = 1 / np.sqrt(2)
    E, _ = p.energy_and_grad(m45, zero, np.zeros(3))
    assert abs(E * SI - (-0.5 * K1 * V_si)) / (0.5 * K1 * V_si) < 1e-6
    # demag of a uniformly magnetised cube ~ 1/3 (reduced units)
    U, _, _ = solve_U(p, mx, zero, 1e-12, precond="direct")
    E, _ = p.energy_and_grad(mx, U, np.zeros(3))
    terms = p.energy_terms(mx, U, np.zeros(3))
    assert abs(terms["demag"] - 1.0 / 3.0) < 0.05
    assert abs(E - (terms["demag"] + terms["anisotropy"] + terms["exchange"])) < 1e-12


def test_fd_gradients_exact_for_quadratic_energy(oracle_c0):
    """Central differences of E are exact for a quadratic energy: every term at fixed U, and the
    full demag energy with U = A^-1 D m recomputed (tests/test_gradients.py:105-277)."""
    from oracle.solvers import solve_U

    p = oracle_c0
    rng = np.random.default_rng(4)
    m = tilted_m0(p.N, seed=5, sigma=0.4)
    B = np.array([0.05, -0.1, 0.3])
    zero = np.zeros(p.N)
    E0, g = p.energy_and_grad(m, zero, B)
    nodes = rng.choice(np.where(p.M_nodal > 0)[0], 6, replace=False)
    eps = 1e-3
    for i in nodes:
        for j in range(3):
            mp, mm = m.copy(), m.copy()
            mp[i, j] += eps
            mm[i, j] -= eps
            fd = (p.energy_and_grad(mp, zero, B)[0] - p.energy_and_grad(mm, zero, B)[0]) / (2 * eps)
            assert abs(fd - g[i, j]) < 1e-9 * max(1.0, abs(g[i, j]))
    # demag: E_d(m) = 1/2 m.G U(m) is quadratic in m with gradient G U(m)
    U = solve_U(p, m, zero, 1e-14, precond="direct")[0]
    gd = (p.G @ U).reshape(-1, 3) / p.V_mag

    def Ed(mv):
        Uv = solve_U(p, mv, zero, 1e-14, precond="direct")[0]
        return 0.5 * np.sum(mv * (p.G @ Uv).reshape(-1, 3)) / p.V_mag

    for i in nodes[:3]:
        for j in range(3):
            mp, mm = m.copy(), m.copy()
            mp[i, j] += eps
            mm[i, j] -= eps
            fd = (Ed(mp) - Ed(mm)) / (2 * eps)
            assert abs(fd - gd[i, j]) < 1e-8 * max(1.0, np.abs(gd).max())


def test_local_grad_is_exchange_plus_anisotropy(oracle_c0):
    p = oracle_c0
    m = tilted_m0(p.N, seed=6, sigma=0.3)
    zero = np.zeros(p.N)
    _, g = p.energy_and_grad(m, zero, np.zeros(3))
    assert np.linalg.norm(p.local_grad(m) - g) <= 1e-13 * np.linalg.norm(g)


def test_nodal_sums(oracle_c0):
    p = oracle_c0
    assert abs(p.node_vols.sum() - p.vol.sum()) < 1e-9 * p.vol.sum()
    assert abs(p.M_nodal.sum() - p.V_mag) < 1e-9 * p.V_mag
    assert np.all(p.inv_M_rel[p.M_nodal == 0] == 0)
