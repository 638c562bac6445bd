"""GPU parity on a small Voronoi polycrystal: per-grain easy axes, a grain-boundary phase with
its own exchange constant and K1 = 0, air; 10 materials in the lookup table.

Bar: element kernels <= 1e-12 relative to the oracle (FP64); a field step of each minimiser
reaches the oracle's energy to 1e-9 of the energy scale.
"""

from __future__ import annotations

import numpy as np
import pytest

from conftest import MU0, oracle_levels, tilted_m0

pytestmark = pytest.mark.gpu

TOL = 1e-12
N_GRAINS = 8


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-300))


def poly_materials(n_grains=N_GRAINS, seed=5):
    from mumag_matrixfree2_amd.mesh import random_easy_axes

    Kd = 1.61**2 / (2 * MU0)
    G = n_grains + 2
    A, K1, Js, k = np.zeros(G), np.zeros(G), np.zeros(G), np.tile([0.0, 0.0, 1.0], (G, 1))
    A[:n_grains], K1[:n_grains], Js[:n_grains] = 7.7e-12 * 1e18 / Kd, 4.3e6 / Kd, 1.0
    k[:n_grains] = random_easy_axes(n_grains, seed)
    A[n_grains], Js[n_grains] = 1.5e-12 * 1e18 / Kd, 1.0
    return A, K1, Js, k


@pytest.fixture(scope="module")
def poly():
    from mumag_matrixfree2_amd.context import MeshContext
    from mumag_matrixfree2_amd.mesh import polycrystal_cube
    from oracle.fem import Problem

    m = polycrystal_cube(30.0, 12, 3, n_grains=N_GRAINS, gb_width=1.5, seed=4)
    mats = poly_materials()
    print(f"[poly] N={m['knt'].shape[0]} E={m['ijk'].shape[0]} mats={np.bincount(m['ijk'][:, 4])}", flush=True)
    ctx = MeshContext(m["knt"], m["ijk"][:, :4], m["ijk"][:, 4], *mats, boundary_mask=m["boundary_mask"])
    ctx.setup_precond("amgcl", max_coarse=200)
    prob = Problem(m["knt"], m["ijk"], *mats, boundary_mask=m["boundary_mask"])
    yield ctx, prob
    ctx.close()


def test_poly_operators(poly):
    ctx, prob = poly
    rng = np.random.default_rng(11)
    x = rng.standard_normal(ctx.N)
    assert rel(ctx.apply_A(x), prob.A @ x) <= TOL
    m = tilted_m0(ctx.N, seed=12, sigma=0.5)
    assert rel(ctx.rhs(m), prob.rhs(m)) <= TOL
    U = rng.standard_normal(ctx.N)
    B = np.array([0.1, -0.2, 0.3])
    E, g = ctx.energy_grad(m, U, B)
    E_ref, g_ref = prob.energy_and_grad(m, U, B)
    assert abs(E - E_ref) <= TOL * abs(E_ref)
    assert rel(g, g_ref) <= TOL
    v = rng.standard_normal((ctx.N, 3))
    assert rel(ctx.local_grad(v), prob.local_grad(v)) <= TOL


@pytest.mark.parametrize("method", ["pcohen_hs", "lbfgs"])
def test_poly_field_step_energy(poly, method):
    from mumag_matrixfree2_amd.context import make_params
    from oracle.minimize import Minimizer
    from oracle.solvers import solve_U

    ctx, prob = poly
    m0 = tilted_m0(ctx.N, seed=13, sigma=0.05)
    B = np.array([0.017452, 0.0, 0.999848]) * (-1.5)
    p = make_params(method, tau_f=1e-10, cg_tol=1e-10)
    ctx.state_set(m0, np.zeros(ctx.N))
    info = ctx.minimize_resident(p, B)
    m, U = ctx.state_get()

    lv = oracle_levels(ctx.amg_levels, "amgcl")
    mn = Minimizer(prob, lambda mm, x0, tol: solve_U(prob, mm, x0, tol, precond="amg", levels=lv, kind="amgcl"),
                   method=method, cg_tol=1e-10)
    m_ref, U_ref, info_ref = mn.minimize(m0, B, U0=np.zeros(ctx.N), tau_f=1e-10)
    terms = prob.energy_terms(m_ref, U_ref, B)
    scale = sum(abs(v) for v in terms.values())
    print(f"[poly {method}] E={info['E']:.15e} E_ref={info_ref['E']:.15e} it={info['iters']}/{info_ref['iters']}",
          flush=True)
    assert info["converged"] and info_ref["converged"]
    assert abs(info["E"] - info_ref["E"]) <= 1e-9 * scale
    assert np.allclose(np.linalg.norm(m, axis=1), 1.0, atol=1e-12)
