"""Batched entry points over independent contexts (a small grain set): the same results as the
per-context calls, per-system tolerances honoured, bad batches rejected.

Bar: energies of a field step equal the sequential ones to 1e-10 relative (the only source of
difference is the order of the FP64 atomic adds on block-shared nodes); solves meet the true
residual tolerance against the oracle's assembled matrix.
"""

from __future__ import annotations

import numpy as np
import pytest

from conftest import ndfeb_materials, tilted_m0

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def grains():
    from mumag_matrixfree2_amd.context import MeshContext
    from mumag_matrixfree2_amd.mesh import grain_set

    gs = grain_set(5, seed=7, t_min=3e3, t_max=3e4)
    A, K1, Js, _ = ndfeb_materials()
    ctxs = []
    for g in gs:
        k = np.array([g["easy_axis"], g["easy_axis"]])
        c = MeshContext(g["knt"], g["ijk"][:, :4], g["ijk"][:, 4], A, K1, Js, k, boundary_mask=g["boundary_mask"])
        c.setup_precond("amgcl", max_coarse=300)
        ctxs.append(c)
    print(f"[grains] N={[c.N for c in ctxs]} shapes={[g['shape'] for g in gs]}", flush=True)
    yield gs, ctxs
    for c in ctxs:
        c.close()


def test_minimize_batched_equals_sequential(grains):
    from mumag_matrixfree2_amd.context import make_params, minimize_batched

    gs, ctxs = grains
    p = make_params("pcohen_hs", tau_f=1e-10, cg_tol=1e-10)
    h = np.array([0.017452, 0.0, 0.999848])
    B = np.array([h * b for b in np.linspace(-0.2, -0.9, len(ctxs))])
    m0 = [tilted_m0(c.N, seed=40 + i, sigma=0.05) for i, c in enumerate(ctxs)]
    seq = []
    for c, m, b in zip(ctxs, m0, B):
        c.state_set(m, np.zeros(c.N))
        seq.append(c.minimize_resident(p, b))
    m_seq = [c.state_get()[0] for c in ctxs]
    for c, m in zip(ctxs, m0):
        c.state_set(m, np.zeros(c.N))
    bat = minimize_batched(ctxs, p, B, n_threads=4)
    for i, (a, b) in enumerate(zip(seq, bat)):
        print(f"[grain {i}] seq E={a['E']:.15e} it={a['iters']} | batched E={b['E']:.15e} it={b['iters']}", flush=True)
        assert a["converged"] and b["converged"]
        assert abs(a["E"] - b["E"]) <= 1e-10 * abs(a["E"])
        assert abs(a["iters"] - b["iters"]) <= 2
        assert np.max(np.abs(ctxs[i].state_get()[0] - m_seq[i])) < 1e-6


def test_solve_U_batched_per_system_tolerance(grains):
    from mumag_matrixfree2_amd.context import solve_U_batched
    from oracle.fem import Problem

    gs, ctxs = grains
    A, K1, Js, _ = ndfeb_materials()
    ms = [tilted_m0(c.N, seed=50 + i, sigma=0.3) for i, c in enumerate(ctxs)]
    tols = np.array([1e-6, 1e-8, 1e-10, 1e-10, 1e-9])[: len(ctxs)]
    Us, its, rrs = solve_U_batched(ctxs, ms, [np.zeros(c.N) for c in ctxs], tols, n_threads=3)
    for i, (g, c) in enumerate(zip(gs, ctxs)):
        k = np.array([g["easy_axis"], g["easy_axis"]])
        prob = Problem(g["knt"], g["ijk"], A, K1, Js, k, boundary_mask=g["boundary_mask"])
        b = prob.rhs(ms[i])
        true_rr = np.linalg.norm(b - prob.A @ Us[i]) / np.linalg.norm(b)
        U1, it1, rr1 = c.solve_U(ms[i], np.zeros(c.N), tols[i])
        print(f"[grain {i}] tol={tols[i]:.0e} it={its[i]} (single {it1}) rr={rrs[i]:.2e} true={true_rr:.2e}",
              flush=True)
        assert true_rr <= tols[i] * 1.001
        assert abs(its[i] - it1) <= 1


def test_bad_batches_are_rejected(grains):
    from mumag_matrixfree2_amd.context import make_params, minimize_batched

    _, ctxs = grains
    p = make_params("pcohen_hs")
    with pytest.raises(RuntimeError, match="twice"):
        minimize_batched([ctxs[0], ctxs[0]], p, np.zeros(3))
    for c in ctxs:
        c.state_set(np.tile([0.0, 0.0, 1.0], (c.N, 1)))
