"""GPU field-step minimiser (pcohen_hs, lbfgs) against the oracle minimiser run with the same
AMG hierarchy, and a short hysteresis sweep against the oracle sweep (switching step)."""

from __future__ import annotations

import numpy as np
import pytest

from conftest import ndfeb_materials, oracle_levels, tilted_m0

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small_pair(small_cube):
    from mumag_matrixfree2_amd.context import MeshContext
    from oracle.fem import Problem
    from oracle.solvers import solve_U

    A, K1, Js, k = ndfeb_materials()
    m = small_cube
    ctx = MeshContext(m["knt"], m["ijk"][:, :4], m["ijk"][:, 4], A, K1, Js, k, boundary_mask=m["boundary_mask"])
    ctx.setup_precond("amgcl")
    p = Problem(m["knt"], m["ijk"], A, K1, Js, k, boundary_mask=m["boundary_mask"])
    lv = oracle_levels(ctx.amg_levels, "amgcl")

    def sU(mm, x0, tol):
        return solve_U(p, mm, x0, tol, precond="amg", levels=lv, kind="amgcl")

    print(f"[small_pair] N={ctx.N} E={ctx.E} levels={len(lv)}", flush=True)
    return ctx, p, sU


@pytest.mark.parametrize("method", ["pcohen_hs", "lbfgs"])
@pytest.mark.parametrize("Bz", [0.5, -0.3])
def test_field_step_matches_oracle(small_pair, method, Bz):
    from mumag_matrixfree2_amd.context import make_params
    from oracle.minimize import Minimizer

    ctx, p, sU = small_pair
    m0 = tilted_m0(ctx.N, seed=21, sigma=0.2)
    B = np.array([0.017452, 0.0, 0.999848]) * Bz
    kw = dict(tau_f=1e-10, eps_a=1e-12, max_iter=500)
    ctx.state_set(m0, np.zeros(ctx.N))
    info = ctx.minimize_resident(make_params(method, cg_tol=1e-10, **kw), B)
    m_gpu, U_gpu = ctx.state_get()
    m_ref, U_ref, info_ref = Minimizer(p, sU, method=method, cg_tol=1e-10).minimize(m0, B, **kw)
    print(f"[{method} Bz={Bz}] gpu it={info['iters']} evals={info['evals']} E={info['E']:.15e} | "
          f"ref it={info_ref['iters']} evals={info_ref['evals']} E={info_ref['E']:.15e}", flush=True)
    terms = p.energy_terms(m_ref, U_ref, B)
    scale = sum(abs(v) for v in terms.values())
    assert info["converged"] and info_ref["converged"]
    assert abs(info["E"] - info_ref["E"]) <= 1e-9 * scale
    assert info["iters"] == info_ref["iters"] and info["evals"] == info_ref["evals"]
    mag = p.M_nodal > 0
    assert np.max(np.abs(m_gpu[mag] - m_ref[mag])) < 1e-6
    assert np.allclose(np.linalg.norm(m_gpu, axis=1), 1.0, atol=1e-13)


def test_sweep_switching_step_matches_oracle(small_pair):
    """Hysteresis sweep +2 T -> -8 T (0.5 T steps, 1 deg tilt): identical switching step."""
    from mumag_matrixfree2_amd.hysteresis_loop import LoopParams, run_hysteresis_loop
    from oracle.minimize import Minimizer, field_values

    ctx, p, sU = small_pair
    Js_ref = 1.61
    h = np.array([0.017452, 0.0, 0.999848])
    params = LoopParams(h_dir=h, B_start=2.0 / Js_ref, B_end=-8.0 / Js_ref, dB=-0.5 / Js_ref, tau_f=1e-10,
                        cg_tol=1e-10, method="pcohen_hs", Js_ref=Js_ref)
    m0 = tilted_m0(ctx.N, seed=0, sigma=0.01)
    res = run_hysteresis_loop(ctx, m0, params)
    J_gpu = res["history"][:, 1]
    mn = Minimizer(p, sU, method="pcohen_hs", cg_tol=1e-10)
    m = m0.copy()
    U = np.zeros(p.N)
    J_ref = []
    for Bmag in field_values(params.B_start, params.B_end, params.dB, False):
        m, U, info = mn.minimize(m, Bmag * h / np.linalg.norm(h), U0=U, tau_f=1e-10)
        J_ref.append(p.J_parallel(m, h))
    J_ref = np.array(J_ref)
    from mumag_matrixfree2_amd.hysteresis_loop import switching_step

    print("J_gpu", np.round(J_gpu, 4).tolist(), "\nJ_ref", np.round(J_ref, 4).tolist(), flush=True)
    assert switching_step(J_gpu) == switching_step(J_ref)
    assert switching_step(J_ref) > 0
    assert np.max(np.abs(J_gpu - J_ref)) < 1e-3


def test_sweep_writes_result_files(small_pair, tmp_path):
    """hysteresis.csv, .mh and a .vtu snapshot from a 3-step sweep: the CSV J_par and E are the
    sweep's, the .mh energy is E Js_ref^2 / (2 mu0) and (mx, my, mz) is the magnetic-volume
    average of m, which for Js = 1 everywhere in the magnet equals J along each axis."""
    from mumag_matrixfree2_amd import io_utils
    from mumag_matrixfree2_amd.hysteresis_loop import LoopParams, run_hysteresis_loop

    ctx, p, sU = small_pair
    Js_ref = 1.61
    h = np.array([0.017452, 0.0, 0.999848])
    params = LoopParams(h_dir=h, B_start=2.0 / Js_ref, B_end=-8.0 / Js_ref, dB=-0.5 / Js_ref, tau_f=1e-10,
                        cg_tol=1e-10, method="pcohen_hs", Js_ref=Js_ref)
    res = run_hysteresis_loop(ctx, tilted_m0(ctx.N, seed=0, sigma=0.01), params, steps=3, out_dir=tmp_path,
                              name="cube", snapshot_every=2)
    header, csv = io_utils.read_hysteresis_csv(tmp_path / "hysteresis.csv")
    hist = res["history"]
    np.testing.assert_allclose(csv[:, 2], hist[:, 1] * Js_ref, rtol=1e-9)
    np.testing.assert_allclose(csv[:, 3], hist[:, 2], rtol=1e-9)
    mh = np.loadtxt(tmp_path / "cube.mh")
    assert mh.shape == (3, 6)
    np.testing.assert_allclose(mh[:, 5], hist[:, 2] * Js_ref**2 / (2 * 4e-7 * np.pi), rtol=1e-9)
    m_fin = res["m"]
    mx_ref = [ctx.j_parallel(m_fin, e) for e in np.eye(3)]
    np.testing.assert_allclose(mh[-1, 2:5], mx_ref, atol=1e-12)
    assert sorted(f.name.split("_B")[0] for f in tmp_path.glob("*.vtu")) == ["state_cfg00000", "state_cfg00002"]
    plog = (tmp_path / "params.log").read_text().splitlines()
    assert plog[0] == "| Parameter | Value | Source |"
    assert "| method | pcohen_hs | default |" in plog and f"| phi_tol | {1e-10 * 0.1} | derived |" in plog


@pytest.mark.parametrize("k", [4, 8])
def test_projected_solve_guess_keeps_the_sweep(small_pair, k):
    """The projected initial guess of the minimiser's Poisson solves changes where each PCG
    starts, not where it stops: a +2 -> -8 T sweep gives the energies of the reference warm
    start to 1e-9 of the energy scale, the same minimiser iterations within 1 per step and the
    same switching step, and no more PCG iterations in all."""
    from mumag_matrixfree2_amd.context import make_params
    from mumag_matrixfree2_amd.hysteresis_loop import switching_step
    from oracle.minimize import field_values

    ctx, p, _ = small_pair
    h = np.array([0.017452, 0.0, 0.999848])
    par = make_params("pcohen_hs", tau_f=1e-10, eps_a=1e-12, cg_tol=1e-10)
    m0 = tilted_m0(ctx.N, seed=0, sigma=0.01)
    runs = {}
    for kk in (0, k):
        ctx.set_solve_guess(kk)
        ctx.state_set(m0, np.zeros(ctx.N))
        rows = []
        for Bmag in field_values(2.0 / 1.61, -8.0 / 1.61, -0.5 / 1.61, False):
            info = ctx.minimize_resident(par, Bmag * h)
            rows.append((info["E"], info["iters"], info["demag_iters"], ctx.state_j_parallel(h)))
        runs[kk] = np.array(rows)
    ctx.set_solve_guess(0)
    a, b = runs[0], runs[k]
    terms = p.energy_terms(*ctx.state_get(), -8.0 / 1.61 * h)
    scale = sum(abs(v) for v in terms.values())
    print(f"[guess {k}] PCG iterations {int(a[:, 2].sum())} -> {int(b[:, 2].sum())}, "
          f"max dE {np.max(np.abs(a[:, 0] - b[:, 0])):.2e}", flush=True)
    assert np.max(np.abs(a[:, 0] - b[:, 0])) <= 1e-9 * scale
    assert np.max(np.abs(a[:, 1] - b[:, 1])) <= 1
    assert switching_step(a[:, 3]) == switching_step(b[:, 3])
    assert b[:, 2].sum() <= a[:, 2].sum()
