"""The multi-core CPU restatement (oracle/cpu_minimizer.cpp, the CPU baseline of bench.py)
against the NumPy oracle it restates: the same field step on the same mesh and AMG hierarchy.

Bar: the same minimiser iterations and line-search evaluations, energies to 1e-12 relative,
magnetisations to 1e-9, PCG iterations within 1 per solve on average (the coarsest solve is a
product with the inverse instead of a dense solve, so the V-cycle differs in rounding).
"""

from __future__ import annotations

import numpy as np
import pytest

from conftest import ndfeb_materials, oracle_levels, tilted_m0


@pytest.fixture(scope="module")
def cpu_pair():
    from mumag_matrixfree2_amd.amg import build_hierarchy
    from mumag_matrixfree2_amd.mesh import config_cube_mesh
    from oracle import cpu
    from oracle.fem import Problem

    if not cpu.LIB.exists():
        cpu.build()
    mesh = config_cube_mesh(12.0, 6, 2)
    A, K1, Js, k = ndfeb_materials(axis=(0.2, 0.0, 1.0))
    p = Problem(mesh["knt"], mesh["ijk"], A, K1, Js, k, boundary_mask=mesh["boundary_mask"])
    levels = build_hierarchy(p.A.tocsr(), max_coarse=100)
    cm = cpu.CpuMinimizer(p, levels)
    yield p, levels, cm
    cm.close()


@pytest.mark.parametrize("Bz", [0.4, -0.8])
def test_cpu_field_step_matches_numpy_oracle(cpu_pair, Bz):
    from oracle.minimize import Minimizer
    from oracle.solvers import solve_U

    p, levels, cm = cpu_pair
    lv = oracle_levels(levels, "amgcl")
    its = []

    def sU(m, x0, tol):
        out = solve_U(p, m, x0, tol, precond="amg", levels=lv, kind="amgcl")
        its.append(out[1])
        return out

    m0 = tilted_m0(p.N, seed=3, sigma=0.1)
    B = np.array([0.017452, 0.0, 0.999848]) * Bz
    m_ref, U_ref, info_ref = Minimizer(p, sU, method="pcohen_hs", cg_tol=1e-10).minimize(m0, B, tau_f=1e-10)
    m_c, U_c, info_c = cm.minimize(m0, B, tau_f=1e-10)
    print(f"[Bz={Bz}] numpy it={info_ref['iters']} nf={info_ref['evals']} icg={info_ref['demag_iters']} "
          f"E={info_ref['E']:.15e} | cpu it={info_c['iters']} nf={info_c['evals']} icg={info_c['demag_iters']} "
          f"E={info_c['E']:.15e}", flush=True)
    assert info_c["converged"] and info_ref["converged"]
    assert info_c["iters"] == info_ref["iters"] and info_c["evals"] == info_ref["evals"]
    assert abs(info_c["E"] - info_ref["E"]) <= 1e-12 * abs(info_ref["E"])
    assert np.max(np.abs(m_c - m_ref)) < 1e-9
    assert abs(info_c["demag_iters"] - info_ref["demag_iters"]) <= len(its)
    assert info_c["preco_iters"] == info_ref["preco_iters"]


def test_cpu_uses_the_host_threads():
    from oracle import cpu

    if not cpu.LIB.exists():
        cpu.build()
    assert cpu.threads() >= 1
