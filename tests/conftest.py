"""Shared fixtures: cube magnet meshes with a graded airbox, NdFeB-like materials, oracle problems."""

from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

MU0 = 4e-7 * np.pi


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def ndfeb_materials(Js=1.61, K1=4.3e6, A_si=7.7e-12, axis=(0.0, 0.0, 1.0), air=True):
    """Reduced lookups for magnet (+ air): the normalisation of src/loop.py:862-877."""
    Kd = Js**2 / (2 * MU0)
    A_red = A_si * 1e18 / Kd
    K1_red = K1 / Kd
    k = np.asarray(axis, dtype=np.float64)
    k = k / np.linalg.norm(k)
    if air:
        return (np.array([A_red, 0.0]), np.array([K1_red, 0.0]), np.array([1.0, 0.0]), np.array([k, k]))
    return np.array([A_red]), np.array([K1_red]), np.array([1.0]), np.array([k])


def tilted_m0(N, seed=0, sigma=0.01):
    rng = np.random.default_rng(seed)
    m = np.zeros((N, 3))
    m[:, 2] = 1.0
    m += sigma * rng.standard_normal((N, 3))
    return m / np.linalg.norm(m, axis=1, keepdims=True)


@pytest.fixture(scope="session")
def cube_c0():
    """20 nm cube (10^3 hexes x 6 tets) + 3 graded air layers, Morton-ordered."""
    from mumag_matrixfree2_amd.mesh import config_cube_mesh

    return config_cube_mesh(20.0, 10, 3)


@pytest.fixture(scope="session")
def small_cube():
    """8 nm cube (4^3 hexes) + 2 air layers: small enough for dense checks."""
    from mumag_matrixfree2_amd.mesh import config_cube_mesh

    return config_cube_mesh(8.0, 4, 2)


@pytest.fixture(scope="session")
def oracle_c0(cube_c0):
    from oracle.fem import Problem

    A, K1, Js, k = ndfeb_materials()
    return Problem(cube_c0["knt"], cube_c0["ijk"], A, K1, Js, k, boundary_mask=cube_c0["boundary_mask"])


def oracle_levels(levels, kind):
    """Oracle V-cycle levels built from a hierarchy's matrices, with the oracle's own smoother
    diagonals: SPAI0 (``src/amg_utils.py:87-111``) for 'amgcl', diag(A) for 'amg'."""
    from oracle.fem import spai0_diagonal

    out = []
    for i, lv in enumerate(levels):
        A = lv["A"]
        d = {"A": A, "M": spai0_diagonal(A) if kind == "amgcl" else A.diagonal()}
        if "P" in lv:
            d["P"], d["R"] = lv["P"], lv["R"]
        if i == len(levels) - 1:
            d["A_dense"] = A.toarray()
        out.append(d)
    return out
