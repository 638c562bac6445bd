"""The C-ABI library loads and exports every entry point include/mumag_hip.h declares.

No compute calls: this runs without a GPU. The ctypes signature table of the Python
binding must cover the same set of names, so a declaration cannot drift from its binding.
"""

from __future__ import annotations

import ctypes as C
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "mumag_hip.h"


def declared_entry_points() -> set[str]:
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"^\s*int\s+(mm_\w+)\s*\(", text, flags=re.M))


def test_header_declares_entry_points():
    names = declared_entry_points()
    assert {"mm_ctx_create", "mm_poisson_apply", "mm_solve_U", "mm_energy_grad", "mm_minimize"} <= names


def test_library_exports_every_declared_symbol():
    from mumag_matrixfree2_amd import _hip

    if not _hip.LIB_PATH.exists():
        pytest.fail(f"{_hip.LIB_PATH} is not built (run __graft_entry__.build())")
    lib = C.CDLL(str(_hip.LIB_PATH))
    missing = [n for n in sorted(declared_entry_points()) if not hasattr(lib, n)]
    assert not missing, f"declared but not exported: {missing}"


def test_binding_table_matches_header():
    from mumag_matrixfree2_amd import _hip

    assert set(_hip.SIGNATURES) == declared_entry_points()


def test_missing_library_raises(monkeypatch):
    """Without the built library every entry point raises: there is no CPU fallback."""
    from mumag_matrixfree2_amd import _hip

    monkeypatch.setattr(_hip, "LIB_PATH", ROOT / "no_such_dir" / "libmumag_hip.so")
    monkeypatch.setattr(_hip, "_lib", None)
    with pytest.raises(RuntimeError, match="not built"):
        _hip.load()


def test_without_a_device_creation_is_an_error():
    """On a host without a GPU, creating a context reports the HIP error instead of computing
    anywhere else (run in a child process; on a GPU host the context simply builds)."""
    import subprocess
    import sys

    code = (
        "import sys; sys.path.insert(0, %r); sys.path.insert(0, %r)\n"
        "from conftest import ndfeb_materials\n"
        "from mumag_matrixfree2_amd.mesh import config_cube_mesh\n"
        "from mumag_matrixfree2_amd.context import MeshContext\n"
        "m = config_cube_mesh(8.0, 4, 2)\n"
        "A, K1, Js, k = ndfeb_materials()\n"
        "try:\n"
        "    MeshContext(m['knt'], m['ijk'][:, :4], m['ijk'][:, 4], A, K1, Js, k, boundary_mask=m['boundary_mask'])\n"
        "    print('BUILT')\n"
        "except RuntimeError as e:\n"
        "    print('ERROR', e)\n" % (str(ROOT), str(ROOT / "tests"))
    )
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    out = r.stdout.strip().splitlines()[-1]
    assert out == "BUILT" or (out.startswith("ERROR") and "mm_ctx_create" in out), out


@pytest.mark.parametrize("bad", [dict(memory=33), dict(memory=-1), dict(ls_max_evals=0), dict(tau_f=float("nan")),
                                 dict(cg_maxiter=0), dict(pc_iters=-1)])
def test_params_check_rejects_out_of_range(bad):
    """The C-ABI rejects parameter structs the device code cannot run (host-only check, no GPU):
    the L-BFGS history holds at most 32 pairs, limits are non-negative, tolerances finite."""
    from mumag_matrixfree2_amd.context import make_params

    with pytest.raises(RuntimeError, match="mm_params_check"):
        make_params("lbfgs", **bad)


def test_params_check_accepts_reference_defaults():
    from mumag_matrixfree2_amd.context import make_params

    for method in ("pcohen_hs", "lbfgs"):
        for mem in (1, 10, 32):
            assert make_params(method, memory=mem).memory == mem
