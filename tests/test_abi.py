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
