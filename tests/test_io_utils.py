"""Result-file formats of the sweep (hysteresis.csv, .mh, .vtu), checked on the host against
the layouts of src/io_utils.py:31-268."""

from __future__ import annotations

import xml.etree.ElementTree as ET

import numpy as np

from mumag_matrixfree2_amd import io_utils
from mumag_matrixfree2_amd.hysteresis_loop import volume_averaged_m


def test_hysteresis_csv_roundtrip(tmp_path):
    f = tmp_path / "hysteresis.csv"
    io_utils.write_hysteresis_header(f)
    io_utils.append_hysteresis_row(f, 0, 2.0, 1.6094, E=-6.34, gnorm=1e-9)
    io_utils.append_hysteresis_row(f, 1, 1.75, -1.2, E=0.5, gnorm=2e-9)
    lines = f.read_text().splitlines()
    assert lines[0] == "config,B_ext_T,J_par_T,E,gnorm"
    assert lines[1] == "0,2.000000000e+00,1.609400000e+00,-6.340000000e+00,1.000000000e-09"
    header, data = io_utils.read_hysteresis_csv(f)
    assert header == ["config", "B_ext_T", "J_par_T", "E", "gnorm"]
    np.testing.assert_allclose(data[:, 1], [2.0, 1.75])


def test_mh_layout(tmp_path):
    rec = np.array([[2.0, 1.6, 0.0, 0.0, 1.0, -1e6], [1.0, 1.5, 0.1, 0.0, 0.99, -5e5]])
    out = io_utils.write_mh(tmp_path / "cube", rec)
    assert out.suffix == ".mh"
    text = out.read_text().splitlines()
    assert text[0] == "# B_ext [T]  J_parallel [T]  mx  my  mz  E [J/m^3] "
    np.testing.assert_array_equal(np.loadtxt(out), rec)


def test_vtu_is_wellformed(tmp_path):
    pts = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 1]], dtype=float)
    tets = np.array([[0, 1, 2, 3], [1, 2, 3, 4]])
    m = np.tile([0.0, 0.0, 1.0], (5, 1))
    out = io_utils.write_vtu_tetra(tmp_path / "s.vtu", pts, tets, point_data={"m": m}, cell_data={"mat": np.array([1, 2])})
    root = ET.parse(out).getroot()
    piece = root.find("UnstructuredGrid/Piece")
    assert piece.get("NumberOfPoints") == "5" and piece.get("NumberOfCells") == "2"
    arrays = {a.get("Name"): a for a in root.iter("DataArray")}
    assert arrays["m"].get("NumberOfComponents") == "3"
    assert [int(v) for v in arrays["types"].text.split()] == [10, 10]
    assert [int(v) for v in arrays["offsets"].text.split()] == [4, 8]


def test_volume_averaged_m_weights_magnetic_volume():
    m = np.array([[0, 0, 1.0], [0, 0, 1.0], [0, 0, 1.0], [0, 0, 1.0], [1.0, 0, 0]])
    conn = np.array([[0, 1, 2, 3], [1, 2, 3, 4]])
    Ve = np.array([1.0, 3.0])
    np.testing.assert_allclose(volume_averaged_m(m, conn, Ve, np.array([True, False])), [0, 0, 1])
    np.testing.assert_allclose(volume_averaged_m(m, conn, Ve, np.array([True, True])), [0.1875, 0, 0.8125])


def test_params_log_layout(tmp_path):
    """params.log: the reference's Markdown table (src/loop.py:1046-1058), sorted keys, no
    '*extrapolate*' rows, phi_tol = min(cg_tol, 0.1 tau_f) derived right after cg_tol."""
    out = io_utils.write_params_log(tmp_path, {"tau_f": 1e-8, "cg_tol": 1e-10, "method": "pcohen_hs",
                                               "phi_extrapolate": True, "B_start": 1.24},
                                    sources={"method": "cli"})
    lines = out.read_text().splitlines()
    assert out.name == "params.log"
    assert lines[:2] == ["| Parameter | Value | Source |", "| :--- | :--- | :--- |"]
    assert lines[2:] == ["| B_start | 1.24 | default |", "| cg_tol | 1e-10 | default |",
                         "| phi_tol | 1e-10 | derived |", "| method | pcohen_hs | cli |", "| tau_f | 1e-08 | default |"]
    out = io_utils.write_params_log(tmp_path, {"tau_f": 1e-10, "cg_tol": 1e-8})
    # the product 0.1 tau_f is printed as Python prints it (1.0000000000000001e-11), as the reference does
    assert f"| phi_tol | {1e-10 * 0.1} | derived |" in out.read_text().splitlines()


def test_nucleation_field_on_a_known_loop():
    """A loop that holds J = 1.6 down to B = 0.5 T and falls linearly (slope 3.2) to -1.6 at
    B = -0.5 T, sampled every 0.25 T from +2 T. By hand: the central difference at B = 0.5 is
    (0.8 - 1.6) / (-0.5) = 1.6 > 1% of the maximum 3.2, while every point above is flat, so the
    nucleation index is 6 (B = 0.5 T); |d2J/dB2| is 6.4 at B = +0.5 and -0.5 and the first of the
    tie (index 6) is reported."""
    B = np.arange(2.0, -2.01, -0.25)
    J = np.clip(1.6 * B / 0.5, -1.6, 1.6)
    nuc = io_utils.nucleation_field(B, J)
    assert nuc == {"index": 6, "B_nuc": 0.5, "index_2nd": 6, "B_nuc_2nd": 0.5}
    assert io_utils.nucleation_field(B, np.full_like(B, 1.6)) is None
