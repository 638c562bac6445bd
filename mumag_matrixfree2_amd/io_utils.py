"""Result files of a hysteresis sweep in the reference's formats.

* ``hysteresis.csv``: header ``config,B_ext_T,J_par_T[,E,gnorm]`` and one row per field step,
  floats as ``%.9e`` (``src/io_utils.py:31-75``);
* ``<name>.mh``: the mammos-mumag magnetisation history, a ``#`` header line and columns
  B_ext [T], J_parallel [T], mx, my, mz, E [J/m^3] as ``%.9e`` (``src/io_utils.py:78-95``);
* ``.vtu``: an ASCII VTK UnstructuredGrid of tetrahedra with point and cell data
  (``src/io_utils.py:139-268``), e.g. the magnetisation and the potential of a state.

Plain host code: the values come from the GPU state (``MeshContext.state_get``,
``state_j_parallel``) and are written here.
"""

from __future__ import annotations

from pathlib import Path
from typing import Mapping, Sequence

import numpy as np


def write_hysteresis_header(csv_path, extra_cols: Sequence[str] = ("E", "gnorm")) -> None:
    cols = ["config", "B_ext_T", "J_par_T", *extra_cols]
    Path(csv_path).write_text(",".join(cols) + "\n", encoding="utf-8")


def append_hysteresis_row(csv_path, config_idx: int, B_T: float, J_par_T: float, E: float | None = None,
                          gnorm: float | None = None) -> None:
    vals = [str(int(config_idx)), f"{B_T:.9e}", f"{J_par_T:.9e}"]
    vals += [f"{v:.9e}" for v in (E, gnorm) if v is not None]
    with Path(csv_path).open("a", encoding="utf-8") as f:
        f.write(",".join(vals) + "\n")


def write_mh(path, records) -> Path:
    """``records`` (n, 6): B_ext [T], J_parallel [T], mx, my, mz, E [J/m^3]; the suffix is set to .mh."""
    rec = np.asarray(records, dtype=np.float64).reshape(-1, 6)
    out = Path(path).with_suffix(".mh")
    with out.open("w", encoding="utf-8") as f:
        f.write("# B_ext [T]  J_parallel [T]  mx  my  mz  E [J/m^3] \n")
        np.savetxt(f, rec, fmt="%.9e")
    return out


def read_hysteresis_csv(csv_path) -> tuple[list[str], np.ndarray]:
    lines = Path(csv_path).read_text(encoding="utf-8").splitlines()
    header = lines[0].split(",")
    data = np.array([[float(x) for x in ln.split(",")] for ln in lines[1:] if ln.strip()])
    return header, data.reshape(-1, len(header))


_VTK_TETRA = 10


def _vtk_type(a: np.ndarray) -> str:
    if np.issubdtype(a.dtype, np.integer):
        return "Int32" if a.dtype.itemsize <= 4 else "Int64"
    return "Float64"


def _data_array(name: str, a: np.ndarray, indent: str) -> list[str]:
    a = np.asarray(a)
    ncomp = 1 if a.ndim == 1 else a.shape[1]
    fmt = "%d" if np.issubdtype(a.dtype, np.integer) else "%.9e"
    body = " ".join(fmt % v for v in a.reshape(-1))
    return [f'{indent}<DataArray type="{_vtk_type(a)}" Name="{name}" NumberOfComponents="{ncomp}" format="ascii">',
            f"{indent}  {body}", f"{indent}</DataArray>"]


def write_vtu_tetra(path, points, tets, point_data: Mapping[str, np.ndarray] | None = None,
                    cell_data: Mapping[str, np.ndarray] | None = None) -> Path:
    """ASCII .vtu with tetrahedral cells (VTK type 10), 0-based connectivity."""
    pts = np.asarray(points, dtype=np.float64).reshape(-1, 3)
    t4 = np.asarray(tets)[:, :4].astype(np.int64)
    n, e = pts.shape[0], t4.shape[0]
    L = ['<?xml version="1.0"?>',
         '<VTKFile type="UnstructuredGrid" version="0.1" byte_order="LittleEndian">',
         "  <UnstructuredGrid>",
         f'    <Piece NumberOfPoints="{n}" NumberOfCells="{e}">']
    if point_data:
        L.append("      <PointData>")
        for k, v in point_data.items():
            L += _data_array(k, np.asarray(v), "        ")
        L.append("      </PointData>")
    if cell_data:
        L.append("      <CellData>")
        for k, v in cell_data.items():
            L += _data_array(k, np.asarray(v), "        ")
        L.append("      </CellData>")
    L.append("      <Points>")
    L += _data_array("Points", pts, "        ")
    L.append("      </Points>")
    L.append("      <Cells>")
    L += _data_array("connectivity", t4.reshape(-1), "        ")
    L += _data_array("offsets", 4 * np.arange(1, e + 1, dtype=np.int64), "        ")
    L += _data_array("types", np.full(e, _VTK_TETRA, dtype=np.int32), "        ")
    L.append("      </Cells>")
    L += ["    </Piece>", "  </UnstructuredGrid>", "</VTKFile>"]
    out = Path(path)
    out.write_text("\n".join(L) + "\n", encoding="utf-8")
    return out
