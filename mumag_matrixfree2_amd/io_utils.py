"""Result files of a hysteresis sweep in the reference's formats.

* ``hysteresis.csv``: header ``config,B_ext_T,J_par_T[,E,gnorm]`` and one row per field step,
  floats as ``%.9e`` (``src/io_utils.py:31-75``);
* ``<name>.mh``: the mammos-mumag magnetisation history, a ``#`` header line and columns
  B_ext [T], J_parallel [T], mx, my, mz, E [J/m^3] as ``%.9e`` (``src/io_utils.py:78-95``);
* ``.vtu``: an ASCII VTK UnstructuredGrid of tetrahedra with point and cell data
  (``src/io_utils.py:139-268``), e.g. the magnetisation and the potential of a state;
* ``params.log``: the run's parameters as a Markdown table ``| Parameter | Value | Source |``,
  keys sorted, with the derived ``phi_tol`` row after ``cg_tol`` (``src/loop.py:1046-1058``).

``nucleation_field`` reads the nucleation field off a sweep as ``src/extract_nucleation.py:8-44``
does (first field where dJ/dB exceeds 1% of its maximum, and the field of the largest
|d2J/dB2|).

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


def write_params_log(out_dir, params: Mapping[str, object], sources: Mapping[str, str] | None = None) -> Path:
    """``params.log`` in ``out_dir``: one row per parameter, sorted by name, keys containing
    'extrapolate' skipped, and after ``cg_tol`` the derived ``phi_tol = min(cg_tol, 0.1 tau_f)``
    (``src/loop.py:1046-1058``). ``sources`` maps a key to its source ('default' otherwise)."""
    sources = sources or {}
    out = Path(out_dir)
    out.mkdir(parents=True, exist_ok=True)
    lines = ["| Parameter | Value | Source |", "| :--- | :--- | :--- |"]
    for k in sorted(params):
        if "extrapolate" in k:
            continue
        lines.append(f"| {k} | {params[k]} | {sources.get(k, 'default')} |")
        if k == "cg_tol":
            phi_tol = float(min(float(params["cg_tol"]), float(params["tau_f"]) * 0.1))
            lines.append(f"| phi_tol | {phi_tol} | derived |")
    path = out / "params.log"
    path.write_text("\n".join(lines) + "\n", encoding="utf-8")
    return path


def nucleation_field(B_T, J_T) -> dict | None:
    """Nucleation field of a sweep (``src/extract_nucleation.py:8-44``): dJ/dB by
    ``np.gradient`` over the field values; the first index where it exceeds 1% of its maximum
    (``index``, ``B_nuc``) and the index of the largest |d2J/dB2| (``index_2nd``, ``B_nuc_2nd``).
    None when the largest dJ/dB is below 1e-3 (no nucleation in the range)."""
    B = np.asarray(B_T, dtype=np.float64)
    J = np.asarray(J_T, dtype=np.float64)
    if B.size < 2:
        return None
    dJ = np.gradient(J, B)
    d2J = np.gradient(dJ, B)
    mx = float(np.max(dJ))
    if mx < 1e-3:
        return None
    idx = np.where(dJ > 0.01 * mx)[0]
    if idx.size == 0:
        return None
    i2 = int(np.argmax(np.abs(d2J)))
    return {"index": int(idx[0]), "B_nuc": float(B[idx[0]]), "index_2nd": i2, "B_nuc_2nd": float(B[i2])}
