"""Hysteresis sweep driver (``src/hysteresis_loop.py``) on the HIP backend.

One device context per mesh; the magnetisation and potential stay resident on the GPU
across field steps (U is carried from step to step as the reference does,
``src/hysteresis_loop.py:383,504-507``); the host reads back per step only the statistics
and the volume-averaged J along h.
"""

from __future__ import annotations

import time
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from .context import MeshContext, make_params


@dataclass
class LoopParams:
    """Sweep and solver settings; field names and defaults of ``src/hysteresis_loop.py:34-125``
    with the CLI defaults of ``src/loop.py`` where those override them (pc_iters 10)."""

    h_dir: np.ndarray
    B_start: float
    B_end: float
    dB: float
    loop: bool = False
    max_iter: int = 2000
    tau_f: float = 1e-8
    eps_a: float = 1e-12
    tau0: float = 1e-2
    ls_eta1: float = 0.1
    ls_eta2: float = 0.1
    ls_C: float = 2.0
    ls_c: float = 0.5
    ls_max_evals: int = 15
    Js_ref: float = 1.0
    cg_maxiter: int = 2000
    cg_tol: float = 1e-8
    poisson_reg: float = 1e-12
    mfinal: float | None = None
    method: str = "pcohen_hs"
    pc_iters: int = 10
    pc_auto: bool = True
    pc_force_eta: float = 0.5
    pc_force_alpha: float = 0.5
    pc_stagnation_nu: float = 0.01
    memory: int = 10
    pc_reg: float = 0.0
    L: int | None = None
    precond_type: str = "amgcl"
    verbose: bool = False
    extra: dict = field(default_factory=dict)


def field_values(H_start: float, H_end: float, dH: float, loop: bool) -> np.ndarray:
    """Field magnitudes of the sweep (``src/hysteresis_loop.py:128-150``)."""
    if dH == 0:
        raise ValueError("dH must be nonzero")
    if (H_end - H_start) * dH < 0:
        dH = -dH
    n = int(np.floor((H_end - H_start) / dH + 1e-12))
    vals = H_start + dH * np.arange(n + 1)
    if abs(vals[-1] - H_end) > 1e-12 * max(1.0, abs(H_end)):
        vals = np.append(vals, H_end)
    if not loop:
        return vals
    return np.concatenate([vals, vals[-2::-1]])


def switching_step(J_par: np.ndarray) -> int:
    """First field step with J_par < 0 (sweeping down from positive saturation); -1 if none."""
    idx = np.where(np.asarray(J_par) < 0)[0]
    return int(idx[0]) if idx.size else -1


def volume_averaged_m(m: np.ndarray, conn: np.ndarray, Ve: np.ndarray, magnetic: np.ndarray) -> np.ndarray:
    """Volume average of the element-mean magnetisation over magnetic elements (Js > 0)
    (``src/hysteresis_loop.py:193-222``)."""
    me = np.asarray(m)[np.asarray(conn)].mean(axis=1)
    w = np.where(magnetic, Ve, 0.0)
    return (w[:, None] * me).sum(axis=0) / (w.sum() + 1e-30)


def _params_table(params: LoopParams) -> dict:
    """The LoopParams fields as the rows of params.log (h_dir as a list, ``extra`` merged in)."""
    from dataclasses import fields

    d = {}
    for f in fields(params):
        v = getattr(params, f.name)
        if f.name == "extra":
            continue
        d[f.name] = np.asarray(v).tolist() if isinstance(v, np.ndarray) else v
    d.update(params.extra)
    return d


def _params_sources(params: LoopParams) -> dict:
    """'default' where a field holds its LoopParams default, 'caller' where it was set."""
    from dataclasses import MISSING, fields

    src = {}
    for f in fields(params):
        if f.name == "extra":
            continue
        default = f.default if f.default is not MISSING else MISSING
        v = getattr(params, f.name)
        same = default is not MISSING and not isinstance(v, np.ndarray) and v == default
        src[f.name] = "default" if same else "caller"
    src.update({k: "caller" for k in params.extra})
    return src


def run_hysteresis_loop(ctx: MeshContext, m0: np.ndarray, params: LoopParams, *, steps: int | None = None,
                        on_step=None, out_dir=None, name: str = "hysteresis", snapshot_every: int = 0) -> dict:
    """Sweep the field along ``params.h_dir`` and minimise at each value.

    Returns a dict with the ``switching_step``, the ``nucleation`` field and index
    (``io_utils.nucleation_field``), ``history`` rows (B_ext [Js_ref units], J_par [Js_ref units], E, gnorm,
    iters, evals, demag_iters, preco_iters, seconds) and the final ``m``, ``U``.

    With ``out_dir``, the result files of ``src/hysteresis_loop.py:455-657`` are written there:
    ``params.log`` (``src/loop.py:1046-1058``), ``hysteresis.csv`` (one row per step), ``<name>.mh`` (B [T], J_par [T], mx, my, mz,
    E [J/m^3]) and, every ``snapshot_every`` steps, a ``state_cfg*.vtu`` with m and U. The
    host then downloads m once per step for the volume-averaged (mx, my, mz).

    A segmented context (a grain set, ``grainset.grain_set_context``) sweeps every grain at
    once (``minimize_segmented``, pcohen_hs): J_par and E of a row are the magnetic-volume
    averages over the grains, iterations and evaluations the largest per grain, and the result
    also holds ``grain_J`` (steps x grains).
    """
    seg = ctx.segments is not None
    if seg and params.method != "pcohen_hs":
        raise ValueError("a segmented context sweeps with pcohen_hs")
    vol = ctx.segment_volumes() if seg else None
    grain_J = []
    h = np.asarray(params.h_dir, dtype=np.float64)
    h = h / (np.linalg.norm(h) + 1e-30)
    if ctx.precond_type != params.precond_type:
        ctx.setup_precond(params.precond_type)
    p = make_params(params.method, max_iter=params.max_iter, tau_f=params.tau_f, eps_a=params.eps_a,
                    tau0=params.tau0, ls_eta1=params.ls_eta1, ls_eta2=params.ls_eta2, ls_C=params.ls_C,
                    ls_c=params.ls_c, ls_max_evals=params.ls_max_evals, pc_iters=params.pc_iters,
                    pc_auto=params.pc_auto, pc_force_eta=params.pc_force_eta,
                    pc_force_alpha=params.pc_force_alpha, pc_stagnation_nu=params.pc_stagnation_nu,
                    pc_reg=params.pc_reg, L=params.L, cg_tol=params.cg_tol, cg_maxiter=params.cg_maxiter,
                    memory=params.memory)
    m = np.asarray(m0, dtype=np.float64)
    m = m / np.linalg.norm(m, axis=1, keepdims=True)
    ctx.state_set(m, np.zeros(ctx.N))
    B_vals = field_values(params.B_start, params.B_end, params.dB, params.loop)
    if steps is not None:
        B_vals = B_vals[:steps]
    hist, mh = [], []
    if out_dir is not None:
        from . import io_utils

        out = Path(out_dir)
        out.mkdir(parents=True, exist_ok=True)
        io_utils.write_params_log(out, _params_table(params), _params_sources(params))
        csv_path = out / "hysteresis.csv"
        io_utils.write_hysteresis_header(csv_path)
        conn, Ve, mag = ctx.element_table()
    t_all = time.perf_counter()
    for k, Bmag in enumerate(B_vals):
        t0 = time.perf_counter()
        if seg:
            infos = ctx.minimize_segmented(p, Bmag * h)
            Jg = ctx.state_j_parallel_segmented(h)
            grain_J.append(Jg)
            J = float(vol @ Jg / vol.sum())
            info = {"E": float(vol @ np.array([i["E"] for i in infos]) / vol.sum()),
                    "gnorm": max(i["gnorm"] for i in infos)}
            for key in ("iters", "evals", "demag_iters", "preco_iters"):
                info[key] = max(i[key] for i in infos)
        else:
            info = ctx.minimize_resident(p, Bmag * h)
            J = ctx.state_j_parallel(h)
        dt = time.perf_counter() - t0
        row = [Bmag, J, info["E"], info["gnorm"], info["iters"], info["evals"], info["demag_iters"],
               info["preco_iters"], dt]
        hist.append(row)
        if out_dir is not None:
            B_T, J_T = Bmag * params.Js_ref, J * params.Js_ref
            m_k, U_k = ctx.state_get()
            mh.append([B_T, J_T, *volume_averaged_m(m_k, conn, Ve, mag),
                       info["E"] * params.Js_ref**2 / (2.0 * 4e-7 * np.pi)])
            io_utils.append_hysteresis_row(csv_path, k, B_T, J_T, info["E"], info["gnorm"])
            if snapshot_every > 0 and k % snapshot_every == 0:
                io_utils.write_vtu_tetra(out / f"state_cfg{k:05d}_B{B_T:+.4e}T.vtu", ctx.knt, conn,
                                         point_data={"m": m_k.astype(np.float32), "U": U_k.astype(np.float32)})
        if params.verbose:
            print(f"step {k:05d} B={Bmag * params.Js_ref:+.6e} T J_par={J * params.Js_ref:+.6e} T "
                  f"E={info['E']:.6e} t={dt:.3f}s it={info['iters']} nf={info['evals']} "
                  f"icg={info['demag_iters']}", flush=True)
        if on_step is not None:
            on_step(k, row)
        if params.mfinal is not None and J <= params.mfinal:
            break
    m, U = ctx.state_get()
    hist = np.asarray(hist)
    if out_dir is not None and mh:
        io_utils.write_mh(Path(out_dir) / name, np.asarray(mh))
    out = {"history": hist, "m": m, "U": U, "seconds": time.perf_counter() - t_all,
           "switching_step": switching_step(hist[:, 1]) if len(hist) else -1}
    if len(hist) > 1:
        from .io_utils import nucleation_field

        out["nucleation"] = nucleation_field(hist[:, 0] * params.Js_ref, hist[:, 1] * params.Js_ref)
    if seg:
        out["grain_J"] = np.asarray(grain_J)
    return out
