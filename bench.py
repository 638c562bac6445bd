"""Hysteresis benchmark: field-steps/s of the FP64 GPU path on the 2.06M-tet cube workload.

Workload (the plain FP64 GPU configuration): 80 nm NdFeB-like cube, 60^3 hex cells x 6 tets
(1,296,000 magnet tets, 226,981 magnet nodes) inside a 5-layer graded airbox (2,058,000 tets,
357,911 nodes), Morton-ordered; Js = 1.61 T, A = 7.7 pJ/m, K1 = 4.3 MJ/m^3, easy axis z;
m0 = z + N(0, 0.01) (seed 0), renormalised; field +2 T -> -8 T in 41 steps along
(sin 1 deg, 0, cos 1 deg); pcohen_hs, tau_f = 1e-10, PCG tol 1e-10 (line-search solves at
min(cg_tol, 0.1 tau_f)); smoothed-aggregation AMG V-cycle with one pre- and one
post-smoothing sweep.

A step is one field step: a full energy minimisation at one field value, with the
magnetisation and potential resident on the GPU. ``--warmup W`` field steps run first on
a separate state (untimed); then ``--steps K`` field steps of the sweep from +2 T are
timed. One JSON line is printed on stdout.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

MU0 = 4e-7 * np.pi
JS, A_SI, K1_SI = 1.61, 7.7e-12, 4.3e6
H_DIR = np.array([np.sin(np.deg2rad(1.0)), 0.0, np.cos(np.deg2rad(1.0))])
HBM_PEAK_GBS = 8000.0


def materials():
    Kd = JS**2 / (2 * MU0)
    A = np.array([A_SI * 1e18 / Kd, 0.0])
    K1 = np.array([K1_SI / Kd, 0.0])
    Js = np.array([1.0, 0.0])
    k = np.array([[0.0, 0.0, 1.0], [0.0, 0.0, 1.0]])
    return A, K1, Js, k


def initial_m(N, seed=0, sigma=0.01):
    rng = np.random.default_rng(seed)
    m = np.zeros((N, 3))
    m[:, 2] = 1.0
    m += sigma * rng.standard_normal((N, 3))
    return m / np.linalg.norm(m, axis=1, keepdims=True)


def apply_algorithmic_bytes(ctx) -> int:
    """Bytes one on-the-fly Poisson apply must move if every stream is touched once:
    packed corner indices (4 B/tet), node lists (4 B/entry), node coordinates (24 B/node),
    x (8 B/node) and y (8 B/node)."""
    return 4 * ctx.E + 4 * ctx.node_list_entries + 40 * ctx.N


def reference_layout_bytes(ctx) -> int:
    """Bytes of the same apply in the reference's stored-J^-T layout: per tet 4 x int32 corners,
    9 J^-T entries and Ve in FP64; per node x read and y written."""
    return ctx.E * (16 + 10 * 8) + ctx.N * 2 * 8


def measured_traffic():
    """HBM bytes per k_apply launch from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary
    (profiles/k_apply_cube2M_fp64.md), or None if absent."""
    f = ROOT / "profiles" / "k_apply_cube2M_fp64.md"
    if not f.exists():
        return None
    for line in f.read_text().splitlines():
        if line.startswith("traffic_bytes_per_launch:"):
            return int(line.split(":")[1])
    return None


def cpu_baseline(mesh, mats, steps_T, budget_s=60.0):
    """The oracle (NumPy/SciPy restatement, kind 'port') on the same mesh, hierarchy and method,
    timed on the host for the first field steps of the sweep."""
    sys.path.insert(0, str(ROOT / "tests"))
    from mumag_matrixfree2_amd.amg import build_hierarchy
    from oracle.fem import Problem, spai0_diagonal
    from oracle.minimize import Minimizer
    from oracle.solvers import solve_U

    A, K1, Js, k = mats
    p = Problem(mesh["knt"], mesh["ijk"], A, K1, Js, k, boundary_mask=mesh["boundary_mask"])
    lv = build_hierarchy(p.A, "amgcl")
    levels = []
    for i, l in enumerate(lv):
        d = {"A": l["A"], "M": spai0_diagonal(l["A"])}
        if "P" in l:
            d["P"], d["R"] = l["P"], l["R"]
        if i == len(lv) - 1:
            d["A_dense"] = l["A"].toarray()
        levels.append(d)

    def sU(m, x0, tol):
        return solve_U(p, m, x0, tol, precond="amg", levels=levels, kind="amgcl")

    mn = Minimizer(p, sU, method="pcohen_hs", cg_tol=1e-10)
    m = initial_m(p.N)
    U = np.zeros(p.N)
    done = 0
    t0 = time.perf_counter()
    for BT in steps_T:
        m, U, info = mn.minimize(m, BT / JS * H_DIR, U0=U, tau_f=1e-10)
        done += 1
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    return done, dt


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=41)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precond", default="amgcl", choices=["amgcl", "amg"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--max-coarse", type=int, default=None)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the hysteresis path runs on a single MI355X here "
              "(element-block partitions would shard across GPUs with a halo exchange per apply); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    from mumag_matrixfree2_amd.context import MeshContext, make_params
    from mumag_matrixfree2_amd.hysteresis_loop import field_values, switching_step
    from mumag_matrixfree2_amd.mesh import config_cube_mesh

    log = lambda *a: print(*a, file=sys.stderr, flush=True)  # noqa: E731
    t0 = time.perf_counter()
    mesh = config_cube_mesh(80.0, 60, 5)
    mats = materials()
    ctx = MeshContext(mesh["knt"], mesh["ijk"][:, :4], mesh["ijk"][:, 4], *mats,
                      boundary_mask=mesh["boundary_mask"])
    ctx.setup_precond(args.precond, max_coarse=args.max_coarse)
    ctx.set_precision(args.precision)
    log(f"[bench] set-up {time.perf_counter() - t0:.1f}s: N={ctx.N} E={ctx.E} "
        f"AMG levels={[l['A'].shape[0] for l in ctx.amg_levels]}")

    p = make_params("pcohen_hs", tau_f=1e-10, eps_a=1e-12, cg_tol=1e-10, max_iter=2000)
    B_vals = field_values(2.0 / JS, -8.0 / JS, -0.25 / JS, False)  # 41 values
    m0 = initial_m(ctx.N)
    steps = min(args.steps, len(B_vals))

    def sweep(precision, timed):
        ctx.set_precision(precision)
        ctx.state_set(m0, np.zeros(ctx.N))
        if timed:
            ctx.apply_timing(True)
        J, E, rows = [], [], []
        t_start = time.perf_counter()
        for k in range(steps):
            ts = time.perf_counter()
            info = ctx.minimize_resident(p, B_vals[k] * H_DIR)
            J.append(ctx.state_j_parallel(H_DIR))
            E.append(info["E"])
            rows.append(time.perf_counter() - ts)
            log(f"[bench:{precision}] step {k:2d} B={B_vals[k] * JS:+.3f} T J={J[-1] * JS:+.5f} T "
                f"E={info['E']:+.10e} it={info['iters']} nf={info['evals']} icg={info['demag_iters']} "
                f"t={rows[-1]:.3f}s")
        return time.perf_counter() - t_start, np.array(J), np.array(E), rows

    # warm-up on a separate state
    ctx.set_precision(args.precision)
    ctx.state_set(m0, np.zeros(ctx.N))
    for k in range(args.warmup):
        ctx.minimize_resident(p, B_vals[k % len(B_vals)] * H_DIR)

    c0 = ctx.counters()
    elapsed, J, Es, rows = sweep(args.precision, timed=True)
    apply_ms, apply_n = ctx.apply_timing(False)
    c1 = ctx.counters()
    pcg_its = c1["pcg_iters"] - c0["pcg_iters"]
    errors = None
    if args.precision == "mixed":
        # reference FP64 sweep in the same run: per-step relative energy error and switching step
        t64, J64, E64, _ = sweep("fp64", timed=False)
        errors = {
            "max_rel_energy_err_vs_fp64": float(np.max(np.abs(Es - E64) / np.maximum(np.abs(E64), 1.0))),
            "switching_step_fp64": switching_step(J64),
            "fp64_field_steps_per_s_same_run": steps / t64,
            "speedup_vs_fp64": t64 / elapsed,
        }

    avg_apply_s = apply_ms / max(apply_n, 1) * 1e-3
    algo = apply_algorithmic_bytes(ctx)
    achieved = algo / avg_apply_s / 1e9
    label = "FP64" if args.precision == "fp64" else "mixed FP32 V-cycle / FP64 PCG"
    result = {
        "metric": f"hysteresis field-steps/s ({label}, 2.06M-tet cube + airbox)",
        "value": steps / elapsed,
        "unit": "field-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f64+f32",
        "data": "synthetic (generated mesh, seeded initial magnetisation)",
        "config": {
            "workload": f"cube80nm_60^3x6tets+5layer_airbox_2.06Mtet_{args.precision}_sweep+2..-8T_41steps",
            "tets": ctx.E, "nodes": ctx.N, "method": "pcohen_hs", "precond": f"SA-AMG V-cycle ({args.precond})",
            "amg_levels": [int(l["A"].shape[0]) for l in ctx.amg_levels],
            "pcg_tol": 1e-10, "tau_f": 1e-10,
        },
        "pcg_iters_per_s": pcg_its / elapsed,
        "switching_step": switching_step(J),
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(),
            "kernel": "k_apply<on_the_fly> (matrix-free Poisson apply)",
            "avg_launch_us": avg_apply_s * 1e6,
            "launches": apply_n,
            "algorithmic_bytes_per_launch": algo,
            "reference_layout_GBps": reference_layout_bytes(ctx) / avg_apply_s / 1e9,
        },
    }

    if errors is not None:
        result["mixed_vs_fp64"] = errors
    if not args.no_cpu_baseline:
        cpu_T = [2.0 - 0.25 * k for k in range(args.cpu_steps)]
        done, dt = cpu_baseline(mesh, mats, cpu_T)
        gpu_same = sum(rows[:done])
        result["cpu_baseline"] = {
            "value": done / dt, "unit": "field-steps/s", "cores": 1, "kind": "port",
            "sample": f"oracle (NumPy/SciPy, single thread) pcohen_hs on the same mesh and AMG hierarchy, "
                      f"first {done} field steps (+2.00 T ...), {dt:.1f}s; GPU on the same steps: "
                      f"{done / max(gpu_same, 1e-12):.2f} field-steps/s",
        }
    print(json.dumps(result), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
