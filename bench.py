"""Hysteresis benchmark: field-steps/s of the GPU path, mixed precision against FP64.

Default workload (``--workload poly25M``, the configuration the metric is quoted on): 200 nm cube
of 27 Voronoi grains (centres from default_rng(4)), 150^3 hex cells x 6 tets (20.25M magnet tets)
in a 5-layer graded airbox (24,576,000 tets, 4,173,281 nodes), Morton-ordered; tets within 2 nm of
a grain-boundary bisector plane form a grain-boundary phase with A = 1.5 pJ/m, K1 = 0; easy axes
per grain uniform on the sphere (default_rng(40)); Js = 1.61 T, A = 7.7 pJ/m, K1 = 4.3 MJ/m^3;
m0 = z + N(0, 0.01) (seed 0), renormalised; 81 field values +2 T -> -8 T (dB = -0.125 T) along
(sin 1 deg, 0, cos 1 deg); pcohen_hs, tau_f = 1e-10, PCG tol 1e-10 (line-search solves at
min(cg_tol, 0.1 tau_f)); smoothed-aggregation AMG V-cycle with one pre- and one post-smoothing
sweep. The headline form is mixed precision (``--precision mixed``); the FP64 form runs the same
field steps in the same process and the line carries the comparison (``mixed_vs_fp64``).

``--workload cube2M``: the plain FP64 GPU configuration, 80 nm cube, 60^3 hex cells x 6 tets +
5-layer graded airbox (2,058,000 tets, 357,911 nodes), 41 field values (dB = -0.25 T).

``--workload grains64``: 64 independent grains (``mesh.grain_set``, default_rng(3)): total tets per
grain log-uniform in [2e4, 4e5] (7,752,942 tets, 1,431,475 nodes in all), cube or sphere by coin
flip, easy axis uniform on the sphere, each in its own 3-layer graded airbox with 2 nm cells; 41
field values; one field step minimises every grain: by default all grains in one segmented
context; ``--grains-form threads``: one context per grain, host threads driving them.

A step is one field step: a full energy minimisation at one field value, with the
magnetisation and potential resident on the GPU. ``--warmup W`` field steps run first on
a separate state (untimed); then the first ``--steps K`` field steps of the sweep from +2 T are
timed (the line's ``config.workload`` names K and the field range). One JSON line is printed on
stdout.
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


A_GB_SI = 1.5e-12
N_GRAINS = 27


def materials():
    Kd = JS**2 / (2 * MU0)
    A = np.array([A_SI * 1e18 / Kd, 0.0])
    K1 = np.array([K1_SI / Kd, 0.0])
    Js = np.array([1.0, 0.0])
    k = np.array([[0.0, 0.0, 1.0], [0.0, 0.0, 1.0]])
    return A, K1, Js, k


def poly_materials(n_grains=N_GRAINS):
    """Material ids 1..n_grains grains, n_grains+1 grain boundary, n_grains+2 air (table index id-1)."""
    from mumag_matrixfree2_amd.mesh import random_easy_axes

    Kd = JS**2 / (2 * MU0)
    G = n_grains + 2
    A = np.zeros(G)
    K1 = np.zeros(G)
    Js = np.zeros(G)
    k = np.zeros((G, 3))
    k[:, 2] = 1.0
    A[:n_grains], K1[:n_grains], Js[:n_grains] = A_SI * 1e18 / Kd, K1_SI / Kd, 1.0
    k[:n_grains] = random_easy_axes(n_grains, 40)
    A[n_grains], Js[n_grains] = A_GB_SI * 1e18 / Kd, 1.0
    return A, K1, Js, k


WORKLOADS = {
    "cube2M": {"name": "cube80nm_60^3x6tets+5layer_airbox_2.06Mtet", "steps": 41, "dB": -0.25},
    "grains64": {"name": "64grains_2e4-4e5tets_cube|sphere+3layer_airbox_7.75Mtet", "steps": 41, "dB": -0.25},
    "poly25M": {"name": "poly200nm_27grains_150^3x6tets+5layer_airbox_24.6Mtet", "steps": 81, "dB": -0.125},
}


def make_workload(name):
    from mumag_matrixfree2_amd.mesh import config_cube_mesh, polycrystal_cube

    if name == "cube2M":
        return config_cube_mesh(80.0, 60, 5), materials()
    return polycrystal_cube(200.0, 150, 5, n_grains=N_GRAINS, gb_width=2.0, seed=4), poly_materials()


def initial_m(N, seed=0, sigma=0.01):
    rng = np.random.default_rng(seed)
    m = np.zeros((N, 3))
    m[:, 2] = 1.0
    m += sigma * rng.standard_normal((N, 3))
    return m / np.linalg.norm(m, axis=1, keepdims=True)


def apply_algorithmic_bytes(ctx, fp32: bool = False) -> int:
    """Bytes one on-the-fly Poisson apply must move if every stream is touched once: node
    coordinates (24 B/node, FP64), x and y (8 B/node each, 4 in the FP32 apply), the blocks' node
    lists (4 B/entry) and the work items: per edge-star cluster 3 words (12 B: slot indices, tet
    and slot masks), or per tet 4 packed corner indices (4 B)."""
    xy = 8 if fp32 else 16
    if ctx.apply_shape["unit"] in ("cluster", "persistent"):
        return 12 * ctx.n_clusters + 4 * ctx.cluster_node_list_entries + (24 + xy) * ctx.N
    return 4 * ctx.E + 4 * ctx.node_list_entries + (24 + xy) * ctx.N


def amg_kwargs(args) -> dict:
    """Options of the host SA set-up from the command line (defaults of ``amg.build_hierarchy``
    where not given)."""
    kw = {"max_coarse": args.max_coarse}
    if args.amg_theta is not None:
        kw["theta"] = tuple(float(t) for t in args.amg_theta.split(","))
    if args.amg_omega is not None:
        kw["omega"] = args.amg_omega
    if args.amg_filter is not None:
        kw["filter_levels"] = () if args.amg_filter == "none" else tuple(int(v) for v in args.amg_filter.split(","))
    return kw


def measured_traffic(workload, cluster_block_nodes, fp32=False):
    """HBM bytes per launch of the PCG's apply (FP64, or FP32 in the FP32-PCG forms) from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary of this workload
    (``profiles/k_apply_cl_<workload>.md``, line ``traffic_bytes_per_launch_<fp64|fp32>_<B>: N`` for
    cluster blocks of B nodes), or None if the summary has no line for this run's kernel and block
    size."""
    f = ROOT / "profiles" / f"k_apply_cl_{workload}.md"
    if not f.exists():
        return None
    key = f"traffic_bytes_per_launch_{'fp32' if fp32 else 'fp64'}_{cluster_block_nodes}"
    for line in f.read_text().splitlines():
        k, sep, val = line.partition(":")
        if sep and k.strip() == key:
            return {"bytes": int(val), "source": str(f.relative_to(ROOT))}
    return None


def cpu_baseline(mesh, mats, B_vals, budget_s, log=None, max_iter=2000):
    """The multi-core CPU restatement of the field step (``oracle/cpu_minimizer.cpp``: pcohen_hs on
    the assembled operators with an SA-AMG V-cycle, OpenMP over all host threads; kind 'port') on
    ``mesh``: field steps of ``B_vals`` in order from +2 T until ``budget_s`` seconds are spent or
    the list ends; ``max_iter`` caps the minimiser iterations of each step. Returns (steps,
    seconds, minimiser iterations, threads, tets)."""
    from mumag_matrixfree2_amd.amg import build_hierarchy
    from oracle import cpu
    from oracle.fem import Problem

    if not cpu.LIB.exists():
        cpu.build()
    A, K1, Js, k = mats
    t0 = time.perf_counter()
    p = Problem(mesh["knt"], mesh["ijk"], A, K1, Js, k, boundary_mask=mesh["boundary_mask"])
    cm = cpu.CpuMinimizer(p, build_hierarchy(p.A, "amgcl"))
    if log is not None:
        log(f"[bench:cpu] operators and hierarchy of {p.N} nodes in {time.perf_counter() - t0:.1f}s, "
            f"{cpu.threads()} threads")
    m, U = initial_m(p.N), np.zeros(p.N)
    steps, its = 0, 0
    t0 = time.perf_counter()
    while steps < len(B_vals) and time.perf_counter() - t0 < budget_s:
        m, U, info = cm.minimize(m, B_vals[steps] * H_DIR, U0=U, tau_f=1e-10, eps_a=1e-12, cg_tol=1e-10,
                                 max_iter=max_iter)
        its += info["iters"]
        steps += 1
        if log is not None:
            log(f"[bench:cpu] step {steps - 1} it={info['iters']} t={time.perf_counter() - t0:.2f}s")
    dt = time.perf_counter() - t0
    cm.close()
    return steps, dt, its, cpu.threads(), int(mesh["ijk"].shape[0])


def segmented_poisson(grains, ctxs, args, log, union=None, reps=3, tol=1e-10):
    """Poisson solves of all grains: one segmented union context (every kernel launch covers
    every grain, per-grain PCG scalars) against the per-grain contexts driven by host threads.
    Both solve the same magnetisations from zero to ``tol``; FP64 V-cycle on both sides."""
    from mumag_matrixfree2_amd.context import solve_U_batched
    from mumag_matrixfree2_amd.grainset import grain_set_context, split

    t0 = time.perf_counter()
    u = union
    if u is None:
        A, K1, Js, _ = materials()
        tables = [(A, K1, Js, np.array([g["easy_axis"], g["easy_axis"]])) for g in grains]
        u = grain_set_context(grains, tables)
        u.setup_precond(args.precond, **amg_kwargs(args))
        log(f"[bench:segmented] union context N={u.N} E={u.E}, AMG levels "
            f"{[l['A'].shape[0] for l in u.amg_levels]}, {time.perf_counter() - t0:.1f}s")
    precisions = [c.precision for c in ctxs] + [u.precision]
    for c in ctxs + [u]:
        c.set_precision("fp64")
    ms = [initial_m(c.N) for c in ctxs]
    zeros = [np.zeros(c.N) for c in ctxs]
    m_u, z_u = np.vstack(ms), np.zeros(u.N)
    solve_U_batched(ctxs, ms, zeros, tol, n_threads=args.threads)
    u.solve_U_segmented(m_u, z_u, tol)
    t0 = time.perf_counter()
    for _ in range(reps):
        Us, its_t, _ = solve_U_batched(ctxs, ms, zeros, tol, n_threads=args.threads)
    t_thr = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        U_u, its_s, rr_s = u.solve_U_segmented(m_u, z_u, tol)
    t_seg = (time.perf_counter() - t0) / reps
    for c, prec in zip(ctxs + [u], precisions):
        c.set_precision(prec)
    diff = max(float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)) for a, b in zip(split(u, U_u), Us))
    out = {"grains": len(ctxs), "tol": tol, "threads": args.threads,
           "threaded_ms_per_all_grain_solve": t_thr * 1e3, "segmented_ms_per_all_grain_solve": t_seg * 1e3,
           "speedup": t_thr / t_seg, "pcg_iters_max_threaded": int(np.max(its_t)),
           "pcg_iters_max_segmented": int(np.max(its_s)), "max_rel_res_segmented": float(np.max(rr_s)),
           "max_rel_U_diff": diff}
    log(f"[bench:segmented] {out}")
    if union is None:
        u.close()
    return out


def run_grains(args, log):
    """The 64-grain set. Form "segmented" (default): the grains concatenated into one segmented
    context (one AMG hierarchy on the union, per-grain scalars), a field step is
    ``minimize_segmented``, one launch per kernel for all grains. Form "threads": one context
    per grain, a field step is ``minimize_batched`` with host threads driving the contexts.
    Unless --no-compare-forms, the other form runs the same sweep and the report gives its rate
    and the largest per-grain differences."""
    from mumag_matrixfree2_amd.autotune import tune_apply
    from mumag_matrixfree2_amd.context import MeshContext, make_params, minimize_batched
    from mumag_matrixfree2_amd.grainset import grain_set_context
    from mumag_matrixfree2_amd.hysteresis_loop import field_values, switching_step
    from mumag_matrixfree2_amd.mesh import grain_set

    wl = WORKLOADS["grains64"]
    t0 = time.perf_counter()
    grains = grain_set(64, seed=3)
    A, K1, Js, _ = materials()
    tables = [(A, K1, Js, np.array([g["easy_axis"], g["easy_axis"]])) for g in grains]
    m0 = [initial_m(g["knt"].shape[0]) for g in grains]
    p = make_params("pcohen_hs", tau_f=1e-10, eps_a=1e-12, cg_tol=1e-10, max_iter=2000)
    B_vals = field_values(2.0 / JS, -8.0 / JS, wl["dB"] / JS, False)
    steps = min(args.steps or wl["steps"], len(B_vals))
    forms = [args.grains_form] + ([] if args.no_compare_forms else
                                  [f for f in ("segmented", "threads") if f != args.grains_form])

    ctxs, union = [], None
    if "threads" in forms or not args.no_segmented or args.fp64_sequential_baseline:
        for i, (g, t) in enumerate(zip(grains, tables)):
            c = MeshContext(g["knt"], g["ijk"][:, :4], g["ijk"][:, 4], *t, boundary_mask=g["boundary_mask"])
            c.setup_precond(args.precond, **amg_kwargs(args))
            c.set_precision(args.precision)
            tune_apply(c)
            ctxs.append(c)
        log(f"[bench] {len(ctxs)} grain contexts set up, {time.perf_counter() - t0:.1f}s")
    if "segmented" in forms:
        union = grain_set_context(grains, tables)
        union.setup_precond(args.precond, **amg_kwargs(args))
        union.set_precision(args.precision)
        shape = tune_apply(union)
        log(f"[bench] union context N={union.N} E={union.E} AMG levels "
            f"{[l['A'].shape[0] for l in union.amg_levels]} apply {shape['unit']}:{shape['param']}, "
            f"{time.perf_counter() - t0:.1f}s")
    E_tot = sum(g["ijk"].shape[0] for g in grains)
    N_tot = sum(g["knt"].shape[0] for g in grains)

    def sweep(form, timed, n_threads=None):
        """(seconds, J per step, per-grain energies per step, minimiser iterations, apply timing)"""
        n_threads = args.threads if n_threads is None else n_threads
        if form == "segmented":
            union.state_set(np.vstack(m0), np.zeros(union.N))
            vol = union.segment_volumes()
            if timed:
                union.apply_timing(True)
        else:
            for c, m in zip(ctxs, m0):
                c.state_set(m, np.zeros(c.N))
                if timed:
                    c.apply_timing(True)
            vol = np.array([c.V_mag for c in ctxs])
        J, Es, its = [], [], 0
        t_start = time.perf_counter()
        for k in range(steps):
            ts = time.perf_counter()
            B = B_vals[k] * H_DIR
            if form == "segmented":
                infos = union.minimize_segmented(p, B)
                Jg = union.state_j_parallel_segmented(H_DIR)
            else:
                infos = minimize_batched(ctxs, p, B, n_threads=n_threads)
                Jg = np.array([c.state_j_parallel(H_DIR) for c in ctxs])
            J.append(float(vol @ Jg / vol.sum()))
            Es.append([i["E"] for i in infos])
            its += sum(i["iters"] for i in infos)
            log(f"[bench:grains {form} t={n_threads}] step {k:2d} B={B_vals[k] * JS:+.3f} T J={J[-1] * JS:+.5f} T "
                f"iters={sum(i['iters'] for i in infos)} max={max(i['iters'] for i in infos)} "
                f"t={time.perf_counter() - ts:.3f}s")
        el = time.perf_counter() - t_start
        timing = None
        if timed:
            timing = [union.apply_timing(False)] if form == "segmented" else [c.apply_timing(False) for c in ctxs]
        return el, np.array(J), np.array(Es), its, timing

    seg = None if args.no_segmented or not ctxs else segmented_poisson(grains, ctxs, args, log, union)
    # warm-up on a separate state
    if args.warmup:
        if args.grains_form == "segmented":
            union.state_set(np.vstack(m0), np.zeros(union.N))
            for k in range(args.warmup):
                union.minimize_segmented(p, B_vals[k % len(B_vals)] * H_DIR)
        else:
            for c, m in zip(ctxs, m0):
                c.state_set(m, np.zeros(c.N))
            for k in range(args.warmup):
                minimize_batched(ctxs, p, B_vals[k % len(B_vals)] * H_DIR, n_threads=args.threads)
    elapsed, J, Es, its, timing = sweep(args.grains_form, timed=True)
    other = None
    if len(forms) > 1:
        t1, J1, Es1, its1, _ = sweep(forms[1], timed=False)
        other = {"form": forms[1], "field_steps_per_s": steps / t1, "speedup_of_headline_form": t1 / elapsed,
                 "minimizer_iters": its1, "max_abs_J_diff": float(np.max(np.abs(J - J1))),
                 "max_rel_grain_energy_diff": float(np.max(np.abs(Es - Es1) / np.maximum(np.abs(Es1), 1.0))),
                 "switching_step": switching_step(J1)}
    seq = None
    if args.fp64_sequential_baseline and ctxs:
        # the plain FP64 form: one grain at a time, each in its own context
        precs = [c.precision for c in ctxs]
        for c in ctxs:
            c.set_precision("fp64")
        t2, J2, Es2, its2, _ = sweep("threads", timed=False, n_threads=1)
        for c, pr in zip(ctxs, precs):
            c.set_precision(pr)
        seq = {"field_steps_per_s": steps / t2, "speedup_of_headline_form": t2 / elapsed, "minimizer_iters": its2,
               "max_abs_J_diff": float(np.max(np.abs(J - J2))),
               "max_rel_grain_energy_diff": float(np.max(np.abs(Es - Es2) / np.maximum(np.abs(Es2), 1.0))),
               "switching_step": switching_step(J2)}
    if args.grains_form == "segmented":
        # the sweep's applies cover shrinking segment subsets; the roofline is taken on
        # back-to-back applies over every grain (HIP events on the context's stream)
        reps = 100
        t_apply = union.bench_apply(reps=reps) * 1e-3 * reps
        n_apply = reps
        algo = apply_algorithmic_bytes(union) * reps
    else:
        algo = sum(apply_algorithmic_bytes(c) * n for c, (_, n) in zip(ctxs, timing))
        t_apply = sum(ms for ms, _ in timing) * 1e-3
        n_apply = sum(n for _, n in timing)
    achieved = algo / max(t_apply, 1e-12) / 1e9
    label = "FP64" if args.precision == "fp64" else "mixed FP32 V-cycle / FP64 PCG"
    result = {
        "metric": f"hysteresis field-steps/s ({label}, {wl['name']})",
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
        "data": "synthetic (generated grain set, seeded initial magnetisation)",
        "config": {"workload": f"{wl['name']}_{args.precision}_sweep_first{steps}of{len(B_vals)}steps_from+2T",
                   "grains": len(grains),
                   "tets": E_tot, "nodes": N_tot, "method": "pcohen_hs", "form": args.grains_form,
                   "threads": args.threads if args.grains_form == "threads" else 1,
                   "precond": f"SA-AMG V-cycle ({args.precond})", "pcg_tol": 1e-10, "tau_f": 1e-10},
        "minimizer_iters": its,
        "switching_step": switching_step(J),
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": None,
            "kernel": ("k_apply_cl over all grains in one launch (back-to-back applies after the sweep)"
                       if args.grains_form == "segmented" else
                       "k_apply over all grains (event time includes overlap with other grains)"),
            "avg_launch_us": t_apply / max(n_apply, 1) * 1e6, "timed_launches": n_apply,
        },
    }
    if other is not None:
        result["other_form"] = other
    if seq is not None:
        result["fp64_sequential"] = seq
    if seg is not None:
        result["segmented_poisson"] = seg
    print(json.dumps(result), flush=True)
    for c in ctxs:
        c.close()
    if union is not None:
        union.close()


def cpu_sample_mesh(workload):
    """The mesh the CPU restatement runs on: the workload's own mesh for the cube; for the 24.6M-tet
    polycrystal (whose assembled operators the CPU restatement cannot hold in a bounded run) the
    same 200 nm, 27-grain polycrystal with its grain-boundary phase and airbox at 40^3 hexes."""
    from mumag_matrixfree2_amd.mesh import config_cube_mesh, polycrystal_cube

    if workload == "cube2M":
        return config_cube_mesh(80.0, 60, 5), materials(), "the same mesh"
    return (polycrystal_cube(200.0, 40, 5, n_grains=N_GRAINS, gb_width=2.0, seed=4), poly_materials(),
            "the same 200 nm 27-grain polycrystal (grain-boundary phase, airbox) meshed at 40^3 hexes x 6 tets")


def host_cpu_description() -> dict:
    """CPU model, sockets, physical cores and logical CPUs of the host (lscpu, /proc/cpuinfo)."""
    import subprocess

    out = {"model": None, "sockets": None, "physical_cores": None, "logical_cpus": os.cpu_count()}
    try:
        txt = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        kv = {ln.split(":", 1)[0].strip(): ln.split(":", 1)[1].strip() for ln in txt.splitlines() if ":" in ln}
        out["model"] = kv.get("Model name")
        if kv.get("Socket(s)", "").isdigit():
            out["sockets"] = int(kv["Socket(s)"])
        if out["sockets"] and kv.get("Core(s) per socket", "").isdigit():
            out["physical_cores"] = out["sockets"] * int(kv["Core(s) per socket"])
    except Exception:
        pass
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", default="poly25M", choices=sorted(WORKLOADS))
    ap.add_argument("--steps", type=int, default=20, help="field steps timed, from +2 T (at most the sweep length)")
    ap.add_argument("--warmup", type=int, default=1, help="untimed field steps first, on a separate state")
    ap.add_argument("--precond", default="amgcl", choices=["amgcl", "amg"])
    ap.add_argument("--method", default="pcohen_hs", choices=["pcohen_hs", "lbfgs"],
                    help="field-step minimiser (cube2M / poly25M; the grain set runs pcohen_hs)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="time budget of the CPU baseline (field steps from the start of the sweep)")
    ap.add_argument("--precision", default=None, choices=["fp64", "mixed", "fp32pcg", "fp32all"],
                    help="headline form: 'fp32pcg' (the default on one mesh: FP32 PCG and V-cycle with FP64 dots "
                         "and FP64 true-residual replacement), 'mixed' (FP32 V-cycle under the FP64 PCG; the "
                         "default on the grain set), 'fp32all' or 'fp64'; the FP64 form runs the same steps for "
                         "the comparison")
    ap.add_argument("--no-fp64-compare", action="store_true", help="mixed headline without the FP64 sweep")
    ap.add_argument("--max-coarse", type=int, default=None)
    ap.add_argument("--amg-theta", default=None,
                    help="strength thresholds of the SA set-up, fine level first, comma-separated (default 0.1,0.05)")
    ap.add_argument("--amg-omega", type=float, default=None,
                    help="prolongator smoothing weight omega / rho(D^-1 A) of the SA set-up (default 1.5)")
    ap.add_argument("--ref-eta", type=float, default=None,
                    help="FP32-PCG forms: replace the recursive residual by the true one after this reduction "
                         "(default: the library's)")
    ap.add_argument("--amg-filter", default=None,
                    help="levels whose prolongator is smoothed with the filtered matrix, comma-separated, or "
                         "'none' (default 0: the fine level)")
    ap.add_argument("--block-nodes", type=int, default=256,
                    help="distinct nodes per element block (16..256): the LDS staging size of the element kernels")
    ap.add_argument("--apply-shape", default="auto",
                    help="'auto' (autotuned at set-up) or unit:param[:cluster block nodes[:scatter]], e.g. "
                         "cluster:256, cluster:256:1024, cluster:256:1024:colour, element:128, persistent:2")
    ap.add_argument("--solve-guess", type=int, default=8,
                    help="initial guess of the minimiser's Poisson solves: projection onto the last K solutions "
                         "(0: the previous solution, as the reference)")
    ap.add_argument("--spmv-lanes", default="auto",
                    help='lanes per row of the V-cycle CSR products: "auto" (timed at set-up) or a JSON list '
                         'of per-level {"A": a, "P": p, "R": r}')
    ap.add_argument("--max-iter", type=int, default=2000,
                    help="minimiser iterations per field step at most (the workload's is 2000; lower only to "
                         "bound profiler runs)")
    ap.add_argument("--calibrate-copies", type=int, default=0,
                    help="after the sweep, launch this many plain copies of N doubles (k_copy_calib), the "
                         "known-size reference of the HBM byte counters")
    ap.add_argument("--guess-compare", action="store_true",
                    help="also rerun the headline sweep with the reference warm start")
    ap.add_argument("--threads", type=int, default=8, help="host threads driving the grains (grains64)")
    ap.add_argument("--grains-form", default="segmented", choices=["segmented", "threads"],
                    help="grains64: one segmented context for all grains, or one context per grain on host threads")
    ap.add_argument("--no-compare-forms", action="store_true",
                    help="grains64: do not run the sweep in the other form")
    ap.add_argument("--fp64-sequential-baseline", action="store_true",
                    help="grains64: also run the sweep in FP64 one grain at a time (one context per grain)")
    ap.add_argument("--no-segmented", action="store_true",
                    help="grains64: skip the segmented vs threaded Poisson solve comparison")
    args = ap.parse_args()
    if args.precision is None:
        args.precision = "mixed" if args.workload == "grains64" else "fp32pcg"
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the hysteresis path runs on a single MI355X here "
              "(element-block partitions would shard across GPUs with a halo exchange per apply); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    from mumag_matrixfree2_amd.context import MeshContext, make_params
    from mumag_matrixfree2_amd.hysteresis_loop import field_values, switching_step
    from mumag_matrixfree2_amd.io_utils import nucleation_field

    log = lambda *a: print(*a, file=sys.stderr, flush=True)  # noqa: E731
    if args.workload == "grains64":
        return run_grains(args, log)
    wl = WORKLOADS[args.workload]
    t0 = time.perf_counter()
    mesh, mats = make_workload(args.workload)
    log(f"[bench] mesh {time.perf_counter() - t0:.1f}s")
    ctx = MeshContext(mesh["knt"], mesh["ijk"][:, :4], mesh["ijk"][:, 4], *mats,
                      boundary_mask=mesh["boundary_mask"], block_nodes=args.block_nodes)
    log(f"[bench] context {time.perf_counter() - t0:.1f}s")
    ctx.setup_precond(args.precond, log=log, **amg_kwargs(args))
    from mumag_matrixfree2_amd.autotune import tune_apply, tune_spmv_lanes

    if args.apply_shape == "auto":
        shape = tune_apply(ctx)
    else:
        parts = args.apply_shape.split(":")
        if len(parts) >= 3:
            ctx.set_cluster_block_nodes(int(parts[2]))
        ctx.set_apply_shape(parts[0], int(parts[1]))
        if len(parts) == 4:
            ctx.set_apply_scatter(parts[3])
        shape = {"unit": parts[0], "param": int(parts[1])}
    log(f"[bench] apply shape {shape}")
    if args.ref_eta is not None:
        ctx.set_refinement(args.ref_eta)
    ctx.set_precision(args.precision)
    if args.spmv_lanes == "auto":
        lanes = tune_spmv_lanes(ctx)
    else:
        for lev, d in enumerate(json.loads(args.spmv_lanes)):
            for which, v in d.items():
                ctx.set_amg_lanes(lev, which, int(v))
        lanes = {"lanes": ctx.amg_lanes()}
    log(f"[bench] V-cycle SpMV lanes {lanes}")
    log(f"[bench] set-up {time.perf_counter() - t0:.1f}s: N={ctx.N} E={ctx.E} "
        f"AMG levels={[l['A'].shape[0] for l in ctx.amg_levels]}")

    p = make_params(args.method, tau_f=1e-10, eps_a=1e-12, cg_tol=1e-10, max_iter=args.max_iter)
    B_vals = field_values(2.0 / JS, -8.0 / JS, wl["dB"] / JS, False)
    m0 = initial_m(ctx.N)
    steps = max(1, min(args.steps, len(B_vals)))

    def sweep(precision, timed, guess=None, keep_states=False):
        ctx.set_precision(precision)
        ctx.set_solve_guess(args.solve_guess if guess is None else guess)
        ctx.state_set(m0, np.zeros(ctx.N))
        if timed:
            ctx.apply_timing(True)
        c0 = ctx.counters()
        J, E, iters, states = [], [], [], []
        t_copy = 0.0  # host copies of the states (kept for the worst-step control), not counted
        t_start = time.perf_counter()
        for k in range(steps):
            if keep_states:
                tc = time.perf_counter()
                states.append(ctx.state_get())
                t_copy += time.perf_counter() - tc
            ts = time.perf_counter()
            info = ctx.minimize_resident(p, B_vals[k] * H_DIR)
            J.append(ctx.state_j_parallel(H_DIR))
            E.append(info["E"])
            iters.append(info["iters"])
            log(f"[bench:{precision}] step {k:2d} B={B_vals[k] * JS:+.3f} T J={J[-1] * JS:+.5f} T "
                f"E={info['E']:+.10e} it={info['iters']} nf={info['evals']} icg={info['demag_iters']} "
                f"t={time.perf_counter() - ts:.3f}s")
        el = time.perf_counter() - t_start - t_copy
        c1 = ctx.counters()
        timing = ctx.apply_timing(False) if timed else None
        return {"seconds": el, "J": np.array(J), "E": np.array(E), "iters": iters,
                "pcg_iters": c1["pcg_iters"] - c0["pcg_iters"], "applies": c1["apply"] - c0["apply"],
                "timing": timing, "states": states}

    # warm-up on a separate state
    ctx.set_precision(args.precision)
    ctx.set_solve_guess(args.solve_guess)
    ctx.state_set(m0, np.zeros(ctx.N))
    for k in range(args.warmup):
        ctx.minimize_resident(p, B_vals[k % len(B_vals)] * H_DIR)

    head = sweep(args.precision, timed=True)
    elapsed = head["seconds"]
    apply_ms, apply_n = head["timing"]
    B_T = B_vals[:steps] * JS

    def nuc(J):
        n = nucleation_field(B_T, J * JS)
        return None if n is None else n["index"]

    compare = None
    if args.precision != "fp64" and not args.no_fp64_compare:
        # the FP64 form on the same steps in the same run: per-step relative energy error (energies in
        # units of Kd V_mag, the denominator floored at 1 since E crosses 0 during the sweep), switching
        # step and nucleation index, the final states' potential, gradient and magnetisation, and the
        # two solvers on the same magnetisation
        B_last = B_vals[steps - 1] * H_DIR
        m_mx, U_mx = ctx.state_get()
        _, g_mx = ctx.energy_grad(m_mx, U_mx, B_last)
        r64 = sweep("fp64", timed=False, keep_states=True)
        m64, U64 = ctx.state_get()
        _, g64 = ctx.energy_grad(m64, U64, B_last)
        U_same64, _, _ = ctx.solve_U(m64, np.zeros(ctx.N), 1e-10)
        ctx.set_precision(args.precision)
        U_samemx, _, _ = ctx.solve_U(m64, np.zeros(ctx.N), 1e-10)

        def rel(a, b):
            return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))

        err = np.abs(head["E"] - r64["E"]) / np.maximum(np.abs(r64["E"]), 1.0)
        # The worst step once more from the FP64 sweep's state before it: the headline form, the FP64
        # form, and the FP64 form with its sums in another order (the colour-sequenced scatter), so
        # the line shows how far FP64 rounding alone moves that step's result
        kw = int(np.argmax(err))
        m_s, U_s = r64["states"][kw]
        scatter0 = ctx.apply_scatter

        def rerun(precision, scatter):
            ctx.set_precision(precision)
            ctx.set_apply_scatter(scatter)
            ctx.set_solve_guess(args.solve_guess)
            ctx.state_set(m_s, U_s)
            info = ctx.minimize_resident(p, B_vals[kw] * H_DIR)
            return info["E"], info["iters"]

        E_a, it_a = rerun("fp64", "atomic")
        E_c, it_c = rerun("fp64", "colour")
        E_m, it_m = rerun(args.precision, "atomic")
        ctx.set_apply_scatter(scatter0)
        ctx.set_precision(args.precision)
        sc = max(abs(E_a), 1.0)
        worst = {"step": kw, "B_T": float(B_vals[kw] * JS), "sweep_rel_energy_err": float(err[kw]),
                 "from_the_fp64_state_before_it": {
                     "mixed_vs_fp64_rel_energy_err": abs(E_m - E_a) / sc,
                     "fp64_vs_fp64_other_sum_order_rel_energy_diff": abs(E_c - E_a) / sc,
                     "minimizer_iters": {"mixed": it_m, "fp64": it_a, "fp64_other_sum_order": it_c}}}
        log(f"[bench] worst step {kw}: {worst}")
        compare = {
            "fp64_field_steps_per_s": steps / r64["seconds"],
            "speedup_vs_fp64": r64["seconds"] / elapsed,
            "max_rel_energy_err": float(err.max()),
            "energy_err_bar": 1e-8,
            "switching_step": [switching_step(head["J"]), switching_step(r64["J"])],
            "nucleation_index": [nuc(head["J"]), nuc(r64["J"])],
            "final_state_rel_err": {"U": rel(U_mx, U64), "grad": rel(g_mx, g64), "m": rel(m_mx, m64)},
            "same_m_potential_rel_err": rel(U_samemx, U_same64),
            "minimizer_iters": [int(sum(head["iters"])), int(sum(r64["iters"]))],
            "pcg_iters": [head["pcg_iters"], r64["pcg_iters"]],
            "worst_step": worst,
        }

    guess_cmp = None
    if args.solve_guess > 0 and args.guess_compare:
        rg = sweep(args.precision, timed=False, guess=0)
        ctx.set_solve_guess(args.solve_guess)
        guess_cmp = {"field_steps_per_s": steps / rg["seconds"], "speedup_of_headline_form": rg["seconds"] / elapsed,
                     "minimizer_iters": int(sum(rg["iters"])), "headline_minimizer_iters": int(sum(head["iters"])),
                     "max_rel_energy_diff": float(np.max(np.abs(head["E"] - rg["E"]) /
                                                         np.maximum(np.abs(rg["E"]), 1.0))),
                     "switching_step": switching_step(rg["J"])}

    if args.calibrate_copies > 0:
        ms = ctx.bench_copy(args.calibrate_copies)
        log(f"[bench] {args.calibrate_copies} copies of {ctx.N} doubles ({16 * ctx.N} bytes each): {ms * 1e3:.2f} us")
    avg_apply_s = apply_ms / max(apply_n, 1) * 1e-3
    algo = apply_algorithmic_bytes(ctx, fp32=args.precision in ("fp32pcg", "fp32all"))
    achieved = algo / avg_apply_s / 1e9
    traffic = measured_traffic(args.workload, ctx.cluster_block_nodes, fp32=args.precision in ("fp32pcg", "fp32all"))
    label = {"fp64": "FP64", "mixed": "mixed: FP32 V-cycle, FP64 PCG, FP64 energies",
             "fp32pcg": "FP32 PCG and V-cycle with FP64 dots and FP64 true-residual refinement, FP64 energies",
             "fp32all": "FP32 PCG and V-cycle with FP64 dots and FP64 true-residual refinement, FP32 energy "
                        "densities with compensated FP64 sums"}[args.precision]
    result = {
        "metric": f"hysteresis field-steps/s ({label}, {wl['name']})",
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
            "workload": (f"{wl['name']}_{args.precision}_sweep_first{steps}of{len(B_vals)}steps_"
                         f"+2T..{B_T[-1]:+.3f}T"),
            "tets": ctx.E, "nodes": ctx.N, "method": args.method, "precond": f"SA-AMG V-cycle ({args.precond})",
            "amg_levels": [int(l["A"].shape[0]) for l in ctx.amg_levels],
            "apply_shape": {"unit": shape["unit"], "param": shape["param"], "block_nodes": ctx.block_nodes,
                            "cluster_block_nodes": ctx.cluster_block_nodes, "scatter": ctx.apply_scatter},
            "spmv_lanes": lanes["lanes"],
            "solve_guess": (f"projection onto the last {args.solve_guess} solutions" if args.solve_guess
                            else "previous solution (reference warm start)"),
            "pcg_tol": 1e-10, "tau_f": 1e-10,
        },
        "pcg_iters_per_s": head["pcg_iters"] / elapsed,
        "switching_step": switching_step(head["J"]),
        "nucleation_index": nuc(head["J"]),
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None if traffic is None else traffic["bytes"],
            "kernel": ("k_apply_cl<float, false, true> (FP32 matrix-free Poisson apply of the PCG, edge-star "
                       "clusters)" if args.precision in ("fp32pcg", "fp32all") else
                       "k_apply_cl<double, false, true> (FP64 matrix-free Poisson apply of the PCG, edge-star "
                       "clusters)" if shape["unit"] != "element" else "k_apply (FP64 matrix-free Poisson apply)"),
            "avg_launch_us": avg_apply_s * 1e6,
            "timed_launches": apply_n, "launches": head["applies"],
            "algorithmic_bytes_per_launch": algo,
            "traffic_source": None if traffic is None else traffic["source"],
        },
    }
    if compare is not None:
        result["mixed_vs_fp64"] = compare
    if guess_cmp is not None:
        result["reference_warm_start"] = guess_cmp
    if not args.no_cpu_baseline and args.method == "pcohen_hs":
        result["cpu_baseline"] = cpu_baseline_line(args, ctx, head, steps, log)
    print(json.dumps(result), flush=True)
    ctx.close()


def cpu_baseline_line(args, ctx, head, steps, log) -> dict:
    """The CPU restatement's rate in field-steps/s of the timed workload. On the cube it runs the
    same mesh and sweep; on the polycrystal it runs the first 40 minimiser iterations of the
    first field step on the 40^3-hex polycrystal (the same physical problem on a mesh whose
    assembled operators it builds in seconds), and its seconds per minimiser iteration are
    scaled to the headline mesh by the tet ratio and multiplied by the minimiser iterations the
    GPU needed per field step on the headline mesh."""
    from mumag_matrixfree2_amd.hysteresis_loop import field_values

    mesh_c, mats_c, what = cpu_sample_mesh(args.workload)
    wl = WORKLOADS[args.workload]
    B_vals = field_values(2.0 / JS, -8.0 / JS, wl["dB"] / JS, False)
    k_max = steps
    if args.workload == "cube2M":
        k_cpu, dt, it_cpu, nthr, E_c = cpu_baseline(mesh_c, mats_c, B_vals[:k_max], args.cpu_seconds, log)
    else:  # a bounded sample: the first minimiser iterations of the first field step
        k_cpu, dt, it_cpu, nthr, E_c = cpu_baseline(mesh_c, mats_c, B_vals[:1], args.cpu_seconds, log, max_iter=40)
    cpu = host_cpu_description()
    gpu_its = head["iters"][:k_cpu]
    if args.workload == "cube2M":
        gpu_t = None
        value = k_cpu / dt
        how = (f"the first {k_cpu} field steps of the same sweep ({it_cpu} minimiser iterations; the GPU took "
               f"{int(sum(gpu_its))}) in {dt:.2f} s")
    else:
        s_per_it = dt / max(it_cpu, 1) * (ctx.E / E_c)
        its_per_step = float(np.mean(head["iters"]))
        value = 1.0 / (s_per_it * its_per_step)
        how = (f"{what} ({E_c} tets): the first {it_cpu} minimiser iterations of the first field step took "
               f"{dt:.2f} s ({dt / max(it_cpu, 1) * 1e3:.1f} ms per iteration); scaled by the tet ratio "
               f"{ctx.E / E_c:.2f} to {s_per_it:.3f} s per iteration and multiplied by the {its_per_step:.1f} "
               f"minimiser iterations per field step of the GPU run on the {ctx.E}-tet mesh (an estimate: PCG "
               f"iterations per solve differ between the meshes)")
    return {"value": value, "unit": "field-steps/s", "cores": nthr, "kind": "port",
            "sample": (f"oracle/cpu_minimizer.cpp (C++/OpenMP restatement of the oracle's pcohen_hs field step on "
                       f"the assembled operators, SA-AMG PCG, {nthr} threads): " + how),
            "host_cpu": cpu}


if __name__ == "__main__":
    main()
