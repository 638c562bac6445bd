"""Autotuner for the work shape of the matrix-free Poisson apply.

The apply (``k_apply`` / ``k_apply_cl``) has two knobs that change its speed but not its
result beyond rounding:

* the unit of work of a thread: one edge-star cluster (up to 6 tets sharing an edge, whose
  8 node slots are read from LDS once) or one tet, and for clusters whether a persistent grid
  walks the blocks with the next block's staging loads in flight;
* the workgroup size (64, 128 or 256 threads per element block), or for the persistent grid
  the workgroups per CU.

A third knob is the number of distinct nodes per block of the cluster partition (256, 512,
1024 or 2048; larger blocks stage fewer node entries and leave fewer shared nodes to the
write-back atomics): ``tune_apply`` rebuilds the cluster partition for each value it tries. The node cap
of the element blocks (the per-tet kernels) is fixed when a context is created;
``tune_block_nodes`` times contexts built with each candidate value.

The timings come from back-to-back applies on device-resident vectors, bracketed by HIP
events on the context's stream (``mm_bench_apply``). The reference has no counterpart: its
apply is a fixed JAX segment-sum or a CSR SpMV (``src/poisson_solve.py:166-171,210-252``).
"""

from __future__ import annotations

SHAPES = (("persistent", 2), ("persistent", 3), ("cluster", 256), ("cluster", 128), ("cluster", 64),
          ("element", 256), ("element", 128), ("element", 64))


CLUSTER_BLOCK_NODES = (512, 1024, 2048)


SCATTER_MODES = ("atomic", "colour")


def tune_apply(ctx, shapes=SHAPES, reps: int = 40, cluster_block_nodes=CLUSTER_BLOCK_NODES,
               scatter_modes=SCATTER_MODES) -> dict:
    """Time every (unit, param) shape on ``ctx`` with 256-node cluster blocks, then 256-thread
    cluster applies on ``cluster_block_nodes``-node blocks (not on segmented contexts); keep the
    fastest on the context; when it is a cluster shape, also time the colour-sequenced scatter on
    it and keep whichever scatter is faster. Returns ``{"unit", "param", "cluster_block_nodes",
    "scatter", "us_per_apply": {"unit:param[:nodes][:scatter]": microseconds}}``."""
    if getattr(ctx, "apply_scatter", "atomic") != "atomic":
        ctx.set_apply_scatter("atomic")
    times = {}
    if getattr(ctx, "cluster_block_nodes", 256) != 256:
        ctx.set_cluster_block_nodes(256)
    for unit, param in shapes:
        ctx.set_apply_shape(unit, param)
        ctx.bench_apply(reps=3)
        times[(unit, param, 256)] = ctx.bench_apply(reps=reps) * 1e3
    if ctx.segments is None:
        for nb in cluster_block_nodes:
            ctx.set_cluster_block_nodes(nb)
            ctx.set_apply_shape("cluster", 256)
            ctx.bench_apply(reps=3)
            times[("cluster", 256, nb)] = ctx.bench_apply(reps=reps) * 1e3
    best = min(times, key=times.get)
    if ctx.segments is None:
        ctx.set_cluster_block_nodes(best[2])
    ctx.set_apply_shape(best[0], best[1])
    us = {f"{u}:{q}" + ("" if nb == 256 else f":{nb}"): round(v, 3) for (u, q, nb), v in times.items()}
    scatter = "atomic"
    if scatter_modes and ctx.segments is None and best[0] == "cluster":
        # the colour-sequenced scatter against the FP64 atomics, on the chosen shape
        key = f"{best[0]}:{best[1]}" + ("" if best[2] == 256 else f":{best[2]}")
        for mode in scatter_modes:
            if mode == "atomic":
                continue
            ctx.set_apply_scatter(mode)
            ctx.bench_apply(reps=3)
            t = ctx.bench_apply(reps=reps) * 1e3
            us[f"{key}:{mode}"] = round(t, 3)
            if t < times[best]:
                scatter, times[best] = mode, t
        ctx.set_apply_scatter(scatter)
    return {"unit": best[0], "param": best[1], "cluster_block_nodes": best[2], "scatter": scatter,
            "us_per_apply": us}


def tune_block_nodes(make_ctx, caps=(128, 192, 256), reps: int = 40):
    """Build a context per block node cap with ``make_ctx(block_nodes)``, tune each with
    ``tune_apply`` and return ``(best context, report)``; the other contexts are closed."""
    best_ctx, best_us, report = None, float("inf"), {}
    for cap in caps:
        ctx = make_ctx(cap)
        res = tune_apply(ctx, reps=reps)
        us = min(res["us_per_apply"].values())
        report[cap] = res
        if us < best_us:
            if best_ctx is not None:
                best_ctx.close()
            best_ctx, best_us = ctx, us
        else:
            ctx.close()
    return best_ctx, {"block_nodes": best_ctx.block_nodes, "shapes": report}


RHS_WIDTHS = (1, 2, 4, 8)


def tune_rhs_width(ctx, widths=RHS_WIDTHS, reps: int = 20) -> dict:
    """Right-hand-side width of a multi-vector solve: time the k-vector apply (one staging of
    the mesh for k node-major vectors, ``mm_bench_apply_multi``) for each k and the best
    single-vector apply of ``ctx``, and return ``{"width": k with the lowest time per vector,
    "us_per_vector": {k: us}, "single_us": us}``. A width of 1 means batching vectors saves
    nothing on this mesh. Runs on 256-node cluster blocks; restores the context's block size."""
    nb = getattr(ctx, "cluster_block_nodes", 256)
    shape = dict(ctx.apply_shape)
    ctx.bench_apply(reps=3)
    single = ctx.bench_apply(reps=reps) * 1e3
    if ctx.segments is None:
        ctx.set_cluster_block_nodes(256)
    ctx.set_apply_shape("cluster", 256)
    per = {}
    for k in widths:
        ctx.bench_apply_multi(k, reps=2)
        per[k] = ctx.bench_apply_multi(k, reps=reps) * 1e3 / k
    if ctx.segments is None:
        ctx.set_cluster_block_nodes(nb)
    ctx.set_apply_shape(shape["unit"], shape["param"])
    best = min(per, key=per.get)
    width = best if per[best] < single else 1
    return {"width": width, "us_per_vector": {k: round(v, 3) for k, v in per.items()}, "single_us": round(single, 3)}


SPMV_LANES = (1, 2, 4, 8, 16, 32, 64)


def tune_spmv_lanes(ctx, lanes=SPMV_LANES, reps: int = 10, rounds: int = 1) -> dict:
    """Lanes per row of every CSR product of the V-cycle (each level's A, P and R), chosen by
    coordinate descent on the measured time of a whole preconditioner application in the
    context's current precision form: for each product in turn, every lane count is timed with
    the others fixed and the fastest kept. Returns ``{"lanes": ctx.amg_lanes(), "us_before",
    "us_after"}``."""
    levels = ctx.amg_levels or []
    ctx.bench_precond(reps=2)
    before = ctx.bench_precond(reps=reps) * 1e3
    best_us = before
    for _ in range(rounds):
        for lev, L in enumerate(levels):
            for which in ("A", "P", "R"):
                if which == "A" and (lev == 0 or L.get("A") is None):
                    continue  # level 0 applies A matrix-free
                if which in ("P", "R") and L.get("P") is None:
                    continue  # the coarsest level has no transfer operators
                if which == "A" and lev == len(levels) - 1:
                    continue  # the coarsest level is solved with its dense inverse
                cur = ctx.amg_lanes()[lev][which]
                times = {}
                for ln in lanes:
                    ctx.set_amg_lanes(lev, which, ln)
                    times[ln] = ctx.bench_precond(reps=reps) * 1e3
                best = min(times, key=times.get)
                if times[best] < times.get(cur, float("inf")):
                    cur = best
                ctx.set_amg_lanes(lev, which, cur)
                best_us = times[cur]
    after = ctx.bench_precond(reps=reps) * 1e3
    return {"lanes": ctx.amg_lanes(), "us_before": round(before, 2), "us_after": round(after, 2),
            "us_best_seen": round(best_us, 2)}

