"""Autotuner for the work shape of the matrix-free Poisson apply.

The apply (``k_apply`` / ``k_apply_cl``) has two knobs that change its speed but not its
result beyond rounding:

* the unit of work of a thread: one edge-star cluster (up to 6 tets sharing an edge, whose
  8 node slots are read from LDS once) or one tet;
* the workgroup size (64, 128 or 256 threads per element block).

A third knob, the number of distinct nodes per element block (the LDS staging size, hence
the elements per block), is fixed when a context is created; ``tune_block_nodes`` times
contexts built with each candidate value.

The timings come from back-to-back applies on device-resident vectors, bracketed by HIP
events on the context's stream (``mm_bench_apply``). The reference has no counterpart: its
apply is a fixed JAX segment-sum or a CSR SpMV (``src/poisson_solve.py:166-171,210-252``).
"""

from __future__ import annotations

SHAPES = ((True, 256), (True, 128), (True, 64), (False, 256), (False, 128), (False, 64))


def _label(clusters: bool, threads: int) -> str:
    return f"{'cluster' if clusters else 'element'}x{threads}"


def tune_apply(ctx, shapes=SHAPES, reps: int = 40) -> dict:
    """Time every (clusters, threads) shape on ``ctx``, keep the fastest set on the context and
    return ``{"clusters", "threads", "us_per_apply": {label: microseconds}}``."""
    times = {}
    for clusters, threads in shapes:
        ctx.set_apply_shape(clusters, threads)
        ctx.bench_apply(reps=3)
        times[(clusters, threads)] = ctx.bench_apply(reps=reps) * 1e3
    best = min(times, key=times.get)
    ctx.set_apply_shape(*best)
    return {"clusters": best[0], "threads": best[1],
            "us_per_apply": {_label(*k): round(v, 3) for k, v in times.items()}}


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
