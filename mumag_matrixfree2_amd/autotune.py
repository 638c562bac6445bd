"""Autotuner for the work shape of the matrix-free Poisson apply.

The apply (``k_apply`` / ``k_apply_cl``) has two knobs that change its speed but not its
result beyond rounding:

* the unit of work of a thread: one edge-star cluster (up to 6 tets sharing an edge, whose
  8 node slots are read from LDS once) or one tet, and for clusters whether a persistent grid
  walks the blocks with the next block's staging loads in flight;
* the workgroup size (64, 128 or 256 threads per element block), or for the persistent grid
  the workgroups per CU.

A third knob, the number of distinct nodes per element block (the LDS staging size, hence
the elements per block), is fixed when a context is created; ``tune_block_nodes`` times
contexts built with each candidate value.

The timings come from back-to-back applies on device-resident vectors, bracketed by HIP
events on the context's stream (``mm_bench_apply``). The reference has no counterpart: its
apply is a fixed JAX segment-sum or a CSR SpMV (``src/poisson_solve.py:166-171,210-252``).
"""

from __future__ import annotations

SHAPES = (("persistent", 2), ("persistent", 3), ("cluster", 256), ("cluster", 128), ("cluster", 64),
          ("element", 256), ("element", 128), ("element", 64))


def tune_apply(ctx, shapes=SHAPES, reps: int = 40) -> dict:
    """Time every (unit, param) shape on ``ctx``, keep the fastest set on the context and return
    ``{"unit", "param", "us_per_apply": {"unit:param": microseconds}}``."""
    times = {}
    for unit, param in shapes:
        ctx.set_apply_shape(unit, param)
        ctx.bench_apply(reps=3)
        times[(unit, param)] = ctx.bench_apply(reps=reps) * 1e3
    best = min(times, key=times.get)
    ctx.set_apply_shape(*best)
    return {"unit": best[0], "param": best[1],
            "us_per_apply": {f"{u}:{q}": round(v, 3) for (u, q), v in times.items()}}


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
