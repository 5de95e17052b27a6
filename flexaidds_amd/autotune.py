"""Launch-shape autotuner for the pose kernel.

The pose kernel keeps one wave per pose and pulls poses from a work queue,
so its launch knobs are the number of resident waves and the variant of the
neighbour scan (fa_ctx_set_launch_shape):

* waves: one per resident slot (2 per SIMD), or fewer — a SIMD then runs
  fewer waves side by side, each of them faster, which can win when a
  launch has few poses and its slowest pose sets the time;
* flags: the per-cell neighbourhood lists (one contiguous stream for the 27
  cells around an atom) or the receptor cell lists (27 short segments).

A wave cap leaves results unchanged; the scan variants find the same
neighbours in the same order and differ only in the order the clash
pre-pass sums its wall terms (the wall of a clash pose, in its last bits).
`tune` times every candidate on the
caller's own workload (a callable that runs one launch and waits for it)
and leaves the fastest on the context; `bench.py` records the choice per
workload.  The reference fixes its launch constants at compile time
(LIB/hip_eval.hip:57-71).
"""
from __future__ import annotations

import time

from . import abi


def candidates(ctx, fractions=(1.0, 0.75, 0.5), flag_sets=(0, abi.FA_SHAPE_CELL_LISTS)):
    """(waves, flags) pairs: wave caps as fractions of the resident slots."""
    full = ctx.launch_waves()
    out = []
    for fr in fractions:
        w = max(1, int(round(full * fr)))
        for fl in flag_sets:
            out.append((w, fl))
    return out


def tune(ctx, run, reps: int = 3, shapes=None, sync=None, tie: float = 0.03) -> dict:
    """Time `run()` under each launch shape (one warm-up, then the best of
    `reps` runs), set the fastest on `ctx` and return
    {"waves", "flags", "ms", "table": [{"waves", "flags", "ms"}, ...]}.
    Shapes within `tie` (relative) of the fastest count as equal and the
    first of them in candidate order (most waves first) is taken: a wave cap
    below the launch's pose count only pays when it is clearly faster, since
    the same launch on a later, costlier population (a GA's converged
    children, fewer cheap clash exits) then runs in two rounds.
    `sync` (default: torch.cuda.synchronize when torch is present) waits for
    the device before and after each timed run."""
    if sync is None:
        try:
            import torch
            sync = torch.cuda.synchronize
        except ImportError:  # pragma: no cover
            sync = lambda: None  # noqa: E731
    shapes = shapes or candidates(ctx)
    table = []
    for waves, flags in shapes:
        ctx.set_launch_shape(waves, flags)
        run()
        sync()
        best = float("inf")
        for _ in range(reps):
            t0 = time.perf_counter()
            run()
            sync()
            best = min(best, time.perf_counter() - t0)
        table.append({"waves": ctx.launch_shape()[0], "flags": flags, "ms": round(best * 1e3, 4)})
    fastest = min(r["ms"] for r in table)
    win = next(r for r in table if r["ms"] <= fastest * (1.0 + tie))
    ctx.set_launch_shape(win["waves"], win["flags"])
    return {"waves": win["waves"], "flags": win["flags"], "ms": win["ms"], "table": table}
