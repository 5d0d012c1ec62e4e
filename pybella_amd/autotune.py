"""Tile-shape autotuner for the 2D solver launches.

The stencil launches tile the grid into 64 x (4 * rows_per_thread) node tiles
(one 256-thread block each).  Taller tiles re-read fewer halo rows and emit
fewer block partials; shorter tiles put more blocks in flight per CU; and the
block count decides whether the grid fits the CUs' block slots in one wave or
leaves a short second wave (a tail).  The best height depends on the grid
size, precision and register use, so it is measured: each
candidate runs the caller's timing function on the live operator and the
fastest wins.  Results are cached per (nx, ny, mode) for the process.
"""
from __future__ import annotations

import time

_cache: dict = {}

CANDIDATES = (2, 3, 4, 5, 6, 8)


def tune_tile(op, run, key=None, reps=1):
    """``run()`` performs one timed unit of work on ``op``; returns (best
    rows_per_thread, {rows: seconds})."""
    import torch
    key = key or (op.nx, op.ny)
    if key in _cache:
        best, table = _cache[key]
        op.set_tile(best)
        return best, table
    table = {}
    for ry in CANDIDATES:
        op.set_tile(ry)
        run()                       # warm (graph capture, first-touch)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            run()
        torch.cuda.synchronize()
        table[ry] = (time.perf_counter() - t0) / reps
    best = min(table, key=table.get)
    op.set_tile(best)
    _cache[key] = (best, table)
    return best, table
