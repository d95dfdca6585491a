"""Autotuner for the fused step's launch shape (rows per wave, planes per wave).

For one grid shape it times each candidate over a few device steps (HIP events
on the solver's stream, after one warm-up launch), keeps the fastest and
stores it in the tiling cache that ``tuning.pick_tiling`` reads. Timing runs
real steps on the solver (which keep advancing it), so tune on a scratch
solver or call ``solver.reset()`` afterwards.
"""

from __future__ import annotations

import itertools

from . import _native
from .tuning import ROW_CHOICES, store_tiling


def candidate_tilings(shape, rows=ROW_CHOICES, planes=(4, 8, 12, 16, 24, 32, 48, 64)):
    nx = shape[0]
    out = []
    for r, p in itertools.product(rows, planes):
        if p <= nx and r <= max(2, shape[1]):
            out.append((r, p))
    return out


def time_tiling(solver, rows: int, planes: int, steps: int = 10) -> float:
    """Milliseconds per step for one launch shape (device time)."""
    import ctypes as C
    solver.set_tiling(rows, planes)
    solver._prepare(steps + 2)
    _native.call("fdtd_step_n", solver._ctx, 1)        # warm-up (code object load)
    solver._advance_host(1)
    ms = C.c_float()
    _native.call("fdtd_time_steps", solver._ctx, steps, 0, C.byref(ms))
    solver._advance_host(steps)
    return ms.value / steps


def autotune(solver, candidates=None, steps: int = 10, store: bool = True, verbose: bool = False):
    """Pick the fastest (rows, planes) for this solver's shape; returns (best, table)."""
    shape = solver.shape
    cands = candidates or candidate_tilings(shape)
    table = []
    for r, p in cands:
        ms = time_tiling(solver, r, p, steps)
        table.append((r, p, ms))
        if verbose:
            print(f"[autotune] {shape} rows={r} planes={p}: {ms * 1e3:.1f} us/step", flush=True)
    best = min(table, key=lambda t: t[2])
    solver.set_tiling(best[0], best[1])
    if store:
        store_tiling(shape, not solver.grid.is_uniform, bool(solver._ade_tensors), best[0], best[1], best[2])
    return (best[0], best[1]), table
