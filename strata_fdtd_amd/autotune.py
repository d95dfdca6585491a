"""Autotuner for the fused step's launch shape (rows per wave, planes per wave,
waves per LDS-sharing j-stack).

For one grid shape it times each candidate over a few device steps (HIP events
on the solver's stream, after one warm-up launch), keeps the fastest and
stores it in the tiling cache that ``tuning.pick_tiling`` reads. Timing runs
real steps on the solver (which keep advancing it), so tune on a scratch
solver or call ``solver.reset()`` afterwards.
"""

from __future__ import annotations

import itertools

from . import _native
from .tuning import ROW_CHOICES, STACK_CHOICES, normalize_tiling, store_tiling


def candidate_tilings(shape, rows=ROW_CHOICES, planes=(1, 2, 4, 8, 12, 16, 24, 32, 48, 64), stacks=STACK_CHOICES):
    nx, ny = shape[0], shape[1]
    out = []
    for r, p, w in itertools.product(rows, planes, stacks):
        if p <= nx and r <= max(2, ny) and (w == 1 or r * w <= ny):
            out.append((r, p, w))
    return out


def time_tiling(solver, tiling, steps: int = 10) -> float:
    """Milliseconds per step for one launch shape (device time)."""
    import ctypes as C
    solver.set_tiling(*normalize_tiling(tiling))
    solver._prepare(steps + 2)
    _native.call("fdtd_step_n", solver._ctx, 1)        # warm-up (code object load)
    solver._advance_host(1)
    ms = C.c_float()
    _native.call("fdtd_time_steps", solver._ctx, steps, 0, C.byref(ms))
    solver._advance_host(steps)
    return ms.value / steps


def autotune(solver, candidates=None, steps: int = 10, store: bool = True, verbose: bool = False):
    """Pick the fastest (rows, planes, stack) for this solver's shape; returns
    (best, table) with table rows (rows, planes, stack, ms)."""
    shape = solver.shape
    cands = [normalize_tiling(t) for t in (candidates or candidate_tilings(shape))]
    table = []
    for t in cands:
        ms = time_tiling(solver, t, steps)
        table.append((*t, ms))
        if verbose:
            print(f"[autotune] {shape} rows={t[0]} planes={t[1]} stack={t[2]}: {ms * 1e3:.1f} us/step", flush=True)
    best = min(table, key=lambda e: e[3])
    solver.set_tiling(*best[:3])
    if store:
        store_tiling(shape, not solver.grid.is_uniform, bool(solver._ade_tensors), best[:3], best[3])
    return tuple(best[:3]), table
