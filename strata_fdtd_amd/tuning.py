"""Launch-shape selection for the fused step kernel.

A launch is described by (rows, planes, stack): a wave owns a 64-wide k
segment of ``rows`` consecutive j rows and marches ``planes`` planes along i;
``stack`` waves of one workgroup own consecutive j-strips and share their halo
rows through LDS (1: independent waves). The
heuristic below gives each of the 256 CUs several resident waves while keeping
the per-strip prologue (one extra plane) and the j-halo rows small. The
autotuner (``strata_fdtd_amd.autotune``) measures candidates on the device and
caches the winner per grid shape; ``pick_tiling`` consults that cache first.
"""

from __future__ import annotations

import json
import os

NUM_CU = 256
ROW_CHOICES = (2, 4, 8, 16)
STACK_CHOICES = (1, 2, 4, 8)
STACK_DEFAULT = 1
_CACHE_ENV = "STRATA_FDTD_AMD_TUNE_CACHE"
_DEFAULT_CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lib", "tune_cache.json")


def cache_path() -> str:
    return os.environ.get(_CACHE_ENV, _DEFAULT_CACHE)


def _load_cache() -> dict:
    try:
        with open(cache_path()) as fh:
            return json.load(fh)
    except (OSError, ValueError):
        return {}


def shape_key(shape, nonuniform: bool, ade: bool) -> str:
    return f"{shape[0]}x{shape[1]}x{shape[2]}:{'nu' if nonuniform else 'u'}:{'ade' if ade else 'plain'}"


def normalize_tiling(t) -> tuple[int, int, int]:
    """(rows, planes) or (rows, planes, stack) -> (rows, planes, stack)."""
    t = tuple(int(v) for v in t)
    if len(t) == 2:
        return t[0], t[1], 1
    if len(t) == 3:
        return t
    raise ValueError(f"tiling must be (rows, planes) or (rows, planes, stack), got {t}")


def heuristic_tiling(shape) -> tuple[int, int, int]:
    """rows=4; the most planes per strip (<= 4) that still leaves >= 8 waves
    per CU. Measured at 512^3 (plain and enclosure): planes 4 is 1% faster
    than 8 and 3% faster than 16; rows 2 and LDS stacks are slower."""
    nx, ny, nz = shape
    rows = 4
    segs = (nz + 63) // 64
    strips_ji = segs * ((ny + rows - 1) // rows)
    planes = 4
    while planes > 2 and strips_ji * ((nx + planes - 1) // planes) < 8 * NUM_CU:
        planes //= 2
    return rows, planes, STACK_DEFAULT


def pick_tiling(shape, nonuniform: bool = False, ade: bool = False) -> tuple[int, int, int]:
    hit = _load_cache().get(shape_key(shape, nonuniform, ade))
    if hit:
        # entries are [rows, planes, stack, ms]
        if len(hit) == 4:
            return int(hit[0]), int(hit[1]), int(hit[2])
    return heuristic_tiling(shape)


def store_tiling(shape, nonuniform: bool, ade: bool, tiling, ms: float) -> None:
    rows, planes, stack = normalize_tiling(tiling)
    cache = _load_cache()
    cache[shape_key(shape, nonuniform, ade)] = [rows, planes, stack, ms]
    path = cache_path()
    os.makedirs(os.path.dirname(path), exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as fh:
        json.dump(cache, fh, indent=1, sort_keys=True)
    os.replace(tmp, path)
