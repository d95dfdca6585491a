"""Launch-shape selection for the fused step kernel.

A launch is described by (rows, planes, stack): a wave owns a 64-wide k
segment of ``rows`` consecutive j rows and marches ``planes`` planes along i;
``stack`` waves of one workgroup own consecutive j-strips and share their halo
rows through LDS (1: independent waves). The
heuristic below picks from measurements: short 2-row strips (many waves) for
solves that fit the Infinity Cache, 4 rows x 4 planes beyond it. The
autotuner (``strata_fdtd_amd.autotune``) measures candidates on the device and
caches the winner per grid shape; ``pick_tiling`` consults that cache first.
"""

from __future__ import annotations

import json
import os

# solves up to this many cells (33 B each: four float32 fields and a flag byte)
# fit the 256 MiB Infinity Cache
CACHE_RESIDENT_CELLS = (256 << 20) // 33
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
    """Measured on MI355X (air boxes, device time per step, best of rows
    2/4 x planes 1/2/4/8):

    * a solve whose fields and flags (33 B/cell) fit the 256 MiB Infinity
      Cache runs fastest with 2-row strips 2 planes long, which give the chip
      several times more waves: 192^3 34.8 us per step (4 rows, 4 planes:
      43.1), 176^3 29.4 (34.2), 128^3 15.0 (20.2), 100^3 9.5 (11.4), 80^3 7.4
      (9.9); grids too small for 2048 such waves take 1 plane (64^3: 5.4 us
      against 9.0);
    * beyond the cache, 4 rows and 4 planes: at 512^3 planes 4 is 1% faster
      than 8 and 3% faster than 16, and rows 2 and LDS stacks are slower;
      224^3 and 256^3 are within 1% of their best there."""
    nx, ny, nz = shape
    segs = (nz + 63) // 64
    if nx * ny * nz <= CACHE_RESIDENT_CELLS:
        rows = 2
        planes = 2 if segs * ((ny + 1) // 2) * ((nx + 1) // 2) >= 2048 else 1
        return rows, planes, STACK_DEFAULT
    return 4, 4, STACK_DEFAULT


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
