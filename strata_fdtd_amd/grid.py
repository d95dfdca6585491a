"""Grid specifications (uniform and nonuniform cell spacing).

Same classes, attributes and arithmetic as the reference's ``core/grid.py``
(UniformGrid :50-145, NonuniformGrid :148-540): the 1D spacing tables built
here are what the MI355X kernels read for nonuniform grids, so they follow the
reference formulas operation by operation (float64, then float32 at the end).
"""

from __future__ import annotations

from collections.abc import Sequence
from dataclasses import dataclass

import numpy as np
from numpy.typing import NDArray


@dataclass
class UniformGrid:
    """Uniform grid: one spacing for all axes (core/grid.py:50-145)."""

    shape: tuple[int, int, int]
    resolution: float

    def __post_init__(self):
        nx, ny, nz = self.shape
        self.shape = (int(nx), int(ny), int(nz))
        r = self.resolution
        self._dx = np.full(nx, r, dtype=np.float64)
        self._dy = np.full(ny, r, dtype=np.float64)
        self._dz = np.full(nz, r, dtype=np.float64)
        self._x_coords = np.arange(nx) * r + r / 2
        self._y_coords = np.arange(ny) * r + r / 2
        self._z_coords = np.arange(nz) * r + r / 2

    dx = property(lambda self: self._dx)
    dy = property(lambda self: self._dy)
    dz = property(lambda self: self._dz)
    x_coords = property(lambda self: self._x_coords)
    y_coords = property(lambda self: self._y_coords)
    z_coords = property(lambda self: self._z_coords)

    @property
    def min_spacing(self) -> float:
        return self.resolution

    @property
    def max_spacing(self) -> float:
        return self.resolution

    @property
    def is_uniform(self) -> bool:
        return True

    def physical_extent(self) -> tuple[float, float, float]:
        return tuple(n * self.resolution for n in self.shape)


class NonuniformGrid:
    """Grid with per-axis cell-centre coordinates (core/grid.py:148-540)."""

    def __init__(self, x_coords, y_coords, z_coords):
        self._x_coords = np.asarray(x_coords, dtype=np.float64).ravel()
        self._y_coords = np.asarray(y_coords, dtype=np.float64).ravel()
        self._z_coords = np.asarray(z_coords, dtype=np.float64).ravel()
        for name, c in (("x", self._x_coords), ("y", self._y_coords), ("z", self._z_coords)):
            if len(c) < 2:
                raise ValueError(f"{name}_coords must have at least 2 points")
            if not np.all(np.diff(c) > 0):
                raise ValueError(f"{name}_coords must be monotonically increasing")
        self._dx = self._compute_spacing(self._x_coords)
        self._dy = self._compute_spacing(self._y_coords)
        self._dz = self._compute_spacing(self._z_coords)
        self._shape = (len(self._x_coords), len(self._y_coords), len(self._z_coords))

    @staticmethod
    def _compute_spacing(coords: NDArray[np.float64]) -> NDArray[np.float64]:
        """spacing[i] = (c[i+1] - c[i-1]) / 2 inside, one-sided at both ends (core/grid.py:219-237)."""
        n = len(coords)
        spacing = np.zeros(n, dtype=np.float64)
        if n > 2:
            spacing[1:-1] = (coords[2:] - coords[:-2]) / 2
        spacing[0] = coords[1] - coords[0]
        spacing[-1] = coords[-1] - coords[-2]
        return spacing

    shape = property(lambda self: self._shape)
    dx = property(lambda self: self._dx)
    dy = property(lambda self: self._dy)
    dz = property(lambda self: self._dz)
    x_coords = property(lambda self: self._x_coords)
    y_coords = property(lambda self: self._y_coords)
    z_coords = property(lambda self: self._z_coords)

    @property
    def min_spacing(self) -> float:
        return float(min(np.min(self._dx), np.min(self._dy), np.min(self._dz)))

    @property
    def max_spacing(self) -> float:
        return float(max(np.max(self._dx), np.max(self._dy), np.max(self._dz)))

    @property
    def is_uniform(self) -> bool:
        return False

    @property
    def stretch_ratio(self) -> tuple[float, float, float]:
        return tuple(float(np.max(d) / np.min(d)) for d in (self._dx, self._dy, self._dz))

    def physical_extent(self) -> tuple[float, float, float]:
        out = []
        for c, d in ((self._x_coords, self._dx), (self._y_coords, self._dy), (self._z_coords, self._dz)):
            out.append(float(c[-1] - c[0] + d[-1] / 2 + d[0] / 2))
        return tuple(out)

    @classmethod
    def from_stretch(cls, shape, base_resolution, stretch_x=1.0, stretch_y=1.0, stretch_z=1.0,
                     center_fine=True) -> NonuniformGrid:
        """Geometric stretch per axis (core/grid.py:324-365)."""
        nx, ny, nz = shape
        return cls(
            x_coords=cls._generate_stretched_coords(nx, base_resolution, stretch_x, center_fine),
            y_coords=cls._generate_stretched_coords(ny, base_resolution, stretch_y, center_fine),
            z_coords=cls._generate_stretched_coords(nz, base_resolution, stretch_z, center_fine),
        )

    @staticmethod
    def _generate_stretched_coords(n, base_dx, stretch, center_fine) -> NDArray[np.float64]:
        """Cell centres growing geometrically from the centre or from 0 (core/grid.py:367-410)."""
        if stretch == 1.0:
            return np.arange(n, dtype=np.float64) * base_dx + base_dx / 2
        if not center_fine:
            sizes = base_dx * (stretch ** np.arange(n, dtype=np.float64))
            return np.cumsum(sizes) - sizes / 2
        n_half = n // 2
        sizes = base_dx * (stretch ** np.arange(n_half, dtype=np.float64))
        positions = np.cumsum(sizes) - sizes[0] / 2
        left = -positions[::-1] - sizes[::-1] / 2
        if n % 2 == 0:
            coords = np.concatenate([left, positions + sizes / 2])
        else:
            right = positions[1:] + sizes[1:] / 2 if n_half > 0 else np.array([])
            coords = np.concatenate([left, np.array([0.0]), right])
        return coords - coords[0] + base_dx / 2

    @classmethod
    def from_regions(cls, x_regions: Sequence, y_regions: Sequence | None = None,
                     z_regions: Sequence | None = None) -> NonuniformGrid:
        """Piecewise-constant resolution regions (core/grid.py:412-496)."""
        y_regions = x_regions if y_regions is None else y_regions
        z_regions = x_regions if z_regions is None else z_regions
        return cls(cls._generate_region_coords(x_regions), cls._generate_region_coords(y_regions),
                   cls._generate_region_coords(z_regions))

    @staticmethod
    def _generate_region_coords(regions) -> NDArray[np.float64]:
        if not regions:
            raise ValueError("At least one region must be specified")
        coords: list[float] = []
        for idx, (start, end, resolution) in enumerate(regions):
            if end <= start:
                raise ValueError(f"Region {idx}: end ({end}) must be > start ({start})")
            if resolution <= 0:
                raise ValueError(f"Region {idx}: resolution must be positive")
            if idx > 0 and abs(start - regions[idx - 1][1]) > 1e-10:
                raise ValueError(f"Regions must be contiguous: region {idx-1} ends at "
                                 f"{regions[idx - 1][1]}, region {idx} starts at {start}")
            n_cells = max(1, int(round((end - start) / resolution)))
            actual = (end - start) / n_cells
            rc = start + (np.arange(n_cells) + 0.5) * actual
            if coords and len(rc) > 0 and abs(rc[0] - coords[-1]) < actual / 2:
                rc = rc[1:]
            coords.extend(rc.tolist())
        return np.array(coords, dtype=np.float64)

    def get_spacing_arrays_for_stencil(self) -> dict[str, NDArray[np.float32]]:
        """float32 inverse spacings for the stencil (core/grid.py:498-532).

        inv_d*_face[i] = 1/(c[i+1]-c[i]) (n-1 entries, velocity faces) and
        inv_d*_cell[i] = 1/spacing[i] (n entries, pressure cells), formed in
        float64 and rounded once to float32.
        """
        out = {}
        for name, c, d in (("x", self._x_coords, self._dx), ("y", self._y_coords, self._dy),
                           ("z", self._z_coords, self._dz)):
            out[f"inv_d{name}_face"] = (1.0 / np.diff(c)).astype(np.float32)
            out[f"inv_d{name}_cell"] = (1.0 / d).astype(np.float32)
        return out

    def __repr__(self) -> str:
        return (f"NonuniformGrid(shape={self.shape}, min_spacing={self.min_spacing:.4g}, "
                f"max_spacing={self.max_spacing:.4g}, "
                f"stretch_ratio={tuple(f'{r:.2f}' for r in self.stretch_ratio)})")


def explicit_core_stretch_coords(n: int, core: int, base_dx: float, ratio: float) -> NDArray[np.float64]:
    """Cell centres with `core` uniform cells of base_dx centred on the axis and
    (n - core)/2 cells on each side growing by `ratio` per cell (the first outer
    cell is base_dx*ratio). ``from_stretch`` cannot express a uniform core, so
    the 320x320x480 nonuniform workload builds its coordinates with this.
    """
    if (n - core) % 2:
        raise ValueError("n - core must be even")
    side = (n - core) // 2
    outer = base_dx * ratio ** np.arange(1, side + 1, dtype=np.float64)
    sizes = np.concatenate([outer[::-1], np.full(core, base_dx), outer])
    edges = np.concatenate([[0.0], np.cumsum(sizes)])
    return (edges[:-1] + edges[1:]) / 2
