"""Membrane sources: pressure (or normal velocity) injected over a plane patch
with a modal weight per cell (core/solver.py:210-757 of the reference).

A membrane lies in the grid plane nearest its centre, normal to one axis. Its
weight at a cell of that plane is the mode shape's magnitude at the cell
centre, normalised to 1 at the largest one: Bessel modes J_m(alpha_mn r / a)
cos(m theta) for a circular membrane clamped at radius a, products of sines
for a rectangular one. Each step, every air cell with a positive weight gets

    field[c] = float32(float64(field[c]) + w(t_n) * weight[c])

after the step's point and plane sources (core/solver.py:2389-2412), where
w is the membrane's own waveform object sampled at the accumulated time.

The weights are computed here on the host in float64 with the reference's
formulas and numpy/scipy operations; the device applies them
(``k_membrane`` in csrc/fused_step.hip). Only the membrane's plane is
evaluated: every operation is elementwise except the normalising maximum,
which runs over the same cells as in the reference's full-grid evaluation.
"""

from __future__ import annotations

import warnings
from dataclasses import dataclass, field
from typing import Literal

import numpy as np
from numpy.typing import NDArray
from scipy.special import j0, jn_zeros, jv

_AXES = {"x": 0, "y": 1, "z": 2}

# n-th zeros of J_m used without calling jn_zeros (core/solver.py:438-449)
_BESSEL_ZEROS = {
    (0, 1): 2.4048255576957727,
    (0, 2): 5.5200781102863115,
    (1, 1): 3.8317059702075125,
    (1, 2): 7.0155866698156190,
    (2, 1): 5.1356223018406826,
    (2, 2): 8.4172441403998649,
}


def _coords(grid):
    return grid.x_coords, grid.y_coords, grid.z_coords


def _normalise(shape: NDArray) -> NDArray:
    """|shape| scaled so its maximum is 1 (all zeros if the maximum is
    1e-10 or less)."""
    shape = np.abs(shape)
    peak = np.max(shape)
    if peak > 1e-10:
        return shape / peak
    return np.zeros_like(shape)


@dataclass
class MembraneSource:
    """Base of the membrane sources (core/solver.py:210-370).

    center: physical (x, y, z) in metres; normal_axis: 'x', 'y' or 'z';
    waveform: any object with ``.waveform(t, dt)``; mode: mode indices;
    injection_type: 'pressure' or 'velocity' (the velocity component normal
    to the membrane)."""

    center: tuple[float, float, float]
    normal_axis: Literal["x", "y", "z"]
    waveform: object
    mode: tuple[int, int] = (0, 1)
    injection_type: Literal["pressure", "velocity"] = "pressure"
    source_type: Literal["membrane"] = field(default="membrane", init=False)
    _cached_weights: NDArray | None = field(default=None, repr=False, init=False)
    _cached_mask: NDArray | None = field(default=None, repr=False, init=False)

    # ---- subclass interface
    def _plane_shape(self, grid, a: NDArray, b: NDArray) -> tuple[NDArray, NDArray]:
        """(inside mask, raw mode shape) on the membrane plane, given the two
        in-plane offsets from the centre (a, b) as 2D arrays."""
        raise NotImplementedError

    # ---- geometry
    def plane_index(self, grid) -> int:
        """Index of the grid plane nearest the centre along the normal."""
        ax = _AXES[self.normal_axis]
        return int(np.argmin(np.abs(_coords(grid)[ax] - self.center[ax])))

    def _offsets(self, grid) -> tuple[NDArray, NDArray]:
        """In-plane offsets of the plane's cell centres from the membrane
        centre: (x - cx, y - cy) for a z normal, (x - cx, z - cz) for y,
        (y - cy, z - cz) for x; 2D arrays in grid index order."""
        ax = _AXES[self.normal_axis]
        (u_ax, v_ax) = [d for d in range(3) if d != ax]
        cu = _coords(grid)[u_ax]
        cv = _coords(grid)[v_ax]
        uu, vv = np.meshgrid(cu, cv, indexing="ij")
        return uu - self.center[u_ax], vv - self.center[v_ax]

    def _check_grid_alignment(self, grid) -> None:
        """Warn when the centre is more than half a cell off its plane
        (core/solver.py:338-369)."""
        ax = _AXES[self.normal_axis]
        coords = _coords(grid)[ax]
        idx = self.plane_index(grid)
        offset = abs(coords[idx] - self.center[ax])
        d = (grid.dx, grid.dy, grid.dz)[ax]
        spacing = d[idx] if hasattr(d, "__getitem__") else grid.min_spacing
        if offset > 0.5 * spacing:
            warnings.warn(
                f"MembraneSource center {self.center} is {offset:.4f}m "
                f"({offset / spacing:.1f} cells) off-grid in {self.normal_axis}-direction. "
                f"Source will be placed at nearest grid plane.", stacklevel=3)

    def _embed(self, grid, plane: NDArray, dtype) -> NDArray:
        out = np.zeros(grid.shape, dtype=dtype)
        idx = self.plane_index(grid)
        sl = [slice(None)] * 3
        sl[_AXES[self.normal_axis]] = idx
        out[tuple(sl)] = plane
        return out

    def get_injection_mask(self, grid) -> NDArray[np.bool_]:
        """True on the membrane's cells of its plane."""
        a, b = self._offsets(grid)
        inside, _ = self._plane_shape(grid, a, b)
        return self._embed(grid, inside, bool)

    def get_injection_weights(self, grid) -> NDArray[np.float64]:
        """Mode-shape weight per cell (0 off the membrane), float64."""
        self._check_grid_alignment(grid)
        a, b = self._offsets(grid)
        inside, raw = self._plane_shape(grid, a, b)
        plane = np.zeros(inside.shape, dtype=np.float64)
        plane[inside] = _normalise(raw[inside])
        return self._embed(grid, plane, np.float64)

    def injection_field(self) -> int:
        """0 for p, else 1 + the normal axis (vx, vy, vz)."""
        return 0 if self.injection_type == "pressure" else 1 + _AXES[self.normal_axis]


@dataclass(kw_only=True)
class CircularMembraneSource(MembraneSource):
    """Circular membrane of radius `radius` (metres) clamped at its edge:
    weight |J_m(alpha_mn r/a) cos(m theta)|, normalised (core/solver.py:372-556).
    mode = (m, n): m >= 0 azimuthal, n >= 1 radial."""

    radius: float
    _alpha: float = field(default=0.0, repr=False, init=False)

    def __post_init__(self) -> None:
        m, n = self.mode
        if m < 0:
            raise ValueError(f"Azimuthal mode m must be non-negative, got {m}")
        if n < 1:
            raise ValueError(f"Radial mode n must be positive, got {n}")
        self._alpha = _BESSEL_ZEROS[(m, n)] if (m, n) in _BESSEL_ZEROS else float(jn_zeros(m, n)[-1])

    def mode_shape(self, r: NDArray, theta: NDArray) -> NDArray:
        """Normalised |Bessel mode| at radii r (metres) and angles theta."""
        m, _ = self.mode
        rho = r / self.radius
        if m == 0:
            shape = j0(self._alpha * rho)
        else:
            shape = jv(m, self._alpha * rho) * np.cos(m * theta)
        return _normalise(np.where(rho <= 1.0, shape, 0.0))

    def _plane_shape(self, grid, a, b):
        r = np.sqrt(a**2 + b**2)
        theta = np.arctan2(b, a)
        inside = r <= self.radius
        m, _ = self.mode
        rho = r / self.radius
        if m == 0:
            raw = j0(self._alpha * rho)
        else:
            raw = jv(m, self._alpha * rho) * np.cos(m * theta)
        return inside, np.where(rho <= 1.0, raw, 0.0)


@dataclass(kw_only=True)
class RectangularMembraneSource(MembraneSource):
    """Rectangular membrane `size` = (width, height) in metres in its plane,
    clamped at all edges: weight |sin(m pi x') sin(n pi y')| with x', y' in
    [0, 1] across it, normalised (core/solver.py:558-757). The in-plane axes
    are (x, y) for a z normal, (x, z) for y, (y, z) for x."""

    size: tuple[float, float]
    mode: tuple[int, int] = (1, 1)

    def __post_init__(self) -> None:
        m, n = self.mode
        if m < 1:
            raise ValueError(f"Mode index m must be >= 1, got {m}")
        if n < 1:
            raise ValueError(f"Mode index n must be >= 1, got {n}")

    def mode_shape(self, u: NDArray, v: NDArray) -> NDArray:
        """Normalised |sin sin| at in-plane offsets u, v (metres)."""
        inside, raw = self._raw(u, v)
        return _normalise(np.where(inside, raw, 0.0))

    def _raw(self, u, v):
        m, n = self.mode
        lx, ly = self.size
        xn = u / lx + 0.5
        yn = v / ly + 0.5
        raw = np.sin(m * np.pi * xn) * np.sin(n * np.pi * yn)
        inside = (np.abs(u) <= lx / 2) & (np.abs(v) <= ly / 2)
        return inside, raw

    def _plane_shape(self, grid, a, b):
        inside, raw = self._raw(a, b)
        return inside, np.where(inside, raw, 0.0)
