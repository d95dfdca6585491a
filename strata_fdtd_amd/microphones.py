"""Microphones: trilinear pressure (and staggered-velocity) gathers recorded on
the device every step, with the reference's Microphone API
(core/solver.py:787-1440: Microphone; add_microphone :1887-1971;
_record_microphones :2441-2518).

Device side (``k_mic_gather``, one thread per microphone, after each step):
for every microphone four 8-corner gathers (p, vx, vy, vz) with float32
weights, summed in float32 in the reference's corner order. Host side: the
gather tables, and the combination of the four recorded values into the
microphone output with the reference's dtype rules:

* a Python float times a numpy float32 scalar is a float32 product of the
  float32-rounded weight (numpy 2 weak scalars), so for positions given as
  Python floats the device sums float32 products of float32 weights; for
  positions given as numpy float64 the weights are float64 scalars and the
  reference's sums are float64, which the device reproduces in double;
* directional output (1 - k) P + k Z (v . d): the first term is float32, the
  dot product and the second term are float64 (float32 times a float64
  direction component), the sum is float64, and ``get_waveform`` rounds to
  float32;
* custom patterns call the user's function on theta = arccos(clip(v.d/|v|))
  per sample, exactly as the reference does per step.

Velocity gathers follow ``_interpolate_velocity`` (:1060-1113): the grid
position shifted by -1/2 along the component's axis, clamped to
[0, n - 1.001], corner indices clamped to n - 1.
"""

from __future__ import annotations

from collections.abc import Callable

import numpy as np
from numpy.typing import NDArray

POLAR_PATTERNS: dict[str, float] = {
    "omni": 0.0,
    "subcardioid": 0.3,
    "cardioid": 0.5,
    "supercardioid": 0.63,
    "hypercardioid": 0.75,
    "figure8": 1.0,
}


def _trilinear(gx: float, gy: float, gz: float):
    i0, j0, k0 = int(gx), int(gy), int(gz)
    fx, fy, fz = gx - i0, gy - j0, gz - k0
    ax, ay, az = 1 - fx, 1 - fy, 1 - fz
    corners = [(i0, j0, k0), (i0 + 1, j0, k0), (i0, j0 + 1, k0), (i0 + 1, j0 + 1, k0),
               (i0, j0, k0 + 1), (i0 + 1, j0, k0 + 1), (i0, j0 + 1, k0 + 1), (i0 + 1, j0 + 1, k0 + 1)]
    weights = [ax * ay * az, fx * ay * az, ax * fy * az, fx * fy * az,
               ax * ay * fz, fx * ay * fz, ax * fy * fz, fx * fy * fz]
    return corners, weights


class Microphone:
    """Virtual microphone at a physical position (metres), trilinear pickup,
    omni / cardioid family / figure-8 / custom polar pattern (see the module
    docstring for the recording arithmetic)."""

    def __init__(self, position: tuple[float, float, float], name: str | None = None,
                 pattern: str | Callable[[float], float] = "omni",
                 direction: tuple[float, float, float] | None = None,
                 up: tuple[float, float, float] | None = None):
        self.position = position
        self.name = name
        if isinstance(pattern, str):
            if pattern not in POLAR_PATTERNS:
                raise ValueError(f"Unknown pattern '{pattern}'. Valid patterns: {list(POLAR_PATTERNS.keys())}")
            self._pattern_name = pattern
            self._pattern_k = POLAR_PATTERNS[pattern]
            self._custom_pattern = None
        elif callable(pattern):
            self._pattern_name = "custom"
            self._pattern_k = None
            self._custom_pattern = pattern
        else:
            raise TypeError(f"pattern must be str or callable, got {type(pattern).__name__}")
        self.pattern = pattern
        d = np.array((1.0, 0.0, 0.0) if direction is None else direction, dtype=np.float64)
        dn = np.linalg.norm(d)
        if dn < 1e-10:
            raise ValueError("direction vector cannot be zero")
        self._direction = d / dn
        u = np.array((0.0, 0.0, 1.0) if up is None else up, dtype=np.float64)
        un = np.linalg.norm(u)
        if un < 1e-10:
            raise ValueError("up vector cannot be zero")
        self._up = u / un
        if abs(np.dot(self._direction, self._up)) > 0.999:
            raise ValueError("direction and up vectors cannot be parallel")
        self._data: list = []
        self._times: list[float] = []
        self._solver_dt: float | None = None
        self._solver_rho: float | None = None
        self._solver_c: float | None = None
        self._grid_position: tuple[float, float, float] | None = None
        self._interpolation_weights: dict | None = None
        self._solver = None  # set by FDTDSolver.add_microphone: device-recorded samples live there

    @property
    def direction(self) -> tuple[float, float, float]:
        return tuple(self._direction)

    def is_directional(self) -> bool:
        return self._pattern_name != "omni"

    def _initialize(self, solver) -> None:
        self._solver_dt = solver.dt
        self._solver_rho = solver.rho
        self._solver_c = solver.c
        gx = self.position[0] / solver.dx
        gy = self.position[1] / solver.dx
        gz = self.position[2] / solver.dx
        self._grid_position = (gx, gy, gz)
        nx, ny, nz = solver.shape
        if not (0 <= gx < nx - 1 and 0 <= gy < ny - 1 and 0 <= gz < nz - 1):
            raise ValueError(
                f"Microphone position {self.position} is outside simulation domain. Valid range: (0, 0, 0) to "
                f"({(nx - 1) * solver.dx:.4f}, {(ny - 1) * solver.dx:.4f}, {(nz - 1) * solver.dx:.4f})")
        corners, weights = _trilinear(gx, gy, gz)
        self._interpolation_weights = {"indices": corners, "weights": weights}

    # ---------------------------------------------------------- device tables
    def _velocity_corners(self, axis: int, shape):
        g = list(self._grid_position)
        g[axis] -= 0.5
        g = [max(0.0, min(v, n - 1.001)) for v, n in zip(g, shape)]
        corners, weights = _trilinear(*g)
        corners = [tuple(min(c, n - 1) for c, n in zip(cc, shape)) for cc in corners]
        return corners, weights

    @staticmethod
    def _sum_in_float64(weights) -> bool:
        """The reference sums ``w * field`` (field float32) in the weights' type:
        numpy float64 weights give float64 sums; Python floats (weak scalars)
        and float32 weights give float32 sums of the float32-rounded weight."""
        kinds = {isinstance(w, np.float64) for w in weights}
        if len(kinds) != 1:
            raise NotImplementedError("mixed float64/float32 interpolation weights")
        if any(isinstance(w, np.floating) and np.dtype(type(w)).itemsize > 8 for w in weights):
            raise NotImplementedError("extended-precision microphone positions")
        return kinds.pop()

    def gather_table(self, shape) -> tuple[NDArray[np.int32], NDArray[np.float64], NDArray[np.int32]]:
        """Flat cell indices int32 [4][8], weights float64 [4][8] and the
        float64-sum flags int32 [4] of the p, vx, vy, vz gathers (a float32
        gather's weights are stored already rounded to float32)."""
        if self._interpolation_weights is None:
            raise RuntimeError("Microphone not initialized. Add to solver first.")
        rows = [(self._interpolation_weights["indices"], self._interpolation_weights["weights"])]
        rows += [self._velocity_corners(ax, shape) for ax in range(3)]
        idx = np.array([[np.ravel_multi_index(c, shape) for c in corners] for corners, _ in rows], dtype=np.int64)
        if idx.max() >= 2**31:
            raise ValueError("microphone gathers need flat indices below 2**31")
        f64 = np.array([self._sum_in_float64(weights) for _, weights in rows], dtype=np.int32)
        w = np.array([[float(x) if wide else float(np.float32(x)) for x in weights]
                      for (_, weights), wide in zip(rows, f64)], dtype=np.float64)
        return idx.astype(np.int32), w, f64

    # -------------------------------------------------------------- recording
    def _combine(self, rec: NDArray[np.float64], f64) -> list:
        """Microphone outputs from device records [n][4] (p, vx, vy, vz, each
        a float32 or float64 sum as `f64` says), with the reference's scalar
        types and arithmetic."""
        vals = [rec[:, g] if f64[g] else rec[:, g].astype(np.float32) for g in range(4)]
        p = vals[0]
        if not self.is_directional():
            return list(p)
        if self._custom_pattern is None:
            k = self._pattern_k
            Z = self._solver_rho * self._solver_c
            d = self._direction
            v_dot_d = vals[1] * d[0] + vals[2] * d[1] + vals[3] * d[2]
            return list((1 - k) * p + k * Z * v_dot_d)
        return [self._compute_directional_output(a, b, c, e) for a, b, c, e in zip(*vals)]

    def _compute_directional_output(self, pressure, vel_x, vel_y, vel_z):
        Z = self._solver_rho * self._solver_c
        v_dot_d = vel_x * self._direction[0] + vel_y * self._direction[1] + vel_z * self._direction[2]
        if self._custom_pattern is not None:
            v_mag = np.sqrt(vel_x**2 + vel_y**2 + vel_z**2)
            if v_mag > 1e-20:
                theta = np.arccos(np.clip(v_dot_d / v_mag, -1.0, 1.0))
            else:
                theta = 0.0
            return pressure * self._custom_pattern(theta)
        k = self._pattern_k
        return (1 - k) * pressure + k * Z * v_dot_d

    def _interp(self, field, corners, weights):
        acc = 0.0
        for (i, j, k), w in zip(corners, weights, strict=True):
            acc += w * field[i, j, k]
        return acc

    def record(self, pressure_field, time: float, vx=None, vy=None, vz=None) -> None:
        """Record one sample from host arrays (the reference's manual path,
        core/solver.py:1004-1058); the solver records on the device."""
        if self._interpolation_weights is None:
            raise RuntimeError("Microphone not initialized. Add to solver first.")
        self._sync()
        w = self._interpolation_weights
        pressure = self._interp(pressure_field, w["indices"], w["weights"])
        if not self.is_directional():
            self._data.append(pressure)
            self._times.append(time)
            return
        if vx is None or vy is None or vz is None:
            raise ValueError("Velocity fields (vx, vy, vz) are required for directional microphones")
        vel = [self._interp(f, *self._velocity_corners(ax, f.shape)) for ax, f in enumerate((vx, vy, vz))]
        self._data.append(self._compute_directional_output(pressure, *vel))
        self._times.append(time)

    def _sync(self) -> None:
        if self._solver is not None:
            self._solver._sync_microphones()

    def get_waveform(self, weighting=None) -> NDArray[np.floating]:
        """Recorded output, float32; weighting "A" or "C" filters it with the
        IEC 61672-1 weighting at the solver's sample rate (core/solver.py:
        1175-1205)."""
        self._sync()
        waveform = np.array(self._data, dtype=np.float32)
        if weighting in ("Z", None) or len(waveform) == 0:
            return waveform
        from .levels import apply_weighting
        return apply_weighting(waveform, self.get_sample_rate(), weighting).astype(np.float32)

    def _levels_input(self, weighting):
        if len(self) == 0:
            raise RuntimeError("No data recorded. Run simulation first.")
        return self.get_waveform(weighting=weighting)

    def get_spl(self, weighting=None, p_ref: float = 2e-5) -> float:
        """20 log10(p_rms / p_ref) of the (weighted) recording (core/solver.py:1226-1258)."""
        from .levels import spl
        return spl(self._levels_input(weighting), p_ref)

    def get_leq(self, weighting=None, duration: float | None = None, p_ref: float = 2e-5) -> float:
        """Equivalent level of the most recent `duration` seconds, or of the
        whole recording (core/solver.py:1260-1296)."""
        from .levels import leq
        return leq(self._levels_input(weighting), self.get_sample_rate(), duration, p_ref)

    def get_time_weighted_level(self, weighting=None, time_constant: str = "fast",
                                p_ref: float = 2e-5) -> NDArray[np.floating]:
        """Fast / slow / impulse exponentially averaged level per sample
        (core/solver.py:1298-1335)."""
        from .levels import time_weighted_level
        return time_weighted_level(self._levels_input(weighting), self.get_sample_rate(), time_constant, p_ref)

    def get_time_axis(self) -> NDArray[np.floating]:
        self._sync()
        return np.array(self._times, dtype=np.float64)

    def get_sample_rate(self) -> float:
        if self._solver_dt is None:
            raise RuntimeError("Microphone not initialized. Add to solver first.")
        return 1.0 / self._solver_dt

    def to_wav(self, filepath: str, sample_rate: int = 44100, normalize: bool = True, bit_depth: int = 16) -> None:
        """Mono WAV export (core/solver.py:1337-1392): linear resampling,
        optional peak normalisation to 0.95, 16- or 32-bit integers."""
        import wave
        self._sync()
        if len(self._data) == 0:
            raise RuntimeError("No data recorded. Run simulation first.")
        if bit_depth not in (16, 32):
            raise ValueError("bit_depth must be 16 or 32")
        waveform = self.get_waveform()
        if self._solver_dt is not None:
            native_sr = 1.0 / self._solver_dt
            if abs(native_sr - sample_rate) > 1.0:
                waveform = self._resample(waveform, native_sr, sample_rate)
        if normalize:
            peak = np.max(np.abs(waveform))
            if peak > 0:
                waveform = waveform / peak * 0.95
        dtype, max_int = (np.int16, 32767) if bit_depth == 16 else (np.int32, 2147483647)
        data = (waveform * max_int).astype(dtype)
        with wave.open(filepath, "wb") as fh:
            fh.setnchannels(1)
            fh.setsampwidth(bit_depth // 8)
            fh.setframerate(sample_rate)
            fh.writeframes(data.tobytes())

    @staticmethod
    def _resample(waveform, src_rate: float, target_rate: int):
        src_len = len(waveform)
        target_len = int(src_len * target_rate / src_rate)
        if target_len == src_len:
            return waveform
        return np.interp(np.linspace(0, src_len - 1, target_len), np.arange(src_len), waveform).astype(np.float32)

    def clear(self) -> None:
        self._sync()
        self._data.clear()
        self._times.clear()

    def __len__(self) -> int:
        self._sync()
        return len(self._data)

    def __repr__(self) -> str:
        name_str = f"'{self.name}'" if self.name else "unnamed"
        dir_str = f", direction={self.direction}" if self.is_directional() else ""
        return (f"Microphone({name_str}, position={self.position}, pattern='{self._pattern_name}'{dir_str}, "
                f"samples={len(self._data)})")
