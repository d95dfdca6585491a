"""Oracle (test infrastructure only): the reference's microphone recording,
restated with numpy scalars so every rounding step is the reference's own.

Follows core/solver.py of the reference:
  Microphone._initialize            :930-961   grid position = position / dx
  Microphone._compute_trilinear_weights :962-1003  corner order and weights
  Microphone.record                 :1004-1058  pressure sum, directional path
  Microphone._interpolate_velocity  :1060-1113  staggered shift, clamping
  Microphone._compute_directional_output :1114-1173  (1-k) P + k Z v.d, custom
  POLAR_PATTERNS                    :787-794

Rounding, as numpy 2 performs it: a Python float times a numpy float32 scalar
casts the Python float to float32 and multiplies in float32; a Python float
plus a float32 scalar stays float32; float32 times a float64 scalar is
float64. Recorded values are turned into float32 by get_waveform.

Pinned by the reference's own known answers (tests/test_microphone.py:100-160:
exact value at a grid point, the average at a cell centre, 0.75/0.25 weights)
and by exactness of trilinear interpolation on fields linear in i, j, k
(tests/test_oracle.py).
"""

from __future__ import annotations

import numpy as np

POLAR_PATTERNS = {"omni": 0.0, "subcardioid": 0.3, "cardioid": 0.5, "supercardioid": 0.63,
                  "hypercardioid": 0.75, "figure8": 1.0}


def trilinear(gx: float, gy: float, gz: float):
    """Corner indices and float64 weights, in the reference's order."""
    i0, j0, k0 = int(gx), int(gy), int(gz)
    i1, j1, k1 = i0 + 1, j0 + 1, k0 + 1
    fx, fy, fz = gx - i0, gy - j0, gz - k0
    weights = [(1 - fx) * (1 - fy) * (1 - fz), fx * (1 - fy) * (1 - fz), (1 - fx) * fy * (1 - fz),
               fx * fy * (1 - fz), (1 - fx) * (1 - fy) * fz, fx * (1 - fy) * fz, (1 - fx) * fy * fz,
               fx * fy * fz]
    indices = [(i0, j0, k0), (i1, j0, k0), (i0, j1, k0), (i1, j1, k0), (i0, j0, k1), (i1, j0, k1),
               (i0, j1, k1), (i1, j1, k1)]
    return indices, weights


def grid_position(position, dx: float):
    return position[0] / dx, position[1] / dx, position[2] / dx


def interp_pressure(p: np.ndarray, g) -> np.float32:
    indices, weights = trilinear(*g)
    acc = 0.0
    for (i, j, k), w in zip(indices, weights):
        acc += w * p[i, j, k]
    return acc


def interp_velocity(v: np.ndarray, g, axis: int) -> np.float32:
    gx, gy, gz = g
    if axis == 0:
        gx -= 0.5
    elif axis == 1:
        gy -= 0.5
    else:
        gz -= 0.5
    nx, ny, nz = v.shape
    gx = max(0.0, min(gx, nx - 1.001))
    gy = max(0.0, min(gy, ny - 1.001))
    gz = max(0.0, min(gz, nz - 1.001))
    indices, weights = trilinear(gx, gy, gz)
    acc = 0.0
    for (i, j, k), w in zip(indices, weights):
        acc += w * v[min(i, nx - 1), min(j, ny - 1), min(k, nz - 1)]
    return acc


def directional(pressure, vel, direction: np.ndarray, pattern, rho: float, c: float):
    """Output of a directional microphone from interpolated p and v."""
    Z = rho * c
    v_dot_d = vel[0] * direction[0] + vel[1] * direction[1] + vel[2] * direction[2]
    if callable(pattern):
        v_mag = np.sqrt(vel[0] ** 2 + vel[1] ** 2 + vel[2] ** 2)
        if v_mag > 1e-20:
            theta = np.arccos(np.clip(v_dot_d / v_mag, -1.0, 1.0))
        else:
            theta = 0.0
        return pressure * pattern(theta)
    k = POLAR_PATTERNS[pattern]
    return (1 - k) * pressure + k * Z * v_dot_d


def record(p, vx, vy, vz, position, dx: float, pattern="omni", direction=(1.0, 0.0, 0.0),
           rho: float = 1.2, c: float = 343.0):
    """One sample of a microphone (before get_waveform's float32 cast)."""
    g = grid_position(position, dx)
    pressure = interp_pressure(p, g)
    if pattern == "omni":
        return pressure
    d = np.asarray(direction, dtype=np.float64)
    d = d / np.linalg.norm(d)
    vel = (interp_velocity(vx, g, 0), interp_velocity(vy, g, 1), interp_velocity(vz, g, 2))
    return directional(pressure, vel, d, pattern, rho, c)
