"""Independent numpy restatement of the reference's host-side precompute.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Each function follows the reference file:line given in its docstring and states
what pins it. The product computes the same tables in ``strata_fdtd_amd``; the
tests compare the two bit for bit and pin both against closed forms.
"""

from __future__ import annotations

import numpy as np


def time_step(min_spacing: float, c: float, courant: float = 0.95) -> float:
    """dt = courant * min_dx / (c * sqrt(3)) in float64 (core/solver.py:1576-1580).

    Pinned by the closed form: 0.95e-3/(343*sqrt 3) = 1.599075e-6 s at 1 mm.
    """
    return courant * (min_spacing / (c * np.sqrt(3)))


def uniform_coefficients(dt: float, dx: float, rho: float, c: float) -> tuple[float, float]:
    """(coeff_v, coeff_p) = (-dt/(rho dx), -rho c^2 dt/dx) in float64 (core/solver.py:1584-1586)."""
    coeff_p = -rho * c**2 * dt / dx
    coeff_v = -dt / (rho * dx)
    return coeff_v, coeff_p


def base_coefficients(dt: float, rho: float, c: float) -> tuple[float, float]:
    """Nonuniform (coeff_v_base, coeff_p_base) = (-dt/rho, -rho c^2 dt) (core/solver.py:1589-1590)."""
    return -dt / rho, -rho * c**2 * dt


def pml_sigma_max(depth: int, order: int, c: float, min_spacing: float) -> float:
    """sigma_max = -(order+1) c ln(1e-6) / (2 depth min_dx) (boundaries/_boundaries.py:182-187).

    Pinned by the closed form: depth 10, order 3, c 343, 1 mm -> 947 744 1/s.
    """
    d = depth * min_spacing
    return -(order + 1) * c * np.log(1e-6) / (2 * d)


def pml_profile_uniform(n: int, depth: int, order: int, sigma_max: float) -> np.ndarray:
    """float32 polynomial profile (boundaries/_boundaries.py:228-255).

    Left cells i < depth: sigma_max * ((depth - i)/depth)**order; right cells
    i >= n - depth: sigma_max * ((i - (n - depth - 1))/depth)**order; each value
    is computed in float64 and stored into a float32 array.
    """
    sigma = np.zeros(n, dtype=np.float32)
    for i in range(depth):
        sigma[i] = sigma_max * (((depth - i) / depth) ** order)
    for i in range(n - depth, n):
        sigma[i] = sigma_max * (((i - (n - depth - 1)) / depth) ** order)
    return sigma


def pml_profile_nonuniform(coords: np.ndarray, spacing: np.ndarray, depth: int, order: int,
                           sigma_max: float) -> np.ndarray:
    """float32 profile from physical distances (boundaries/_boundaries.py:257-315)."""
    n = len(coords)
    sigma = np.zeros(n, dtype=np.float32)
    d = depth
    if d >= n // 2:
        d = max(1, n // 4)
    left_thickness = float(np.sum(spacing[:d]))
    right_thickness = float(np.sum(spacing[-d:]))
    left_interface = coords[d] - spacing[d] / 2 if d < n else coords[-1]
    right_interface = coords[n - d - 1] + spacing[n - d - 1] / 2 if n - d > 0 else coords[0]
    for i in range(d):
        dist = left_interface - coords[i]
        nd = min(1.0, dist / left_thickness) if left_thickness > 0 else 0.0
        sigma[i] = sigma_max * (nd ** order)
    for i in range(n - d, n):
        dist = coords[i] - right_interface
        nd = min(1.0, dist / right_thickness) if right_thickness > 0 else 0.0
        sigma[i] = sigma_max * (nd ** order)
    return sigma


def pml_decay(sigma: np.ndarray, dt: float) -> np.ndarray:
    """float32 exp(-sigma*dt) as the native initialize_pml forms it
    (_kernels/src/pml.cpp:13-45): float sigma times float dt, then std::exp on
    float. Evaluated by the C++ oracle so the libm expf is the same one the
    reference's build uses.

    (The reference's numpy fallback, boundaries/_boundaries.py:340-377, forms
    the table in float64 because ``solver.dt`` is a numpy float64 scalar; the
    native form is the one the parity target names.)
    """
    import ctypes as C
    from . import cpu
    sig = np.ascontiguousarray(sigma, dtype=np.float32)
    out = np.empty_like(sig)
    fp = C.POINTER(C.c_float)
    cpu._lib("oracle").orc_pml_decay(sig.ctypes.data_as(fp), C.c_int(len(sig)), C.c_float(float(dt)),
                                     out.ctypes.data_as(fp))
    return out


def stretched_coords(n: int, base_dx: float, stretch: float, center_fine: bool = True) -> np.ndarray:
    """Cell centres of NonuniformGrid.from_stretch (core/grid.py:367-410)."""
    if stretch == 1.0:
        return np.arange(n, dtype=np.float64) * base_dx + base_dx / 2
    if center_fine:
        n_half = n // 2
        sizes = base_dx * (stretch ** np.arange(n_half, dtype=np.float64))
        positions = np.cumsum(sizes) - sizes[0] / 2
        if n % 2 == 0:
            left = -positions[::-1] - sizes[::-1] / 2
            right = positions + sizes / 2
            coords = np.concatenate([left, right])
        else:
            left = -positions[::-1] - sizes[::-1] / 2
            center = np.array([0.0])
            right = positions[1:] + sizes[1:] / 2 if n_half > 0 else np.array([])
            coords = np.concatenate([left, center, right])
        coords = coords - coords[0] + base_dx / 2
    else:
        sizes = base_dx * (stretch ** np.arange(n, dtype=np.float64))
        coords = np.cumsum(sizes) - sizes / 2
    return coords


def cell_spacing(coords: np.ndarray) -> np.ndarray:
    """Half-sum of neighbour distances, one-sided at the ends (core/grid.py:219-237)."""
    n = len(coords)
    s = np.zeros(n, dtype=np.float64)
    for i in range(1, n - 1):
        s[i] = (coords[i + 1] - coords[i - 1]) / 2
    s[0] = coords[1] - coords[0]
    s[-1] = coords[-1] - coords[-2]
    return s


def stencil_arrays(xc: np.ndarray, yc: np.ndarray, zc: np.ndarray) -> dict[str, np.ndarray]:
    """float32 inverse face/cell spacings (core/grid.py:498-532)."""
    out = {}
    for name, c in (("x", xc), ("y", yc), ("z", zc)):
        out[f"inv_d{name}_face"] = (1.0 / np.diff(c)).astype(np.float32)
        out[f"inv_d{name}_cell"] = (1.0 / cell_spacing(c)).astype(np.float32)
    return out


def gaussian_pulse(t: np.ndarray, frequency: float, bandwidth: float | None = None,
                   amplitude: float = 1.0) -> np.ndarray:
    """A exp(-(t-t0)^2/(2 s^2)) sin(2 pi f (t-t0)), s = 1/(pi bw), t0 = 4 s (core/solver.py:152-207).

    Pinned by the closed form at t = t0 (value 0) and at a quarter period after t0.
    """
    if bandwidth is None:
        bandwidth = 2.0 * frequency
    sigma = 1.0 / (np.pi * bandwidth)
    t0 = 4.0 * sigma
    envelope = np.exp(-((t - t0) ** 2) / (2 * sigma**2))
    carrier = np.sin(2 * np.pi * frequency * (t - t0))
    return amplitude * envelope * carrier


def accumulated_times(dt: float, n: int, t_start: float = 0.0) -> np.ndarray:
    """t_0 = t_start, t_{m+1} = t_m + dt accumulated in float64 (core/solver.py:2072).

    The source is evaluated at the accumulated time, not at m*dt
    (core/solver.py:2416).
    """
    t = np.empty(n, dtype=np.float64)
    acc = float(t_start)
    for m in range(n):
        t[m] = acc
        acc += dt
    return t


def waveform_table(t: np.ndarray, frequency: float, bandwidth=None, amplitude=1.0) -> np.ndarray:
    """Waveform sampled one time value per call, as ``_inject_sources`` does."""
    return np.array([gaussian_pulse(np.array([tv]), frequency, bandwidth, amplitude)[0] for tv in t])


def position_to_index(pos, shape, dx):
    """Grid index of a source/probe position (core/solver.py:1815-1838, 1862-1883).

    A 3-tuple containing any float below max(shape) is read as metres and
    rounded to the nearest index; otherwise it is used as an index.
    """
    if isinstance(pos, tuple) and len(pos) == 3:
        if any(isinstance(p, float) and p < max(shape) for p in pos):
            return tuple(int(round(p / dx)) for p in pos)
    return pos


def debye_coefficients(tau: float, delta_chi: float, dt: float) -> tuple[float, float]:
    """(alpha, beta) = ((tau/dt)/(1+tau/dt), dchi/(1+tau/dt)) (materials/base.py:171-176)."""
    tau_dt = tau / dt
    denom = 1 + tau_dt
    return tau_dt / denom, delta_chi / denom


def lorentz_coefficients(omega_0: float, gamma: float, delta_chi: float, dt: float):
    """(a, b, d) of materials/base.py:178-188."""
    dt2 = dt * dt
    c = 1 + gamma * dt / 2
    a = (2 - omega_0**2 * dt2) / c
    b = -(1 - gamma * dt / 2) / c
    d = delta_chi * omega_0**2 * dt2 / c
    return a, b, d


def native_density_correction(dt: float, rho_inf: float, dx: float) -> float:
    """coeff = -dt / rho_inf * inv_dx evaluated in float32 (_kernels/src/ade.cpp:247)."""
    f = np.float32
    return float((f(-f(dt)) / f(rho_inf)) * f(1.0 / dx))


def native_modulus_correction(dt: float, K_inf: float) -> float:
    """coeff = -K_inf * dt evaluated in float32 (_kernels/src/ade.cpp:422)."""
    f = np.float32
    return float(f(-f(K_inf)) * f(dt))

