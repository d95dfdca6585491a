"""Frequency-dependent materials: Debye/Lorentz poles (materials/base.py:38-555).

What reaches the step kernel: the pole parameters, their FDTD recurrence
coefficients, the instantaneous density/modulus, and the material library
constants the workloads use (materials/library.py:31-60). Around them, the
material interface of materials/base.py:215-505 (AcousticMaterial): the
frequency response a material's poles imply, for checking a material before
a run.
"""

from __future__ import annotations

from abc import ABC, abstractmethod
from dataclasses import dataclass, field
from enum import Enum
from typing import Literal

import numpy as np


class PoleType(Enum):
    DEBYE = "debye"
    LORENTZ = "lorentz"


@dataclass
class Pole:
    """One dispersion pole targeting density or bulk modulus (materials/base.py:55-212)."""

    pole_type: PoleType
    delta_chi: float
    target: Literal["density", "modulus"]
    tau: float | None = None
    omega_0: float | None = None
    gamma: float | None = None

    def __post_init__(self):
        if self.pole_type == PoleType.DEBYE:
            if self.tau is None:
                raise ValueError("Debye poles require tau parameter")
            if self.tau <= 0:
                raise ValueError("tau must be positive")
        elif self.pole_type == PoleType.LORENTZ:
            if self.omega_0 is None or self.gamma is None:
                raise ValueError("Lorentz poles require omega_0 and gamma parameters")
            if self.omega_0 <= 0:
                raise ValueError("omega_0 must be positive")
            if self.gamma < 0:
                raise ValueError("gamma must be non-negative")
        if self.target not in ("density", "modulus"):
            raise ValueError("target must be 'density' or 'modulus'")

    def susceptibility(self, omega):
        omega = np.asarray(omega)
        if self.pole_type == PoleType.DEBYE:
            return self.delta_chi / (1 + 1j * omega * self.tau)
        return self.delta_chi * self.omega_0**2 / (self.omega_0**2 - omega**2 + 1j * self.gamma * omega)

    def fdtd_coefficients(self, dt: float) -> tuple[float, ...]:
        """Debye (alpha, beta); Lorentz (a, b, d) — materials/base.py:157-188."""
        if self.pole_type == PoleType.DEBYE:
            tau_dt = self.tau / dt
            denom = 1 + tau_dt
            return (tau_dt / denom, self.delta_chi / denom)
        w0, g = self.omega_0, self.gamma
        dt2 = dt * dt
        c = 1 + g * dt / 2
        a = (2 - w0**2 * dt2) / c
        b = -(1 - g * dt / 2) / c
        d = self.delta_chi * w0**2 * dt2 / c
        return (a, b, d)

    @property
    def is_debye(self) -> bool:
        return self.pole_type == PoleType.DEBYE

    @property
    def is_lorentz(self) -> bool:
        return self.pole_type == PoleType.LORENTZ


class AcousticMaterial(ABC):
    """A material given by its instantaneous density rho_inf, bulk modulus
    K_inf and dispersion poles (materials/base.py:215-505). Density poles add
    their susceptibilities to rho(w) = rho_inf (1 + sum chi(w)), modulus poles
    to K(w) = K_inf (1 + sum chi(w)); the plane-wave quantities follow from
    these two: k = w sqrt(rho / K), Z = sqrt(rho K), phase speed w / Re k,
    attenuation Im k (Np/m; times 20 / ln 10 in dB/m; negative for a lossy
    material, as the poles' susceptibilities use the e^{+iwt} convention),
    normal-incidence reflection (Z - Z') / (Z + Z') and absorption 1 - |R|^2."""

    name: str = "unnamed"

    @property
    @abstractmethod
    def rho_inf(self) -> float:
        ...

    @property
    @abstractmethod
    def K_inf(self) -> float:
        ...

    @property
    @abstractmethod
    def poles(self) -> list["Pole"]:
        ...

    @property
    def density_poles(self) -> list["Pole"]:
        return [p for p in self.poles if p.target == "density"]

    @property
    def modulus_poles(self) -> list["Pole"]:
        return [p for p in self.poles if p.target == "modulus"]

    @property
    def c_inf(self) -> float:
        return float(np.sqrt(self.K_inf / self.rho_inf))

    @property
    def Z_inf(self) -> float:
        return self.rho_inf * self.c_inf

    def _factor(self, omega, poles):
        w = np.asarray(omega)
        total = np.ones_like(w, dtype=complex)
        for pole in poles:
            total += pole.susceptibility(w)
        return total

    def effective_density(self, omega):
        return self.rho_inf * self._factor(omega, self.density_poles)

    def effective_modulus(self, omega):
        return self.K_inf * self._factor(omega, self.modulus_poles)

    def wavenumber(self, omega):
        w = np.asarray(omega)
        return w * np.sqrt(self.effective_density(w) / self.effective_modulus(w))

    def phase_velocity(self, omega):
        return np.real(np.asarray(omega) / self.wavenumber(omega))

    def group_velocity(self, omega, domega: float = 1.0):
        """Central difference dw / Re dk over a step of domega."""
        w = np.asarray(omega)
        return domega / np.real(self.wavenumber(w + domega / 2) - self.wavenumber(w - domega / 2))

    def impedance(self, omega):
        w = np.asarray(omega)
        return np.sqrt(self.effective_density(w) * self.effective_modulus(w))

    def attenuation(self, omega):
        return np.imag(self.wavenumber(omega))

    def attenuation_db_per_meter(self, omega):
        return 20.0 / np.log(10.0) * self.attenuation(omega)

    def reflection_coefficient(self, omega, Z_other: complex):
        Z = self.impedance(omega)
        return (Z - Z_other) / (Z + Z_other)

    def absorption_coefficient(self, omega, Z_backing: complex):
        return 1 - np.abs(self.reflection_coefficient(omega, Z_backing)) ** 2

    def n_poles(self) -> int:
        return len(self.poles)

    def n_density_poles(self) -> int:
        return len(self.density_poles)

    def n_modulus_poles(self) -> int:
        return len(self.modulus_poles)

    def summary(self) -> str:
        out = [f"Material: {self.name}",
               f"  \u03c1\u221e = {self.rho_inf:.2f} kg/m\u00b3",
               f"  K\u221e = {self.K_inf:.2e} Pa",
               f"  c\u221e = {self.c_inf:.1f} m/s",
               f"  Z\u221e = {self.Z_inf:.1f} Pa\u00b7s/m",
               f"  Poles: {self.n_poles()} ({self.n_density_poles()} density, {self.n_modulus_poles()} modulus)"]
        out += [f"    [{i}] {p}" for i, p in enumerate(self.poles)]
        return "\n".join(out)

    def __repr__(self) -> str:
        return f"{type(self).__name__}(name='{self.name}', poles={self.n_poles()})"


@dataclass
class SimpleMaterial(AcousticMaterial):
    """Constant rho/c material with optional poles (materials/base.py:508-555)."""

    name: str = "unnamed"
    _rho: float = field(default=1.2, repr=False)
    _c: float = field(default=343.0, repr=False)
    _poles: list[Pole] = field(default_factory=list, repr=False)

    @property
    def rho_inf(self) -> float:
        return self._rho

    @property
    def K_inf(self) -> float:
        return self._rho * self._c**2

    @property
    def poles(self) -> list[Pole]:
        return self._poles

    @classmethod
    def from_impedance(cls, name: str, rho: float, Z: float) -> "SimpleMaterial":
        """A material of density rho and acoustic impedance Z (c = Z / rho),
        materials/base.py:543-555."""
        return cls(name=name, _rho=rho, _c=Z / rho)


Material = SimpleMaterial

AIR_20C = SimpleMaterial(name="air_20C", _rho=1.204, _c=343.0)
AIR_25C = SimpleMaterial(name="air_25C", _rho=1.184, _c=346.3)
AIR_15C = SimpleMaterial(name="air_15C", _rho=1.225, _c=340.3)
WATER_20C = SimpleMaterial(name="water_20C", _rho=998.2, _c=1482.0)


def density_correction_coeff(dt: float, rho_inf: float, inv_dx: float) -> np.float32:
    """Velocity-correction coefficient as the native kernel forms it in float32:
    ``-dt / rho_inf * inv_dx`` (_kernels/src/ade.cpp:247). The solver passes
    ``1.0 / self.dx`` (the minimum spacing) on uniform and nonuniform grids
    alike (core/solver.py:3295-3305), so nonuniform grids get no per-face
    spacing in this correction."""
    f = np.float32
    return f((f(-f(dt)) / f(rho_inf)) * f(inv_dx))


def modulus_correction_coeff(dt: float, K_inf: float) -> np.float32:
    """Pressure-correction coefficient ``-K_inf * dt`` in float32 (_kernels/src/ade.cpp:422)."""
    f = np.float32
    return f(f(-f(K_inf)) * f(dt))
