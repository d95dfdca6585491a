"""Material models whose ADE poles feed the step kernel (materials/porous.py,
materials/solids.py) and the reference's material library
(materials/library.py): the poles are formed with the reference's numpy
expressions, so a library material gives the reference's poles bit for bit
(checked against poles the reference computed, tests/golden/material_poles.npz).

PorousMaterial: Johnson-Champoux-Allard rigid-frame fluid. The normalised
dynamic density G(w) = 1 + s phi / (i w rho_f a) sqrt(1 + i w 4 a^2 eta rho_f /
(s^2 L^2 phi^2)) and modulus K(w)/K_inf = g / (g - (g - 1) / G'(w)) with
G'(w) = 1 + 8 eta / (i w Pr rho_f L'^2) sqrt(1 + i w Pr rho_f L'^2 / (16 eta))
are fitted, each, by num_poles // 2 (density) and the rest (modulus) Debye
poles with relaxation times log-spaced over [1/(100 w_max), 100/w_min]
(w_v = s phi / (rho_f a), w_t = eta / (rho_f Pr L'^2)): real and imaginary
parts stacked, least squares over 100 log-spaced w in [10, 1e6] rad/s; poles
of strength below 1e-6 dropped. rho_inf = rho_f a / phi, K_inf = g P0 / phi.

ViscoelasticSolid: an effective fluid of density rho and bulk modulus
E / (3 (1 - 2 nu)), relaxation strength 2 eta at tau_0 = 1 / (2 pi f_ref),
spread evenly over num_poles modulus poles at tau_0 10^[-1, 1] (one pole at
tau_0 when num_poles is 1). DampingMaterial: the same with strength
min(3 eta, 0.9), nu = 0.49 and 5 poles by default.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
from numpy.typing import NDArray

from .materials import AIR_15C, AIR_20C, AIR_25C, WATER_20C, AcousticMaterial, Pole, PoleType

AIR_DENSITY = 1.204
AIR_VISCOSITY = 1.81e-5
AIR_PRANDTL = 0.71
AIR_GAMMA = 1.4
AIR_PRESSURE = 101325


def _debye_fit(omega, tau_values, target) -> NDArray[np.floating]:
    """Least-squares Debye strengths for target(w) - 1 (real and imaginary
    parts stacked)."""
    A = np.zeros((len(omega), len(tau_values)), dtype=complex)
    for j, tau in enumerate(tau_values):
        A[:, j] = 1 / (1 + 1j * omega * tau)
    A_stacked = np.vstack([np.real(A), np.imag(A)])
    rhs = np.concatenate([np.real(target) - 1, np.imag(target)])
    return np.linalg.lstsq(A_stacked, rhs, rcond=None)[0]


@dataclass(repr=False)
class PorousMaterial(AcousticMaterial):
    name: str = "unnamed"
    flow_resistivity: float = 10000.0
    porosity: float = 0.95
    tortuosity: float = 1.05
    viscous_length: float = 200e-6
    thermal_length: float = 400e-6
    thickness: float | None = None
    num_poles: int = 6
    rho_fluid: float = AIR_DENSITY
    eta: float = AIR_VISCOSITY
    Pr: float = AIR_PRANDTL
    gamma: float = AIR_GAMMA
    P0: float = AIR_PRESSURE
    _poles_cache: list | None = field(default=None, repr=False)

    def __post_init__(self):
        if self.flow_resistivity <= 0:
            raise ValueError("flow_resistivity must be positive")
        if not 0 < self.porosity <= 1:
            raise ValueError("porosity must be in (0, 1]")
        if self.tortuosity < 1:
            raise ValueError("tortuosity must be >= 1")
        if self.viscous_length <= 0:
            raise ValueError("viscous_length must be positive")
        if self.thermal_length <= 0:
            raise ValueError("thermal_length must be positive")
        if self.num_poles < 2:
            raise ValueError("num_poles must be >= 2")

    @property
    def rho_inf(self) -> float:
        return self.rho_fluid * self.tortuosity / self.porosity

    @property
    def K_inf(self) -> float:
        return self.gamma * self.P0 / self.porosity

    @property
    def omega_viscous(self) -> float:
        return self.flow_resistivity * self.porosity / (self.rho_fluid * self.tortuosity)

    @property
    def omega_thermal(self) -> float:
        return self.eta / (self.rho_fluid * self.Pr * self.thermal_length**2)

    def jca_density_normalized(self, omega):
        w = np.asarray(omega) + 1e-20
        s, phi, a, L = self.flow_resistivity, self.porosity, self.tortuosity, self.viscous_length
        wt = 4 * a**2 * self.eta * self.rho_fluid / (s**2 * L**2 * phi**2)
        return 1 + s * phi / (1j * w * self.rho_fluid * a) * np.sqrt(1 + 1j * w * wt)

    def jca_modulus_normalized(self, omega):
        w = np.asarray(omega) + 1e-20
        g, eta, Pr, rf, Lp = self.gamma, self.eta, self.Pr, self.rho_fluid, self.thermal_length
        wt = Pr * rf * Lp**2 / (16 * eta)
        Gp = 1 + 8 * eta / (1j * w * Pr * rf * Lp**2) * np.sqrt(1 + 1j * w * wt)
        return g / (g - (g - 1) / Gp)

    def jca_density(self, omega):
        return self.rho_inf * self.jca_density_normalized(omega)

    def jca_modulus(self, omega):
        return self.K_inf * self.jca_modulus_normalized(omega)

    @property
    def poles(self) -> list:
        if self._poles_cache is None:
            w_lo = 0.01 * min(self.omega_viscous, self.omega_thermal)
            w_hi = 100 * max(self.omega_viscous, self.omega_thermal)
            n_d = self.num_poles // 2
            out = []
            omega = np.logspace(1, 6, 100)
            for n, target, fn in ((n_d, "density", self.jca_density_normalized),
                                  (self.num_poles - n_d, "modulus", self.jca_modulus_normalized)):
                taus = np.logspace(np.log10(1 / w_hi), np.log10(1 / w_lo), n)
                for tau, dchi in zip(taus, _debye_fit(omega, taus, fn(omega)), strict=True):
                    if abs(dchi) > 1e-6:
                        out.append(Pole(pole_type=PoleType.DEBYE, delta_chi=dchi, tau=tau, target=target))
            self._poles_cache = out
        return self._poles_cache


@dataclass(repr=False)
class ViscoelasticSolid(AcousticMaterial):
    name: str = "unnamed"
    density: float = 700.0
    youngs_modulus: float = 3.5e9
    loss_factor: float = 0.03
    poissons_ratio: float = 0.3
    ref_frequency: float = 1000.0
    num_poles: int = 3
    _poles_cache: list | None = field(default=None, repr=False)

    def __post_init__(self):
        if self.density <= 0:
            raise ValueError("density must be positive")
        if self.youngs_modulus <= 0:
            raise ValueError("youngs_modulus must be positive")
        if self.loss_factor < 0:
            raise ValueError("loss_factor must be non-negative")
        if not -1 < self.poissons_ratio < 0.5:
            raise ValueError("poissons_ratio must be in (-1, 0.5)")
        if self.ref_frequency <= 0:
            raise ValueError("ref_frequency must be positive")

    @property
    def bulk_modulus(self) -> float:
        return self.youngs_modulus / (3 * (1 - 2 * self.poissons_ratio))

    @property
    def shear_modulus(self) -> float:
        return self.youngs_modulus / (2 * (1 + self.poissons_ratio))

    @property
    def rho_inf(self) -> float:
        return self.density

    @property
    def K_inf(self) -> float:
        return self.bulk_modulus

    @property
    def longitudinal_modulus(self) -> float:
        return self.bulk_modulus + 4 * self.shear_modulus / 3

    @property
    def c_longitudinal(self) -> float:
        return np.sqrt(self.longitudinal_modulus / self.density)

    @property
    def relaxation_time(self) -> float:
        return 1 / (2 * np.pi * self.ref_frequency)

    @property
    def relaxation_strength(self) -> float:
        return 2 * self.loss_factor

    @property
    def poles(self) -> list:
        if self._poles_cache is None:
            tau0, total = self.relaxation_time, self.relaxation_strength
            if self.num_poles == 1:
                out = [Pole(pole_type=PoleType.DEBYE, delta_chi=total, tau=tau0, target="modulus")]
            else:
                taus = tau0 * np.logspace(-1, 1, self.num_poles)
                weights = np.ones(len(taus)) / len(taus)
                out = [Pole(pole_type=PoleType.DEBYE, delta_chi=total * w, tau=tau, target="modulus")
                       for tau, w in zip(taus, weights, strict=True) if w > 1e-6]
            self._poles_cache = out
        return self._poles_cache

    def loss_factor_spectrum(self, omega):
        K = self.effective_modulus(omega)
        return np.imag(K) / np.real(K)

    def quality_factor(self, omega):
        return 1 / (self.loss_factor_spectrum(omega) + 1e-20)


@dataclass(repr=False)
class DampingMaterial(ViscoelasticSolid):
    poissons_ratio: float = 0.49
    num_poles: int = 5

    @property
    def relaxation_strength(self) -> float:
        return min(3 * self.loss_factor, 0.9)


def _porous(name, s, phi, a, L, Lp):
    return PorousMaterial(name=name, flow_resistivity=s, porosity=phi, tortuosity=a, viscous_length=L,
                          thermal_length=Lp, num_poles=6)


def _solid(cls, name, rho, E, eta, nu, n):
    return cls(name=name, density=rho, youngs_modulus=E, loss_factor=eta, poissons_ratio=nu, num_poles=n)


# The library's parameter table (materials/library.py:63-372), category by
# category: name, flow resistivity, porosity, tortuosity, viscous and
# thermal lengths; or density, Young's modulus, loss factor, Poisson's ratio.
MATERIALS = {
    "fluids": {m.name: m for m in (AIR_20C, AIR_25C, AIR_15C, WATER_20C)},
    "fiberglass": {m.name: m for m in (
        _porous("fiberglass_24kg", 8000, 0.99, 1.02, 200e-6, 400e-6),
        _porous("fiberglass_32kg", 14000, 0.98, 1.03, 150e-6, 300e-6),
        _porous("fiberglass_48kg", 25000, 0.98, 1.04, 100e-6, 200e-6),
        _porous("fiberglass_96kg", 60000, 0.96, 1.06, 60e-6, 120e-6))},
    "mineral_wool": {m.name: m for m in (
        _porous("mineral_wool_40kg", 20000, 0.97, 1.04, 80e-6, 160e-6),
        _porous("mineral_wool_60kg", 35000, 0.96, 1.05, 60e-6, 120e-6),
        _porous("mineral_wool_100kg", 70000, 0.94, 1.08, 40e-6, 80e-6))},
    "foam": {m.name: m for m in (
        _porous("acoustic_foam_open", 12000, 0.97, 1.10, 150e-6, 300e-6),
        _porous("acoustic_foam_dense", 30000, 0.95, 1.15, 80e-6, 160e-6),
        _porous("polyurethane_foam", 15000, 0.98, 1.08, 120e-6, 240e-6))},
    "fibrous": {m.name: m for m in (
        _porous("polyester_batting", 5000, 0.99, 1.01, 300e-6, 600e-6),
        _porous("wool_felt", 30000, 0.95, 1.30, 80e-6, 160e-6),
        _porous("cotton_batting", 8000, 0.98, 1.05, 200e-6, 400e-6))},
    "mdf": {m.name: m for m in (
        _solid(ViscoelasticSolid, "MDF_light", 600, 2.8e9, 0.035, 0.25, 3),
        _solid(ViscoelasticSolid, "MDF_medium", 720, 3.5e9, 0.030, 0.25, 3),
        _solid(ViscoelasticSolid, "MDF_high", 850, 4.2e9, 0.025, 0.25, 3))},
    "plywood": {m.name: m for m in (
        _solid(ViscoelasticSolid, "plywood_birch", 680, 12.0e9, 0.015, 0.35, 3),
        _solid(ViscoelasticSolid, "plywood_marine", 700, 10.0e9, 0.020, 0.30, 3),
        _solid(ViscoelasticSolid, "plywood_softwood", 550, 8.0e9, 0.020, 0.35, 3))},
    "other_wood": {m.name: m for m in (
        _solid(ViscoelasticSolid, "particle_board", 650, 2.5e9, 0.040, 0.25, 3),
        _solid(ViscoelasticSolid, "OSB", 620, 4.5e9, 0.025, 0.30, 3),
        _solid(ViscoelasticSolid, "hardwood_oak", 750, 12.5e9, 0.010, 0.35, 3))},
    "damping": {m.name: m for m in (
        _solid(DampingMaterial, "butyl_rubber", 1200, 1.0e6, 0.40, 0.49, 5),
        _solid(DampingMaterial, "neoprene", 1300, 2.0e6, 0.15, 0.49, 5),
        _solid(DampingMaterial, "sorbothane", 1100, 0.5e6, 0.50, 0.49, 5),
        _solid(DampingMaterial, "EPDM_rubber", 1150, 3.0e6, 0.12, 0.49, 5),
        _solid(DampingMaterial, "silicone_rubber", 1100, 1.5e6, 0.08, 0.49, 5))},
}


def get_material(name: str):
    """A library material by name, case-insensitive (materials/library.py:433-454)."""
    for cat in MATERIALS.values():
        for key, mat in cat.items():
            if key.lower() == name.lower():
                return mat
    raise KeyError(f"Material '{name}' not found. Use list_materials() to see available materials.")


def list_materials(category: str | None = None) -> list:
    if category is not None:
        if category not in MATERIALS:
            raise KeyError(f"Unknown category '{category}'. Available: {list(MATERIALS.keys())}")
        return list(MATERIALS[category].keys())
    return [k for cat in MATERIALS.values() for k in cat]


def list_categories() -> list:
    return list(MATERIALS.keys())
