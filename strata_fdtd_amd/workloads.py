"""Seeded workload definitions used by bench.py and the tests.

Each builder returns a ``Workload``: a callable that constructs a configured
``FDTDSolver`` (backend="mi355x") plus the step count and a description. The
scenes use only the public API (grids, SDF geometry, materials, PML, sources,
probes), so they are also usage examples.

  plain_100      100^3, 1 mm, AIR_20C, PML 10, 1 kHz pulse at (50,50,50),
                 probe (75,50,50), 1000 steps.
  enclosure_512  512^3, 1 mm, PML 16; rigid SDF box enclosure 300x200x200 cells
                 with 20-cell walls, a radius-12 / length-60 cylindrical port
                 through the -x wall, a radius-40 sphere of ADE absorber (two
                 seeded Debye poles: one density, one modulus) inside; 500 Hz
                 pulse inside the box; 8 seeded probes; 4000 steps.
  resonators_nu  320x320x480 nonuniform: 200 uniform 1 mm cells centred per
                 axis, stretched by 1.05 per cell outwards; PML 16; a 6x6 array
                 of Helmholtz resonators (cavity radius 10, depth 30 cells, neck
                 radius 4, length 4 cells) cut into a rigid plate; 2 kHz plane
                 source; 16 probes; 3000 steps.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Callable

import numpy as np

from .boundaries import PML
from .geometry import Box, Cylinder, Difference, MaterializedGeometry, Sphere, Union
from .grid import NonuniformGrid, UniformGrid, explicit_core_stretch_coords
from .materials import AIR_20C, Pole, PoleType, SimpleMaterial
from .solver import FDTDSolver
from .sources import GaussianPulse


@dataclass
class Workload:
    name: str
    shape: tuple[int, int, int]
    steps: int
    build: Callable[..., FDTDSolver]
    description: str

    @property
    def cells(self) -> int:
        return int(np.prod(self.shape))


def _air(solver_kw):
    kw = dict(c=AIR_20C._c, rho=AIR_20C.rho_inf, backend="mi355x")
    kw.update(solver_kw)
    return kw


def plain_100(**solver_kw) -> Workload:
    def build(**kw) -> FDTDSolver:
        s = FDTDSolver(grid=UniformGrid((100, 100, 100), 1e-3), **_air({**solver_kw, **kw}))
        s.add_boundary(PML(depth=10))
        s.add_source(GaussianPulse(position=(50, 50, 50), frequency=1000.0, amplitude=1.0))
        s.add_probe("probe", (75, 50, 50))
        return s
    return Workload("plain_100", (100, 100, 100), 1000, build,
                    "UniformGrid 100^3 1 mm, AIR_20C, PML 10, 1 kHz pulse, 1 probe, 1000 steps")


def enclosure_512_scene(n: int = 512, seed: int = 2024):
    """Geometry, material, source and probe positions of the 512^3 enclosure
    workload (scaled proportionally when n != 512, for tests)."""
    rng = np.random.default_rng(seed)
    dx = 1e-3
    sc = n / 512.0
    L = lambda cells: cells * sc * dx  # noqa: E731
    c0 = np.array([n / 2 * dx] * 3)
    outer_size = np.array([L(300), L(200), L(200)])
    wall = L(20)
    outer = Box(center=tuple(c0), size=tuple(outer_size))
    inner = Box(center=tuple(c0), size=tuple(outer_size - 2 * wall))
    shell = Difference(outer, inner)
    domain = Box(center=tuple(c0), size=(n * dx * 2,) * 3)
    x_wall_lo = c0[0] - outer_size[0] / 2
    port = Cylinder(p1=(x_wall_lo - L(20), c0[1], c0[2]), p2=(x_wall_lo + L(40), c0[1], c0[2]),
                    radius=L(12))
    air = Union(Difference(domain, shell), port)
    poles = [
        Pole(PoleType.DEBYE, float(rng.uniform(0.1, 0.5)), "density", tau=float(rng.uniform(1e-5, 1e-4))),
        Pole(PoleType.DEBYE, float(rng.uniform(0.1, 0.5)), "modulus", tau=float(rng.uniform(1e-5, 1e-4))),
    ]
    absorber = SimpleMaterial(name="absorber", _rho=AIR_20C.rho_inf, _c=AIR_20C._c, _poles=poles)
    sphere = Sphere(center=(c0[0] + L(44), c0[1], c0[2]), radius=L(40))
    geo = MaterializedGeometry(air)
    geo.set_material(sphere, absorber)
    cav_lo = (c0 - outer_size / 2 + wall) / dx
    cav_hi = (c0 + outer_size / 2 - wall) / dx
    src = tuple(int(v) for v in (cav_lo + 0.25 * (cav_hi - cav_lo)))
    probes = [tuple(int(v) for v in rng.uniform(cav_lo + 2, cav_hi - 2)) for _ in range(8)]
    return geo, src, probes


def enclosure_512(n: int = 512, steps: int = 4000, **solver_kw) -> Workload:
    def build(**kw) -> FDTDSolver:
        geo, src, probes = enclosure_512_scene(n)
        s = FDTDSolver(grid=UniformGrid((n, n, n), 1e-3), **_air({**solver_kw, **kw}))
        s.set_geometry(geo)
        s.add_boundary(PML(depth=max(4, round(16 * n / 512))))
        s.add_source(GaussianPulse(position=src, frequency=500.0))
        for q, p in enumerate(probes):
            s.add_probe(f"probe{q}", p)
        return s
    return Workload("enclosure_512" if n == 512 else f"enclosure_{n}", (n, n, n), steps, build,
                    f"UniformGrid {n}^3 1 mm, PML 16, rigid SDF enclosure 300x200x200 (walls 20) + "
                    "port r12 L60 + ADE sphere r40 (2 Debye poles), 500 Hz pulse, 8 probes")


def resonators_nu_scene(shape=(320, 320, 480), core=200, seed=7, n_side=6, pitch=30, plate=40,
                        cavity_radius=10, cavity_depth=30, neck_radius=4, neck_length=4):
    """Grid, air geometry, probes and plane-source index of the resonator-plate
    workload. The defaults are the 320x320x480 workload; the keyword sizes (in
    cells of the 1 mm core) scale the plate and its n_side x n_side array of
    resonators down for tests."""
    rng = np.random.default_rng(seed)
    xs = [explicit_core_stretch_coords(n, min(core, n - (n - core) % 2), 1e-3, 1.05) for n in shape]
    grid = NonuniformGrid(*xs)
    zc = xs[2]
    kc = shape[2] // 2
    mm = lambda cells: cells / 1000.0  # noqa: E731  (cells of 1 mm, rounded as a literal 4e-3 is)
    plate_lo = zc[kc] - 0.5e-3          # plate occupies cells [kc, kc+plate) along z
    plate_hi = zc[kc + plate - 1] + 0.5e-3
    slab = Box(min_corner=(-1.0, -1.0, plate_lo), max_corner=(2.0, 2.0, plate_hi))
    holes = []
    xc, yc = xs[0], xs[1]
    half = n_side * pitch // 2
    i0, j0 = shape[0] // 2 - half, shape[1] // 2 - half
    for a in range(n_side):
        for b in range(n_side):
            cx = xc[i0 + pitch * a + pitch // 2]
            cy = yc[j0 + pitch * b + pitch // 2]
            neck = Cylinder(p1=(cx, cy, plate_lo - 1e-3), p2=(cx, cy, plate_lo + mm(neck_length)),
                            radius=mm(neck_radius))
            cav = Cylinder(p1=(cx, cy, plate_lo + mm(neck_length)),
                           p2=(cx, cy, plate_lo + mm(neck_length + cavity_depth)), radius=mm(cavity_radius))
            holes += [cav, neck]
    solid = Difference(slab, *holes)
    domain = Box(min_corner=(-1.0, -1.0, -1.0), max_corner=(2.0, 2.0, 50.0))
    air = Difference(domain, solid)
    probes = []
    k_lo = max(1, kc - 60)
    for _ in range(16):
        i = int(rng.integers(shape[0] // 2 - half, shape[0] // 2 + half))
        j = int(rng.integers(shape[1] // 2 - half, shape[1] // 2 + half))
        k = int(rng.integers(k_lo, kc - 5))
        probes.append((i, j, k))
    src_index = max(1, kc - 80)
    return grid, air, probes, src_index


def resonators_nu(shape=(320, 320, 480), steps: int = 3000, **solver_kw) -> Workload:
    def build(**kw) -> FDTDSolver:
        grid, air, probes, src_index = resonators_nu_scene(shape)
        s = FDTDSolver(grid=grid, **_air({**solver_kw, **kw}))
        s.set_geometry(air)
        s.add_boundary(PML(depth=16))
        s.add_source(GaussianPulse(position={"axis": 2, "index": src_index}, frequency=2000.0,
                                   source_type="plane"))
        for q, p in enumerate(probes):
            s.add_probe(f"probe{q}", p)
        return s
    return Workload("resonators_nu", tuple(shape), steps, build,
                    "NonuniformGrid 320x320x480 (1 mm core, x1.05 stretch), PML 16, 6x6 Helmholtz "
                    "resonators in a rigid plate, 2 kHz plane source, 16 probes")
