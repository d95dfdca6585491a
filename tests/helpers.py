"""Shared scenario builders for the parity tests.

A ``Scenario`` describes one solve (grid, medium, PML, geometry, materials,
sources, probes). ``oracle_problem`` builds the CPU oracle's description of it
with the oracle's own host tables (``oracle.tables``), independently of the
product package; ``make_solver`` builds the same solve through the product API
(``strata_fdtd_amd.FDTDSolver(backend="mi355x")``).
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from oracle import cpu, tables as T


@dataclass
class Scenario:
    shape: tuple[int, int, int]
    dx: float = 1e-3
    c: float = 343.0
    rho: float = 1.204
    courant: float = 0.95
    pml_depth: int = 0
    pml_axes: tuple = ("x", "y", "z")
    air: np.ndarray | None = None               # bool, True = air
    material: np.ndarray | None = None          # uint8 ids
    materials: dict = field(default_factory=dict)  # id -> list of pole specs
    sources: list = field(default_factory=list)    # ("point", (i,j,k), f) / ("plane", axis, idx, f)
    probes: list = field(default_factory=list)     # (i,j,k)
    coords: tuple | None = None                 # nonuniform (xc, yc, zc)
    # id -> material object (rho_inf, K_inf, Debye poles) registered as it is,
    # in place of a pole-spec list
    material_objects: dict = field(default_factory=dict)

    @property
    def nonuniform(self) -> bool:
        return self.coords is not None


# pole spec: dict(kind="debye"|"lorentz", target="density"|"modulus", tau=, dchi=, w0=, gamma=)


def _spacing_min(sc: Scenario) -> float:
    if sc.coords is None:
        return sc.dx
    return float(min(T.cell_spacing(c).min() for c in sc.coords))


def oracle_problem(sc: Scenario, n_steps: int) -> cpu.Problem:
    nx, ny, nz = sc.shape
    min_dx = _spacing_min(sc)
    dt = T.time_step(min_dx, sc.c, sc.courant)
    if sc.nonuniform:
        assert tuple(len(c) for c in sc.coords) == tuple(sc.shape), "coords do not match the shape"
        cv, cp = T.base_coefficients(dt, sc.rho, sc.c)
        st = T.stencil_arrays(*sc.coords)
        inv_face = (st["inv_dx_face"], st["inv_dy_face"], st["inv_dz_face"])
        inv_cell = (st["inv_dx_cell"], st["inv_dy_cell"], st["inv_dz_cell"])
    else:
        cv, cp = T.uniform_coefficients(dt, sc.dx, sc.rho, sc.c)
        inv_face = inv_cell = (None, None, None)
    decay = [None, None, None]
    if sc.pml_depth:
        sm = T.pml_sigma_max(sc.pml_depth, 3, sc.c, min_dx)
        for ax, name in enumerate("xyz"):
            if name not in sc.pml_axes:
                continue
            if sc.nonuniform:
                c = sc.coords[ax]
                sig = T.pml_profile_nonuniform(c, T.cell_spacing(c), sc.pml_depth, 3, sm)
            else:
                sig = T.pml_profile_uniform(sc.shape[ax], sc.pml_depth, 3, sm)
            decay[ax] = T.pml_decay(sig, dt)
    air = np.ones(sc.shape, np.uint8) if sc.air is None else sc.air.astype(np.uint8)
    times = T.accumulated_times(dt, n_steps)
    srcs = []
    for s in sc.sources:
        if s[0] == "point":
            w = T.waveform_table(times, s[2])
            srcs.append((0, int(np.ravel_multi_index(s[1], sc.shape)), 0, w))
        else:
            w = T.waveform_table(times, s[3])
            srcs.append((1, s[1], s[2], w))
    probes = [int(np.ravel_multi_index(p, sc.shape)) for p in sc.probes]
    poles = []
    rho_inf = np.zeros(256, np.float32)
    K_inf = np.zeros(256, np.float32)
    specs = dict(sc.materials)
    props = {mid: (sc.rho, sc.rho * sc.c**2) for mid in sc.materials}
    for mid, obj in sc.material_objects.items():
        specs[mid] = [dict(kind="debye", target=p.target, tau=p.tau, dchi=p.delta_chi) for p in obj.poles]
        props[mid] = (obj.rho_inf, obj.K_inf)
    for mid, plist in specs.items():
        rho_m, K_m = props[mid]
        rho_inf[mid] = rho_m
        K_inf[mid] = K_m
        for pl in plist:
            tgt = 0 if pl["target"] == "density" else 1
            if tgt == 0:
                corr = T.native_density_correction(dt, rho_m, min_dx)
            else:
                corr = T.native_modulus_correction(dt, K_m)
            if pl["kind"] == "debye":
                a, b = T.debye_coefficients(pl["tau"], pl["dchi"], dt)
                poles.append(cpu.Pole(mid, tgt, 0, a, b, 0.0, corr))
            else:
                a, b, d = T.lorentz_coefficients(pl["w0"], pl["gamma"], pl["dchi"], dt)
                poles.append(cpu.Pole(mid, tgt, 1, a, b, d, corr))
    # native ordering of the pole list: materials in registration order, poles in order
    return cpu.Problem(
        shape=sc.shape, cv=cv, cp=cp, air=air, decay=tuple(decay), nonuniform=sc.nonuniform,
        inv_face=inv_face, inv_cell=inv_cell, sources=srcs, probes=probes, poles=poles,
        material=None if sc.material is None else sc.material.astype(np.uint8),
        ade_inv_dx=1.0 / min_dx, rho_inf=rho_inf, K_inf=K_inf, dt=dt)


def make_solver(sc: Scenario, tiling=None):
    import strata_fdtd_amd as sfa
    if sc.nonuniform:
        grid = sfa.NonuniformGrid(*sc.coords)
    else:
        grid = sfa.UniformGrid(sc.shape, sc.dx)
    s = sfa.FDTDSolver(grid=grid, c=sc.c, rho=sc.rho, courant=sc.courant, backend="mi355x",
                       tiling=tiling)
    if sc.air is not None:
        s.set_geometry(sc.air.astype(bool))
    if sc.pml_depth:
        s.add_boundary(sfa.PML(depth=sc.pml_depth, axis=tuple(sc.pml_axes)))
    for mid, plist in sc.materials.items():
        poles = []
        for pl in plist:
            if pl["kind"] == "debye":
                poles.append(sfa.Pole(sfa.PoleType.DEBYE, pl["dchi"], pl["target"], tau=pl["tau"]))
            else:
                poles.append(sfa.Pole(sfa.PoleType.LORENTZ, pl["dchi"], pl["target"], omega_0=pl["w0"],
                                      gamma=pl["gamma"]))
        mat = sfa.SimpleMaterial(name=f"m{mid}", _rho=sc.rho, _c=sc.c, _poles=poles)
        s.register_material(mat, material_id=mid)
        s.set_material_region(sc.material == mid, mid)
    for mid, obj in sc.material_objects.items():
        s.register_material(obj, material_id=mid)
        s.set_material_region(sc.material == mid, mid)
    for src in sc.sources:
        if src[0] == "point":
            s.add_source(sfa.GaussianPulse(position=tuple(int(v) for v in src[1]), frequency=src[2]))
        else:
            s.add_source(sfa.GaussianPulse(position={"axis": src[1], "index": src[2]}, frequency=src[3],
                                           source_type="plane"))
    for q, p in enumerate(sc.probes):
        s.add_probe(f"probe{q}", tuple(int(v) for v in p))
    return s


def rel_l2(a, b) -> float:
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / den) if den > 0 else float(np.linalg.norm(a - b))


def box_enclosure(shape, lo, hi, wall, port=None):
    """Air mask: everything outside the shell is air, the shell is solid, the
    cavity is air; `port` = (axis, center_j, center_k, radius) drills a
    cylindrical hole through the wall on the low side of `axis`."""
    air = np.ones(shape, bool)
    lo = np.array(lo)
    hi = np.array(hi)
    air[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]] = False
    ilo = lo + wall
    ihi = hi - wall
    air[ilo[0]:ihi[0], ilo[1]:ihi[1], ilo[2]:ihi[2]] = True
    if port is not None:
        ax, cb, cc, r = port
        others = [a for a in range(3) if a != ax]
        g = np.indices(shape)
        d2 = (g[others[0]] - cb) ** 2 + (g[others[1]] - cc) ** 2
        span = (g[ax] >= lo[ax]) & (g[ax] < ilo[ax])
        air[(d2 <= r * r) & span] = True
    return air


def seeded_scenario(seed: int, shape=(24, 20, 28), pml=4, geometry=True, ade=False, plane=False,
                    nonuniform=False) -> Scenario:
    rng = np.random.default_rng(seed)
    nx, ny, nz = shape
    air = None
    if geometry:
        air = np.ones(shape, bool)
        i0, j0, k0 = nx // 3, ny // 3, nz // 3
        air[i0:i0 + 4, j0:j0 + 3, k0:k0 + 5] = False
        air[rng.integers(0, nx, 20), rng.integers(0, ny, 20), rng.integers(0, nz, 20)] = False
    material = None
    materials = {}
    if ade:
        material = np.zeros(shape, np.uint8)
        material[nx // 2:nx // 2 + 6, 2:8, nz // 2:nz // 2 + 7] = 1
        material[2:6, ny // 2:ny // 2 + 5, 3:9] = 2
        if air is not None:
            material[~air] = 0
        materials = {
            1: [dict(kind="debye", target="density", tau=2e-5, dchi=0.5),
                dict(kind="debye", target="modulus", tau=1e-5, dchi=0.3)],
            2: [dict(kind="lorentz", target="modulus", w0=2 * np.pi * 3000, gamma=500.0, dchi=0.4),
                dict(kind="lorentz", target="density", w0=2 * np.pi * 2000, gamma=300.0, dchi=0.2)],
        }
    coords = None
    if nonuniform:
        coords = tuple(T.stretched_coords(n, 1e-3, r) for n, r in zip(shape, (1.04, 1.0, 1.06)))
        if tuple(len(c) for c in coords) != tuple(shape):
            raise ValueError(f"from_stretch yields {tuple(len(c) for c in coords)} cells for {shape} "
                             "(odd counts lose one cell); use even sizes on stretched axes")
    src_pos = tuple(int(v) for v in (nx // 2 + 1, ny // 2, nz // 2 - 1))
    sources = [("point", src_pos, 8000.0)]
    if plane:
        sources.append(("plane", 2, 3, 6000.0))
    probes = [tuple(int(x) for x in (rng.integers(0, nx), rng.integers(0, ny), rng.integers(0, nz)))
              for _ in range(3)]
    probes.append(src_pos)
    return Scenario(shape=shape, pml_depth=pml, air=air, material=material, materials=materials,
                    sources=sources, probes=probes, coords=coords)


ADE_POLE_DTYPE = np.dtype([("material_id", "<i4"), ("target", "<i4"), ("lorentz", "<i4"), ("field_index", "<i4"),
                           ("a", "<f4"), ("b", "<f4"), ("d", "<f4"), ("reserved", "<f4")])


def ade_phase_case(seed: int, shape=(13, 17, 70)):
    """Random inputs of the per-phase ADE functions: fields, J arrays (two
    Debye and two Lorentz rows), material ids 0-3 and a pole list that mixes
    both kinds and targets with out-of-order field indices."""
    rng = np.random.default_rng(seed)
    f = {k: rng.standard_normal(shape).astype(np.float32) for k in ("p", "vx", "vy", "vz", "div")}
    f["Jd"] = rng.standard_normal((2,) + shape).astype(np.float32)
    f["Jl"] = rng.standard_normal((2,) + shape).astype(np.float32)
    f["Jlp"] = rng.standard_normal((2,) + shape).astype(np.float32)
    mat = rng.integers(0, 4, shape).astype(np.uint8)
    mat[:, :, 20:40] = 2          # a block of one material, so faces with both cells in it are common
    poles = np.zeros(6, dtype=ADE_POLE_DTYPE)
    rows = [(1, 0, 0, 1, 0.91, 0.05, 0.0), (2, 1, 0, 0, 0.83, 0.11, 0.0), (2, 0, 1, 1, 1.93, -0.95, 0.013),
            (3, 1, 1, 0, 1.87, -0.91, 0.021), (2, 0, 0, 0, 0.77, 0.09, 0.0), (1, 1, 1, 1, 1.95, -0.97, 0.008)]
    for q, r in enumerate(rows):
        poles[q] = r + (0.0,)
    rho_inf = (1.0 + rng.random(256)).astype(np.float32)
    K_inf = (1.0e5 + 1.0e5 * rng.random(256)).astype(np.float32)
    return f, mat, poles, rho_inf, K_inf
