"""Host-side precompute of the product package vs the oracle (no GPU needed).

The product builds every O(n) table the kernels read (time step, coefficients,
PML decays, nonuniform spacings, waveform tables, pole coefficients) and the
voxelised masks. Each is compared bit for bit with the oracle's independent
restatement, and the voxeliser with the direct (reference-style) evaluation.
"""

import numpy as np
import pytest

from oracle import tables as T
from strata_fdtd_amd import (AIR_20C, PML, Box, Cylinder, Difference, GaussianPulse,
                             HelmholtzResonator, Intersection, MaterializedGeometry, NonuniformGrid,
                             Pole, PoleType, SimpleMaterial, Sphere, UniformGrid, Union)
from strata_fdtd_amd.geometry import _mask
from strata_fdtd_amd.materials import density_correction_coeff, modulus_correction_coeff
from strata_fdtd_amd.sources import accumulated_times, waveform_table


class _FakeSolver:
    """The attributes PML.initialize reads, without a GPU."""

    def __init__(self, grid, c=343.0, rho=1.204, courant=0.95):
        self.grid = grid
        self.shape = grid.shape
        self.c = c
        self.dx = grid.min_spacing
        self.dt = courant * (grid.min_spacing / (c * np.sqrt(3)))


def test_time_step_and_coefficients_match_oracle():
    s = _FakeSolver(UniformGrid((10, 10, 10), 1e-3))
    assert s.dt == T.time_step(1e-3, 343.0)
    cv, cp = T.uniform_coefficients(s.dt, 1e-3, 1.204, 343.0)
    assert np.float32(cv) == np.float32(-s.dt / (1.204 * 1e-3))
    assert np.float32(cp) == np.float32(-1.204 * 343.0**2 * s.dt / 1e-3)


@pytest.mark.parametrize("depth,axes", [(10, "all"), (4, "x"), (6, ("y", "z"))])
def test_pml_decay_tables_match_oracle(oracle_lib, depth, axes):
    grid = UniformGrid((40, 36, 44), 1e-3)
    s = _FakeSolver(grid)
    pml = PML(depth=depth, axis=axes)
    pml.initialize(s)
    tabs = pml.decay_tables(s.dt, grid.shape)
    sm = T.pml_sigma_max(depth, 3, 343.0, 1e-3)
    for ax, name in enumerate("xyz"):
        if name in pml.axes:
            want = T.pml_decay(T.pml_profile_uniform(grid.shape[ax], depth, 3, sm), s.dt)
        else:
            want = np.ones(grid.shape[ax], np.float32)
        assert tabs[ax].dtype == np.float32
        assert np.array_equal(tabs[ax], want), name
        assert np.all(tabs[ax][depth:-depth] == 1.0)


def test_nonuniform_tables_match_oracle(oracle_lib):
    grid = NonuniformGrid.from_stretch((30, 34, 40), 1e-3, stretch_x=1.04, stretch_z=1.06)
    want = T.stencil_arrays(T.stretched_coords(30, 1e-3, 1.04), T.stretched_coords(34, 1e-3, 1.0),
                            T.stretched_coords(40, 1e-3, 1.06))
    got = grid.get_spacing_arrays_for_stencil()
    for k, v in want.items():
        assert np.array_equal(got[k], v), k
    s = _FakeSolver(grid)
    pml = PML(depth=5)
    pml.initialize(s)
    tabs = pml.decay_tables(s.dt, grid.shape)
    sm = T.pml_sigma_max(5, 3, 343.0, grid.min_spacing)
    for ax, c in enumerate((grid.x_coords, grid.y_coords, grid.z_coords)):
        prof = T.pml_profile_nonuniform(c, T.cell_spacing(c), 5, 3, sm)
        assert np.array_equal(tabs[ax], T.pml_decay(prof, s.dt))


def test_from_stretch_odd_count_keeps_reference_quirk():
    """core/grid.py:386-398 yields n-1 centres for an odd n with center_fine."""
    g = NonuniformGrid.from_stretch((21, 20, 20), 1e-3, stretch_x=1.05)
    assert g.shape == (20, 20, 20)
    assert np.array_equal(g.x_coords, T.stretched_coords(21, 1e-3, 1.05))


def test_waveform_table_matches_oracle():
    dt = T.time_step(1e-3, 343.0)
    for f in (200.0, 1000.0, 4000.0):
        src = GaussianPulse(position=(1, 1, 1), frequency=f)
        t = accumulated_times(dt, 3000)
        assert np.array_equal(t, T.accumulated_times(dt, 3000))
        assert np.array_equal(waveform_table(src, t, dt), T.waveform_table(t, f))


def test_material_from_impedance():
    from strata_fdtd_amd import SimpleMaterial
    m = SimpleMaterial.from_impedance("water-like", rho=998.2, Z=998.2 * 1482.0)
    assert m.name == "water-like" and m.rho_inf == 998.2
    assert abs(m.c_inf - 1482.0) < 1e-9 and abs(m.K_inf - 998.2 * 1482.0**2) < 1e-3


def test_material_frequency_response_closed_forms():
    """AcousticMaterial (materials/base.py:215-505): a lossless material has
    k = w / c, Z = rho c, no attenuation, reflection (Z - Z') / (Z + Z');
    a Debye density pole gives rho(w) = rho (1 + dchi / (1 + i w tau)), with
    the low- and high-frequency limits rho (1 + dchi) and rho, and loss in
    between (Im k < 0: the poles' susceptibilities use the e^{+iwt}
    convention, under which a decaying wave has a negative Im k)."""
    from strata_fdtd_amd import SimpleMaterial
    from strata_fdtd_amd.materials import AcousticMaterial
    w = 2 * np.pi * np.array([100.0, 1000.0, 10000.0])
    air = SimpleMaterial(name="air", _rho=1.2, _c=343.0)
    assert isinstance(air, AcousticMaterial) and air.n_poles() == 0
    np.testing.assert_allclose(air.wavenumber(w), w / 343.0, rtol=1e-15)
    np.testing.assert_allclose(air.phase_velocity(w), 343.0, rtol=1e-15)
    np.testing.assert_allclose(air.group_velocity(w), 343.0, rtol=1e-9)
    np.testing.assert_allclose(air.impedance(w), 1.2 * 343.0, rtol=1e-15)
    assert np.all(air.attenuation(w) == 0) and air.Z_inf == 1.2 * 343.0
    assert np.allclose(air.reflection_coefficient(w, 1.2 * 343.0), 0) and np.allclose(
        air.absorption_coefficient(w, 1.2 * 343.0), 1)
    np.testing.assert_allclose(air.reflection_coefficient(w, 3 * 411.6), (411.6 - 3 * 411.6) / (4 * 411.6))
    tau, dchi = 1e-4, 0.5
    m = SimpleMaterial(name="debye", _rho=1.2, _c=343.0, _poles=[Pole(PoleType.DEBYE, dchi, "density", tau=tau)])
    assert m.n_density_poles() == 1 and m.n_modulus_poles() == 0
    np.testing.assert_allclose(m.effective_density(w), 1.2 * (1 + dchi / (1 + 1j * w * tau)), rtol=1e-15)
    np.testing.assert_allclose(m.effective_modulus(w), m.K_inf, rtol=1e-15)
    np.testing.assert_allclose(m.effective_density(1e-3), 1.2 * (1 + dchi), rtol=1e-6)
    np.testing.assert_allclose(m.effective_density(1e12), 1.2, rtol=1e-6)
    assert np.all(m.attenuation(w) < 0)
    np.testing.assert_allclose(m.attenuation_db_per_meter(w), 20 / np.log(10) * m.attenuation(w), rtol=1e-15)
    assert "Poles: 1 (1 density, 0 modulus)" in m.summary()


def test_pole_and_correction_coefficients_match_oracle():
    dt = T.time_step(1e-3, 343.0)
    p = Pole(PoleType.DEBYE, 0.3, "density", tau=2e-5)
    assert p.fdtd_coefficients(dt) == T.debye_coefficients(2e-5, 0.3, dt)
    q = Pole(PoleType.LORENTZ, 0.4, "modulus", omega_0=2e4, gamma=300.0)
    assert q.fdtd_coefficients(dt) == T.lorentz_coefficients(2e4, 300.0, 0.4, dt)
    assert float(density_correction_coeff(dt, 1.204, 1e3)) == T.native_density_correction(dt, 1.204, 1e-3)
    assert float(modulus_correction_coeff(dt, 1.204 * 343.0**2)) == T.native_modulus_correction(
        dt, 1.204 * 343.0**2)


# ------------------------------------------------------------------ voxeliser
def _random_tree(rng, depth=0):
    def r(a, b, n=None, d=3):
        return np.round(rng.uniform(a, b, n), d)
    t = rng.integers(0, 6 if depth < 2 else 3)
    c = r(0.01, 0.08, 3, 4)
    if t == 0:
        return Box(center=tuple(c), size=tuple(r(0.005, 0.05, 3)))
    if t == 1:
        return Sphere(center=tuple(c), radius=float(r(0.005, 0.03)))
    if t == 2:
        d = np.zeros(3)
        d[rng.integers(0, 3)] = r(0.01, 0.05)
        if rng.random() < 0.2:
            d = r(-0.03, 0.03, 3)  # oblique axis: direct evaluation path
        return Cylinder(p1=tuple(c), p2=tuple(c + d), radius=float(r(0.003, 0.02)))
    kids = [_random_tree(rng, depth + 1) for _ in range(rng.integers(1, 4))]
    return (Union, Intersection, lambda *k: Difference(k[0], *k[1:]))[t - 3](*kids)


def test_exact_voxeliser_equals_direct_evaluation():
    """Boolean CSG voxelisation == sdf <= 0 evaluated at every cell, including
    cells exactly on a surface (grid-aligned coordinates create ties)."""
    rng = np.random.default_rng(0)
    grid = UniformGrid((64, 56, 48), 1e-3)
    ties = 0
    for _ in range(40):
        node = _random_tree(rng)
        for strict in (False, True):
            assert np.array_equal(_mask(node, grid, strict), node.voxelize_direct(grid, None, strict))
        ties += int((node.voxelize_direct(grid, None, False) != node.voxelize_direct(grid, None, True)).sum())
    assert ties > 0


def test_voxeliser_slabs_equal_single_pass():
    node = Difference(Box(center=(0.03, 0.03, 0.03), size=(0.05, 0.04, 0.05)),
                      Sphere(center=(0.03, 0.03, 0.03), radius=0.012))
    grid = UniformGrid((48, 40, 44), 1e-3)
    assert np.array_equal(node.voxelize_direct(grid, 2000), node.voxelize_direct(grid, None))


def test_voxelize_known_answers():
    """Spot checks from the reference tests (tests/test_sdf.py:128-167, 341-356)."""
    grid = UniformGrid((100, 100, 100), 1e-3)
    box = Box(center=(0.05, 0.05, 0.05), size=(0.02, 0.02, 0.02))
    m = box.voxelize(grid)
    assert m[50, 50, 50] and not m[0, 0, 0] and not m[99, 99, 99]
    assert m.sum() == 20**3
    cyl = Cylinder(p1=(0.05, 0.05, 0.04), p2=(0.05, 0.05, 0.06), radius=0.005)
    mc = cyl.voxelize(grid)
    assert mc[50, 50, 50] and not mc[0, 0, 0]
    sph = Sphere(center=(0.05, 0.05, 0.05), radius=0.01)
    ms = sph.voxelize(grid)
    ii, jj, kk = np.nonzero(ms)
    d = np.sqrt(((ii + 0.5) * 1e-3 - 0.05) ** 2 + ((jj + 0.5) * 1e-3 - 0.05) ** 2 + ((kk + 0.5) * 1e-3 - 0.05) ** 2)
    assert d.max() <= 0.01 + 1e-12
    nug = NonuniformGrid.from_stretch((50, 50, 50), 1e-3, stretch_z=1.05)
    mn = box.voxelize(nug)
    assert mn.any() and not mn.all()


def test_materialized_geometry_known_answers():
    """tests/test_material_geometry.py:86-157: regions only claim air, later ones win."""
    grid = UniformGrid((10, 10, 10), 1e-2)
    geo = MaterializedGeometry(Box(center=(0.05, 0.05, 0.05), size=(0.1, 0.1, 0.1)))
    m1 = SimpleMaterial(name="m1", _rho=50.0, _c=300.0)
    m2 = SimpleMaterial(name="m2", _rho=60.0, _c=310.0)
    geo.set_material(Box(center=(0.03, 0.05, 0.05), size=(0.06, 0.1, 0.1)), m1)
    geo.set_material(Box(center=(0.05, 0.05, 0.05), size=(0.04, 0.08, 0.08)), m2)
    mask, ids = geo.voxelize_with_materials(grid)
    assert mask.all()
    assert ids[5, 5, 5] == 2 and ids[0, 5, 5] == 1 and ids[9, 0, 0] == 0
    assert geo.get_material_table()[0] is AIR_20C


def test_helmholtz_resonator_union():
    res = HelmholtzResonator(position=(0.03, 0.03, 0.03), cavity_shape="cylinder",
                             cavity_size=(0.01, 0.02), neck_length=0.005, neck_radius=0.003)
    grid = UniformGrid((60, 60, 60), 1e-3)
    assert np.array_equal(res.voxelize(grid), res.voxelize_direct(grid, None))


# ------------------------------------------------------------ device voxeliser
def _axis_aligned_tree(rng):
    while True:
        node = _random_tree(rng)
        from strata_fdtd_amd.voxelize_device import supports
        if supports(node):
            return node


@pytest.mark.parametrize("nonuniform", [False, True])
def test_device_voxeliser_programs_restate_host_voxeliser(nonuniform):
    """The postfix programs the device voxeliser runs, evaluated by a numpy
    restatement of the kernel, give the host voxeliser's mask (itself pinned
    to the reference's direct sdf <= 0 evaluation) on random CSG trees,
    surface ties included; Difference flips the subtrahend to a strict test."""
    from strata_fdtd_amd.voxelize_device import compile_programs, evaluate_host
    rng = np.random.default_rng(5)
    grid = (NonuniformGrid.from_stretch((64, 56, 48), 1e-3, stretch_z=1.03) if nonuniform
            else UniformGrid((64, 56, 48), 1e-3))
    for _ in range(30):
        node = _axis_aligned_tree(rng)
        air, _ = evaluate_host(compile_programs(node), grid.x_coords, grid.y_coords, grid.z_coords)
        assert np.array_equal(air, node.voxelize(grid))


def test_device_voxeliser_programs_material_ids():
    from strata_fdtd_amd.voxelize_device import compile_programs, evaluate_host
    rng = np.random.default_rng(9)
    grid = UniformGrid((40, 44, 36), 1e-3)
    for _ in range(6):
        geo = MaterializedGeometry(_axis_aligned_tree(rng))
        for q in range(3):
            geo.set_material(_axis_aligned_tree(rng), SimpleMaterial(name=f"m{q}", _rho=1.0 + q, _c=300.0))
        mask, ids = geo.voxelize_with_materials(grid)
        ps = compile_programs(geo.geometry, [(r, mid) for r, _m, mid in geo._material_regions])
        air, ids_p = evaluate_host(ps, grid.x_coords, grid.y_coords, grid.z_coords)
        assert np.array_equal(air, mask) and np.array_equal(ids_p, ids)


def test_device_voxeliser_rejects_oblique_cylinder():
    from strata_fdtd_amd.voxelize_device import Unsupported, compile_programs, supports
    cyl = Cylinder(p1=(0.01, 0.01, 0.01), p2=(0.02, 0.03, 0.04), radius=0.004)
    assert not supports(Union(Box(center=(0.02, 0.02, 0.02), size=(0.01, 0.01, 0.01)), cyl))
    with pytest.raises(Unsupported):
        compile_programs(cyl)
    assert supports(Cylinder(p1=(0.01, 0.01, 0.04), p2=(0.01, 0.01, 0.01), radius=0.004))


def test_voxelize_entry_point_validates_programs():
    """fdtd_voxelize rejects malformed programs before touching the device."""
    from strata_fdtd_amd import _native
    lib = _native.load()
    fake = 4096  # never dereferenced: validation fails first

    def run(ops, start, ids, params):
        ops = np.asarray(ops, np.int32).reshape(-1, 2)
        start = np.asarray(start, np.int32)
        ids = np.asarray(ids, np.int32)
        params = np.asarray(params, np.float64)
        rc = lib.fdtd_voxelize(ops.ctypes.data, len(ops), start.ctypes.data, ids.ctypes.data, len(start) - 1,
                               params.ctypes.data, len(params), fake, 1 << 20, fake, fake, fake, 4, 4, 4,
                               fake, fake, None)
        return rc, lib.fdtd_last_error().decode()

    box = [0.0] * 6
    assert run([(12, 0)], [0, 1], [0], box)[1].endswith("unknown opcode")
    # transforms: balanced, at most 8 deep, scale factors in [1e-6, 1e6]
    one = [1.0] * 6
    assert "restore every transformed point" in run([(9, 0), (2, 0)], [0, 2], [0], one)[1]
    assert "transform stack underflow" in run([(2, 0), (11, 0)], [0, 2], [0], one)[1]
    assert "scale factors" in run([(10, 0), (2, 0), (11, 0)], [0, 3], [0], [0.0] * 6)[1]
    assert "nested deeper than 8" in run([(9, 0)] * 9 + [(2, 0)] + [(11, 0)] * 9, [0, 19], [0], one)[1]
    assert "underflow" in run([(2, 0), (5, 0)], [0, 2], [0], box)[1]
    assert "out of range" in run([(2, 3)], [0, 1], [0], box)[1]
    assert "exactly one value" in run([(2, 0), (2, 0)], [0, 2], [0], box)[1]
    cyl = [0, 0, 0, 0.6, 0.8, 0, 1.0, 0.1, 0]
    assert "signed unit axis" in run([(4, 0)], [0, 1], [0], cyl)[1]
    assert "material ids" in run([(2, 0), (2, 0)], [0, 1, 2], [0, 0], box)[1]


def test_tiling_heuristic_follows_the_cache_boundary():
    """2-row strips for solves whose 33 B/cell fit the 256 MiB Infinity Cache
    (2 planes, 1 when the grid is too small for 2048 such waves); 4 rows x 4
    planes beyond."""
    from strata_fdtd_amd.tuning import CACHE_RESIDENT_CELLS, heuristic_tiling
    assert CACHE_RESIDENT_CELLS == (256 << 20) // 33
    assert heuristic_tiling((64, 64, 64)) == (2, 1, 1)
    assert heuristic_tiling((100, 100, 100)) == (2, 2, 1)
    assert heuristic_tiling((192, 192, 192)) == (2, 2, 1)
    assert heuristic_tiling((224, 224, 224)) == (4, 4, 1)
    assert heuristic_tiling((512, 512, 512)) == (4, 4, 1)
    assert heuristic_tiling((320, 320, 480)) == (4, 4, 1)
    # the boundary itself: one cell more leaves the cache
    n = CACHE_RESIDENT_CELLS
    assert heuristic_tiling((1, 1, n))[0] == 2 and heuristic_tiling((1, 1, n + 1))[0] == 4
    # the autotuner's candidates include the heuristic's choice
    from strata_fdtd_amd.autotune import candidate_tilings
    for shape in ((64, 64, 64), (100, 100, 100), (192, 192, 192), (512, 512, 512)):
        assert heuristic_tiling(shape) in candidate_tilings(shape)


def test_table_source_routing():
    """Membranes and z-normal planes are table sources (z planes stay in the
    sweep on grids of 2^31 cells or more); points and x/y planes are not."""
    from strata_fdtd_amd.solver import _table_source
    from strata_fdtd_amd import CircularMembraneSource
    plane = lambda ax: GaussianPulse(position={"axis": ax, "index": 3}, frequency=1e3, source_type="plane")
    point = GaussianPulse(position=(1, 2, 3), frequency=1e3)
    mem = CircularMembraneSource(center=(0.01, 0.01, 0.01), radius=3e-3, normal_axis="z",
                                 waveform=GaussianPulse(position=(0, 0, 0), frequency=1e3))
    small, huge = 10 ** 6, 2 ** 31
    assert not _table_source(point, small)
    assert not _table_source(plane(0), small) and not _table_source(plane(1), small)
    assert _table_source(plane(2), small) and not _table_source(plane(2), huge)
    assert _table_source(mem, small) and _table_source(mem, huge)


def test_bench_names_the_sweep_instance_rocprof_reports():
    import bench
    cr = (256 << 20) // 33
    assert bench.kernel_name("k_fused_step", 4, 1, True, 512 ** 3) == "k_fused_step<4,false,uint32_t,true>"
    assert bench.kernel_name("k_fused_step", 2, 1, True, cr) == "k_fused_step<2,false,uint32_t,false>"
    assert bench.kernel_name("k_fused_step", 4, 1, False, 1 << 30) == "k_fused_step<4,true,long,true>"
    assert bench.kernel_name("k_fused_stack", 4, 2, True, 512 ** 3) == "k_fused_stack<4,false,2>"


def test_bench_traffic_only_for_the_profiled_instance():
    """roofline.traffic comes from the committed PMC summary only when that
    summary names the workload and the kernel instance the run launches."""
    import bench
    k = bench.kernel_name("k_fused_step", 4, 1, True, 512 ** 3)
    nbytes, note = bench.traffic_from_profiles("enclosure_512", k)
    assert nbytes and nbytes > 4.4e9 and "FETCH_SIZE" in note
    nbytes, note = bench.traffic_from_profiles("enclosure_512", bench.kernel_name("k_fused_step", 2, 1, True, 512 ** 3))
    assert nbytes is None and "this run launches" in note
    assert bench.traffic_from_profiles("no_such_workload", k)[0] is None
