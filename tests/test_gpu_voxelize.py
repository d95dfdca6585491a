"""Device voxeliser vs the reference's masks and the host voxeliser, bit for bit.

The masks and material ids the reference's own sdf.py wrote
(tests/golden/voxel_masks.npz, also checked against the host voxeliser in
tests/test_voxel_golden.py) must come out of the HIP kernel exactly. Beyond
those: random CSG trees on uniform and nonuniform grids (surface ties
included), material regions, and the full-size scenes of the 512^3 enclosure
and the 320x320x480 nonuniform resonator plate, against the host voxeliser.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _random_tree(rng, depth=0):
    from strata_fdtd_amd import Box, Cylinder, Difference, Intersection, Sphere, Union

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
        d[rng.integers(0, 3)] = r(0.01, 0.05) * (1 if rng.random() < 0.5 else -1)
        return Cylinder(p1=tuple(c), p2=tuple(c + d), radius=float(r(0.003, 0.02)))
    kids = [_random_tree(rng, depth + 1) for _ in range(rng.integers(1, 4))]
    return (Union, Intersection, lambda *k: Difference(k[0], *k[1:]))[t - 3](*kids)


def _device(geometry, grid):
    from strata_fdtd_amd.voxelize_device import voxelize_device
    air, ids = voxelize_device(geometry, grid)
    return air.cpu().numpy().astype(bool), None if ids is None else ids.cpu().numpy()


@pytest.mark.parametrize("nonuniform", [False, True])
def test_translated_scaled_trees_equal_host(nonuniform):
    """Translations and scales run in the device programs: random translated
    and scaled trees (cones and surface ties included) and material regions
    give the host voxeliser's masks and ids bit for bit."""
    import strata_fdtd_amd as sfa
    from test_transforms import translated_scaled_tree
    rng = np.random.default_rng(23)
    grid = (sfa.NonuniformGrid.from_stretch((50, 52, 48), 1e-3, stretch_z=1.03) if nonuniform
            else sfa.UniformGrid((50, 52, 48), 1e-3))
    print(f"[voxelize] translated/scaled trees nonuniform={nonuniform}", flush=True)
    for _ in range(30):
        node = sfa.Union(translated_scaled_tree(rng), translated_scaled_tree(rng))
        dev, _ = _device(node, grid)
        host = node.voxelize(grid)
        assert np.array_equal(dev, host), int((dev != host).sum())
    geo = sfa.MaterializedGeometry(translated_scaled_tree(rng))
    for q in range(3):
        geo.set_material(translated_scaled_tree(rng), sfa.SimpleMaterial(name=f"m{q}", _rho=1.0 + q, _c=300.0))
    mask, ids = geo.voxelize_with_materials(grid)
    dmask, dids = _device(geo, grid)
    assert np.array_equal(dmask, mask) and np.array_equal(dids, ids)


@pytest.mark.parametrize("nonuniform", [False, True])
def test_random_trees_equal_host(nonuniform):
    from strata_fdtd_amd import NonuniformGrid, UniformGrid
    rng = np.random.default_rng(21)
    grid = (NonuniformGrid.from_stretch((64, 56, 48), 1e-3, stretch_z=1.03) if nonuniform
            else UniformGrid((64, 56, 48), 1e-3))
    print(f"[voxelize] random trees nonuniform={nonuniform}", flush=True)
    ties = 0
    for _ in range(60):
        node = _random_tree(rng)
        host = node.voxelize(grid)
        dev, _ = _device(node, grid)
        assert np.array_equal(dev, host), int((dev != host).sum())
        ties += int((node.voxelize_direct(grid, None, True) != host).sum())
    assert ties > 0  # cells exactly on a surface were exercised


def test_material_regions_equal_host():
    from strata_fdtd_amd import MaterializedGeometry, SimpleMaterial, UniformGrid
    rng = np.random.default_rng(22)
    grid = UniformGrid((48, 40, 52), 1e-3)
    for _ in range(8):
        geo = MaterializedGeometry(_random_tree(rng))
        for q in range(4):
            geo.set_material(_random_tree(rng), SimpleMaterial(name=f"m{q}", _rho=1.0 + q, _c=300.0))
        mask, ids = geo.voxelize_with_materials(grid)
        dmask, dids = _device(geo, grid)
        assert np.array_equal(dmask, mask)
        assert np.array_equal(dids, ids)


def test_enclosure_512_scene_equals_host():
    from strata_fdtd_amd.workloads import enclosure_512_scene
    from strata_fdtd_amd import UniformGrid
    geo, *_ = enclosure_512_scene(512)
    grid = UniformGrid((512, 512, 512), 1e-3)
    print("[voxelize] enclosure_512 scene", flush=True)
    mask, ids = geo.voxelize_with_materials(grid)
    dmask, dids = _device(geo, grid)
    assert np.array_equal(dmask, mask)
    assert np.array_equal(dids, ids)
    assert 0 < mask.sum() < mask.size and ids.any()


def test_resonator_plate_scene_equals_host():
    from strata_fdtd_amd.workloads import resonators_nu_scene
    grid, geo, *_ = resonators_nu_scene()
    print("[voxelize] resonators_nu scene", flush=True)
    host = geo.voxelize(grid)
    dev, _ = _device(geo, grid)
    assert np.array_equal(dev, host)


def test_solver_uses_device_voxeliser_by_default(monkeypatch):
    from strata_fdtd_amd import FDTDSolver, MaterializedGeometry, UniformGrid
    from strata_fdtd_amd import voxelize_device as VD
    from strata_fdtd_amd.workloads import enclosure_512_scene
    calls = []
    real = VD.voxelize_device
    monkeypatch.setattr(VD, "voxelize_device", lambda *a, **k: calls.append(1) or real(*a, **k))
    geo, *_ = enclosure_512_scene(96)
    assert isinstance(geo, MaterializedGeometry)
    a = FDTDSolver(grid=UniformGrid((96, 96, 96), 1e-3), rho=1.204, backend="mi355x")
    b = FDTDSolver(grid=UniformGrid((96, 96, 96), 1e-3), rho=1.204, backend="mi355x", voxelizer="host")
    a.set_geometry(geo)
    assert len(calls) == 1
    b.set_geometry(geo)
    assert len(calls) == 1
    assert np.array_equal(a.geometry, b.geometry)
    assert np.array_equal(a._material_id, b._material_id)


def test_device_voxeliser_equals_reference_masks():
    """Every scene of tests/golden/voxel_masks.npz (masks and material ids
    written by the reference's own sdf.py / material_assignment.py): the
    device voxeliser reproduces them bit for bit."""
    from strata_fdtd_amd.voxelize_device import Unsupported
    from test_voxel_golden import CASES, product_grid, product_namespace, scene_tree
    ns = product_namespace()
    print(f"[voxelize] {len(CASES)} reference-mask cases", flush=True)
    host_only = 0
    for name, case in sorted(CASES.items()):
        obj = scene_tree.build(case["tree"], ns)
        tree_text = scene_tree.dumps(case["tree"])
        if (name.startswith("cone_oblique") or '"rotate' in tree_text or '"smooth_' in tree_text
                or '"horn"' in tree_text or '"spiral_horn"' in tree_text or '"helical_tube"' in tree_text):
            # tilted axes and rotations go through BLAS in the reference,
            # smooth CSG needs distances rather than inside tests, and horns
            # numpy's exp / cosh / cos: the device voxeliser refuses them and
            # a solver falls back to the host
            with pytest.raises(Unsupported):
                _device(obj, product_grid(case))
            host_only += 1
            continue
        dmask, dids = _device(obj, product_grid(case))
        assert np.array_equal(dmask, case["mask"]), (name, int((dmask != case["mask"]).sum()))
        if "ids" in case:
            assert np.array_equal(dids, case["ids"]), name
    # 3 tilted cone trees, 7 trees with rotations, 9 smooth trees, the 2
    # transformed material scenes (each with a smooth region), 16 horn trees
    # and 6 swept-volume scenes; the other 8 transformed trees (translations
    # and scales only) run on the device
    assert host_only == 3 + 7 + 9 + 2 + 16 + 6
