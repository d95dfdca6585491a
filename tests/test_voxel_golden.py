"""Host voxeliser against masks written by the reference's own SDF code.

tests/golden/voxel_masks.npz holds air masks and material ids that the
reference's ``geometry/sdf.py`` (SDFPrimitive.voxelize :99-127),
``geometry/resonator.py`` and ``geometry/material_assignment.py``
(voxelize_with_materials :133-168) produced for scenes recorded as data trees
(tests/golden/make_voxel_golden.py): the 512^3 enclosure scene at 64^3, a
scaled-down nonuniform resonator plate, three serving requests, random CSG
trees whose millimetre-rounded parameters put many cell centres exactly on a
surface, random material regions, Helmholtz resonators, cones, transformed
trees (translations, scales, rotations), smooth CSG, horns, spiral horns and
helical tubes. The product's
voxeliser must reproduce every mask and id bit for bit. The device voxeliser
is held to the same fixture in tests/test_gpu_voxelize.py.
"""

import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import scene_tree  # noqa: E402

FIXTURE = os.path.join(HERE, "golden", "voxel_masks.npz")


def load_cases():
    g = np.load(FIXTURE, allow_pickle=False)
    out = {}
    for name in g["names"]:
        name = str(name)
        shape = tuple(int(v) for v in g[f"{name}__shape"])
        case = dict(tree=scene_tree.loads(str(g[f"{name}__tree"])), shape=shape,
                    mask=np.unpackbits(g[f"{name}__mask"])[: int(np.prod(shape))].reshape(shape).astype(bool))
        if f"{name}__ids" in g:
            case["ids"] = g[f"{name}__ids"]
        if f"{name}__x" in g:
            case["coords"] = tuple(g[f"{name}__{a}"] for a in "xyz")
        else:
            case["dx"] = float(g[f"{name}__dx"])
        out[name] = case
    return out


CASES = load_cases()


def product_namespace():
    import strata_fdtd_amd as sfa
    return dict(Box=sfa.Box, Sphere=sfa.Sphere, Cylinder=sfa.Cylinder, Cone=sfa.Cone, Union=sfa.Union,
                Intersection=sfa.Intersection, Difference=sfa.Difference, SmoothUnion=sfa.SmoothUnion,
                SmoothIntersection=sfa.SmoothIntersection, SmoothDifference=sfa.SmoothDifference,
                Translate=sfa.Translate, Scale=sfa.Scale, Rotate=sfa.Rotate, Horn=sfa.Horn,
                SpiralHorn=sfa.SpiralHorn, HelicalTube=sfa.HelicalTube,
                HelmholtzResonator=sfa.HelmholtzResonator, MaterializedGeometry=sfa.MaterializedGeometry,
                material=sfa.SimpleMaterial(name="m", _rho=1.2, _c=340.0))


def product_grid(case):
    import strata_fdtd_amd as sfa
    if "coords" in case:
        return sfa.NonuniformGrid(*case["coords"])
    return sfa.UniformGrid(case["shape"], case["dx"])


def test_fixture_covers_the_scenes():
    for prefix in ("enclosure_64", "plate_nu", "request_", "tree_u", "tree_nu", "materials_", "helmholtz_"):
        assert any(n.startswith(prefix) for n in CASES), prefix
    assert sum(n.startswith("tree_") for n in CASES) >= 30


@pytest.mark.parametrize("name", sorted(CASES))
def test_host_voxeliser_equals_reference_masks(name):
    case = CASES[name]
    obj = scene_tree.build(case["tree"], product_namespace())
    grid = product_grid(case)
    if "ids" in case:
        mask, ids = obj.voxelize_with_materials(grid)
        assert np.array_equal(ids, case["ids"])
    else:
        mask = obj.voxelize(grid)
    assert np.array_equal(mask, case["mask"]), int((mask != case["mask"]).sum())


def test_random_trees_exercise_surface_ties():
    """Cells whose centre lies exactly on a surface (sdf == 0) decide the
    '<= 0' rule; the fixture's trees must contain such cells."""
    ns = product_namespace()
    ties = 0
    for name, case in CASES.items():
        if not name.startswith("tree_"):
            continue
        obj = scene_tree.build(case["tree"], ns)
        grid = product_grid(case)
        ties += int((obj.voxelize_direct(grid, None, False) != obj.voxelize_direct(grid, None, True)).sum())
    assert ties > 100


def test_workload_scenes_are_the_recorded_ones():
    """The workload builders still produce the scenes the fixture recorded, so
    the reference masks pin the benchmark and serving scenes themselves."""
    from strata_fdtd_amd.serving import sweep_requests
    from strata_fdtd_amd.workloads import enclosure_512_scene, resonators_nu_scene
    geo, *_ = enclosure_512_scene(64)
    assert scene_tree.to_tree(geo) == CASES["enclosure_64"]["tree"]
    grid, air, *_ = resonators_nu_scene(shape=(96, 96, 128), core=60, n_side=3, pitch=24, plate=24,
                                        cavity_radius=8, cavity_depth=14, neck_radius=3, neck_length=4)
    assert scene_tree.to_tree(air) == CASES["plate_nu"]["tree"]
    assert all(np.array_equal(a, b) for a, b in zip((grid.x_coords, grid.y_coords, grid.z_coords),
                                                    CASES["plate_nu"]["coords"]))
    for q, req in enumerate(sweep_requests(n=3, seed=77, dims=(64, 80))):
        assert scene_tree.to_tree(req.geometry) == CASES[f"request_{q}"]["tree"]
