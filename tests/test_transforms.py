"""Transforms and smooth CSG (geometry/sdf.py Translate / Scale / Rotate
:1188-1391, rotation_matrix :1151-1185, SmoothUnion / SmoothIntersection /
SmoothDifference :736-936) against the known answers of the reference's own
tests (tests/test_transformations.py, tests/test_sdf_csg.py) and closed forms.
Their voxel masks are pinned bit for bit to masks the reference wrote
(tests/test_voxel_golden.py, cases xform_*, smooth_*)."""

import numpy as np
import pytest

import strata_fdtd_amd as sfa


def test_rotation_matrix_known_answers():
    R = sfa.rotation_matrix(np.array([0, 0, 1]), np.pi / 2)
    np.testing.assert_allclose(R @ np.array([1, 0, 0]), [0, 1, 0], atol=1e-10)
    R = sfa.rotation_matrix(np.array([1, 0, 0]), np.pi / 2)
    np.testing.assert_allclose(R @ np.array([0, 1, 0]), [0, 0, 1], atol=1e-10)
    R = sfa.rotation_matrix(np.array([1, 1, 1]), np.pi / 3)
    np.testing.assert_allclose(R @ R.T, np.eye(3), atol=1e-10)
    assert abs(np.linalg.det(R) - 1.0) < 1e-10
    # a third of a turn about (1,1,1) permutes the axes: x -> y -> z
    R2 = sfa.rotation_matrix(np.array([1, 1, 1]), 2 * np.pi / 3)
    np.testing.assert_allclose(R2 @ np.array([1, 0, 0]), [0, 1, 0], atol=1e-12)
    np.testing.assert_allclose(R2 @ R2 @ R2, np.eye(3), atol=1e-12)
    np.testing.assert_allclose(R @ R @ R, sfa.rotation_matrix(np.array([1, 1, 1]), np.pi), atol=1e-12)
    np.testing.assert_allclose(sfa.rotation_matrix(np.array([2, 0, 0]), np.pi / 2),
                               sfa.rotation_matrix(np.array([1, 0, 0]), np.pi / 2), atol=1e-10)
    np.testing.assert_allclose(sfa.rotation_matrix(np.array([1, 0, 0]), 0.0), np.eye(3), atol=1e-10)


def test_translate_scale_rotate_known_answers():
    box = sfa.Box(center=(0, 0, 0), size=(0.02, 0.02, 0.02))
    off = np.array([0.05, 0.05, 0.05])
    t = sfa.Translate(box, off)
    assert box.sdf(np.zeros((1, 3)))[0] < 0 and t.sdf(np.zeros((1, 3)))[0] > 0
    np.testing.assert_allclose(t.sdf(off[None]), box.sdf(np.zeros((1, 3))), atol=1e-10)
    lo, hi = t.bounding_box
    np.testing.assert_allclose(lo, np.array([-0.01] * 3) + off, atol=1e-10)
    np.testing.assert_allclose(hi, np.array([0.01] * 3) + off, atol=1e-10)
    pts = np.random.default_rng(1).uniform(-0.05, 0.1, (50, 3))
    assert np.array_equal(box.translate((0.05, 0.05, 0.05)).sdf(pts), t.sdf(pts))
    with pytest.raises(ValueError):
        sfa.Translate(box, (1.0, 2.0))

    s = sfa.Scale(box, 2.0)
    assert abs(s.sdf(np.array([[0.02, 0, 0]]))[0]) < 1e-10          # surface moved to 2x
    np.testing.assert_allclose(s.sdf(pts), 2.0 * box.sdf(pts / 2.0), atol=1e-12)
    lo, hi = sfa.Scale(box, (2.0, 1.0, 0.5)).bounding_box
    np.testing.assert_allclose(lo, np.array([-0.01] * 3) * [2.0, 1.0, 0.5], atol=1e-10)
    np.testing.assert_allclose(hi, np.array([0.01] * 3) * [2.0, 1.0, 0.5], atol=1e-10)
    nu = sfa.Scale(box, (2.0, 1.0, 0.5))
    np.testing.assert_allclose(nu.sdf(pts), box.sdf(pts / [2.0, 1.0, 0.5]) * 0.5, atol=1e-15)
    for bad in (-1.0, 0.0):
        with pytest.raises(ValueError):
            sfa.Scale(box, bad)

    long = sfa.Box(center=(0, 0, 0), size=(0.04, 0.02, 0.02))
    r = sfa.Rotate(long, axis=(0, 0, 1), angle=np.pi / 2)
    assert abs(r.sdf(np.array([[0.0, 0.02, 0.0]]))[0]) < 1e-10
    lo, hi = r.bounding_box
    np.testing.assert_allclose(lo, [-0.01, -0.02, -0.01], atol=1e-10)
    np.testing.assert_allclose(hi, [0.01, 0.02, 0.01], atol=1e-10)
    np.testing.assert_allclose(sfa.Rotate(long, matrix=sfa.rotation_matrix((0, 0, 1), np.pi / 2)).sdf(pts),
                               r.sdf(pts), atol=0)
    for ax, f in ((0, "rotate_x"), (1, "rotate_y"), (2, "rotate_z")):
        e = np.zeros(3)
        e[ax] = 1.0
        assert np.array_equal(getattr(long, f)(0.3).sdf(pts), sfa.Rotate(long, axis=tuple(e), angle=0.3).sdf(pts))
    with pytest.raises(ValueError):
        sfa.Rotate(long)
    with pytest.raises(ValueError):
        sfa.Rotate(long, matrix=np.eye(2))
    moved = long.rotate_z(np.pi / 2).translate((0.05, 0.05, 0.0))
    assert abs(moved.sdf(np.array([[0.05, 0.07, 0.0]]))[0]) < 1e-10


def test_smooth_csg_closed_forms():
    a = sfa.Sphere(center=(0, 0, 0), radius=0.01)
    b = sfa.Sphere(center=(0.015, 0, 0), radius=0.01)
    k = 0.004
    pts = np.random.default_rng(2).uniform(-0.03, 0.04, (200, 3))
    d1, d2 = a.sdf(pts), b.sdf(pts)
    su, si, sd = sfa.SmoothUnion(a, b, k), sfa.SmoothIntersection(a, b, k), sfa.SmoothDifference(a, b, k)
    # The reference blends all three with h = clip(0.5 + 0.5 (d2 - d1) / k):
    # away from the blend (|d2 - d1| >= k) h is 0 or 1 and each returns the
    # smaller of its two distances exactly (for the intersection and the
    # difference too: the blend term is added, but h does not switch to the
    # larger distance); in the blend the union subtracts k h (1 - h) and the
    # others add it, at most k / 4 at d1 == d2.
    far = np.abs(d1 - d2) >= k
    assert far.sum() > 50 and (~far).sum() > 0
    assert np.array_equal(su.sdf(pts)[far], np.minimum(d1, d2)[far])
    assert np.array_equal(si.sdf(pts)[far], np.minimum(d1, d2)[far])
    far_d = np.abs(-d2 - d1) >= k
    assert np.array_equal(sd.sdf(pts)[far_d], np.minimum(d1, -d2)[far_d])
    lo = np.minimum(d1, d2)
    assert np.all(su.sdf(pts) <= lo + 1e-18) and np.all(su.sdf(pts) >= lo - k / 4 - 1e-18)
    assert np.all(si.sdf(pts) >= lo - 1e-18) and np.all(si.sdf(pts) <= np.maximum(d1, d2) + k / 4 + 1e-18)
    mid = np.array([[0.0075, 0.0, 0.0]])                            # d1 == d2
    np.testing.assert_allclose(su.sdf(mid), a.sdf(mid) - k / 4, rtol=0, atol=1e-15)
    np.testing.assert_allclose(si.sdf(mid), a.sdf(mid) + k / 4, rtol=0, atol=1e-15)
    for cls in (sfa.SmoothUnion, sfa.SmoothIntersection, sfa.SmoothDifference):
        with pytest.raises(ValueError):
            cls(a, b, 0.0)
        assert isinstance(cls(a, b, k), sfa.CSGNode)
    assert isinstance(sfa.Union(a, b), sfa.CSGNode)


def test_device_voxeliser_transform_support():
    """The device voxeliser evaluates translations and scales (factors in
    [1e-6, 1e6], nested at most 8 deep); rotations and smooth CSG go to the
    host voxeliser (voxelizer="auto")."""
    from strata_fdtd_amd.voxelize_device import Unsupported, compile_programs, supports
    box = sfa.Box(center=(0.02, 0.02, 0.02), size=(0.01, 0.01, 0.01))
    assert supports(sfa.Union(box, box.translate((0.001, 0, 0)), box.scale(1.1), box.scale((0.9, 1.0, 1.2))))
    for node in (box.rotate_z(0.2), sfa.SmoothUnion(box, box.translate((0.005, 0, 0)), 0.002),
                 box.scale(1e-7), box.translate((0.001, 0, 0)).scale(2.0).rotate_x(0.1)):
        assert not supports(sfa.Union(box, node))
        with pytest.raises(Unsupported):
            compile_programs(sfa.Difference(box, node))
    deep = box
    for _ in range(8):
        deep = deep.translate((1e-4, 0, 0))
    assert supports(deep)
    with pytest.raises(Unsupported):
        compile_programs(deep.scale(1.01))


def translated_scaled_tree(rng):
    """A random tree (cones and surface ties included) under one to three
    translations (whole millimetres) and uniform or per-axis scales."""
    import os
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
    from make_voxel_golden import cone_tree, random_tree
    child = cone_tree(rng, ties=True) if rng.random() < 0.3 else random_tree(rng, ties=bool(rng.random() < 0.5))
    for _ in range(rng.integers(1, 4)):
        if rng.random() < 0.5:
            child = child.translate(tuple(float(v) for v in rng.integers(-6, 7, 3) * 1e-3))
        elif rng.random() < 0.5:
            # Scale(...) rather than .scale(): a Scale's own `scale` attribute
            # (its factors) hides the method, in the reference as here
            child = sfa.Scale(child, float(np.round(rng.uniform(0.6, 1.4), 2)))
        else:
            child = sfa.Scale(child, tuple(float(v) for v in np.round(rng.uniform(0.7, 1.3, 3), 2)))
    return child


@pytest.mark.parametrize("nonuniform", [False, True])
def test_device_programs_with_transforms_restate_host_voxeliser(nonuniform):
    """Programs of translated and scaled trees (cones included), evaluated by
    the numpy restatement of the kernel, give the host voxeliser's masks
    (pinned to the reference's masks) on random trees, surface ties
    included, with material regions."""
    from strata_fdtd_amd.voxelize_device import compile_programs, evaluate_host
    rng = np.random.default_rng(17)
    grid = (sfa.NonuniformGrid.from_stretch((50, 52, 48), 1e-3, stretch_z=1.03) if nonuniform
            else sfa.UniformGrid((50, 52, 48), 1e-3))

    def xtree():
        return translated_scaled_tree(rng)

    checked = 0
    for _ in range(25):
        node = sfa.Union(xtree(), xtree()) if rng.random() < 0.5 else sfa.Difference(xtree(), xtree())
        air, _ = evaluate_host(compile_programs(node), grid.x_coords, grid.y_coords, grid.z_coords)
        want = node.voxelize(grid)
        assert np.array_equal(air, want)
        checked += int(0 < want.sum() < want.size)
    assert checked >= 10
    geo = sfa.MaterializedGeometry(xtree())
    for q in range(3):
        geo.set_material(xtree(), sfa.SimpleMaterial(name=f"m{q}", _rho=1.0 + q, _c=300.0))
    mask, ids = geo.voxelize_with_materials(grid)
    ps = compile_programs(geo.geometry, [(r, mid) for r, _m, mid in geo._material_regions])
    air, ids_p = evaluate_host(ps, grid.x_coords, grid.y_coords, grid.z_coords)
    assert np.array_equal(air, mask) and np.array_equal(ids_p, ids)


@pytest.mark.parametrize("profile", ["conical", "exponential", "hyperbolic", "tractrix"])
def test_horn_profiles_known_answers(profile):
    """Horn (geometry/sdf.py:937-1150): every profile runs from the throat
    radius to the mouth radius, grows monotonically, and the cutoff has its
    closed form; on the axis inside the horn the distance is minus the
    profile radius, beyond the mouth it is the distance to the mouth rim's
    plane combined with the radial excess (reference tests/test_sdf.py
    TestHorn: radius_at_position, cutoff_frequency)."""
    h = sfa.Horn((0.0, 0.0, 0.1), (0.0, 0.0, 0.0), 0.01, 0.03, profile)
    r = h.radius_at_position(np.linspace(0, 1, 11))
    np.testing.assert_allclose(r[0], 0.01, rtol=1e-12)
    np.testing.assert_allclose(r[-1], 0.03, rtol=1e-12)
    assert np.all(np.diff(r) > 0)
    c, L = 343.0, 0.1
    want = {"conical": c / (2 * np.pi * L),
            "exponential": c * np.log(9.0) / (2 * L) / (2 * np.pi),
            "hyperbolic": c * np.arccosh(3.0) / L / (2 * np.pi),
            "tractrix": c / (4 * L / np.arccos(np.power(1 / 9.0, 0.25)))}[profile]
    np.testing.assert_allclose(h.cutoff_frequency, want, rtol=1e-12)
    d = h.sdf(np.array([[0.0, 0.0, 0.05], [0.0, 0.0, 0.1], [0.0, 0.0, -0.02]]))
    np.testing.assert_allclose(d[0], -h.radius_at_position(0.5), rtol=1e-12)
    np.testing.assert_allclose(d[1], -0.01, rtol=1e-12)
    np.testing.assert_allclose(d[2], np.sqrt(0.03**2 + 0.02**2), rtol=1e-12)
    ctor = {"conical": sfa.conical_horn, "exponential": sfa.exponential_horn,
            "hyperbolic": sfa.hyperbolic_horn, "tractrix": sfa.tractrix_horn}[profile]
    assert ctor((0, 0, 0.1), (0, 0, 0), 0.01, 0.03).profile == profile
    with pytest.raises(ValueError):
        sfa.Horn((0, 0, 0), (0, 0, 0), 0.01, 0.03, profile)
    with pytest.raises(ValueError):
        sfa.Horn((0, 0, 0), (0, 0, 0.1), 0.0, 0.03, profile)
    with pytest.raises(ValueError):
        sfa.Horn((0, 0, 0), (0, 0, 0.1), 0.01, 0.03, "parabolic")


def test_paths_closed_forms():
    """geometry/paths.py restated: spiral and helix samples, their lengths,
    and polyline distances on cases with closed forms."""
    from strata_fdtd_amd.paths import HelixPath, SpiralPath, distance_to_helix, distance_to_polyline
    sp = SpiralPath(center=(0.1, 0.2), inner_radius=0.01, outer_radius=0.04, turns=2.0)
    pts = sp.sample_points(101)
    r = np.linalg.norm(pts - [0.1, 0.2], axis=1)
    np.testing.assert_allclose([r[0], r[-1]], [0.01, 0.04], rtol=1e-12)
    assert np.all(np.diff(r) > 0)
    np.testing.assert_allclose(r[50], np.sqrt(0.01 * 0.04), rtol=1e-12)       # logarithmic: geometric mean
    ar = SpiralPath((0.0, 0.0), 0.01, 0.04, 2.0, "archimedean").sample_points(101)
    np.testing.assert_allclose(np.linalg.norm(ar[50]), 0.025, rtol=1e-12)     # Archimedean: arithmetic mean
    hx = HelixPath(center=(0.0, 0.0, 0.0), radius=0.05, pitch=0.02, turns=3.0)
    np.testing.assert_allclose(hx.total_length, 3 * np.hypot(2 * np.pi * 0.05, 0.02), rtol=1e-15)
    assert hx.total_height == pytest.approx(0.06)
    dense = np.sum(np.linalg.norm(np.diff(hx.sample_points(20000), axis=0), axis=1))
    np.testing.assert_allclose(dense, hx.total_length, rtol=1e-6)
    seg = np.array([[0.0, 0.0], [1.0, 0.0], [1.0, 1.0]])
    q = np.array([[0.5, 0.5], [2.0, 0.5], [-1.0, 0.0], [1.0, 2.0]])
    np.testing.assert_allclose(distance_to_polyline(q, seg), [0.5, 1.0, 1.0, 1.0], rtol=1e-15)
    d = distance_to_helix(np.array([[0.0, 0.0, 0.03]]), (0.0, 0.0, 0.0), 0.05, 0.02, 3.0)
    # on the axis: between the helix radius and its chords' distance r cos(dtheta / 2)
    half = 0.5 * 2 * np.pi * 3.0 / (150 - 1)
    assert 0.05 * np.cos(half) - 1e-9 <= d[0] <= 0.05
    for bad in (dict(inner_radius=0.0), dict(outer_radius=0.005), dict(turns=0.0)):
        with pytest.raises(ValueError):
            SpiralPath(**{**dict(center=(0, 0), inner_radius=0.01, outer_radius=0.04, turns=1.0), **bad})
    with pytest.raises(ValueError):
        distance_to_polyline(q, np.zeros((1, 2)))
    with pytest.raises(ValueError):
        sfa.SpiralHorn((0, 0), 0.01, 0.02, 0.01, 0.03, 1.0, 0.05, profile="tractrix")
    with pytest.raises(ValueError):
        sfa.HelicalTube((0, 0, 0), 0.05, 0.0, 0.01, 0.02, 1.0)
