"""Write voxel-mask fixtures from the reference's own SDF code.

    python tests/golden/make_voxel_golden.py

Run in a container that holds the reference checkout (at /root/reference, or
STRATA_REF_SRC). It imports the reference's ``strata_fdtd.geometry.sdf``
(SDFPrimitive.voxelize :99-127 and the Box/Sphere/Cylinder/CSG formulas),
``geometry/resonator.py`` (HelmholtzResonator) and
``geometry/material_assignment.py`` (MaterializedGeometry.voxelize_with_materials
:133-168), with ``h5py`` (absent here, needed only by the reference's HDF5
writer) replaced by an empty module. Each scene is built once with the
product's classes, written down as a data tree (tests/golden/scene_tree.py),
rebuilt from that tree with the reference's classes and voxelised by the
reference. ``voxel_masks.npz`` then holds, per case: the tree (JSON text), the
grid (shape and spacing, or the cell-centre coordinates of a nonuniform grid),
the packed air mask and, for material scenes, the material ids. Data only; the
reference is not needed to read it.

Cases:
  enclosure_64       the 512^3 enclosure workload scaled to 64^3 (shell, port,
                     ADE sphere region)
  plate_nu           the nonuniform resonator-plate workload scaled down (96x96x128,
                     3x3 resonators, half of them in the stretched cells)
  request_{0,1,2}    three serving requests with axes drawn from 64 and 80
  tree_u{NN}         random CSG trees with millimetre-rounded parameters on a
                     1 mm grid; the even ones centred on cell centres with
                     whole-millimetre sizes (many cell centres exactly on a surface)
  tree_nu{NN}        random CSG trees on a stretched grid
  materials_{N}      random geometry plus three random material regions
  helmholtz_{N}      Helmholtz resonators (box, sphere, cylinder cavities)
  cone_{NN}          CSG trees of cones (frustums, full cones) along grid axes,
                     with boxes and spheres; the even ones with apexes on cell
                     centres and whole-millimetre lengths and radii
  cone_nu{NN}        cone trees on the stretched grid
  cone_oblique{NN}   cone trees with tilted axes (host voxeliser only)
  xform_{NN}         translated, scaled (uniform and per axis) and rotated
                     trees, some nested; the even ones from trees with ties
  xform_nu{NN}       transformed trees on the stretched grid
  smooth_{NN}        smooth unions, intersections and differences
  horn_{NN}          horns of the four flare profiles along grid axes, alone,
                     cut out of a box or joined with a sphere; the even ones
                     with throats on cell centres
  horn_nu{NN}        horn trees on the stretched grid
  horn_oblique{NN}   horns with tilted axes
  swept_{NN}         spiral horns (logarithmic and Archimedean spirals, three
                     flare profiles) and helical tubes (tapered and not)
  swept_union        two of them cut out of a box
  xform_materials_{N} material regions that are transformed and smooth trees
                     (transforms and smooth CSG: host voxeliser only)
Cylinders and the cones other than the oblique ones are axis-aligned: their
np.dot projection sums exact zeros, so the reference's values do not depend on
the BLAS build. The oblique cones' and horns' masks and the rotated trees'
masks (a matrix product per point) do, and hold for the numpy/BLAS build of this image,
the one the tests run with.
"""

from __future__ import annotations

import os
import sys
import types

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

import scene_tree  # noqa: E402

REF_SRC = os.environ.get("STRATA_REF_SRC", "/root/reference/src")
OUT = os.path.join(HERE, "voxel_masks.npz")


def reference_namespace() -> dict:
    sys.modules.setdefault("h5py", types.ModuleType("h5py"))
    sys.path.insert(0, REF_SRC)
    from strata_fdtd.geometry import sdf
    from strata_fdtd.geometry.material_assignment import MaterializedGeometry
    from strata_fdtd.geometry.resonator import HelmholtzResonator
    from strata_fdtd.core.grid import NonuniformGrid, UniformGrid
    ns = dict(Box=sdf.Box, Sphere=sdf.Sphere, Cylinder=sdf.Cylinder, Cone=sdf.Cone, Union=sdf.Union,
              Intersection=sdf.Intersection, Difference=sdf.Difference, SmoothUnion=sdf.SmoothUnion,
              SmoothIntersection=sdf.SmoothIntersection, SmoothDifference=sdf.SmoothDifference,
              Translate=sdf.Translate, Scale=sdf.Scale, Rotate=sdf.Rotate, Horn=sdf.Horn,
              SpiralHorn=sdf.SpiralHorn, HelicalTube=sdf.HelicalTube,
              HelmholtzResonator=HelmholtzResonator, MaterializedGeometry=MaterializedGeometry,
              material=object(), UniformGrid=UniformGrid, NonuniformGrid=NonuniformGrid)
    assert sdf.__file__.startswith(REF_SRC), sdf.__file__
    return ns


def random_tree(rng, depth=0, ties=False):
    """Product-class CSG tree with axis-aligned, millimetre-rounded parameters.
    With ``ties`` the centres sit on cell centres of a 1 mm grid (computed as
    the grid computes them, i*1e-3 + 0.5e-3) and sizes and radii are whole
    millimetres, so faces and spheres pass exactly through cell centres."""
    from strata_fdtd_amd import Box, Cylinder, Difference, Intersection, Sphere, Union

    def r(a, b, n=None, d=3):
        return np.round(rng.uniform(a, b, n), d)

    def centre():
        if ties:
            return rng.integers(8, 44, 3) * 1e-3 + 1e-3 / 2
        return r(0.01, 0.08, 3, 4)
    t = rng.integers(0, 6 if depth < 2 else 3)
    c = centre()
    if t == 0:
        return Box(center=tuple(c), size=tuple(r(0.005, 0.05, 3)))
    if t == 1:
        return Sphere(center=tuple(c), radius=float(r(0.005, 0.03)))
    if t == 2:
        d = np.zeros(3)
        d[rng.integers(0, 3)] = r(0.01, 0.05) * (1 if rng.random() < 0.5 else -1)
        return Cylinder(p1=tuple(c), p2=tuple(c + d), radius=float(r(0.003, 0.02)))
    kids = [random_tree(rng, depth + 1, ties) for _ in range(rng.integers(1, 4))]
    return (Union, Intersection, lambda *k: Difference(k[0], *k[1:]))[t - 3](*kids)


def cone_tree(rng, ties=False, oblique=False, depth=0):
    """Product-class tree around cones (geometry/sdf.py Cone :436-543):
    frustums and full cones (one radius 0) along a grid axis, either sign,
    combined with boxes, spheres and other cones. With ``ties`` the apex
    ends sit on cell centres of a 1 mm grid and lengths and radii are whole
    millimetres, so cap planes and the radius interpolation land exactly on
    cell centres. ``oblique`` tilts the cones' axes."""
    from strata_fdtd_amd import Box, Cone, Difference, Intersection, Sphere, Union

    def cone():
        if ties:
            p1 = rng.integers(10, 40, 3) * 1e-3 + 1e-3 / 2
        else:
            p1 = np.round(rng.uniform(0.012, 0.04, 3), 4)
        if oblique:
            d = np.round(rng.uniform(-0.02, 0.02, 3), 3)
            d[rng.integers(0, 3)] = np.round(rng.uniform(0.012, 0.025), 3)
        else:
            d = np.zeros(3)
            ln = (rng.integers(8, 24) * 1e-3) if ties else np.round(rng.uniform(0.008, 0.024), 4)
            d[rng.integers(0, 3)] = ln * (1 if rng.random() < 0.5 else -1)
        if ties:
            r1, r2 = (float(v) * 1e-3 for v in rng.integers(0, 12, 2))
        else:
            r1, r2 = (float(v) for v in np.round(rng.uniform(0.0, 0.012, 2), 4))
        if rng.random() < 0.3:
            (r1, r2) = (r1, 0.0) if rng.random() < 0.5 else (0.0, r2)
        if r1 == 0 and r2 == 0:
            r1 = 0.006
        return Cone(p1=tuple(p1), p2=tuple(p1 + d), r1=r1, r2=r2)

    t = rng.integers(0, 6 if depth < 1 else 3)
    if t in (0, 1, 2):
        if t == 2 and depth > 0:
            return Sphere(center=tuple(np.round(rng.uniform(0.015, 0.035, 3), 4)),
                          radius=float(np.round(rng.uniform(0.004, 0.01), 4)))
        return cone()
    kids = [cone_tree(rng, ties, oblique, depth + 1) for _ in range(rng.integers(2, 4))]
    if t == 5:
        outer = Box(center=(0.025, 0.025, 0.025), size=(0.03, 0.028, 0.032))
        return Difference(outer, *kids)
    return (Union, Intersection)[t - 3](*kids)


def xform_tree(rng, ties=False, depth=0):
    """Product-class transforms (geometry/sdf.py Translate / Scale / Rotate
    :1188-1391) of random trees: whole-millimetre translations (with ``ties``
    the child's faces on cell centres stay there or move to cell faces),
    uniform and per-axis scales, rotations by quarter turns about grid axes
    and by random angles about random axes, nested up to two deep."""
    from strata_fdtd_amd import Rotate, Scale, Translate, Union
    t = rng.integers(0, 5)
    child = xform_tree(rng, ties, depth + 1) if depth < 1 and rng.random() < 0.4 else random_tree(rng, ties=ties)
    if t == 0:
        return Translate(child, tuple(float(v) for v in rng.integers(-6, 7, 3) * 1e-3))
    if t == 1:
        return Scale(child, float(np.round(rng.uniform(0.6, 1.4), 2)))
    if t == 2:
        return Scale(child, tuple(float(v) for v in np.round(rng.uniform(0.7, 1.3, 3), 2)))
    if t == 3:
        axis = np.zeros(3)
        axis[rng.integers(0, 3)] = 1.0
        # about the grid's centre: translate there, rotate, translate back
        c = np.array([0.025, 0.026, 0.024])
        inner = Translate(child, tuple(-c))
        return Translate(Rotate(inner, axis=tuple(axis), angle=float(rng.integers(1, 4)) * np.pi / 2), tuple(c))
    axis = np.round(rng.normal(size=3), 3)
    return Union(Rotate(child, axis=tuple(float(a) for a in axis), angle=float(np.round(rng.uniform(-0.6, 0.6), 3))))


def horn_tree(rng, ties=False, oblique=False):
    """Product-class horns (geometry/sdf.py Horn :937-1150) of the four flare
    profiles along a grid axis (either direction), alone, as a cavity cut out
    of a box, or joined with a sphere. With ``ties`` the throat sits on a cell
    centre and the length is whole millimetres; ``oblique`` tilts the axis."""
    from strata_fdtd_amd import Box, Difference, Sphere, Union
    from strata_fdtd_amd.geometry import Horn
    if ties:
        throat = rng.integers(14, 36, 3) * 1e-3 + 1e-3 / 2
    else:
        throat = np.round(rng.uniform(0.015, 0.035, 3), 4)
    if oblique:
        d = np.round(rng.uniform(-0.012, 0.012, 3), 3)
        d[rng.integers(0, 3)] = np.round(rng.uniform(0.01, 0.016), 3)
    else:
        d = np.zeros(3)
        ln = (rng.integers(8, 16) * 1e-3) if ties else np.round(rng.uniform(0.008, 0.016), 4)
        d[rng.integers(0, 3)] = ln * (1 if rng.random() < 0.5 else -1)
    rt = float(np.round(rng.uniform(0.002, 0.005), 4))
    rm = float(np.round(rt * rng.uniform(1.3, 2.5), 4))
    horn = Horn(tuple(throat), tuple(throat + d), rt, rm, Horn.PROFILES[rng.integers(0, 4)])
    t = rng.integers(0, 3)
    if t == 0:
        return horn
    if t == 1:
        return Difference(Box(center=(0.025, 0.026, 0.024), size=(0.04, 0.04, 0.04)), horn)
    return Union(horn, Sphere(center=tuple(throat), radius=float(np.round(rt * 1.5, 4))))


def smooth_tree(rng, ties=False):
    """Smooth union / intersection / difference (geometry/sdf.py :736-936) of
    two random trees, blend radius 1-6 mm."""
    from strata_fdtd_amd import SmoothDifference, SmoothIntersection, SmoothUnion
    cls = (SmoothUnion, SmoothIntersection, SmoothDifference)[rng.integers(0, 3)]
    return cls(random_tree(rng, ties=ties), random_tree(rng, ties=ties), float(np.round(rng.uniform(0.001, 0.006), 4)))


def nontrivial_tree(rng, grid, ties=False, tries=200, make=None):
    """A random tree whose mask is neither empty nor full on ``grid`` (the
    product's voxeliser only picks the cases; the fixture holds the
    reference's masks)."""
    for _ in range(tries):
        node = make(rng) if make else random_tree(rng, ties=ties)
        frac = node.voxelize(grid).mean()
        if 0.002 < frac < 0.998:
            return node
    raise RuntimeError("no non-trivial tree drawn")


def cases():
    """(name, product geometry object, grid spec) triples."""
    from strata_fdtd_amd import HelmholtzResonator, MaterializedGeometry, SimpleMaterial
    from strata_fdtd_amd.serving import sweep_requests
    from strata_fdtd_amd.workloads import enclosure_512_scene, resonators_nu_scene
    out = []
    geo, *_ = enclosure_512_scene(64)
    out.append(("enclosure_64", geo, dict(shape=(64, 64, 64), dx=1e-3)))
    grid, air, *_ = resonators_nu_scene(shape=(96, 96, 128), core=60, n_side=3, pitch=24, plate=24,
                                        cavity_radius=8, cavity_depth=14, neck_radius=3, neck_length=4)
    out.append(("plate_nu", air, dict(coords=(grid.x_coords, grid.y_coords, grid.z_coords))))
    for q, req in enumerate(sweep_requests(n=3, seed=77, dims=(64, 80))):
        out.append((f"request_{q}", req.geometry, dict(shape=req.shape, dx=req.resolution)))
    from strata_fdtd_amd import NonuniformGrid, UniformGrid
    rng = np.random.default_rng(2025)
    ug = UniformGrid((64, 56, 48), 1e-3)
    for q in range(24):
        out.append((f"tree_u{q:02d}", nontrivial_tree(rng, ug, ties=q % 2 == 0), dict(shape=(64, 56, 48), dx=1e-3)))
    nug = NonuniformGrid.from_stretch((64, 56, 48), 1e-3, stretch_z=1.03, stretch_x=1.02)
    for q in range(8):
        out.append((f"tree_nu{q:02d}", nontrivial_tree(rng, nug),
                    dict(coords=(nug.x_coords, nug.y_coords, nug.z_coords))))
    mg = UniformGrid((48, 40, 52), 1e-3)
    for q in range(4):
        geo = MaterializedGeometry(nontrivial_tree(rng, mg))
        for m in range(3):
            geo.set_material(nontrivial_tree(rng, mg, ties=True), SimpleMaterial(name=f"m{m}", _rho=1.2 + m, _c=340.0))
        out.append((f"materials_{q}", geo, dict(shape=(48, 40, 52), dx=1e-3)))
    for q, (shape, size) in enumerate((("box", (0.012, 0.01, 0.014)), ("sphere", (0.008,)),
                                       ("cylinder", (0.006, 0.016)))):
        res = HelmholtzResonator(position=(0.02, 0.021, 0.018), cavity_shape=shape, cavity_size=size,
                                 neck_length=0.006, neck_radius=0.003,
                                 neck_direction=((0.0, 0.0, 1.0), (1.0, 0.0, 0.0), (0.0, -1.0, 0.0))[q])
        out.append((f"helmholtz_{q}", res, dict(shape=(40, 44, 48), dx=1e-3)))
    # cones (own random stream, so the cases above stay as they were)
    crng = np.random.default_rng(4242)
    cg = UniformGrid((50, 52, 48), 1e-3)
    for q in range(10):
        out.append((f"cone_{q:02d}", nontrivial_tree(crng, cg, make=lambda g, t=(q % 2 == 0): cone_tree(g, ties=t)),
                    dict(shape=(50, 52, 48), dx=1e-3)))
    for q in range(3):
        out.append((f"cone_nu{q:02d}", nontrivial_tree(crng, nug, make=lambda g: cone_tree(g)),
                    dict(coords=(nug.x_coords, nug.y_coords, nug.z_coords))))
    for q in range(3):
        out.append((f"cone_oblique{q:02d}", nontrivial_tree(crng, cg, make=lambda g: cone_tree(g, oblique=True)),
                    dict(shape=(50, 52, 48), dx=1e-3)))
    # transforms and smooth CSG (own random stream)
    xrng = np.random.default_rng(5151)
    for q in range(12):
        out.append((f"xform_{q:02d}", nontrivial_tree(xrng, cg, make=lambda g, t=(q % 2 == 0): xform_tree(g, ties=t)),
                    dict(shape=(50, 52, 48), dx=1e-3)))
    for q in range(3):
        out.append((f"xform_nu{q:02d}", nontrivial_tree(xrng, nug, make=lambda g: xform_tree(g)),
                    dict(coords=(nug.x_coords, nug.y_coords, nug.z_coords))))
    for q in range(9):
        out.append((f"smooth_{q:02d}", nontrivial_tree(xrng, cg, make=lambda g, t=(q % 2 == 0): smooth_tree(g, ties=t)),
                    dict(shape=(50, 52, 48), dx=1e-3)))
    hrng = np.random.default_rng(6161)
    for q in range(12):
        out.append((f"horn_{q:02d}", nontrivial_tree(hrng, cg, make=lambda g, t=(q % 2 == 0): horn_tree(g, ties=t)),
                    dict(shape=(50, 52, 48), dx=1e-3)))
    for q in range(2):
        out.append((f"horn_nu{q:02d}", nontrivial_tree(hrng, nug, make=lambda g: horn_tree(g)),
                    dict(coords=(nug.x_coords, nug.y_coords, nug.z_coords))))
    for q in range(2):
        out.append((f"horn_oblique{q:02d}", nontrivial_tree(hrng, cg, make=lambda g: horn_tree(g, oblique=True)),
                    dict(shape=(50, 52, 48), dx=1e-3)))
    from strata_fdtd_amd.geometry import Box, Difference, HelicalTube, SpiralHorn
    sg = UniformGrid((40, 42, 24), 1e-3)
    swept = [SpiralHorn((0.02, 0.021), 0.002, 0.005, 0.006, 0.016, 1.5, 0.012, "exponential", "logarithmic"),
             SpiralHorn((0.0205, 0.0215), 0.0015, 0.004, 0.005, 0.015, 2.0, 0.02, "linear", "archimedean"),
             SpiralHorn((0.019, 0.02), 0.002, 0.003, 0.007, 0.016, 1.25, 0.006, "hyperbolic", "logarithmic"),
             HelicalTube((0.02, 0.021, 0.002), 0.012, 0.002, 0.004, 0.008, 2.5),
             HelicalTube((0.0205, 0.0205, 0.0005), 0.01, 0.003, 0.003, 0.01, 1.5)]
    for q, node in enumerate(swept):
        out.append((f"swept_{q:02d}", node, dict(shape=(40, 42, 24), dx=1e-3)))
    out.append(("swept_union", Difference(Box(center=(0.02, 0.021, 0.012), size=(0.036, 0.038, 0.02)), swept[0],
                                          swept[3]), dict(shape=(40, 42, 24), dx=1e-3)))
    for q in range(2):
        geo = MaterializedGeometry(nontrivial_tree(xrng, cg))
        geo.set_material(nontrivial_tree(xrng, cg, make=lambda g: xform_tree(g)), SimpleMaterial(name="x", _rho=1.5, _c=300.0))
        geo.set_material(nontrivial_tree(xrng, cg, make=lambda g: smooth_tree(g)), SimpleMaterial(name="s", _rho=2.0, _c=320.0))
        out.append((f"xform_materials_{q}", geo, dict(shape=(50, 52, 48), dx=1e-3)))
    return out


def ref_grid(ns, spec):
    if "coords" in spec:
        return ns["NonuniformGrid"](*[np.asarray(c, np.float64) for c in spec["coords"]])
    return ns["UniformGrid"](tuple(spec["shape"]), spec["dx"])


def main() -> None:
    ns = reference_namespace()
    arrays = {}
    names = []
    for name, obj, spec in cases():
        tree = scene_tree.to_tree(obj)
        ref_obj = scene_tree.build(tree, ns)
        grid = ref_grid(ns, spec)
        if tree[0] == "materialized":
            mask, ids = ref_obj.voxelize_with_materials(grid)
            arrays[f"{name}__ids"] = np.asarray(ids, np.uint8)
        else:
            mask = ref_obj.voxelize(grid)
        mask = np.asarray(mask, bool)
        arrays[f"{name}__tree"] = np.array(scene_tree.dumps(tree))
        arrays[f"{name}__shape"] = np.array(mask.shape, np.int64)
        arrays[f"{name}__mask"] = np.packbits(mask.ravel())
        if "coords" in spec:
            for a, c in zip("xyz", spec["coords"]):
                arrays[f"{name}__{a}"] = np.asarray(c, np.float64)
        else:
            arrays[f"{name}__dx"] = np.array(spec["dx"], np.float64)
        names.append(name)
        extra = f", material cells {int((arrays[name + '__ids'] > 0).sum())}" if tree[0] == "materialized" else ""
        print(f"{name}: {mask.shape}, air {int(mask.sum())}/{mask.size}{extra}", flush=True)
    arrays["names"] = np.array(names)
    np.savez_compressed(OUT, **arrays)
    print(f"wrote {OUT} ({os.path.getsize(OUT)} bytes)")


if __name__ == "__main__":
    main()
