"""Geometry scenes as plain data, so one scene can be built with two SDF class
sets (the product's and, when writing fixtures, the reference's).

A tree is a JSON-able list:

  ["box", center, size]              Box(center=..., size=...)
  ["sphere", center, radius]
  ["cylinder", p1, p2, radius]
  ["cone", p1, p2, r1, r2]
  ["horn", throat, mouth, throat_radius, mouth_radius, profile]
  ["spiral_horn", center2, throat_r, mouth_r, inner_r, outer_r, turns, height, profile, spiral_type]
  ["helical_tube", center, helix_r, tube_r_start, tube_r_end, pitch, turns]
  ["union", [child, ...]]
  ["intersection", [child, ...]]
  ["difference", base, [child, ...]]
  ["smooth_union" | "smooth_intersection" | "smooth_difference", a, b, radius]
  ["translate", child, offset]
  ["scale", child, factors]          Scale(child, (sx, sy, sz))
  ["rotate", child, axis, angle]     Rotate(child, axis=..., angle=...)
  ["rotate_matrix", child, matrix]   Rotate(child, matrix=...) (row-major 9-list)
  ["helmholtz", position, cavity_shape, cavity_size, neck_length, neck_radius, neck_direction]
  ["materialized", geometry, [region, ...]]   MaterializedGeometry + set_material per region

Floats go through ``float`` and JSON's shortest repr, which round-trips every
float64 exactly, so both class sets see bit-identical parameters. Boxes are
recorded by the centre and size the product stored; a box built from corners
computes them as ``(lo + hi) / 2`` and ``hi - lo`` in either implementation.
"""

from __future__ import annotations

import json

import numpy as np


def _v(a) -> list:
    return [float(x) for x in np.asarray(a, dtype=np.float64).ravel()]


def to_tree(obj) -> list:
    """Describe a product geometry object (strata_fdtd_amd.geometry)."""
    from strata_fdtd_amd import geometry as G
    if isinstance(obj, G.MaterializedGeometry):
        return ["materialized", to_tree(obj.geometry), [to_tree(r) for r, _m, _id in obj._material_regions]]
    if isinstance(obj, G.HelmholtzResonator):
        return ["helmholtz", _v(obj.position), obj.cavity_shape, [float(s) for s in obj.cavity_size],
                float(obj.neck_length), float(obj.neck_radius), _v(obj.neck_direction)]
    if isinstance(obj, G.Box):
        return ["box", _v(obj.center), _v(obj.size)]
    if isinstance(obj, G.Sphere):
        return ["sphere", _v(obj.center), float(obj.radius)]
    if isinstance(obj, G.Cylinder):
        return ["cylinder", _v(obj.p1), _v(obj.p2), float(obj.radius)]
    if isinstance(obj, G.Cone):
        return ["cone", _v(obj.p1), _v(obj.p2), float(obj.r1), float(obj.r2)]
    if isinstance(obj, G.Horn):
        return ["horn", _v(obj.throat_position), _v(obj.mouth_position), float(obj.throat_radius),
                float(obj.mouth_radius), obj.profile]
    if isinstance(obj, G.SpiralHorn):
        return ["spiral_horn", _v(obj.center_2d), obj.throat_radius, obj.mouth_radius, obj.inner_spiral_radius,
                obj.outer_spiral_radius, obj.turns, obj.height, obj.profile, obj.spiral_type]
    if isinstance(obj, G.HelicalTube):
        return ["helical_tube", _v(obj.center), obj.helix_radius, obj.tube_radius_start, obj.tube_radius_end,
                obj.pitch, obj.turns]
    if isinstance(obj, G.Union):
        return ["union", [to_tree(c) for c in obj.children]]
    if isinstance(obj, G.Intersection):
        return ["intersection", [to_tree(c) for c in obj.children]]
    if isinstance(obj, G.Difference):
        return ["difference", to_tree(obj.base), [to_tree(c) for c in obj.subtracted]]
    if isinstance(obj, G.SmoothUnion):
        return ["smooth_union", to_tree(obj.a), to_tree(obj.b), float(obj.radius)]
    if isinstance(obj, G.SmoothIntersection):
        return ["smooth_intersection", to_tree(obj.a), to_tree(obj.b), float(obj.radius)]
    if isinstance(obj, G.SmoothDifference):
        return ["smooth_difference", to_tree(obj.base), to_tree(obj.subtracted), float(obj.radius)]
    if isinstance(obj, G.Translate):
        return ["translate", to_tree(obj.child), _v(obj.offset)]
    if isinstance(obj, G.Scale):
        return ["scale", to_tree(obj.child), _v(obj.scale)]
    if isinstance(obj, G.Rotate):
        if obj.axis is None:
            return ["rotate_matrix", to_tree(obj.child), _v(obj.matrix)]
        return ["rotate", to_tree(obj.child), _v(obj.axis), float(obj.angle)]
    raise TypeError(f"no tree form for {type(obj).__name__}")


def build(tree, ns: dict):
    """Instantiate a tree with the classes in ``ns`` (Box, Sphere, Cylinder, Cone, Horn,
    Union, Intersection, Difference, SmoothUnion, SmoothIntersection,
    SmoothDifference, Translate, Scale, Rotate, HelmholtzResonator, MaterializedGeometry;
    ``material`` is the object handed to set_material)."""
    kind = tree[0]
    if kind == "box":
        return ns["Box"](center=tuple(tree[1]), size=tuple(tree[2]))
    if kind == "sphere":
        return ns["Sphere"](center=tuple(tree[1]), radius=tree[2])
    if kind == "cylinder":
        return ns["Cylinder"](p1=tuple(tree[1]), p2=tuple(tree[2]), radius=tree[3])
    if kind == "cone":
        return ns["Cone"](p1=tuple(tree[1]), p2=tuple(tree[2]), r1=tree[3], r2=tree[4])
    if kind == "horn":
        return ns["Horn"](tuple(tree[1]), tuple(tree[2]), tree[3], tree[4], tree[5])
    if kind == "spiral_horn":
        return ns["SpiralHorn"](tuple(tree[1]), *tree[2:8], profile=tree[8], spiral_type=tree[9])
    if kind == "helical_tube":
        return ns["HelicalTube"](tuple(tree[1]), *tree[2:7])
    if kind == "union":
        return ns["Union"](*[build(c, ns) for c in tree[1]])
    if kind == "intersection":
        return ns["Intersection"](*[build(c, ns) for c in tree[1]])
    if kind == "difference":
        return ns["Difference"](build(tree[1], ns), *[build(c, ns) for c in tree[2]])
    if kind in ("smooth_union", "smooth_intersection", "smooth_difference"):
        cls = {"smooth_union": "SmoothUnion", "smooth_intersection": "SmoothIntersection",
               "smooth_difference": "SmoothDifference"}[kind]
        return ns[cls](build(tree[1], ns), build(tree[2], ns), tree[3])
    if kind == "translate":
        return ns["Translate"](build(tree[1], ns), tuple(tree[2]))
    if kind == "scale":
        return ns["Scale"](build(tree[1], ns), tuple(tree[2]))
    if kind == "rotate":
        return ns["Rotate"](build(tree[1], ns), axis=tuple(tree[2]), angle=tree[3])
    if kind == "rotate_matrix":
        return ns["Rotate"](build(tree[1], ns), matrix=np.array(tree[2], dtype=np.float64).reshape(3, 3))
    if kind == "helmholtz":
        return ns["HelmholtzResonator"](position=tuple(tree[1]), cavity_shape=tree[2],
                                        cavity_size=tuple(tree[3]), neck_length=tree[4],
                                        neck_radius=tree[5], neck_direction=tuple(tree[6]))
    if kind == "materialized":
        geo = ns["MaterializedGeometry"](build(tree[1], ns))
        for r in tree[2]:
            geo.set_material(build(r, ns), ns["material"])
        return geo
    raise ValueError(f"unknown node {kind!r}")


def dumps(tree) -> str:
    return json.dumps(tree)


def loads(text: str):
    return json.loads(text)
