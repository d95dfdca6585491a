"""SDF geometry and voxelisation into the solver's air mask / material ids.

The signed-distance formulas, CSG rules and voxelisation rule follow the
reference's ``geometry/sdf.py`` (SDFPrimitive.voxelize :99-127, Box :274-294,
Sphere :320-333, Cylinder :382-425, Cone :436-543,
Union/Intersection/Difference :566-722, the smooth forms :736-936, Horn
:937-1150, rotation_matrix and Translate/Scale/Rotate :1151-1391, SpiralHorn
and HelicalTube :1392-1723 (with geometry/paths.py), the horn constructors
:1724-1820),
``geometry/resonator.py`` (HelmholtzResonator :137-207) and
``geometry/material_assignment.py`` (MaterializedGeometry :48-205). The masks
must match the reference bit for bit, so every formula uses the same float64
numpy operations in the same order (norms via ``np.linalg.norm``, the cylinder
projection via ``np.dot``): a point is air when ``sdf <= 0``.

Fast exact voxelisation. For the primitives and CSG nodes defined here the
mask is computed as a boolean combination instead of a float min/max tree:
``min(d1, d2) <= 0`` iff ``d1 <= 0 or d2 <= 0``; ``max(d1, d2) <= 0`` iff both;
``max(b, -s) <= 0`` iff ``b <= 0 and not (s < 0)`` (and the strict forms dually).
Box masks are separable per axis (all three |x - c| - s/2 <= 0, the same float64
values the reference computes; differences of millimetre coordinates are never
small enough to underflow the norm). Sphere and cylinder formulas are evaluated
with the reference's expressions only on the cells inside their bounding box
plus a two-cell margin; outside it the distance is positive by construction.
Every value is therefore the reference's own float64 value, so the mask is
identical; the tests compare it with the direct single-pass evaluation. A
translation is the child's mask on shifted coordinates (p - offset per axis,
the same values). Scales, rotations and smooth CSG are evaluated directly (in
slabs of whole i-planes), and so is any other SDF type; trees whose distances
pass through a BLAS product (rotations, oblique cylinders and cones) over
every cell at once, as the reference evaluates them.

The reference's smooth intersection and difference blend with its smooth
minimum's weight h = clip(0.5 + 0.5 (d2 - d1) / k, 0, 1) and add the blend
term, so away from the blend they return the smaller distance, as a union
does; they are restated as written, since the masks must be the reference's.
"""

from __future__ import annotations

from abc import ABC, abstractmethod
from dataclasses import dataclass, field

import numpy as np
from numpy.typing import NDArray

# Cells evaluated per slab (rounded to whole planes). 2**22 points keep the
# float64 temporaries of the most complex CSG tree here under ~1 GB.
VOXEL_SLAB_POINTS = 1 << 22


class SDFPrimitive(ABC):
    """Signed distance primitive: negative inside. Voxelised inside = air."""

    @abstractmethod
    def sdf(self, points: NDArray[np.floating]) -> NDArray[np.floating]:
        ...

    @property
    @abstractmethod
    def bounding_box(self) -> tuple[NDArray[np.floating], NDArray[np.floating]]:
        ...

    def contains(self, points):
        return self.sdf(points) <= 0

    def voxelize(self, grid, slab_points: int | None = VOXEL_SLAB_POINTS) -> NDArray[np.bool_]:
        """Boolean mask, True where sdf(cell centre) <= 0 (geometry/sdf.py:99-127)."""
        return _mask(self, grid, strict=False)

    def voxelize_direct(self, grid, slab_points: int | None = VOXEL_SLAB_POINTS,
                        strict: bool = False) -> NDArray[np.bool_]:
        """The reference's evaluation: sdf at every cell centre, then <= 0
        (< 0 when strict). ``slab_points=None`` evaluates all cells at once."""
        nx, ny, nz = grid.shape
        return _direct(self, grid.x_coords, grid.y_coords, grid.z_coords, slab_points, strict)

    # transforms (geometry/sdf.py:128-231)
    def translate(self, offset) -> "SDFPrimitive":
        return Translate(self, offset)

    def scale(self, factor) -> "SDFPrimitive":
        return Scale(self, factor)

    def rotate(self, axis, angle: float) -> "SDFPrimitive":
        return Rotate(self, axis=axis, angle=angle)

    def rotate_x(self, angle: float) -> "SDFPrimitive":
        return self.rotate((1, 0, 0), angle)

    def rotate_y(self, angle: float) -> "SDFPrimitive":
        return self.rotate((0, 1, 0), angle)

    def rotate_z(self, angle: float) -> "SDFPrimitive":
        return self.rotate((0, 0, 1), angle)


def _direct(node, xc, yc, zc, slab_points, strict) -> NDArray[np.bool_]:
    nx, ny, nz = len(xc), len(yc), len(zc)
    plane = ny * nz
    step = nx if slab_points is None else max(1, slab_points // max(plane, 1))
    mask = np.empty((nx, ny, nz), dtype=bool)
    for i0 in range(0, nx, step):
        i1 = min(nx, i0 + step)
        X, Y, Z = np.meshgrid(xc[i0:i1], yc, zc, indexing="ij")
        points = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
        d = node.sdf(points).reshape((i1 - i0, ny, nz))
        mask[i0:i1] = (d < 0) if strict else (d <= 0)
    return mask


def _index_range(coords, lo, hi, margin):
    a = int(np.searchsorted(coords, lo - margin, side="left"))
    b = int(np.searchsorted(coords, hi + margin, side="right"))
    return a, b


def _mask(node, grid, strict: bool) -> NDArray[np.bool_]:
    """Exact mask of ``sdf <= 0`` (``< 0`` when strict) for a CSG tree."""
    xc, yc, zc = grid.x_coords, grid.y_coords, grid.z_coords
    shape = (len(xc), len(yc), len(zc))
    if isinstance(node, Box):
        half = node.size / 2
        ms = []
        for ax, c in enumerate((xc, yc, zc)):
            q = np.abs(c - node.center[ax]) - half[ax]
            ms.append((q < 0) if strict else (q <= 0))
        return ms[0][:, None, None] & ms[1][None, :, None] & ms[2][None, None, :]
    if isinstance(node, (Sphere, Cylinder, Cone)):
        # an oblique axis goes through np.dot's BLAS kernel, whose arithmetic
        # may depend on a point's place in the array: evaluate every cell at
        # once, as the reference's voxelize does
        if isinstance(node, (Cylinder, Cone)) and np.count_nonzero(node.axis_normalized) != 1:
            return _direct(node, xc, yc, zc, None if np.prod(shape) <= (1 << 25) else VOXEL_SLAB_POINTS,
                           strict)
        lo, hi = node.bounding_box
        margin = 2.0 * max(float(np.max(np.diff(c))) if len(c) > 1 else 0.0 for c in (xc, yc, zc))
        out = np.zeros(shape, dtype=bool)
        rng = [_index_range(c, lo[ax], hi[ax], margin) for ax, c in enumerate((xc, yc, zc))]
        if any(b <= a for a, b in rng):
            return out
        (a0, b0), (a1, b1), (a2, b2) = rng
        out[a0:b0, a1:b1, a2:b2] = _direct(node, xc[a0:b0], yc[a1:b1], zc[a2:b2], VOXEL_SLAB_POINTS, strict)
        return out
    if isinstance(node, Union):
        out = np.zeros(shape, dtype=bool)
        for ch in node.children:
            out |= _mask(ch, grid, strict)
        return out
    if isinstance(node, Intersection):
        out = np.ones(shape, dtype=bool)
        for ch in node.children:
            out &= _mask(ch, grid, strict)
        return out
    if isinstance(node, Difference):
        out = _mask(node.base, grid, strict)
        for sub in node.subtracted:
            # max(b, -s) <= 0  <=>  b <= 0 and s >= 0  <=>  b <= 0 and not (s < 0)
            # max(b, -s) <  0  <=>  b <  0 and s >  0  <=>  b <  0 and not (s <= 0)
            out &= ~_mask(sub, grid, not strict)
        return out
    if isinstance(node, HelmholtzResonator):
        return _mask(node._combined, grid, strict)
    if isinstance(node, Translate):
        # p - offset per axis: the child sees the shifted cell-centre
        # coordinates, the same float64 values
        return _mask(node.child, _Coords(xc - node.offset[0], yc - node.offset[1], zc - node.offset[2]), strict)
    # scales, rotations and smooth CSG: the reference's evaluation of the
    # distances; trees whose values go through BLAS (rotations, oblique axes)
    # over every cell at once, as the reference's voxelize does
    whole = _blas_dependent(node) and np.prod(shape) <= (1 << 25)
    return _direct(node, xc, yc, zc, None if whole else VOXEL_SLAB_POINTS, strict)


class _Coords:
    """Cell-centre coordinates per axis (the part of a grid _mask reads)."""

    def __init__(self, xc, yc, zc):
        self.x_coords, self.y_coords, self.z_coords = xc, yc, zc


def _blas_dependent(node) -> bool:
    """Whether a tree's distances pass through a BLAS product whose rounding
    may depend on a point's place in the array."""
    if isinstance(node, Rotate):
        return True
    if isinstance(node, (Cylinder, Cone)):
        return np.count_nonzero(node.axis_normalized) != 1
    if isinstance(node, Horn):
        return np.count_nonzero(node._axis) != 1
    if isinstance(node, (SpiralHorn, HelicalTube)):
        return True         # np.dot against the path's segments
    kids = []
    if isinstance(node, (Union, Intersection)):
        kids = list(node.children)
    elif isinstance(node, Difference):
        kids = [node.base, *node.subtracted]
    elif isinstance(node, (SmoothUnion, SmoothIntersection)):
        kids = [node.a, node.b]
    elif isinstance(node, SmoothDifference):
        kids = [node.base, node.subtracted]
    elif isinstance(node, Transform):
        kids = [node.child]
    elif isinstance(node, HelmholtzResonator):
        kids = [node._combined]
    return any(_blas_dependent(k) for k in kids)


def _points(points) -> NDArray[np.float64]:
    points = np.asarray(points, dtype=np.float64)
    if points.ndim != 2 or points.shape[1] != 3:
        raise ValueError(f"points must be Nx3 array, got shape {points.shape}")
    return points


class Box(SDFPrimitive):
    """Axis-aligned box from (center, size) or (min_corner, max_corner)."""

    def __init__(self, center=None, size=None, min_corner=None, max_corner=None):
        if center is not None and size is not None:
            self.center = np.array(center, dtype=np.float64)
            self.size = np.array(size, dtype=np.float64)
        elif min_corner is not None and max_corner is not None:
            lo = np.array(min_corner, dtype=np.float64)
            hi = np.array(max_corner, dtype=np.float64)
            self.center = (lo + hi) / 2
            self.size = hi - lo
        else:
            raise ValueError("Must provide either (center, size) or (min_corner, max_corner)")
        if np.any(self.size <= 0):
            raise ValueError(f"Box size must be positive, got {self.size}")

    def sdf(self, points):
        points = _points(points)
        q = np.abs(points - self.center) - self.size / 2
        outside = np.linalg.norm(np.maximum(q, 0), axis=1)
        inside = np.minimum(np.max(q, axis=1), 0)
        return outside + inside

    @property
    def bounding_box(self):
        h = self.size / 2
        return self.center - h, self.center + h


class Sphere(SDFPrimitive):
    def __init__(self, center, radius: float):
        self.center = np.array(center, dtype=np.float64)
        self.radius = float(radius)
        if self.radius <= 0:
            raise ValueError(f"Sphere radius must be positive, got {self.radius}")

    def sdf(self, points):
        points = _points(points)
        return np.linalg.norm(points - self.center, axis=1) - self.radius

    @property
    def bounding_box(self):
        r = np.array([self.radius] * 3)
        return self.center - r, self.center + r


class Cylinder(SDFPrimitive):
    """Capped cylinder between p1 and p2 (geometry/sdf.py:345-435)."""

    def __init__(self, p1, p2, radius: float):
        self.p1 = np.array(p1, dtype=np.float64)
        self.p2 = np.array(p2, dtype=np.float64)
        self.radius = float(radius)
        if self.radius <= 0:
            raise ValueError(f"Cylinder radius must be positive, got {self.radius}")
        self.axis = self.p2 - self.p1
        self.length = np.linalg.norm(self.axis)
        if self.length == 0:
            raise ValueError("Cylinder endpoints must be distinct")
        self.axis_normalized = self.axis / self.length

    def sdf(self, points):
        points = _points(points)
        v = points - self.p1
        t = np.dot(v, self.axis_normalized)
        t_clamped = np.clip(t, 0, self.length)
        closest = self.p1 + t_clamped[:, np.newaxis] * self.axis_normalized
        radial = np.linalg.norm(points - closest, axis=1)
        axial = np.maximum(0, np.abs(t - self.length / 2) - self.length / 2)
        inside_body = (t >= 0) & (t <= self.length)
        return np.where(inside_body, radial - self.radius,
                        np.sqrt((radial - self.radius) ** 2 + axial**2))

    @property
    def bounding_box(self):
        r = np.array([self.radius] * 3)
        return np.minimum(self.p1, self.p2) - r, np.maximum(self.p1, self.p2) + r


class Cone(SDFPrimitive):
    """Truncated cone (frustum) between p1 (radius r1) and p2 (radius r2),
    geometry/sdf.py:436-543: the radius interpolated linearly along the axis,
    the radial distance inside the body and, beyond the caps, the radial
    and axial excesses combined as sqrt(a^2 + b^2)."""

    def __init__(self, p1, p2, r1: float, r2: float):
        self.p1 = np.array(p1, dtype=np.float64)
        self.p2 = np.array(p2, dtype=np.float64)
        self.r1 = float(r1)
        self.r2 = float(r2)
        if self.r1 < 0 or self.r2 < 0:
            raise ValueError(f"Cone radii must be non-negative, got r1={r1}, r2={r2}")
        if self.r1 == 0 and self.r2 == 0:
            raise ValueError("Cone cannot have both radii zero")
        self.axis = self.p2 - self.p1
        self.length = np.linalg.norm(self.axis)
        if self.length == 0:
            raise ValueError("Cone endpoints must be distinct")
        self.axis_normalized = self.axis / self.length

    def sdf(self, points):
        points = _points(points)
        v = points - self.p1
        t = np.dot(v, self.axis_normalized) / self.length
        t_clamped = np.clip(t, 0, 1)
        radius_at_t = self.r1 + t_clamped * (self.r2 - self.r1)
        closest = self.p1 + (t_clamped * self.length)[:, np.newaxis] * self.axis_normalized
        radial = np.linalg.norm(points - closest, axis=1)
        axial = np.maximum(0, np.abs(t - 0.5) - 0.5) * self.length
        inside_body = (t >= 0) & (t <= 1)
        return np.where(inside_body, radial - radius_at_t,
                        np.sqrt((radial - radius_at_t) ** 2 + axial**2))

    @property
    def bounding_box(self):
        r = np.array([max(self.r1, self.r2)] * 3)
        return np.minimum(self.p1, self.p2) - r, np.maximum(self.p1, self.p2) + r


class CSGNode(SDFPrimitive):
    """Base class of the CSG operations (geometry/sdf.py:544-547)."""


class Union(CSGNode):
    def __init__(self, *children: SDFPrimitive):
        self.children = children

    def sdf(self, points):
        if not self.children:
            return np.full(len(points), np.inf, dtype=np.float64)
        return np.minimum.reduce([c.sdf(points) for c in self.children])

    @property
    def bounding_box(self):
        if not self.children:
            return np.array([np.inf] * 3), np.array([-np.inf] * 3)
        return (np.minimum.reduce([c.bounding_box[0] for c in self.children]),
                np.maximum.reduce([c.bounding_box[1] for c in self.children]))


class Intersection(CSGNode):
    def __init__(self, *children: SDFPrimitive):
        self.children = children

    def sdf(self, points):
        if not self.children:
            return np.full(len(points), -np.inf, dtype=np.float64)
        return np.maximum.reduce([c.sdf(points) for c in self.children])

    @property
    def bounding_box(self):
        if not self.children:
            return np.array([-np.inf] * 3), np.array([np.inf] * 3)
        lo = np.maximum.reduce([c.bounding_box[0] for c in self.children])
        hi = np.minimum.reduce([c.bounding_box[1] for c in self.children])
        if np.any(lo > hi):
            return np.array([np.inf] * 3), np.array([-np.inf] * 3)
        return lo, hi


class Difference(CSGNode):
    """base minus subtracted: max(base, -s1, -s2, ...) applied one at a time."""

    def __init__(self, base: SDFPrimitive, *subtracted: SDFPrimitive):
        self.base = base
        self.subtracted = subtracted

    def sdf(self, points):
        result = self.base.sdf(points)
        for shape in self.subtracted:
            result = np.maximum(result, -shape.sdf(points))
        return result

    @property
    def bounding_box(self):
        return self.base.bounding_box


# ---- smooth CSG (geometry/sdf.py:736-936) ----------------------------------
# Polynomial smooth minimum / maximum of two distances with blend radius k:
# h = clip(0.5 + 0.5 (d2 - d1) / k, 0, 1), then d2 (1 - h) + d1 h -+ k h (1 - h)
# (minus for the union, plus for the intersection and for the difference,
# whose second distance is the negated subtracted shape). Evaluated with the
# reference's expressions, term for term.
class _SmoothPair(CSGNode):
    _sign = 1.0

    def _check(self, radius):
        if radius <= 0:
            raise ValueError(f"radius must be positive, got {radius}")
        self.radius = radius

    def _blend(self, d1, d2):
        h = np.clip(0.5 + 0.5 * (d2 - d1) / self.radius, 0, 1)
        if self._sign < 0:
            return d2 * (1 - h) + d1 * h - self.radius * h * (1 - h)
        return d2 * (1 - h) + d1 * h + self.radius * h * (1 - h)


class SmoothUnion(_SmoothPair):
    _sign = -1.0

    def __init__(self, a: SDFPrimitive, b: SDFPrimitive, radius: float):
        self._check(radius)
        self.a, self.b = a, b

    def sdf(self, points):
        return self._blend(self.a.sdf(points), self.b.sdf(points))

    @property
    def bounding_box(self):
        (lo_a, hi_a), (lo_b, hi_b) = self.a.bounding_box, self.b.bounding_box
        e = np.array([self.radius] * 3)
        return np.minimum(lo_a, lo_b) - e, np.maximum(hi_a, hi_b) + e


class SmoothIntersection(_SmoothPair):
    def __init__(self, a: SDFPrimitive, b: SDFPrimitive, radius: float):
        self._check(radius)
        self.a, self.b = a, b

    def sdf(self, points):
        return self._blend(self.a.sdf(points), self.b.sdf(points))

    @property
    def bounding_box(self):
        (lo_a, hi_a), (lo_b, hi_b) = self.a.bounding_box, self.b.bounding_box
        e = np.array([self.radius] * 3)
        lo = np.maximum(lo_a, lo_b) + e
        hi = np.minimum(hi_a, hi_b) - e
        if np.any(lo > hi):
            return np.array([np.inf] * 3), np.array([-np.inf] * 3)
        return lo, hi


class SmoothDifference(_SmoothPair):
    def __init__(self, base: SDFPrimitive, subtracted: SDFPrimitive, radius: float):
        self._check(radius)
        self.base, self.subtracted = base, subtracted

    def sdf(self, points):
        return self._blend(self.base.sdf(points), -self.subtracted.sdf(points))

    @property
    def bounding_box(self):
        lo, hi = self.base.bounding_box
        e = np.array([self.radius] * 3)
        return lo - e, hi + e


# ---- horns (geometry/sdf.py:937-1150, 1724-1820) -----------------------------
class Horn(SDFPrimitive):
    """Surface of revolution from a throat (radius rt) to a mouth (radius rm)
    whose radius follows a flare profile along the axis, x = t * L:
    conical rt (1 + m t) with m = sqrt(Am/A0) - 1; exponential rt exp(a x)
    with a = ln(Am/A0) / 2L; hyperbolic rt cosh(a x) with a =
    arccosh(sqrt(Am/A0)) / L; tractrix rt / cos(x/R)^2 with R = L /
    arccos((A0/Am)^(1/4)); A = pi r^2. The distance is the radial excess over
    the profile radius (at the axial position clamped to the horn) between
    the end planes and the hypotenuse with the distance past the nearer end
    plane beyond them."""

    PROFILES = ("conical", "exponential", "hyperbolic", "tractrix")

    def __init__(self, throat_position, mouth_position, throat_radius: float, mouth_radius: float,
                 profile: str = "exponential"):
        if profile not in self.PROFILES:
            raise ValueError(f"profile must be one of {set(self.PROFILES)}, got '{profile}'")
        self.throat_position = np.array(throat_position, dtype=np.float64)
        self.mouth_position = np.array(mouth_position, dtype=np.float64)
        self.throat_radius = float(throat_radius)
        self.mouth_radius = float(mouth_radius)
        self.profile = profile
        if self.throat_radius <= 0 or self.mouth_radius <= 0:
            raise ValueError(f"Horn radii must be positive, got throat={throat_radius}, mouth={mouth_radius}")
        axis = self.mouth_position - self.throat_position
        self._length = np.linalg.norm(axis)
        if self._length == 0:
            raise ValueError("Horn throat and mouth positions must be distinct")
        self._axis = axis / self._length
        A0 = np.pi * self.throat_radius**2
        Am = np.pi * self.mouth_radius**2
        L = self._length
        if profile == "conical":
            self._m = np.sqrt(Am / A0) - 1
        elif profile == "exponential":
            self._alpha = np.log(Am / A0) / (2 * L)
        elif profile == "hyperbolic":
            self._alpha = np.arccosh(np.sqrt(Am / A0)) / L
        else:
            self._R = L / np.arccos(np.power(A0 / Am, 0.25))

    def radius_at_position(self, t):
        t = np.asarray(t)
        x = t * self._length
        if self.profile == "conical":
            return self.throat_radius * (1 + self._m * t)
        if self.profile == "exponential":
            return self.throat_radius * np.exp(self._alpha * x)
        if self.profile == "hyperbolic":
            return self.throat_radius * np.cosh(self._alpha * x)
        return self.throat_radius / np.cos(x / self._R) ** 2

    @property
    def cutoff_frequency(self) -> float:
        """Flare cutoff for c = 343 m/s: c / (2 pi L) conical, c a / (2 pi)
        exponential and hyperbolic, c / (4 R) tractrix."""
        c = 343.0
        if self.profile == "conical":
            return c / (2 * np.pi * self._length)
        if self.profile in ("exponential", "hyperbolic"):
            return c * self._alpha / (2 * np.pi)
        return c / (4 * self._R)

    def sdf(self, points):
        points = _points(points)
        v = points - self.throat_position
        t = np.dot(v, self._axis)
        tn = np.clip(t / self._length, 0, 1)
        r = np.linalg.norm(v - np.outer(t, self._axis), axis=1)
        radial = r - self.radius_at_position(tn)
        beyond = np.maximum(-t, t - self._length)
        inside_body = (t >= 0) & (t <= self._length)
        return np.where(inside_body, radial, np.sqrt(radial**2 + beyond**2))

    @property
    def bounding_box(self):
        r = np.array([self.mouth_radius] * 3)
        return (np.minimum(self.throat_position, self.mouth_position) - r,
                np.maximum(self.throat_position, self.mouth_position) + r)


def conical_horn(throat_pos, mouth_pos, throat_r: float, mouth_r: float) -> Horn:
    return Horn(throat_pos, mouth_pos, throat_r, mouth_r, "conical")


def exponential_horn(throat_pos, mouth_pos, throat_r: float, mouth_r: float) -> Horn:
    return Horn(throat_pos, mouth_pos, throat_r, mouth_r, "exponential")


def hyperbolic_horn(throat_pos, mouth_pos, throat_r: float, mouth_r: float) -> Horn:
    return Horn(throat_pos, mouth_pos, throat_r, mouth_r, "hyperbolic")


def tractrix_horn(throat_pos, mouth_pos, throat_r: float, mouth_r: float) -> Horn:
    return Horn(throat_pos, mouth_pos, throat_r, mouth_r, "tractrix")


# ---- swept volumes (geometry/sdf.py:1392-1723) --------------------------------
class SpiralHorn(SDFPrimitive):
    """A horn swept along a spiral in the xy plane and extruded over z in
    [0, height] (geometry/sdf.py:1392-1578): the spiral sampled at max(200,
    100 turns) points, the cross-section radius at sample n from the profile
    (exponential rt (rm/rt)^t, linear, or hyperbolic rt / (1 - t (1 - rt/rm))
    at t = n/(N-1)); the distance is the planar distance to the sampled spiral
    minus the radius at the nearest sample, combined with the z excess beyond
    [0, height]."""

    def __init__(self, center, throat_radius: float, mouth_radius: float, inner_spiral_radius: float,
                 outer_spiral_radius: float, turns: float, height: float, profile: str = "exponential",
                 spiral_type: str = "logarithmic"):
        from .paths import SpiralPath
        self.center_2d = np.array(center, dtype=np.float64)
        self.throat_radius = float(throat_radius)
        self.mouth_radius = float(mouth_radius)
        self.inner_spiral_radius = float(inner_spiral_radius)
        self.outer_spiral_radius = float(outer_spiral_radius)
        self.turns = float(turns)
        self.height = float(height)
        self.profile = profile
        self.spiral_type = spiral_type
        if self.throat_radius <= 0 or self.mouth_radius <= 0:
            raise ValueError("Throat and mouth radii must be positive")
        if self.height <= 0:
            raise ValueError("Height must be positive")
        if self.profile not in ("exponential", "linear", "hyperbolic"):
            raise ValueError("profile must be 'exponential', 'linear', or 'hyperbolic'")
        self.spiral = SpiralPath(center=self.center_2d, inner_radius=self.inner_spiral_radius,
                                 outer_radius=self.outer_spiral_radius, turns=self.turns,
                                 spiral_type=self.spiral_type)
        n = max(200, int(100 * self.turns))
        self.path_2d = self.spiral.sample_points(n)
        t = np.linspace(0, 1, n)
        if self.profile == "exponential":
            self.radii = self.throat_radius * (self.mouth_radius / self.throat_radius) ** t
        elif self.profile == "linear":
            self.radii = self.throat_radius + (self.mouth_radius - self.throat_radius) * t
        else:
            self.radii = self.throat_radius / (1 - t * (1 - self.throat_radius / self.mouth_radius))

    def sdf(self, points):
        from .paths import distance_to_polyline, nearest_sample
        points = _points(points)
        xy = points[:, :2]
        z = points[:, 2]
        radial = distance_to_polyline(xy, self.path_2d) - self.radii[nearest_sample(xy, self.path_2d)]
        z_dist = np.maximum(0, np.maximum(-z, z - self.height))
        inside = (z >= 0) & (z <= self.height)
        return np.where(inside, radial, np.sqrt(radial**2 + z_dist**2))

    @property
    def bounding_box(self):
        r = self.outer_spiral_radius + max(self.throat_radius, self.mouth_radius)
        return (np.array([self.center_2d[0] - r, self.center_2d[1] - r, 0.0]),
                np.array([self.center_2d[0] + r, self.center_2d[1] + r, self.height]))


class HelicalTube(SDFPrimitive):
    """A tube swept along a helix, its radius tapering linearly from start to
    end (geometry/sdf.py:1579-1723): the helix sampled at max(200, 100 turns)
    points; the distance is the distance to the sampled helix minus the tube
    radius at the nearest sample."""

    def __init__(self, center, helix_radius: float, tube_radius_start: float, tube_radius_end: float,
                 pitch: float, turns: float):
        from .paths import HelixPath
        self.center = np.array(center, dtype=np.float64)
        self.helix_radius = float(helix_radius)
        self.tube_radius_start = float(tube_radius_start)
        self.tube_radius_end = float(tube_radius_end)
        self.pitch = float(pitch)
        self.turns = float(turns)
        if self.helix_radius <= 0:
            raise ValueError("helix_radius must be positive")
        if self.tube_radius_start <= 0 or self.tube_radius_end <= 0:
            raise ValueError("tube radii must be positive")
        if self.pitch <= 0:
            raise ValueError("pitch must be positive")
        if self.turns <= 0:
            raise ValueError("turns must be positive")
        self.helix = HelixPath(center=self.center, radius=self.helix_radius, pitch=self.pitch, turns=self.turns)
        n = max(200, int(100 * self.turns))
        self.path_3d = self.helix.sample_points(n)
        t = np.linspace(0, 1, n)
        self.tube_radii = self.tube_radius_start + (self.tube_radius_end - self.tube_radius_start) * t

    def sdf(self, points):
        from .paths import distance_to_polyline, nearest_sample
        points = _points(points)
        return distance_to_polyline(points, self.path_3d) - self.tube_radii[nearest_sample(points, self.path_3d)]

    @property
    def bounding_box(self):
        # the reference's box: the helix's height only along z
        r = self.helix_radius + max(self.tube_radius_start, self.tube_radius_end)
        return (np.array([self.center[0] - r, self.center[1] - r, self.center[2]]),
                np.array([self.center[0] + r, self.center[1] + r, self.center[2] + self.helix.total_height]))


# ---- transforms (geometry/sdf.py:1151-1391) ---------------------------------
def rotation_matrix(axis, angle: float) -> NDArray[np.floating]:
    """Rodrigues' rotation matrix I + sin(a) K + (1 - cos(a)) K K for the unit
    axis (K its cross-product matrix), formed with the reference's numpy
    expressions (geometry/sdf.py:1151-1185)."""
    u = np.asarray(axis, dtype=np.float64)
    u = u / np.linalg.norm(u)
    K = np.array([[0, -u[2], u[1]], [u[2], 0, -u[0]], [-u[1], u[0], 0]])
    return np.eye(3) + np.sin(angle) * K + (1 - np.cos(angle)) * (K @ K)


class Transform(SDFPrimitive):
    """A child primitive evaluated in transformed coordinates."""

    def __init__(self, child: SDFPrimitive):
        self.child = child


class Translate(Transform):
    """child moved by `offset`: sdf(p) = child.sdf(p - offset)."""

    def __init__(self, child: SDFPrimitive, offset):
        super().__init__(child)
        self.offset = np.asarray(offset, dtype=np.float64)
        if self.offset.shape != (3,):
            raise ValueError(f"offset must be a 3-element vector, got shape {self.offset.shape}")

    def sdf(self, points):
        return self.child.sdf(points - self.offset)

    @property
    def bounding_box(self):
        lo, hi = self.child.bounding_box
        return lo + self.offset, hi + self.offset


class Scale(Transform):
    """child scaled by a factor (scalar or per axis): sdf(p) =
    child.sdf(p / s) times s (uniform) or times min(s) (not a true distance)."""

    def __init__(self, child: SDFPrimitive, scale):
        super().__init__(child)
        if np.isscalar(scale):
            self.scale = np.array([scale, scale, scale], dtype=np.float64)
            self.uniform = True
        else:
            self.scale = np.asarray(scale, dtype=np.float64)
            self.uniform = np.allclose(self.scale, self.scale[0])
        if self.scale.shape != (3,):
            raise ValueError(f"scale must be scalar or 3-element vector, got shape {self.scale.shape}")
        if np.any(self.scale <= 0):
            raise ValueError(f"scale factors must be positive, got {self.scale}")

    def sdf(self, points):
        d = self.child.sdf(points / self.scale)
        if self.uniform:
            return d * self.scale[0]
        return d * np.min(self.scale)

    @property
    def bounding_box(self):
        lo, hi = self.child.bounding_box
        return lo * self.scale, hi * self.scale


class Rotate(Transform):
    """child rotated by `matrix` (or by `angle` about `axis`): sdf(p) =
    child.sdf(p R), the inverse rotation applied as a product with the
    transpose of the inverse, as the reference writes it."""

    def __init__(self, child: SDFPrimitive, axis=None, angle: float | None = None, matrix=None):
        super().__init__(child)
        self.axis, self.angle = (None, None) if matrix is not None else (axis, angle)
        if matrix is not None:
            self.matrix = np.asarray(matrix, dtype=np.float64)
            if self.matrix.shape != (3, 3):
                raise ValueError(f"matrix must be 3x3, got shape {self.matrix.shape}")
        elif axis is not None and angle is not None:
            self.matrix = rotation_matrix(np.asarray(axis), angle)
        else:
            raise ValueError("Must provide either (axis, angle) or matrix")
        self.inv_matrix = self.matrix.T

    def sdf(self, points):
        return self.child.sdf(points @ self.inv_matrix.T)

    @property
    def bounding_box(self):
        lo, hi = self.child.bounding_box
        corners = np.array([[x, y, z] for x in (lo[0], hi[0]) for y in (lo[1], hi[1]) for z in (lo[2], hi[2])])
        rc = corners @ self.matrix.T
        return rc.min(axis=0), rc.max(axis=0)


@dataclass
class HelmholtzResonator(SDFPrimitive):
    """Cavity + cylindrical neck (geometry/resonator.py:59-298)."""

    position: NDArray[np.floating] | tuple[float, float, float]
    cavity_shape: str = "box"
    cavity_size: tuple[float, ...] = (0.03, 0.03, 0.03)
    neck_length: float = 0.01
    neck_radius: float = 0.005
    neck_direction: NDArray[np.floating] | tuple[float, float, float] = field(
        default_factory=lambda: np.array([0.0, 0.0, 1.0]))
    opening_type: str = "unflanged"

    def __post_init__(self):
        self.position = np.asarray(self.position, dtype=np.float64)
        self.neck_direction = np.asarray(self.neck_direction, dtype=np.float64)
        norm = np.linalg.norm(self.neck_direction)
        if norm < 1e-10:
            raise ValueError("neck_direction cannot be zero vector")
        self.neck_direction = self.neck_direction / norm
        if self.neck_length <= 0:
            raise ValueError(f"neck_length must be positive, got {self.neck_length}")
        if self.neck_radius <= 0:
            raise ValueError(f"neck_radius must be positive, got {self.neck_radius}")
        if self.cavity_shape == "box":
            if len(self.cavity_size) != 3 or any(s <= 0 for s in self.cavity_size):
                raise ValueError("Box cavity requires 3 positive sizes")
            self._cavity = Box(center=tuple(self.position), size=self.cavity_size)
            half = np.array(self.cavity_size, dtype=np.float64) / 2
            t = np.min(half / (np.abs(self.neck_direction) + 1e-10))
            start = self.position + self.neck_direction * t
        elif self.cavity_shape == "sphere":
            if len(self.cavity_size) != 1 or self.cavity_size[0] <= 0:
                raise ValueError("Sphere cavity requires 1 positive radius")
            self._cavity = Sphere(center=tuple(self.position), radius=self.cavity_size[0])
            start = self.position + self.neck_direction * self.cavity_size[0]
        elif self.cavity_shape == "cylinder":
            if len(self.cavity_size) != 2 or any(s <= 0 for s in self.cavity_size):
                raise ValueError("Cylinder cavity requires (radius, length) > 0")
            radius, length = self.cavity_size
            axis = self.neck_direction * length
            self._cavity = Cylinder(p1=tuple(self.position - axis / 2),
                                    p2=tuple(self.position + axis / 2), radius=radius)
            start = self.position + self.neck_direction * (length / 2)
        else:
            raise ValueError(f"Unknown cavity_shape: {self.cavity_shape}")
        end = start + self.neck_direction * self.neck_length
        self._neck = Cylinder(p1=tuple(start), p2=tuple(end), radius=self.neck_radius)
        self._combined = Union(self._cavity, self._neck)

    def sdf(self, points):
        return self._combined.sdf(points)

    @property
    def bounding_box(self):
        return self._combined.bounding_box


@dataclass
class MaterializedGeometry:
    """Geometry plus ordered material regions (material_assignment.py:48-205).

    Later regions override earlier ones; a region only claims air cells.
    """

    geometry: SDFPrimitive
    default_material: object | None = None

    def __post_init__(self):
        self._material_regions: list[tuple[SDFPrimitive, object, int]] = []
        self._material_id_counter = 1
        if self.default_material is None:
            from .materials import AIR_20C
            self.default_material = AIR_20C

    def set_material(self, region: SDFPrimitive, material) -> int:
        if self._material_id_counter > 255:
            raise ValueError("Maximum of 255 material regions exceeded")
        mid = self._material_id_counter
        self._material_id_counter += 1
        self._material_regions.append((region, material, mid))
        return mid

    def voxelize_with_materials(self, grid) -> tuple[NDArray[np.bool_], NDArray[np.uint8]]:
        mask = self.geometry.voxelize(grid)
        ids = np.zeros(grid.shape, dtype=np.uint8)
        for region, _m, mid in self._material_regions:
            ids[mask & region.voxelize(grid)] = mid
        return mask, ids

    def get_material_table(self) -> dict:
        table = {0: self.default_material}
        for _r, m, mid in self._material_regions:
            table[mid] = m
        return table

    def material_count(self) -> int:
        return len(self._material_regions) + 1
