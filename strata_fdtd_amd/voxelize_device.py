"""Device voxeliser: SDF trees compiled to postfix programs for the HIP kernel
(``csrc/voxelize.hip``, C ABI ``fdtd_voxelize``).

The kernel evaluates every primitive with the reference's float64 formulas
(geometry/sdf.py: Box :274-294, Sphere :320-333, Cylinder :382-425, Cone
:436-543) at each
cell centre and combines them as booleans (Union :566-585 -> or,
Intersection :629-648 -> and, Difference :704-722 -> and-not with the
subtrahend's test made strict, since ``max(b, -s) <= 0`` iff ``b <= 0`` and not
``s < 0``). Material regions follow MaterializedGeometry.voxelize_with_materials
(geometry/material_assignment.py:133-168): later regions override earlier ones,
only inside air.

Supported: Box, Sphere, Cylinder and Cone whose axis is parallel to a grid
axis, Union, Intersection, Difference and HelmholtzResonator built from those,
and Translate and Scale (geometry/sdf.py :1199-1301) of any of them: the
kernel evaluates the subtree at p - offset or p / scale, the coordinates numpy
forms, and a scale's positive factor (limited to [1e-6, 1e6]) keeps the sign
of the distance. For an oblique cylinder or cone, and for a rotation (a
matrix product per point), the reference's values go through BLAS, whose
summation order is not fixed; smooth CSG needs distances, not inside tests.
Such trees raise ``Unsupported`` and the caller uses the host voxeliser.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import _native
from .geometry import Box, Cone, Cylinder, Difference, HelmholtzResonator, Intersection, Scale, Sphere, Translate, Union

(VOX_FALSE, VOX_TRUE, VOX_BOX, VOX_SPH, VOX_CYL, VOX_OR, VOX_AND, VOX_ANDNOT, VOX_CONE, VOX_TRANSLATE, VOX_SCALE,
 VOX_POP) = range(12)
VOX_STRICT = 16
MAX_DEPTH = 32
MAX_XFORM = 8            # transforms nested per program
SCALE_RANGE = (1e-6, 1e6)


class Unsupported(NotImplementedError):
    """The tree holds a node the device voxeliser does not evaluate exactly."""


@dataclass
class ProgramSet:
    ops: np.ndarray      # int32 [n_ops, 2]
    start: np.ndarray    # int32 [n_prog + 1]
    ids: np.ndarray      # int32 [n_prog] (entry 0 unused)
    params: np.ndarray   # float64

    @property
    def n_prog(self) -> int:
        return len(self.start) - 1


def _cyl_axis(node) -> int:
    n = node.axis_normalized
    nz = np.flatnonzero(n)
    if len(nz) != 1 or abs(float(n[nz[0]])) != 1.0:
        kind = type(node).__name__.lower()
        raise Unsupported(f"device voxeliser: {kind} axis must be parallel to a grid axis")
    return int(nz[0])


def _emit(node, strict: bool, ops: list, params: list) -> int:
    """Append the postfix code of `node` (test sdf < 0 if strict, else <= 0);
    returns the stack depth it needs."""
    s = VOX_STRICT if strict else 0
    if isinstance(node, Box):
        ops.append((VOX_BOX | s, len(params)))
        half = node.size / 2
        params.extend([*map(float, node.center), *map(float, half)])
        return 1
    if isinstance(node, Sphere):
        ops.append((VOX_SPH | s, len(params)))
        params.extend([*map(float, node.center), float(node.radius)])
        return 1
    if isinstance(node, Cylinder):
        ax = _cyl_axis(node)
        ops.append((VOX_CYL | s, len(params)))
        params.extend([*map(float, node.p1), *map(float, node.axis_normalized), float(node.length),
                       float(node.radius), float(ax)])
        return 1
    if isinstance(node, Cone):
        ax = _cyl_axis(node)
        ops.append((VOX_CONE | s, len(params)))
        params.extend([*map(float, node.p1), *map(float, node.axis_normalized), float(node.length),
                       float(node.r1), float(node.r2), float(ax)])
        return 1
    if isinstance(node, HelmholtzResonator):
        return _emit(node._combined, strict, ops, params)
    if isinstance(node, (Translate, Scale)):
        if isinstance(node, Scale):
            if not np.all((node.scale >= SCALE_RANGE[0]) & (node.scale <= SCALE_RANGE[1])):
                raise Unsupported("device voxeliser: scale factors outside [1e-6, 1e6]")
            ops.append((VOX_SCALE, len(params)))
            params.extend(map(float, node.scale))
        else:
            ops.append((VOX_TRANSLATE, len(params)))
            params.extend(map(float, node.offset))
        depth = _emit(node.child, strict, ops, params)
        ops.append((VOX_POP, 0))
        return depth
    if isinstance(node, (Union, Intersection)):
        if not node.children:
            # min over nothing is +inf (never inside); max over nothing is -inf (always)
            ops.append((VOX_FALSE if isinstance(node, Union) else VOX_TRUE, 0))
            return 1
        comb = VOX_OR if isinstance(node, Union) else VOX_AND
        depth = _emit(node.children[0], strict, ops, params)
        for ch in node.children[1:]:
            depth = max(depth, 1 + _emit(ch, strict, ops, params))
            ops.append((comb, 0))
        return depth
    if isinstance(node, Difference):
        depth = _emit(node.base, strict, ops, params)
        for sub in node.subtracted:
            depth = max(depth, 1 + _emit(sub, not strict, ops, params))
            ops.append((VOX_ANDNOT, 0))
        return depth
    raise Unsupported(f"device voxeliser: {type(node).__name__} is not supported")


def compile_programs(geometry, regions=()) -> ProgramSet:
    """Program 0: the air mask of `geometry`; program m: region m with its id.
    `regions` is a sequence of (SDF node, material id)."""
    ops: list = []
    params: list = []
    start = [0]
    ids = []
    for node, mid in [(geometry, 0), *regions]:
        first = len(ops)
        if _emit(node, False, ops, params) > MAX_DEPTH:
            raise Unsupported("device voxeliser: CSG tree deeper than 32")
        nest = 0
        for op, _arg in ops[first:]:
            nest += 1 if op in (VOX_TRANSLATE, VOX_SCALE) else (-1 if op == VOX_POP else 0)
            if nest > MAX_XFORM:
                raise Unsupported("device voxeliser: transforms nested deeper than 8")
        start.append(len(ops))
        ids.append(int(mid))
    return ProgramSet(np.asarray(ops, dtype=np.int32).reshape(-1, 2), np.asarray(start, dtype=np.int32),
                      np.asarray(ids, dtype=np.int32), np.asarray(params, dtype=np.float64))


def supports(geometry) -> bool:
    try:
        _emit(geometry, False, [], [])
        return True
    except Unsupported:
        return False


def _regions_of(geometry):
    from .geometry import MaterializedGeometry
    if isinstance(geometry, MaterializedGeometry):
        return geometry.geometry, [(r, mid) for r, _m, mid in geometry._material_regions]
    return geometry, []


def voxelize_device(geometry, grid, device=None, stream=None):
    """Air mask (uint8 0/1) and material ids (uint8, or None without regions)
    as device tensors, computed on the GPU. `geometry` is an SDF tree or a
    MaterializedGeometry; raises Unsupported for trees the kernel does not
    evaluate exactly."""
    import torch
    tree, regions = _regions_of(geometry)
    ps = compile_programs(tree, regions)
    dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
    st = stream if stream is not None else torch.cuda.current_stream(dev)
    nx, ny, nz = grid.shape
    coords = [torch.from_numpy(np.ascontiguousarray(c, dtype=np.float64)).to(dev)
              for c in (grid.x_coords, grid.y_coords, grid.z_coords)]
    lib = _native.load()
    need = int(lib.fdtd_voxelize_scratch_bytes(len(ps.ops), ps.n_prog, len(ps.params)))
    scratch = torch.empty(max(need, 8), dtype=torch.uint8, device=dev)
    air = torch.empty((nx, ny, nz), dtype=torch.uint8, device=dev)
    ids = torch.empty((nx, ny, nz), dtype=torch.uint8, device=dev) if ps.n_prog > 1 else None
    ops = np.ascontiguousarray(ps.ops)
    _native.call("fdtd_voxelize", ops.ctypes.data, len(ops), ps.start.ctypes.data, ps.ids.ctypes.data, ps.n_prog,
                 ps.params.ctypes.data if len(ps.params) else None, len(ps.params), _native.ptr(scratch),
                 scratch.numel(), *[_native.ptr(c) for c in coords], nx, ny, nz, _native.ptr(air),
                 _native.ptr(ids), st.cuda_stream)
    # keep the inputs alive until the kernel has run on `st`
    for t in (scratch, *coords):
        t.record_stream(st)
    return air, ids


def evaluate_host(ps: ProgramSet, xc, yc, zc):
    """Numpy restatement of the kernel's program semantics (test helper):
    returns (air bool, ids uint8) over the grid given by the coordinates."""
    X, Y, Z = np.meshgrid(np.asarray(xc, np.float64), np.asarray(yc, np.float64), np.asarray(zc, np.float64),
                          indexing="ij")

    def norm3(a, b, c):
        return np.sqrt((a * a + b * b) + c * c)

    def run(m):
        nonlocal X, Y, Z
        stack = []
        saved = []
        for op, arg in ps.ops[ps.start[m]:ps.start[m + 1]]:
            strict = bool(op & VOX_STRICT)
            code = int(op) & ~VOX_STRICT
            le = (lambda d: d < 0) if strict else (lambda d: d <= 0)
            P = ps.params[arg:arg + 10]
            if code in (VOX_TRANSLATE, VOX_SCALE):
                saved.append((X, Y, Z))
                if code == VOX_TRANSLATE:
                    X, Y, Z = X - P[0], Y - P[1], Z - P[2]
                else:
                    X, Y, Z = X / P[0], Y / P[1], Z / P[2]
                continue
            if code == VOX_POP:
                X, Y, Z = saved.pop()
                continue
            if code == VOX_CONE:
                ax = int(P[9])
                v = (X - P[0], Y - P[1], Z - P[2])
                na = P[3 + ax]
                L, r1, r2 = P[6], P[7], P[8]
                t = (v[ax] if na > 0 else -v[ax]) / L
                tc = np.where(t > 0, t, 0.0)
                tc = np.where(tc < 1.0, tc, 1.0)
                rt = r1 + tc * (r2 - r1)
                sl = tc * L
                c = [np.full_like(X, P[d]) if d != ax else P[d] + sl * na for d in range(3)]
                radial = norm3(X - c[0], Y - c[1], Z - c[2])
                ad = np.abs(t - 0.5) - 0.5
                axial = np.where(0.0 >= ad, 0.0, ad) * L
                body = (t >= 0) & (t <= 1.0)
                rr = radial - rt
                stack.append(le(np.where(body, rr, np.sqrt(rr * rr + axial * axial))))
                continue
            if code == VOX_BOX:
                bit = le(np.abs(X - P[0]) - P[3]) & le(np.abs(Y - P[1]) - P[4]) & le(np.abs(Z - P[2]) - P[5])
            elif code == VOX_SPH:
                bit = le(norm3(X - P[0], Y - P[1], Z - P[2]) - P[3])
            elif code == VOX_CYL:
                ax = int(P[8])
                v = (X - P[0], Y - P[1], Z - P[2])
                na = P[3 + ax]
                t = v[ax] if na > 0 else -v[ax]
                L, r = P[6], P[7]
                tc = np.where(t > 0, t, 0.0)
                tc = np.where(tc < L, tc, L)
                c = [np.full_like(X, P[d]) if d != ax else P[d] + tc * na for d in range(3)]
                radial = norm3(X - c[0], Y - c[1], Z - c[2])
                ad = np.abs(t - L / 2) - L / 2
                axial = np.where(0.0 >= ad, 0.0, ad)
                body = (t >= 0) & (t <= L)
                rr = radial - r
                bit = le(np.where(body, rr, np.sqrt(rr * rr + axial * axial)))
            elif code in (VOX_TRUE, VOX_FALSE):
                bit = np.full(X.shape, code == VOX_TRUE)
            else:
                b = stack.pop()
                a = stack.pop()
                bit = (a | b) if code == VOX_OR else ((a & b) if code == VOX_AND else (a & ~b))
            stack.append(bit)
        assert len(stack) == 1
        return stack[0]

    air = run(0)
    ids = np.zeros(X.shape, dtype=np.uint8)
    for m in range(1, ps.n_prog):
        ids[air & run(m)] = ps.ids[m]
    return air, ids
