"""ctypes bindings for the C++ oracle and for the reference's own native kernels.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

``Problem`` is a plain description of one solve (shape, coefficients, tables,
sources, probes, ADE poles). ``run(problem, n_steps, impl=...)`` advances it
with one of three CPU implementations:

* ``"oracle"``  — ``fdtd_oracle.cpp`` in float32 (the parity oracle);
* ``"oracle64"`` — the same step sequence in float64 (the "error vs FP64" column);
* ``"ref"``     — the reference's own ``_kernels/src/*.cpp`` compiled from source
  into ``oracle/_ref/libstrata_ref.so`` and composed in the solver's python step
  ordering by ``ref_harness.cpp`` (available only where the reference checkout
  was present at build time).
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "_build", "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libstrata_ref.so")

_fp = C.POINTER(C.c_float)
_dp = C.POINTER(C.c_double)
_u8p = C.POINTER(C.c_uint8)
_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)


class _Pole(C.Structure):
    _fields_ = [("material_id", C.c_int32), ("target", C.c_int32), ("lorentz", C.c_int32),
                ("pad", C.c_int32), ("a", C.c_double), ("b", C.c_double), ("d", C.c_double),
                ("corr", C.c_double)]


class _Boundary(C.Structure):
    _fields_ = [("kind", C.c_int32), ("axis", C.c_int32), ("side", C.c_int32), ("pad", C.c_int32),
                ("coeff", C.c_double), ("refl", C.c_double), ("omr", C.c_double), ("state_off", C.c_int64)]


class _Problem(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
                ("nonuniform", C.c_int32), ("cv", C.c_double), ("cp", C.c_double),
                ("inv_face", _fp * 3), ("inv_cell", _fp * 3), ("air", _u8p),
                ("decay", _fp * 3), ("n_src", C.c_int32), ("src_kind", _i32p),
                ("src_a", _i64p), ("src_b", _i64p), ("src_wave", _dp), ("wave_len", C.c_int32),
                ("n_probe", C.c_int32), ("probe_idx", _i64p), ("n_pole", C.c_int32),
                ("poles", C.POINTER(_Pole)), ("material", _u8p), ("ade_inv_dx", C.c_double),
                ("src_weight", C.POINTER(_dp)), ("n_bnd", C.c_int32), ("bnd", C.POINTER(_Boundary)),
                ("bnd_state", C.c_void_p)]


class _Material(C.Structure):
    _fields_ = [("n", C.c_int32), ("rho_inf", _fp), ("K_inf", _fp), ("dt", C.c_double)]


@dataclass
class Pole:
    material_id: int
    target: int          # 0 density, 1 modulus
    lorentz: int         # 0 Debye, 1 Lorentz
    a: float
    b: float
    d: float = 0.0
    corr: float = 0.0    # float32-formed correction coefficient (see tables.native_*)


@dataclass
class Problem:
    shape: tuple[int, int, int]
    cv: float
    cp: float
    air: np.ndarray                      # uint8/bool, shape
    decay: tuple = (None, None, None)    # float32 tables or None
    nonuniform: bool = False
    inv_face: tuple = (None, None, None)
    inv_cell: tuple = (None, None, None)
    sources: list = field(default_factory=list)  # (kind, a, b, wave float64[][, weights float64 shape])
    probes: list = field(default_factory=list)   # flat indices
    poles: list = field(default_factory=list)    # Pole
    material: np.ndarray | None = None
    ade_inv_dx: float = 0.0
    rho_inf: np.ndarray | None = None    # per material id (reference harness)
    K_inf: np.ndarray | None = None
    dt: float = 0.0
    # boundary planes after the PML, before the sources, in order:
    # (kind 0 Mur ABC | 1 radiation impedance, axis, side, coeff, refl, omr)
    boundaries: list = field(default_factory=list)
    # their previous-step planes per run dtype, kept across calls (zeroed by a
    # call that starts at step 0)
    bnd_state: dict = field(default_factory=dict)


def boundary_layout(shape, boundaries):
    """Element offset of each boundary entry's planes in the state buffer and
    the total size: a Mur entry holds two cross-sections of its axis, a
    radiation entry one."""
    offs, total = [], 0
    for kind, axis, *_ in boundaries:
        cross = int(np.prod(shape)) // shape[axis]
        offs.append(total)
        total += cross * (2 if kind == 0 else 1)
    return offs, total


def build(ref: bool = True) -> None:
    """Compile the oracle (and, where the reference checkout exists, oracle/_ref)."""
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)
    if ref and os.path.isdir("/root/reference/src/strata_fdtd/_kernels"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


_libs: dict[str, C.CDLL] = {}


def _lib(which: str) -> C.CDLL:
    if which not in _libs:
        path = ORACLE_SO if which == "oracle" else REF_SO
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} is not built (run oracle.cpu.build())")
        _libs[which] = C.CDLL(path)
    return _libs[which]


def have_ref() -> bool:
    return os.path.exists(REF_SO)


def _ptr(a, t):
    return None if a is None else a.ctypes.data_as(t)


class _Packed:
    """Keeps numpy buffers alive while the ctypes struct points into them."""

    def __init__(self, pb: Problem):
        nx, ny, nz = pb.shape
        self.keep = []

        def f32(a):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=np.float32)
            self.keep.append(a)
            return a

        s = _Problem()
        s.nx, s.ny, s.nz = nx, ny, nz
        s.nonuniform = 1 if pb.nonuniform else 0
        s.cv, s.cp = pb.cv, pb.cp
        for ax in range(3):
            s.inv_face[ax] = _ptr(f32(pb.inv_face[ax]), _fp)
            s.inv_cell[ax] = _ptr(f32(pb.inv_cell[ax]), _fp)
            s.decay[ax] = _ptr(f32(pb.decay[ax]), _fp)
        air = np.ascontiguousarray(pb.air, dtype=np.uint8)
        self.keep.append(air)
        s.air = _ptr(air, _u8p)
        ns = len(pb.sources)
        wl = max([len(src[3]) for src in pb.sources], default=1)
        kinds = np.array([src[0] for src in pb.sources] or [0], dtype=np.int32)
        sa = np.array([src[1] for src in pb.sources] or [0], dtype=np.int64)
        sb = np.array([src[2] for src in pb.sources] or [0], dtype=np.int64)
        wave = np.zeros((max(ns, 1), wl), dtype=np.float64)
        for q, src in enumerate(pb.sources):
            wave[q, :len(src[3])] = src[3]
        self.keep += [kinds, sa, sb, wave]
        s.n_src, s.wave_len = ns, wl
        s.src_kind, s.src_a, s.src_b = _ptr(kinds, _i32p), _ptr(sa, _i64p), _ptr(sb, _i64p)
        s.src_wave = _ptr(wave, _dp)
        wptrs = (_dp * max(ns, 1))()
        for q, src in enumerate(pb.sources):
            if src[0] == 2:
                w = np.ascontiguousarray(src[4], dtype=np.float64)
                assert w.shape == tuple(pb.shape)
                self.keep.append(w)
                wptrs[q] = w.ctypes.data_as(_dp)
        self.keep.append(wptrs)
        s.src_weight = C.cast(wptrs, C.POINTER(_dp))
        pr = np.array(pb.probes or [0], dtype=np.int64)
        self.keep.append(pr)
        s.n_probe = len(pb.probes)
        s.probe_idx = _ptr(pr, _i64p)
        s.n_pole = len(pb.poles)
        if pb.poles:
            arr = (_Pole * len(pb.poles))()
            for q, p in enumerate(pb.poles):
                arr[q] = _Pole(p.material_id, p.target, p.lorentz, 0, p.a, p.b, p.d, p.corr)
            self.keep.append(arr)
            s.poles = C.cast(arr, C.POINTER(_Pole))
            mat = np.ascontiguousarray(pb.material, dtype=np.uint8)
            self.keep.append(mat)
            s.material = _ptr(mat, _u8p)
        s.ade_inv_dx = pb.ade_inv_dx
        s.n_bnd = len(pb.boundaries)
        if pb.boundaries:
            offs, _ = boundary_layout(pb.shape, pb.boundaries)
            arr = (_Boundary * len(pb.boundaries))()
            for q, (bd, off) in enumerate(zip(pb.boundaries, offs)):
                kind, axis, side, coeff, refl, omr = bd
                arr[q] = _Boundary(kind, axis, side, 0, coeff, refl, omr, off)
            self.keep.append(arr)
            s.bnd = C.cast(arr, C.POINTER(_Boundary))
        self.s = s
        m = _Material()
        if pb.rho_inf is not None:
            ri = np.ascontiguousarray(pb.rho_inf, dtype=np.float32)
            ki = np.ascontiguousarray(pb.K_inf, dtype=np.float32)
            self.keep += [ri, ki]
            m.n, m.rho_inf, m.K_inf = len(ri), _ptr(ri, _fp), _ptr(ki, _fp)
        m.dt = pb.dt
        self.m = m


def new_state(shape, dtype=np.float32):
    return {k: np.zeros(shape, dtype=dtype) for k in ("p", "vx", "vy", "vz")}


def run(pb: Problem, n_steps: int, state: dict | None = None, impl: str = "oracle",
        step0: int = 0) -> tuple[dict, np.ndarray]:
    """Advance ``state`` (created zeroed if None) by ``n_steps``; returns (state, trace).

    The ADE pole accumulators live only inside one call (they start at zero),
    so a problem with poles must run from step 0 in a single call: continuing
    it with step0 > 0 would restart its poles and is refused."""
    if pb.poles and step0 != 0:
        raise ValueError("a problem with ADE poles runs from step 0 in one call "
                         "(pole state does not carry across calls)")
    dtype = np.float64 if impl == "oracle64" else np.float32
    if state is None:
        state = new_state(pb.shape, dtype)
    for k in ("p", "vx", "vy", "vz"):
        assert state[k].dtype == dtype and state[k].flags.c_contiguous
    pk = _Packed(pb)
    if pb.boundaries:
        if impl == "ref":
            raise ValueError("the reference kernels have no boundary planes (they are python-only there)")
        _, total = boundary_layout(pb.shape, pb.boundaries)
        st = pb.bnd_state.get(dtype)
        if st is None or st.size != max(total, 1) or step0 == 0:
            st = pb.bnd_state[dtype] = np.zeros(max(total, 1), dtype=dtype)
        pk.s.bnd_state = st.ctypes.data
    trace = np.zeros((n_steps, max(len(pb.probes), 1)), dtype=np.float32)
    args = [state[k] for k in ("p", "vx", "vy", "vz")]
    if impl == "oracle":
        fn = _lib("oracle").orc_run_f32
        rc = fn(C.byref(pk.s), n_steps, step0, *[_ptr(a, _fp) for a in args], _ptr(trace, _fp))
    elif impl == "oracle64":
        fn = _lib("oracle").orc_run_f64
        rc = fn(C.byref(pk.s), n_steps, step0, *[_ptr(a, _dp) for a in args], _ptr(trace, _fp))
    elif impl == "ref":
        fn = _lib("ref").ref_run
        rc = fn(C.byref(pk.s), C.byref(pk.m), n_steps, step0, *[_ptr(a, _fp) for a in args],
                _ptr(trace, _fp))
    else:
        raise ValueError(impl)
    if rc != 0:
        raise RuntimeError(f"{impl} run failed: {rc}")
    return state, trace[:, :len(pb.probes)]


def boundary_planes(pb: Problem, p: np.ndarray, reset: bool = False) -> None:
    """Apply the problem's boundary planes once to the float32 field p (in
    place), continuing their plane state (zeroed first when ``reset``)."""
    assert p.dtype == np.float32 and p.flags.c_contiguous and p.shape == tuple(pb.shape)
    pk = _Packed(pb)
    _, total = boundary_layout(pb.shape, pb.boundaries)
    st = pb.bnd_state.get(np.float32)
    if st is None or st.size != max(total, 1) or reset:
        st = pb.bnd_state[np.float32] = np.zeros(max(total, 1), dtype=np.float32)
    pk.s.bnd_state = st.ctypes.data
    _lib("oracle").orc_boundary_planes_f32(C.byref(pk.s), _ptr(p, _fp))


# ---- single-phase helpers (float32, in place) --------------------------------
def phase(name: str, *arrays, impl: str = "oracle", **kw):
    """Call one phase of the oracle (``orc_*``) or of the reference (``ref_*``)."""
    lib = _lib(impl)
    fn = getattr(lib, ("orc_" if impl == "oracle" else "ref_") + name)
    conv = []
    for a in arrays:
        if a is None:
            conv.append(None)
        elif isinstance(a, np.ndarray):
            conv.append(a.ctypes.data_as(_u8p if a.dtype in (np.uint8, np.bool_) else _fp))
        elif isinstance(a, float):
            conv.append(C.c_float(a))
        else:
            conv.append(C.c_int(int(a)))
    return fn(*conv)


def num_threads(impl: str = "oracle") -> int:
    lib = _lib(impl)
    return int((lib.orc_num_threads if impl == "oracle" else lib.ref_num_threads)())
