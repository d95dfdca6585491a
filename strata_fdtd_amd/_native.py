"""ctypes binding of ``libfdtd_mi355x.so`` (the C ABI in include/fdtd_mi355x.h).

The library is built in-tree by ``build.py`` (hipcc, gfx950). There is no
fallback: if the library is missing or no MI355X is visible, using the backend
raises.
"""

from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_lib", "libfdtd_mi355x.so")

_vp = C.c_void_p
_i = C.c_int
_f = C.c_float
_i64 = C.c_int64

class BoundaryPlane(C.Structure):
    """fdtd_boundary_plane of the C ABI."""
    _fields_ = [("kind", C.c_int32), ("axis", C.c_int32), ("side", C.c_int32), ("pad", C.c_int32),
                ("coeff", C.c_double), ("refl", C.c_double), ("omr", C.c_double), ("state_off", C.c_int64)]


_SIGS = {
    "fdtd_last_error": (C.c_char_p, []),
    "fdtd_version": (C.c_char_p, []),
    "fdtd_device_count": (_i, [C.POINTER(_i)]),
    "fdtd_ctx_create": (_i, [_i, _vp, C.POINTER(_vp)]),
    "fdtd_ctx_destroy": (_i, [_vp]),
    "fdtd_ctx_set_grid": (_i, [_vp, _i, _i, _i, _i, _f, _f, _vp, _vp, _vp, _vp, _vp, _vp]),
    "fdtd_ctx_set_fields": (_i, [_vp] + [_vp] * 8),
    "fdtd_ctx_set_flags": (_i, [_vp, _vp]),
    "fdtd_ctx_set_pml": (_i, [_vp, _vp, _vp, _vp]),
    "fdtd_ctx_set_boundary_planes": (_i, [_vp, _i, _vp, _vp, _i64]),
    "fdtd_ctx_set_sources": (_i, [_vp, _i, _vp, _vp, _vp, _vp, _i64]),
    "fdtd_ctx_set_probes": (_i, [_vp, _i, _vp, _vp, _i64]),
    "fdtd_ctx_set_membranes": (_i, [_vp, _i, _vp, _vp, _i, _vp, _vp, _vp, _vp, _i, _vp, _i, _vp, _i64]),
    "fdtd_ctx_set_microphones": (_i, [_vp, _i, _vp, _vp, _vp, _vp, _i64]),
    "fdtd_ctx_set_counters": (_i, [_vp, _vp, _vp]),
    "fdtd_ctx_set_ade": (_i, [_vp, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _f]),
    "fdtd_ctx_set_tiling": (_i, [_vp, _i, _i, _i]),
    "fdtd_step_n": (_i, [_vp, _i]),
    "fdtd_graph_run": (_i, [_vp, _i, _i]),
    "fdtd_time_steps": (_i, [_vp, _i, _i, C.POINTER(_f)]),
    "fdtd_stream_ceiling": (_i, [_vp, _i, C.POINTER(_f)]),
    "fdtd_stream_ceiling_march": (_i, [_vp, _i, C.POINTER(_f)]),
    "fdtd_ctx_energy": (_i, [_vp, _vp]),
    "fdtd_ctx_centred_velocity": (_i, [_vp, _vp, _vp, _vp]),
    "fdtd_time_sweeps": (_i, [_vp, _i, C.POINTER(_f), C.POINTER(_f)]),
    "fdtd_batch_create": (_i, [_i, _vp, _i, C.POINTER(_vp)]),
    "fdtd_batch_destroy": (_i, [_vp]),
    "fdtd_batch_set": (_i, [_vp, _vp, _vp, _i]),
    "fdtd_batch_step_n": (_i, [_vp, _i, _i]),
    "fdtd_batch_info": (_i, [_vp, C.POINTER(_i), C.POINTER(_i)]),
    "fdtd_param_step": (_i, [_i] + [_vp] * 8 + [_i, _i, _i] + [_vp] * 8
                        + [C.c_double, C.c_double, _i64, C.c_double, _i, _vp, _vp, _vp]),
    "fdtd_adjoint_step": (_i, [_i] + [_vp] * 12 + [_i, _i, _i] + [_vp] * 8
                          + [C.c_double, C.c_double, _i] + [_vp] * 7),
    "fdtd_param_run": (_i, [_i, _vp, _i, _i64, _i, _i, _i, _i, _i] + [_vp] * 8
                       + [C.c_double, C.c_double, _i64, _vp, _i, _vp, _vp, _vp]),
    "fdtd_adjoint_run": (_i, [_i, _vp, _vp, _i64, _i64, _i, _i, _i] + [_vp] * 8
                         + [C.c_double, C.c_double, _i] + [_vp] * 7),
    "fdtd_design_history_bytes": (_i64, [_vp, _vp, _i64]),
    "fdtd_design_forward": (_i, [_vp, _vp, _vp, _i, _i, _i, _f, _f, _vp, _vp, _vp, _vp, _vp, _i, _vp, _vp,
                                 _i64, _i64, _i, _vp, _vp, _i64, _i64, _i, _vp]),
    "fdtd_design_adjoint": (_i, [_vp, _vp, _vp, _i, _i, _i, _f, _f, _vp, _vp, _vp, _i, _vp, _vp, _i64, _i64, _i64,
                                 _vp, _vp, _i64, _i64, _vp, _vp, _vp, _i, _vp]),
    "fdtd_design_face_count": (_i64, [_vp, _vp]),
    "fdtd_update_velocity": (_i, [_vp, _vp, _vp, _vp, _i, _i, _i, _f, _vp]),
    "fdtd_update_velocity_nonuniform": (_i, [_vp, _vp, _vp, _vp, _i, _i, _i, _f, _vp, _vp, _vp, _vp]),
    "fdtd_apply_rigid": (_i, [_vp, _vp, _vp, _vp, _i, _i, _i, _vp]),
    "fdtd_update_pressure": (_i, [_vp, _vp, _vp, _vp, _vp, _i, _i, _i, _f, _vp]),
    "fdtd_update_pressure_nonuniform": (_i, [_vp, _vp, _vp, _vp, _vp, _i, _i, _i, _f, _vp, _vp, _vp, _vp]),
    "fdtd_apply_pml_velocity": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _i, _vp]),
    "fdtd_apply_pml_pressure": (_i, [_vp, _vp, _vp, _vp, _i, _i, _i, _vp]),
    "fdtd_compute_divergence": (_i, [_vp, _vp, _vp, _vp, _i, _i, _i, _f, _vp, _vp, _vp, _vp]),
    "fdtd_build_flags": (_i, [_vp, _vp, _vp, _i, _i, _i, _vp]),
    "fdtd_update_ade_density_debye": (_i, [_vp, _vp, _vp, _vp, _i, _i, _i, _i, _vp]),
    "fdtd_update_ade_modulus_debye": (_i, [_vp, _vp, _vp, _vp, _i, _i, _i, _i, _vp]),
    "fdtd_update_ade_density_lorentz": (_i, [_vp, _vp, _vp, _vp, _vp, _i, _i, _i, _i, _vp]),
    "fdtd_update_ade_modulus_lorentz": (_i, [_vp, _vp, _vp, _vp, _vp, _i, _i, _i, _i, _vp]),
    "fdtd_apply_ade_velocity_correction": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _vp, _f, _f, _i, _i, _i,
                                                _vp]),
    "fdtd_apply_ade_pressure_correction": (_i, [_vp, _vp, _vp, _vp, _vp, _i, _vp, _f, _i, _i, _i, _vp]),
    "fdtd_voxelize_scratch_bytes": (_i64, [_i, _i, _i]),
    "fdtd_voxelize": (_i, [_vp, _i, _vp, _vp, _i, _vp, _i, _vp, _i64, _vp, _vp, _vp, _i, _i, _i, _vp, _vp, _vp]),
    "fdtd_voxelize_static": (_i, [_vp, _i, _vp, _vp, _i, _vp, _i, _vp, _vp, _vp, _i, _i, _i, _vp, _vp, _vp]),
}

_lib: C.CDLL | None = None


class NativeError(RuntimeError):
    """A C-ABI call returned a negative status."""


def declared_symbols() -> list[str]:
    return list(_SIGS)


def load(path: str = LIB_PATH) -> C.CDLL:
    """Load the library (once) and attach signatures. Raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(
            f"MI355X backend library not built: {path} is missing "
            "(run `python -c 'import __graft_entry__ as g; g.build()'` or "
            "`python -m strata_fdtd_amd.build`)")
    lib = C.CDLL(path)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def call(name: str, *args) -> int:
    """Call a status-returning entry point; raise NativeError on failure."""
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc != 0:
        msg = lib.fdtd_last_error().decode(errors="replace")
        raise NativeError(f"{name} failed ({rc}): {msg}")
    return rc


def ptr(t) -> int | None:
    """Device pointer of a torch tensor (None passes NULL)."""
    return None if t is None else t.data_ptr()
