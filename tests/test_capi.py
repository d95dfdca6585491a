"""The C-ABI library loads and exports every entry point include/*.h declares
(no compute: runs without a GPU)."""

import ctypes
import glob
import os
import re

from conftest import ROOT


def _declared():
    names = []
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(fdtd_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    from strata_fdtd_amd import _native
    lib = ctypes.CDLL(_native.LIB_PATH)
    declared = _declared()
    assert len(declared) >= 20
    missing = [n for n in declared if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_declarations():
    from strata_fdtd_amd import _native
    assert set(_declared()) == set(_native.declared_symbols())


def test_version_and_error_string_without_gpu():
    from strata_fdtd_amd import _native
    lib = _native.load()
    assert b"gfx950" in lib.fdtd_version()
    # a failing call on a null context returns an error and sets the message
    rc = lib.fdtd_step_n(None, 1)
    assert rc < 0
    assert b"context" in lib.fdtd_last_error()


def test_availability_checks():
    """The reference's has_native_kernels / has_gpu_support: the built library
    loads; a GPU is reported only where torch sees one."""
    import torch

    import strata_fdtd_amd as S
    assert S.has_native_kernels() is True
    assert S.has_gpu_support() == torch.cuda.is_available()


def test_design_calls_refuse_rows_past_their_tables_without_gpu():
    """fdtd_design_forward / fdtd_design_adjoint check the trace, waveform and
    g rows a call would touch against steps_len on the host, before any
    launch (the pointers below are never dereferenced)."""
    from strata_fdtd_amd import _native
    lib = _native.load()
    fake = [ctypes.c_void_p(0x10000 + 0x100 * q) for q in range(8)]
    sets = (ctypes.c_void_p * 8)(*fake)
    dec = (ctypes.c_void_p * 3)(*fake[:3])
    lo = (ctypes.c_int32 * 3)(2, 2, 2)
    hi = (ctypes.c_int32 * 3)(10, 10, 10)
    src = (ctypes.c_int32 * 3)(5, 5, 5)
    common = (sets, fake[0], dec, 16, 16, 16, -0.01, -30.0, lo, hi, None)
    # forward: steps [5, 15) against 10 rows
    rc = lib.fdtd_design_forward(*common, src, fake[1], 1, fake[2], fake[3], 10, 5, 10, None, None, 0, 0, 2, None)
    assert rc == -4
    # adjoint: steps [0, 12) against 10 rows of g
    rc = lib.fdtd_design_adjoint(*common, 1, fake[2], fake[4], 10, 0, 12, fake[5], fake[6], 0, 12, fake[7],
                                 fake[1], fake[3], 4, None)
    assert rc == -4
