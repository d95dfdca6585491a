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
