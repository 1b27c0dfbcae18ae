"""The C ABI library builds, loads and exports every entry point that
include/fmm3d.h declares; without a device it fails with a status code
instead of aborting."""
import ctypes as C
import os
import re

import pytest

from fmm3d import build, lib, synth

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    with open(os.path.join(REPO, "include", "fmm3d.h")) as f:
        text = f.read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(fmm3d_\w+)\s*\(", text, re.M)))


def test_library_exports_header():
    if not os.path.exists(lib.LIB_PATH):
        build.build_hip()
    L = lib.load()
    names = declared()
    assert len(names) >= 9
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(lib.EXPORTS)


def test_create_reports_errors_without_aborting():
    L = lib.load()
    g = synth.make_grid((11, 12, 13), 100.0, (2.0, 2.0), nv=(6, 6, 7))
    h = C.c_void_p()
    bad = lib.grid_c(g)
    bad.nnr = 1
    assert L.fmm3d_create(C.byref(bad), 0, C.byref(h)) == lib.E_ARG
    assert b"3 nodes" in L.fmm3d_last_error(None)
    rc = L.fmm3d_create(C.byref(lib.grid_c(g)), 0, C.byref(h))
    if rc == lib.OK:            # a device is present
        assert L.fmm3d_destroy(h) == lib.OK
    else:
        assert rc == lib.E_DEVICE
        assert L.fmm3d_last_error(None)
