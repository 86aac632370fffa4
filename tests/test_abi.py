"""The C ABI library: builds, loads and exports exactly what include/npbmz.h
declares; argument validation runs before any device work (no GPU needed)."""
import ctypes
import os
import re

import pytest

import npbapps_amd as P
from npbapps_amd import native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    hdr = open(os.path.join(ROOT, "include", "npbmz.h")).read()
    return set(re.findall(r"\b(npbmz_\w+)\s*\(", hdr))


def test_header_and_binding_list_the_same_entry_points():
    assert _declared() == set(native.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(native.LIB_PATH)
    for name in _declared():
        assert hasattr(lib, name), name


def test_create_validates_before_touching_a_device():
    lib = native.load()
    z, p = P.zones_for("BT", "S")
    ctx = native.Context.__new__(native.Context)
    bad = native.Problem(0, 3, 2, 2, 24, 24, 6, None, None, None, None, None, None, None,
                         0.01, 0.0)
    h = ctypes.c_void_p()
    rc = lib.npbmz_create(ctypes.byref(bad), 0, ctypes.byref(h))
    assert rc == -1 and b"nzones" in lib.npbmz_last_error(None)
    ok = [(ctypes.c_int32 * 4)(*v) for v in (z.nx, z.ny, z.nz, z.west, z.east, z.south, z.north)]
    q = native.Problem(0, 4, 2, 2, 24, 24, 6, *ok, -1.0, 0.0)
    rc = lib.npbmz_create(ctypes.byref(q), 0, ctypes.byref(h))
    assert rc == -1 and b"dt" in lib.npbmz_last_error(None)
    mism = list(z.ny)
    mism[1] += 1
    ok[1] = (ctypes.c_int32 * 4)(*mism)
    q = native.Problem(0, 4, 2, 2, 24, 24, 6, *ok, 0.01, 0.0)
    rc = lib.npbmz_create(ctypes.byref(q), 0, ctypes.byref(h))
    assert rc == -1 and b"neighbours" in lib.npbmz_last_error(None)
    del ctx


def test_missing_library_fails_loudly(tmp_path):
    native_mod = native
    with pytest.raises(native_mod.NpbmzError, match="not built"):
        old = native_mod._lib
        native_mod._lib = None
        try:
            native_mod.load(str(tmp_path / "absent.so"))
        finally:
            native_mod._lib = old
