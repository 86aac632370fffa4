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


def test_binding_constants_match_the_header():
    """Every NPBMZ_SCHED_*, NPBMZ_PHASE_* and NPBMZ_FLAG_* value of the header
    is the binding's constant of the same name."""
    hdr = open(os.path.join(ROOT, "include", "npbmz.h")).read()
    defs = dict(re.findall(r"#define NPBMZ_((?:SCHED|PHASE|FLAG)_\w+)\s+\(?(\w+)", hdr))
    assert defs
    for name, val in defs.items():
        assert hasattr(native, name), name
        assert getattr(native, name) == int(val, 0), (name, val)


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


def test_create_validates_lu_problems():
    """LU needs 0 < omega < 2 and equal zones (src/LU/zone_setup.jl:22-33)."""
    lib = native.load()
    z, p = P.zones_for("LU", "S")
    n = z.count
    arrs = [list(v) for v in (z.nx, z.ny, z.nz, z.west, z.east, z.south, z.north)]

    def create(arrs, omega):
        c = [(ctypes.c_int32 * n)(*v) for v in arrs]
        q = native.Problem(native.LU, n, 4, 4, 24, 24, 6, *c, 0.5, omega)
        h = ctypes.c_void_p()
        return lib.npbmz_create(ctypes.byref(q), 0, ctypes.byref(h))

    assert create(arrs, 2.0) == -1 and b"omega" in lib.npbmz_last_error(None)
    odd = [list(v) for v in arrs]
    odd[2][3] += 1
    assert create(odd, 1.2) == -1
    assert b"same size" in lib.npbmz_last_error(None) or b"neighbours" in lib.npbmz_last_error(None)


def test_missing_library_fails_loudly(tmp_path):
    native_mod = native
    with pytest.raises(native_mod.NpbmzError, match="not built"):
        old = native_mod._lib
        native_mod._lib = None
        try:
            native_mod.load(str(tmp_path / "absent.so"))
        finally:
            native_mod._lib = old


def test_header_is_plain_c_and_the_c_harness_compiles(tmp_path):
    """include/npbmz.h is usable from C11 (the language a cgo, JNI or Julia
    ccall binding sees): the C caller compiles against it with warnings as
    errors."""
    import subprocess
    src = os.path.join(ROOT, "npbapps.jl_amd", "harness", "npbmz_run.c")
    r = subprocess.run(["gcc", "-std=c11", "-Wall", "-Wextra", "-Werror", "-pedantic",
                        "-Wno-vla", "-I", os.path.join(ROOT, "include"), "-c", src, "-o",
                        str(tmp_path / "h.o")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_harness_verification_tables_match_verify_py(tmp_path):
    """The C harness's class tables (harness/verify_tables.h, committed) are
    exactly what the build generates from verify.py."""
    import importlib.util
    spec = importlib.util.spec_from_file_location(
        "_npbmz_build", os.path.join(ROOT, "npbapps.jl_amd", "build.py"))
    build = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(build)
    out = build.write_verify_tables(str(tmp_path / "t.h"))
    committed = os.path.join(ROOT, "npbapps.jl_amd", "harness", "verify_tables.h")
    assert open(out).read() == open(committed).read()


@pytest.mark.parametrize("bench", ["bt", "sp", "lu"])
def test_c_harness_zone_split_matches_the_python_one(bench):
    """The C caller's zone split (src/BT/zone_setup.jl, LU rounding from
    src/LU/zone_setup.jl) against npbapps_amd.zones for classes S, W, A, B;
    its zones mode prints the table without touching a device."""
    import json
    import subprocess
    from npbapps_amd import build
    if not os.path.exists(native.LIB_PATH):
        pytest.skip("libnpbmz is not built")
    exe = build.build_harness()
    for cls in "SWAB":
        out = subprocess.run([exe, bench, cls, "zones"], capture_output=True, text=True,
                             check=True).stdout
        got = json.loads(out)
        z, _ = P.zones_for(bench.upper(), cls)
        for key in ("nx", "ny", "nz", "west", "east", "south", "north"):
            assert got[key] == list(getattr(z, key)), (bench, cls, key)
