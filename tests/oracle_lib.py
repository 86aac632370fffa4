"""ctypes binding of the CPU oracle (oracle/liboracle_npb.so).

Test infrastructure only: the oracle is the checker the GPU path is compared
with (and bench.py's reported CPU baseline), never part of the product.
"""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "liboracle_npb.so")

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(ORACLE_LIB):
        build()
    L = ctypes.CDLL(ORACLE_LIB)
    vp, i, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    dp = ctypes.POINTER(d)
    L.oracle_bt_create.restype = vp
    L.oracle_bt_create.argtypes = [i, i, i, i, i, d, d]
    L.oracle_bt_destroy.argtypes = [vp]
    L.oracle_bt_destroy.restype = None
    L.oracle_bt_set_threads.argtypes = [vp, i]
    L.oracle_bt_nzones.argtypes = [vp]
    L.oracle_bt_zone.argtypes = [vp, i, ctypes.POINTER(i)]
    L.oracle_bt_phase.argtypes = [vp, i]
    L.oracle_bt_step.argtypes = [vp, i]
    L.oracle_bt_norms.argtypes = [vp, dp, dp]
    L.oracle_bt_get.argtypes = [vp, i, i, dp]
    L.oracle_bt_set.argtypes = [vp, i, i, dp]
    L.oracle_bt_consts.argtypes = [vp, dp, i]
    L.oracle_binvcrhs.argtypes = [dp, dp, dp, i]
    L.oracle_binvcrhs.restype = None
    L.oracle_bt_line_blocks.argtypes = [vp, i, i, i, i, dp, dp, dp]
    L.oracle_bt_line_blocks.restype = None
    for f in ("oracle_bt_set_threads", "oracle_bt_phase", "oracle_bt_step", "oracle_bt_norms",
              "oracle_bt_get", "oracle_bt_set", "oracle_bt_zone"):
        getattr(L, f).restype = None
    _lib = L
    return L


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class OracleBT:
    """The BT-MZ CPU restatement for one problem (all zones on one rank)."""

    U, FORCING, RHS = 0, 1, 2

    def __init__(self, x_zones, y_zones, gx, gy, gz, ratio, dt, threads=None):
        self.L = load()
        self.h = ctypes.c_void_p(self.L.oracle_bt_create(x_zones, y_zones, gx, gy, gz,
                                                         float(ratio), float(dt)))
        self.L.oracle_bt_set_threads(self.h, threads or min(8, os.cpu_count() or 1))
        self.dt = dt
        n = self.L.oracle_bt_nzones(self.h)
        z7 = (ctypes.c_int * 7)()
        self.zones = []
        for z in range(n):
            self.L.oracle_bt_zone(self.h, z, z7)
            self.zones.append(tuple(z7))

    @classmethod
    def for_class(cls, params, **kw):
        p = params
        return cls(p.x_zones, p.y_zones, p.gx, p.gy, p.gz, p.ratio, p.dt, **kw)

    def close(self):
        if self.h:
            self.L.oracle_bt_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def dims(self, z):
        return self.zones[z][:3]

    def get(self, z, field):
        nx, ny, nz = self.dims(z)
        out = np.zeros((nz, ny, nx, 5))
        self.L.oracle_bt_get(self.h, z, field, _dp(out))
        return out

    def set(self, z, field, arr):
        arr = np.ascontiguousarray(arr, dtype=np.float64)
        self.L.oracle_bt_set(self.h, z, field, _dp(arr))

    def step(self, n):
        self.L.oracle_bt_step(self.h, n)

    def phase(self, p):
        self.L.oracle_bt_phase(self.h, p)

    def norms(self):
        xcr = np.zeros(5)
        xce = np.zeros(5)
        self.L.oracle_bt_norms(self.h, _dp(xcr), _dp(xce))
        return xcr, xce

    def line_blocks(self, z, d, a, b):
        nl = self.dims(z)[d]
        A = np.zeros((nl, 5, 5))
        B = np.zeros((nl, 5, 5))
        C = np.zeros((nl, 5, 5))
        self.L.oracle_bt_line_blocks(self.h, z, d, a, b, _dp(A), _dp(B), _dp(C))
        return A, B, C


def binvcrhs(B, C, r, with_c=True):
    B = np.array(B, dtype=np.float64, order="C")
    C = np.array(C, dtype=np.float64, order="C")
    r = np.array(r, dtype=np.float64)
    load().oracle_binvcrhs(_dp(B), _dp(C), _dp(r), 1 if with_c else 0)
    return B, C, r
