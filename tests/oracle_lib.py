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
    for b in ("bt", "sp"):
        f = lambda n: getattr(L, f"oracle_{b}_{n}")
        f("create").restype = vp
        f("create").argtypes = [i, i, i, i, i, d, d]
        f("destroy").argtypes = [vp]
        f("set_threads").argtypes = [vp, i]
        f("nzones").argtypes = [vp]
        f("phase").argtypes = [vp, i]
        f("step").argtypes = [vp, i]
        f("norms").argtypes = [vp, dp, dp]
        f("get").argtypes = [vp, i, i, dp]
        f("set").argtypes = [vp, i, i, dp]
        for n in ("destroy", "set_threads", "phase", "step", "norms", "get", "set"):
            f(n).restype = None
    L.oracle_bt_zone.argtypes = [vp, i, ctypes.POINTER(i)]
    L.oracle_bt_zone.restype = None
    L.oracle_bt_consts.argtypes = [vp, dp, i]
    L.oracle_binvcrhs.argtypes = [dp, dp, dp, i]
    L.oracle_binvcrhs.restype = None
    L.oracle_bt_line_blocks.argtypes = [vp, i, i, i, i, dp, dp, dp]
    L.oracle_bt_line_blocks.restype = None
    L.oracle_sp_line_lhs.argtypes = [vp, i, i, i, i, dp]
    L.oracle_sp_line_lhs.restype = None
    L.oracle_lu_create.restype = vp
    L.oracle_lu_create.argtypes = [i, i, i, i, i, d, d, i]
    for n, a in (("destroy", [vp]), ("set_threads", [vp, i]), ("phase", [vp, i]),
                 ("step", [vp, i]), ("norms", [vp, dp, dp, dp, dp]), ("get", [vp, i, i, dp]),
                 ("set", [vp, i, i, dp]), ("zone", [vp, i, ctypes.POINTER(i)]),
                 ("point_blocks", [vp, i, i, i, i, i, dp]), ("set_exact", [vp, i])):
        getattr(L, f"oracle_lu_{n}").argtypes = a
        getattr(L, f"oracle_lu_{n}").restype = None
    L.oracle_lu_nzones.argtypes = [vp]
    for b in ("bt", "sp", "lu"):
        getattr(L, f"oracle_{b}_dual").argtypes = [vp, i, i, dp, dp, dp, dp]
        getattr(L, f"oracle_{b}_dual").restype = None
    _lib = L
    return L


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class _OracleMZ:
    """A CPU restatement of one multi-zone benchmark (all zones on one rank)."""

    U, FORCING, RHS = 0, 1, 2
    PREFIX = None

    def __init__(self, x_zones, y_zones, gx, gy, gz, ratio, dt, threads=None):
        self.L = load()
        self._f = lambda n: getattr(self.L, f"oracle_{self.PREFIX}_{n}")
        self.h = ctypes.c_void_p(self._f("create")(x_zones, y_zones, gx, gy, gz, float(ratio),
                                                   float(dt)))
        self._f("set_threads")(self.h, threads or min(8, os.cpu_count() or 1))
        self.dt = dt
        # zone geometry from the shared zone split (same formula for BT and SP)
        import npbapps_amd as P
        zs = P.zone_setup(x_zones, y_zones, gx, gy, gz, ratio)
        self.zones = [(zs.nx[z], zs.ny[z], zs.nz[z], zs.west[z], zs.east[z], zs.south[z],
                       zs.north[z]) for z in range(zs.count)]

    @classmethod
    def for_class(cls, params, **kw):
        p = params
        return cls(p.x_zones, p.y_zones, p.gx, p.gy, p.gz, p.ratio, p.dt, **kw)

    def close(self):
        if self.h:
            self._f("destroy")(self.h)
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
        self._f("get")(self.h, z, field, _dp(out))
        return out

    def set(self, z, field, arr):
        arr = np.ascontiguousarray(arr, dtype=np.float64)
        self._f("set")(self.h, z, field, _dp(arr))

    def step(self, n):
        self._f("step")(self.h, n)

    def phase(self, p):
        self._f("phase")(self.h, p)

    def norms(self):
        xcr = np.zeros(5)
        xce = np.zeros(5)
        self._f("norms")(self.h, _dp(xcr), _dp(xce))
        return xcr, xce


    def dual_gradient(self, nsteps, w, vu, vf):
        """Forward-mode dual numbers through nsteps steps from the current
        state: returns (loss, d) with loss = sum(w*u_N**2) and d[i] the
        derivative of the loss along (vu[i] on u0, vf[i] on the forcing).
        Arrays in the whole-run layout (every zone's [k][j][i][m], zone after
        zone); the state is left unchanged."""
        vu = np.ascontiguousarray(np.atleast_2d(vu), dtype=np.float64)
        vf = np.ascontiguousarray(np.atleast_2d(vf), dtype=np.float64)
        w = np.ascontiguousarray(w, dtype=np.float64)
        nd = vu.shape[0]
        out = np.zeros(1 + nd)
        self._f("dual")(self.h, int(nsteps), nd, _dp(w), _dp(vu), _dp(vf), _dp(out))
        return out[0], out[1:]


class OracleBT(_OracleMZ):
    PREFIX = "bt"

    def line_blocks(self, z, d, a, b):
        nl = self.dims(z)[d]
        A = np.zeros((nl, 5, 5))
        B = np.zeros((nl, 5, 5))
        C = np.zeros((nl, 5, 5))
        self.L.oracle_bt_line_blocks(self.h, z, d, a, b, _dp(A), _dp(B), _dp(C))
        return A, B, C


class OracleSP(_OracleMZ):
    PREFIX = "sp"

    def line_lhs(self, z, d, a, b):
        nl = self.dims(z)[d]
        out = np.zeros((nl, 15))
        self.L.oracle_sp_line_lhs(self.h, z, d, a, b, _dp(out))
        return out


class OracleLU:
    """The CPU restatement of LU-MZ (oracle/lu_oracle.cpp).  flags: bit 0
    exchanges planes k = 2..nz-1 (NPB-MZ) instead of the reference's
    k = 1..nz-2; bit 1 uses NPB's signs in buts' back substitution."""

    U, FORCING, RHS = 0, 1, 2
    NPB_EXCH, NPB_BUTS = 1, 2

    def __init__(self, x_zones, y_zones, gx, gy, gz, dt, omega=1.2, flags=0, threads=None):
        self.L = load()
        self.h = ctypes.c_void_p(self.L.oracle_lu_create(x_zones, y_zones, gx, gy, gz, float(dt),
                                                         float(omega), int(flags)))
        self.L.oracle_lu_set_threads(self.h, threads or min(8, os.cpu_count() or 1))
        self.dt, self.omega, self.flags = dt, omega, flags
        self.zones = []
        for z in range(self.L.oracle_lu_nzones(self.h)):
            out = (ctypes.c_int * 7)()
            self.L.oracle_lu_zone(self.h, z, out)
            self.zones.append(tuple(out))

    @classmethod
    def for_class(cls, params, **kw):
        p = params
        return cls(p.x_zones, p.y_zones, p.gx, p.gy, p.gz, p.dt, **kw)

    def close(self):
        if self.h:
            self.L.oracle_lu_destroy(self.h)
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
        self.L.oracle_lu_get(self.h, z, field, _dp(out))
        return out

    def set(self, z, field, arr):
        arr = np.ascontiguousarray(arr, dtype=np.float64)
        self.L.oracle_lu_set(self.h, z, field, _dp(arr))

    def step(self, n):
        self.L.oracle_lu_step(self.h, n)

    def phase(self, p):
        self.L.oracle_lu_phase(self.h, p)

    def norms(self, per_zone=False):
        xcr, xce, xci = np.zeros(5), np.zeros(5), np.zeros(1)
        zo = np.zeros((len(self.zones), 11))
        self.L.oracle_lu_norms(self.h, _dp(xcr), _dp(xce), _dp(xci), _dp(zo))
        if per_zone:
            return xcr, xce, xci[0], zo
        return xcr, xce, xci[0]

    def point_blocks(self, z, upper, i, j, k):
        """d and the k-, j-, i-neighbour blocks of 1-based point (i, j, k)."""
        out = np.zeros((4, 5, 5))
        self.L.oracle_lu_point_blocks(self.h, z, int(upper), i, j, k, _dp(out))
        return out

    def set_exact(self, z):
        self.L.oracle_lu_set_exact(self.h, z)

    def dual_gradient(self, nsteps, w, vu, vf):
        """Forward-mode dual numbers through nsteps iterations from the
        current u and frct (rsd recomputed as rhs(u) first): returns (loss, d)
        with loss = sum(w*u_N**2) and d[i] its derivative along (vu[i] on u,
        vf[i] on frct), every point.  Arrays in the whole-run layout; the state
        is left unchanged."""
        vu = np.ascontiguousarray(np.atleast_2d(vu), dtype=np.float64)
        vf = np.ascontiguousarray(np.atleast_2d(vf), dtype=np.float64)
        w = np.ascontiguousarray(w, dtype=np.float64)
        nd = vu.shape[0]
        out = np.zeros(1 + nd)
        self.L.oracle_lu_dual(self.h, int(nsteps), nd, _dp(w), _dp(vu), _dp(vf), _dp(out))
        return out[0], out[1:]


def oracle_for(bench):
    return {"BT": OracleBT, "SP": OracleSP, "LU": OracleLU}[bench]


def binvcrhs(B, C, r, with_c=True):
    B = np.array(B, dtype=np.float64, order="C")
    C = np.array(C, dtype=np.float64, order="C")
    r = np.array(r, dtype=np.float64)
    load().oracle_binvcrhs(_dp(B), _dp(C), _dp(r), 1 if with_c else 0)
    return B, C, r
