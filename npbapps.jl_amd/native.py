"""ctypes binding of libnpbmz (the C ABI declared in include/npbmz.h).

The library is loaded from this package directory only.  There is no CPU
fallback: every entry point raises NpbmzError when the library is missing or
a call fails, so nothing silently runs anywhere but on the GPU.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnpbmz.so")

BT, SP, LU = 0, 1, 2
BENCH = {"BT": BT, "SP": SP, "LU": LU}

FLAG_STRICT_FP = 1
FLAG_LU_NPB_EXCH = 2
FIELD_U, FIELD_FORCING, FIELD_RHS = 0, 1, 2
ORDER_MIJK, ORDER_IJKM = 0, 1
MODE_PERZONE, MODE_BATCHED, MODE_GRAPH = 0, 1, 2
MODES = {"perzone": MODE_PERZONE, "batched": MODE_BATCHED, "graph": MODE_GRAPH}
PHASE_EXCH, PHASE_RHS, PHASE_XSOLVE, PHASE_YSOLVE, PHASE_ZSOLVE, PHASE_ADD = range(6)
# LU's phases (exch and rhs as above)
PHASE_SSOR_LOWER, PHASE_SSOR_UPPER, PHASE_SSOR_UPDATE = 2, 3, 4
SCHED_LU_SWEEP, SCHED_LU_THREADS, SCHED_SOLVE_GRID, SCHED_SOLVE_PAIR, SCHED_SOLVE_TWIST = 0, 1, 2, 3, 4
SCHED_LU_FORWARD = 5
SCHED_SOLVE_ROWS = 6
SCHED_LU_TILES, SCHED_LU_TILE_BATCH, SCHED_LU_SPIN = 7, 8, 9
SCHED_LU_HANDOFF = 10
SCHED_RHS_EXCH = 11

# Every symbol include/npbmz.h declares.
EXPORTS = (
    "npbmz_create", "npbmz_destroy", "npbmz_last_error", "npbmz_put_field",
    "npbmz_get_field", "npbmz_init_exact", "npbmz_step", "npbmz_run_phase", "npbmz_phase_times",
    "npbmz_norms", "npbmz_zone_norms", "npbmz_tune", "npbmz_describe", "npbmz_size",
    "npbmz_probe_copy", "npbmz_set_schedule", "npbmz_adjoint", "npbmz_loss",
    "npbmz_norms_noalloc",
)


class NpbmzError(RuntimeError):
    """A libnpbmz call failed (negative return code) or the library is absent."""


class Problem(ctypes.Structure):
    _fields_ = [
        ("bench", ctypes.c_int32), ("nzones", ctypes.c_int32),
        ("x_zones", ctypes.c_int32), ("y_zones", ctypes.c_int32),
        ("gx_size", ctypes.c_int32), ("gy_size", ctypes.c_int32), ("gz_size", ctypes.c_int32),
        ("nx", ctypes.POINTER(ctypes.c_int32)), ("ny", ctypes.POINTER(ctypes.c_int32)),
        ("nz", ctypes.POINTER(ctypes.c_int32)),
        ("iz_west", ctypes.POINTER(ctypes.c_int32)), ("iz_east", ctypes.POINTER(ctypes.c_int32)),
        ("iz_south", ctypes.POINTER(ctypes.c_int32)), ("iz_north", ctypes.POINTER(ctypes.c_int32)),
        ("dt", ctypes.c_double), ("omega", ctypes.c_double),
    ]


_lib = None


def load(path=LIB_PATH):
    """Load libnpbmz and declare its prototypes; raise NpbmzError if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise NpbmzError(f"{path} is not built; run __graft_entry__.build() first")
    lib = ctypes.CDLL(path)
    P = ctypes.c_void_p
    i32, i64, dbl = ctypes.c_int32, ctypes.c_int64, ctypes.c_double
    dp = ctypes.POINTER(dbl)
    i64p = ctypes.POINTER(i64)
    lib.npbmz_create.argtypes = [ctypes.POINTER(Problem), i32, ctypes.POINTER(P)]
    lib.npbmz_destroy.argtypes = [P]
    lib.npbmz_destroy.restype = None
    lib.npbmz_last_error.argtypes = [P]
    lib.npbmz_last_error.restype = ctypes.c_char_p
    for f in ("npbmz_put_field", "npbmz_get_field"):
        getattr(lib, f).argtypes = [P, i32, i32, dp, i64p, i64p, i32]
    lib.npbmz_init_exact.argtypes = [P]
    lib.npbmz_step.argtypes = [P, i32, i32, dp]
    lib.npbmz_run_phase.argtypes = [P, i32, i32]
    lib.npbmz_phase_times.argtypes = [P, i32, dp]
    lib.npbmz_norms.argtypes = [P, dp, dp, dp]
    lib.npbmz_zone_norms.argtypes = [P, dp, dp]
    lib.npbmz_tune.argtypes = [P, i32]
    lib.npbmz_describe.argtypes = [P, ctypes.c_char_p, i32]
    lib.npbmz_size.argtypes = [P, ctypes.POINTER(i32), i64p]
    lib.npbmz_probe_copy.argtypes = [P, i64, i32, dp]
    lib.npbmz_set_schedule.argtypes = [P, i32, i32]
    lib.npbmz_adjoint.argtypes = [P, i32, i32, dp, dp, dp, dp, dp, dp]
    lib.npbmz_loss.argtypes = [P, dp, dp]
    lib.npbmz_norms_noalloc.argtypes = [P, P, P, P, dp, dp, dp]
    for f in EXPORTS:
        if f not in ("npbmz_destroy", "npbmz_last_error"):
            getattr(lib, f).restype = ctypes.c_int
    _lib = lib
    return lib


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _i32(a):
    return (ctypes.c_int32 * len(a))(*[int(x) for x in a])


class Context:
    """One libnpbmz context: a multi-zone problem resident in HBM."""

    def __init__(self, bench, zones, dt, omega=0.0, strict=False, npb_exchange=False):
        """bench "BT", "SP" or "LU"; zones from zones.zone_setup; strict selects
        the kernels built without FMA contraction; npb_exchange (LU only)
        exchanges planes k = 2..nz-1 as NPB-MZ does instead of the
        reference's k = 1..nz-2."""
        lib = load()
        self._lib = lib
        self.bench = BENCH[bench] if isinstance(bench, str) else int(bench)
        self.zones = zones
        self.dt = float(dt)
        n = zones.count
        self._keep = [_i32(zones.nx), _i32(zones.ny), _i32(zones.nz), _i32(zones.west),
                      _i32(zones.east), _i32(zones.south), _i32(zones.north)]
        p = Problem(self.bench, n, zones.x_zones, zones.y_zones, zones.gx, zones.gy, zones.gz,
                    *self._keep, float(dt), float(omega))
        h = ctypes.c_void_p()
        flags = (FLAG_STRICT_FP if strict else 0) | (FLAG_LU_NPB_EXCH if npb_exchange else 0)
        rc = lib.npbmz_create(ctypes.byref(p), flags, ctypes.byref(h))
        if rc != 0:
            raise NpbmzError(f"npbmz_create -> {rc}: {lib.npbmz_last_error(None).decode()}")
        self._h = h

    def _check(self, rc, what):
        if rc != 0:
            msg = self._lib.npbmz_last_error(self._h).decode()
            raise NpbmzError(f"{what} -> {rc}: {msg}")

    def close(self):
        if getattr(self, "_h", None):
            self._lib.npbmz_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- fields ------------------------------------------------------------
    @staticmethod
    def _check_extent(host, lo, hi):
        n = 1
        for a, b in zip(lo, hi):
            n *= int(b) - int(a) + 1
        if host.size != n:
            raise NpbmzError(f"host array has {host.size} elements, bounds describe {n}")

    def put_field(self, zone, field, host, lo, hi, order=ORDER_MIJK):
        host = np.ascontiguousarray(host, dtype=np.float64)
        self._check_extent(host, lo, hi)
        lo_a = (ctypes.c_int64 * 4)(*lo)
        hi_a = (ctypes.c_int64 * 4)(*hi)
        self._check(self._lib.npbmz_put_field(self._h, zone, field, _dptr(host), lo_a, hi_a,
                                              order), "npbmz_put_field")

    def get_field(self, zone, field, host, lo, hi, order=ORDER_MIJK):
        assert host.dtype == np.float64 and host.flags["C_CONTIGUOUS"]
        self._check_extent(host, lo, hi)
        lo_a = (ctypes.c_int64 * 4)(*lo)
        hi_a = (ctypes.c_int64 * 4)(*hi)
        self._check(self._lib.npbmz_get_field(self._h, zone, field, _dptr(host), lo_a, hi_a,
                                              order), "npbmz_get_field")
        return host

    def get_dense(self, zone, field):
        """Zone field as a (nz, ny, nx, 5) array over the zone's grid points
        (the reference's [m,i,j,k] column-major order)."""
        nx, ny, nz = self.zones.nx[zone], self.zones.ny[zone], self.zones.nz[zone]
        org = 1 if self.bench == LU else 0
        out = np.zeros((nz, ny, nx, 5))
        return self.get_field(zone, field, out, (1, org, org, org),
                              (5, nx - 1 + org, ny - 1 + org, nz - 1 + org))

    def put_dense(self, zone, field, arr):
        nx, ny, nz = self.zones.nx[zone], self.zones.ny[zone], self.zones.nz[zone]
        org = 1 if self.bench == LU else 0
        arr = np.ascontiguousarray(arr, dtype=np.float64).reshape(nz, ny, nx, 5)
        self.put_field(zone, field, arr, (1, org, org, org),
                       (5, nx - 1 + org, ny - 1 + org, nz - 1 + org))

    @property
    def size(self):
        """Values in a whole-run field: every zone's nx*ny*nz*5, zone after zone."""
        return 5 * sum(int(a) * int(b) * int(c)
                       for a, b, c in zip(self.zones.nx, self.zones.ny, self.zones.nz))

    def get_all(self, field):
        """Every zone's field, zone after zone, each (nz, ny, nx, 5) flattened."""
        return np.concatenate([self.get_dense(z, field).ravel() for z in range(self.zones.count)])

    def put_all(self, field, flat):
        flat = np.ascontiguousarray(flat, dtype=np.float64)
        off = 0
        for z in range(self.zones.count):
            n = 5 * self.zones.nx[z] * self.zones.ny[z] * self.zones.nz[z]
            self.put_dense(z, field, flat[off:off + n])
            off += n

    # -- compute -------------------------------------------------------------
    def init_exact(self):
        self._check(self._lib.npbmz_init_exact(self._h), "npbmz_init_exact")

    def step(self, nsteps, mode="graph"):
        """Run nsteps timesteps; returns the HIP-event time in seconds."""
        ms = ctypes.c_double(0.0)
        m = MODES[mode] if isinstance(mode, str) else int(mode)
        self._check(self._lib.npbmz_step(self._h, int(nsteps), m, ctypes.byref(ms)), "npbmz_step")
        return ms.value * 1e-3

    def run_phase(self, phase, mode="batched"):
        m = MODES[mode] if isinstance(mode, str) else int(mode)
        self._check(self._lib.npbmz_run_phase(self._h, int(phase), m), "npbmz_run_phase")

    def phase_times(self, nsteps):
        """Seconds spent in each phase summed over nsteps batched timesteps,
        from HIP events on the context's stream, indexed by PHASE_*: BT/SP
        exch, rhs, x, y, z(+add), add; LU exch, rhs, lower, upper, update."""
        ms = np.zeros(6)
        self._check(self._lib.npbmz_phase_times(self._h, int(nsteps), _dptr(ms)),
                    "npbmz_phase_times")
        return ms * 1e-3

    def norms(self):
        xcr = np.zeros(5)
        xce = np.zeros(5)
        xci = np.zeros(1)
        self._check(self._lib.npbmz_norms(self._h, _dptr(xcr), _dptr(xce), _dptr(xci)),
                    "npbmz_norms")
        return xcr, xce, float(xci[0])

    def norms_noalloc(self, u_ptr, forcing_ptr, rhs_ptr):
        """npbmz_norms on caller-owned device fields (device addresses, the
        whole-run layout), with no allocation in the library."""
        xcr = np.zeros(5)
        xce = np.zeros(5)
        xci = np.zeros(1)
        self._check(self._lib.npbmz_norms_noalloc(self._h, u_ptr, forcing_ptr, rhs_ptr,
                                                  _dptr(xcr), _dptr(xce), _dptr(xci)),
                    "npbmz_norms_noalloc")
        return xcr, xce, float(xci[0])

    def zone_norms(self):
        n = self.zones.count
        r = np.zeros((n, 5))
        e = np.zeros((n, 5))
        self._check(self._lib.npbmz_zone_norms(self._h, _dptr(r), _dptr(e)), "npbmz_zone_norms")
        return r, e

    def tune(self, budget_ms=2000):
        self._check(self._lib.npbmz_tune(self._h, int(budget_ms)), "npbmz_tune")

    def set_schedule(self, which, value):
        self._check(self._lib.npbmz_set_schedule(self._h, int(which), int(value)),
                    "npbmz_set_schedule")

    def loss(self, w):
        """sum(w * u**2) over every grid point of the current u (get_all layout)."""
        w = np.ascontiguousarray(w, dtype=np.float64)
        assert w.size == self.size
        out = ctypes.c_double(0.0)
        self._check(self._lib.npbmz_loss(self._h, _dptr(w), ctypes.byref(out)), "npbmz_loss")
        return out.value

    def adjoint(self, nsteps, ckpt_every, w):
        """loss = sum(w*u_N**2) after nsteps steps from the current state, and
        its gradients w.r.t. the initial u and the forcing (LU: frct, whose
        rsd = rhs(u0) is recomputed first), both in get_all layout, with u
        kept every ckpt_every steps.  Returns a dict
        with loss, grad_u0, grad_forcing and the forward / reverse times in
        seconds; u is left at its initial value."""
        w = np.ascontiguousarray(w, dtype=np.float64)
        assert w.size == self.size
        gu = np.zeros(self.size)
        gf = np.zeros(self.size)
        loss, tf, ta = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        self._check(self._lib.npbmz_adjoint(self._h, int(nsteps), int(ckpt_every), _dptr(w),
                                            _dptr(gu), _dptr(gf), ctypes.byref(loss),
                                            ctypes.byref(tf), ctypes.byref(ta)), "npbmz_adjoint")
        return {"loss": loss.value, "grad_u0": gu, "grad_forcing": gf,
                "fwd_s": tf.value * 1e-3, "adj_s": ta.value * 1e-3}

    def probe_copy(self, nbytes, reps=20):
        """Seconds per launch of a plain nbytes device-to-device copy kernel."""
        ms = ctypes.c_double(0.0)
        self._check(self._lib.npbmz_probe_copy(self._h, int(nbytes), int(reps), ctypes.byref(ms)),
                    "npbmz_probe_copy")
        return ms.value * 1e-3

    def describe(self):
        buf = ctypes.create_string_buffer(512)
        self._check(self._lib.npbmz_describe(self._h, buf, 512), "npbmz_describe")
        return buf.value.decode()
