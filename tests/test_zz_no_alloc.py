"""npbmz_norms_noalloc: the norms on caller-owned device fields, with the
kernels' partial sums in the library's static scratch instead of context
memory.  Checked through the C ABI against npbmz_norms on the same state,
bit for bit, on two calls in a row (the second on a smaller problem, so the
scratch holds the first call's larger contents), for BT and LU; and a
problem beyond the scratch fails cleanly."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class _Hip:
    """Device buffers from the HIP runtime libnpbmz itself links (the test is
    the caller that owns them; the library allocates nothing)."""

    def __init__(self):
        self.h = ctypes.CDLL("libamdhip64.so.7")
        self.h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                     ctypes.c_int]
        self.h.hipFree.argtypes = [ctypes.c_void_p]
        self.h.hipDeviceSynchronize.argtypes = []
        self.live = []

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = ctypes.c_void_p()
        assert self.h.hipMalloc(ctypes.byref(p), max(a.nbytes, 8)) == 0
        self.live.append(p)
        assert self.h.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return p.value

    def get(self, ptr, n):
        out = np.empty(n)
        assert self.h.hipMemcpy(out.ctypes.data, ctypes.c_void_p(ptr), out.nbytes, 2) == 0
        return out

    def free(self):
        assert self.h.hipDeviceSynchronize() == 0
        for p in self.live:
            self.h.hipFree(p)
        self.live = []


def _state(gpu, bench, cls, steps):
    zones, p = gpu.zones_for(bench, cls)
    ctx = gpu.Context(bench, zones, p.dt, omega=1.2, npb_exchange=bench == "LU")
    ctx.init_exact()
    ctx.step(steps, "graph")
    return ctx


def _check_bt(gpu, hip, cls):
    ctx = _state(gpu, "BT", cls, 5)
    u, f = ctx.get_all(0), ctx.get_all(1)
    ref = ctx.norms()  # leaves compute_rhs of u in the context's rhs
    rhs_ref = ctx.get_all(2)
    du, df, dr = hip.put(u), hip.put(f), hip.put(np.zeros(u.size))
    got = ctx.norms_noalloc(du, df, dr)
    assert np.array_equal(got[0], ref[0]), (cls, got[0], ref[0])
    assert np.array_equal(got[1], ref[1]), (cls, got[1], ref[1])
    assert got[2] == 0.0
    assert np.array_equal(hip.get(dr, u.size), rhs_ref)
    # the caller's u and forcing are inputs only
    assert np.array_equal(hip.get(du, u.size), u) and np.array_equal(hip.get(df, f.size), f)
    ctx.close()


def test_noalloc_norms_match_npbmz_norms(gpu):
    hip = _Hip()
    # two calls in a row, the second on a smaller problem
    _check_bt(gpu, hip, "W")
    _check_bt(gpu, hip, "S")
    ctx = _state(gpu, "LU", "S", 3)
    u, rsd = ctx.get_all(0), ctx.get_all(2)
    ref = ctx.norms()
    du, dr = hip.put(u), hip.put(rsd)
    got = ctx.norms_noalloc(du, None, dr)
    assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])
    assert got[2] == ref[2]
    assert np.array_equal(hip.get(dr, rsd.size), rsd)
    ctx.close()
    hip.free()


def test_noalloc_norms_reject_problems_beyond_the_scratch(gpu):
    hip = _Hip()
    zones, p = gpu.zones_for("BT", "C")
    ctx = gpu.Context("BT", zones, p.dt)
    d = hip.put(np.zeros(8))
    with pytest.raises(gpu.NpbmzError, match="too large for the no-allocation scratch"):
        ctx.norms_noalloc(d, d, d)
    with pytest.raises(gpu.NpbmzError, match="null argument"):
        ctx.norms_noalloc(None, d, d)
    ctx.close()
    # a problem that fits still works after the rejection
    _check_bt(gpu, hip, "S")
    hip.free()
