"""SP-MZ on the GPU against the CPU oracle, through the C ABI: strict-FP
kernels bit for bit; FMA kernels within 1e-12 per step, 1e-10 at the end of a
run, and verified against the reference's tables."""
import numpy as np
import pytest

from oracle_lib import OracleSP

pytestmark = pytest.mark.gpu


def _ctx(P, cls, strict):
    zones, p = P.zones_for("SP", cls)
    return P.Context("SP", zones, p.dt, strict=strict), zones, p


def _interior(a):
    return a[1:-1, 1:-1, 1:-1]


@pytest.mark.parametrize("cls", ["S", "W"])
def test_init_exact_is_bit_exact(gpu, cls):
    ctx, zones, p = _ctx(gpu, cls, True)
    o = OracleSP.for_class(p)
    ctx.init_exact()
    for z in range(zones.count):
        assert np.array_equal(ctx.get_dense(z, 0), o.get(z, 0)), z
        assert np.array_equal(ctx.get_dense(z, 1), o.get(z, 1)), z
    ctx.close()


@pytest.mark.parametrize("cls", ["S", "W"])
@pytest.mark.parametrize("mode", ["perzone", "batched"])
def test_each_phase_is_bit_exact(gpu, cls, mode):
    """exch_qbc (with SP's mirrored y faces), compute_rhs+txinvr,
    x_solve+ninvr, y_solve+pinvr, z_solve+tzetar and add."""
    ctx, zones, p = _ctx(gpu, cls, True)
    o = OracleSP.for_class(p)
    o.step(3)
    for z in range(zones.count):
        ctx.put_dense(z, 0, o.get(z, 0))
        ctx.put_dense(z, 1, o.get(z, 1))
    for ph in range(6):
        o.phase(ph)
        ctx.run_phase(ph, mode)
        for z in range(zones.count):
            if ph in (0, 5):
                assert np.array_equal(ctx.get_dense(z, 0), o.get(z, 0)), (ph, z)
            else:
                g, r = _interior(ctx.get_dense(z, 2)), _interior(o.get(z, 2))
                assert np.array_equal(g, r), (ph, z, np.abs(g - r).max())
    ctx.close()


def test_full_run_strict_is_bit_exact_class_S(gpu):
    ctx, zones, p = _ctx(gpu, "S", True)
    o = OracleSP.for_class(p)
    ctx.init_exact()
    ctx.step(p.niter, "graph")
    o.step(p.niter)
    for z in range(zones.count):
        assert np.array_equal(ctx.get_dense(z, 0), o.get(z, 0)), z
    ctx.close()


@pytest.mark.parametrize("cls", ["S", "W"])
def test_fma_kernels_track_oracle_per_step(gpu, cls):
    ctx, zones, p = _ctx(gpu, cls, False)
    o = OracleSP.for_class(p)
    ctx.init_exact()
    for s in range(10):
        ctx.step(1, "batched")
        o.step(1)
        gxcr, gxce, _ = ctx.norms()
        oxcr, oxce = o.norms()
        assert np.abs(gxcr / oxcr - 1).max() < 1e-12, s
        assert np.abs(gxce / oxce - 1).max() < 1e-12, s
    ctx.close()


@pytest.mark.parametrize("cls", ["S", "W", "A"])
def test_fma_run_verifies_and_tracks_oracle(gpu, cls):
    """SP-MZ class A (4x4 zones of 32x32x16, 400 steps at dt=0.0015) is the
    per-zone-launch SP configuration of the benchmark."""
    r = gpu.go("SP", cls, mode="graph", verbose=False)
    assert r["verified"], r["rel"]
    zones, p = gpu.zones_for("SP", cls)
    o = OracleSP.for_class(p)
    o.step(p.niter)
    oxcr, oxce = o.norms()
    assert np.abs(np.array(r["xcr"]) / oxcr - 1).max() < 1e-10
    assert np.abs(np.array(r["xce"]) / oxce - 1).max() < 1e-10


def test_perzone_and_graph_agree_class_A(gpu):
    res = []
    for mode in ("perzone", "graph"):
        ctx, zones, p = _ctx(gpu, "A", False)
        ctx.init_exact()
        ctx.step(5, mode)
        res.append([ctx.get_dense(z, 0) for z in range(zones.count)])
        ctx.close()
    for a, b in zip(*res):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("cls", ["S", "W", "A"])
def test_two_sided_solves_track_the_one_sided_ones(gpu, cls):
    """The two-sided pentadiagonal solves (two waves per 64 lines meeting in
    a 4x4 interface system, FMA only) against the one-wave solves: 5 steps
    from the same state agree to 1e-12 of the field's size."""
    N = gpu.native
    zones, p = gpu.zones_for("SP", cls)
    out = []
    for twist in (0, 7):
        ctx = gpu.Context("SP", zones, p.dt)
        ctx.set_schedule(N.SCHED_SOLVE_TWIST, twist)
        ctx.init_exact()
        ctx.step(5, "graph")
        out.append(ctx.get_all(0))
        ctx.close()
    a, b = out
    assert np.abs(a - b).max() <= 1e-12 * np.abs(a).max()
