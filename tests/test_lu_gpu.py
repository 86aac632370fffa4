"""LU-MZ on the GPU against the CPU oracle, through the C ABI.

Strict-FP kernels (no FMA contraction) reproduce the oracle bit for bit: the
initial state, every SSOR phase under both sweep schedules and both exchange
plane ranges, and whole runs.  The FMA kernels keep the norms within 1e-12
per step and 1e-10 at the end of a run; with NPB-MZ's exchange planes every
run passes the reference's verification tables, and with the reference's own
planes the GPU matches the reference as written (which does not verify).
"""
import numpy as np
import pytest

from oracle_lib import OracleLU

pytestmark = pytest.mark.gpu

U, FRCT, RSD = 0, 1, 2


def _ctx(P, cls, strict, npb=True):
    zones, p = P.zones_for("LU", cls)
    ctx = P.Context("LU", zones, p.dt, omega=1.2, strict=strict, npb_exchange=npb)
    return ctx, zones, p


def _oracle(p, npb=True):
    return OracleLU.for_class(p, flags=OracleLU.NPB_EXCH if npb else 0)


def _load(ctx, o, zones):
    for z in range(zones.count):
        for f in (U, FRCT, RSD):
            ctx.put_dense(z, f, o.get(z, f))


@pytest.mark.parametrize("cls", ["S", "W"])
def test_init_exact_is_bit_exact(gpu, cls):
    """setbv + setiv + erhs + rhs on the device equal the host routines."""
    ctx, zones, p = _ctx(gpu, cls, True)
    o = _oracle(p)
    ctx.init_exact()
    for z in range(zones.count):
        for f in (U, FRCT, RSD):
            assert np.array_equal(ctx.get_dense(z, f), o.get(z, f)), (z, f)
    ctx.close()


# (name, sweep, threads, forwarding through LDS: 0 none, 1 rsd, 2 rsd and u, 3 (the default)
# the padded-halo sweep, which strict arithmetic runs as 2)
SCHEDULES = [("plane", 0, 256, 3), ("zone", 1, 128, 3), ("zone", 1, 256, 3), ("zone", 1, 512, 3),
             ("zone-rsd-u-lds", 1, 512, 2), ("zone-rsd-lds", 1, 512, 1), ("zone-global", 1, 512, 0)]


@pytest.mark.parametrize("cls", ["S", "W"])
@pytest.mark.parametrize("npb", [False, True])
@pytest.mark.parametrize("mode", ["perzone", "batched"])
@pytest.mark.parametrize("sched", SCHEDULES, ids=[f"{s[0]}{s[2]}" for s in SCHEDULES])
def test_each_phase_is_bit_exact(gpu, cls, npb, mode, sched):
    """exch_qbc, the lower sweep (rsd *= dt, jacld, blts), the upper sweep
    (jacu, buts), the update and rhs, one at a time from the oracle's state a
    few iterations into the run, under every sweep schedule, with each
    hyperplane of a zone sweep handed to the next through LDS or read back
    from global memory."""
    N = gpu.native
    ctx, zones, p = _ctx(gpu, cls, True, npb)
    ctx.set_schedule(N.SCHED_LU_SWEEP, sched[1])
    ctx.set_schedule(N.SCHED_LU_THREADS, sched[2])
    ctx.set_schedule(N.SCHED_LU_FORWARD, sched[3])
    o = _oracle(p, npb)
    o.step(3)
    _load(ctx, o, zones)
    # GPU phase -> oracle phases (the GPU folds the dt scaling into the lower sweep)
    steps = [(N.PHASE_EXCH, [0], U), (N.PHASE_SSOR_LOWER, [1, 2], RSD),
             (N.PHASE_SSOR_UPPER, [3], RSD), (N.PHASE_SSOR_UPDATE, [4], U),
             (N.PHASE_RHS, [5], RSD)]
    for ph, ophs, f in steps:
        for q in ophs:
            o.phase(q)
        ctx.run_phase(ph, mode)
        for z in range(zones.count):
            g, r = ctx.get_dense(z, f), o.get(z, f)
            assert np.array_equal(g, r), (ph, z, np.abs(g - r).max())
    ctx.close()


@pytest.mark.parametrize("npb", [False, True])
def test_full_run_strict_is_bit_exact_class_S(gpu, npb):
    ctx, zones, p = _ctx(gpu, "S", True, npb)
    o = _oracle(p, npb)
    ctx.init_exact()
    ctx.step(p.niter, "graph")
    o.step(p.niter)
    for z in range(zones.count):
        assert np.array_equal(ctx.get_dense(z, U), o.get(z, U)), z
        assert np.array_equal(ctx.get_dense(z, RSD), o.get(z, RSD)), z
    gxcr, gxce, gxci = ctx.norms()
    oxcr, oxce, oxci = o.norms()
    assert np.abs(gxcr / oxcr - 1).max() < 1e-13
    assert np.abs(gxce / oxce - 1).max() < 1e-13
    assert abs(gxci / oxci - 1) < 1e-13
    ctx.close()


@pytest.mark.parametrize("cls,fwd", [("S", None), ("W", None), ("S", 3), ("W", 3)])
def test_fma_kernels_track_oracle_per_step(gpu, cls, fwd):
    """The FMA kernels' norms within 1e-12 of the oracle's after each of ten
    iterations, with the default schedule (at S and W a workgroup per zone
    reading global memory) and with the padded-halo sweep (fwd = 3)."""
    ctx, zones, p = _ctx(gpu, cls, False)
    if fwd is not None:
        ctx.set_schedule(gpu.native.SCHED_LU_FORWARD, fwd)
    o = _oracle(p)
    ctx.init_exact()
    for s in range(10):
        ctx.step(1, "batched")
        o.step(1)
        gxcr, gxce, gxci = ctx.norms()
        oxcr, oxce, oxci = o.norms()
        assert np.abs(gxcr / oxcr - 1).max() < 1e-12, s
        assert np.abs(gxce / oxce - 1).max() < 1e-12, s
        assert abs(gxci / oxci - 1) < 1e-12, s
    ctx.close()


@pytest.mark.parametrize("cls", ["S", "W", "A"])
def test_fma_run_verifies_and_tracks_oracle(gpu, cls):
    """With NPB-MZ's exchange planes: the class table passes and the norms
    stay within 1e-10 of the oracle's.  LU-MZ class A (4x4 zones of
    32x32x16, 250 iterations at dt = 2.0) is the benchmark's LU case."""
    r = gpu.go("LU", cls, mode="graph", verbose=False, npb_exchange=True)
    assert r["verified"], r["rel"]
    zones, p = gpu.zones_for("LU", cls)
    o = _oracle(p)
    o.step(p.niter)
    oxcr, oxce, oxci = o.norms()
    assert np.abs(np.array(r["xcr"]) / oxcr - 1).max() < 1e-10
    assert np.abs(np.array(r["xce"]) / oxce - 1).max() < 1e-10
    assert abs(r["xci"] / oxci - 1) < 1e-10


def test_reference_exchange_run_matches_reference_as_written(gpu):
    """With the reference's exchange planes the GPU reproduces the reference
    as written: the oracle's norms to 1e-10, and, like it, no verification."""
    r = gpu.go("LU", "W", mode="graph", verbose=False, npb_exchange=False)
    assert not r["verified"]
    zones, p = gpu.zones_for("LU", "W")
    o = _oracle(p, npb=False)
    o.step(p.niter)
    oxcr, oxce, oxci = o.norms()
    assert np.abs(np.array(r["xcr"]) / oxcr - 1).max() < 1e-10
    assert np.abs(np.array(r["xce"]) / oxce - 1).max() < 1e-10
    assert abs(r["xci"] / oxci - 1) < 1e-10


@pytest.mark.parametrize("npb,halo", [(False, False), (True, False), (False, True), (True, True)])
def test_class_A_default_schedule_tracks_oracle_per_iteration(gpu, npb, halo):
    """Class A's default schedule (FMA arithmetic, pipelined tiles of the
    zones' padded hyperplanes) and the one-workgroup padded-halo zone sweep,
    with either exchange's planes: the norms within 1e-12 of the oracle's
    after each of five iterations."""
    ctx, zones, p = _ctx(gpu, "A", False, npb)
    if halo:
        ctx.set_schedule(gpu.native.SCHED_LU_SWEEP, 1)
        assert "lds-halo" in ctx.describe()
    else:
        assert "pipelined-tiles tiles=2" in ctx.describe()
    o = OracleLU.for_class(p, flags=OracleLU.NPB_EXCH if npb else 0, threads=16)
    ctx.init_exact()
    for s in range(5):
        ctx.step(1, "batched")
        o.step(1)
        gxcr, gxce, gxci = ctx.norms()
        oxcr, oxce, oxci = o.norms()
        assert np.abs(gxcr / oxcr - 1).max() < 1e-12, s
        assert np.abs(gxce / oxce - 1).max() < 1e-12, s
        assert abs(gxci / oxci - 1) < 1e-12, s
    ctx.close()


@pytest.mark.parametrize("strict", [True, False])
def test_schedules_and_modes_agree_class_A(gpu, strict):
    """Every schedule computes each point with the same operations in the
    same order: without contraction all give the same bits; with it each
    kernel instantiation picks its own fused multiply-adds, so they agree to
    rounding."""
    N = gpu.native
    res = []
    for sweep, nt, mode, fwd in ((0, 256, "batched", 3), (1, 256, "batched", 3),
                                 (1, 512, "graph", 3), (1, 128, "perzone", 3),
                                 (0, 256, "graph", 3), (1, 512, "batched", 0),
                                 (1, 512, "batched", 1), (1, 512, "batched", 2)):
        ctx, zones, p = _ctx(gpu, "A", strict)
        ctx.set_schedule(N.SCHED_LU_SWEEP, sweep)
        ctx.set_schedule(N.SCHED_LU_THREADS, nt)
        ctx.set_schedule(N.SCHED_LU_FORWARD, fwd)
        ctx.init_exact()
        ctx.step(4, mode)
        res.append([ctx.get_dense(z, U) for z in range(zones.count)])
        ctx.close()
    for other in res[1:]:
        for a, b in zip(res[0], other):
            if strict:
                assert np.array_equal(a, b)
            else:
                assert np.abs(a - b).max() <= 1e-13 * np.abs(a).max()


def test_tune_restores_the_state(gpu):
    ctx, zones, p = _ctx(gpu, "W", False)
    ctx.init_exact()
    ctx.step(2, "batched")
    before = [(ctx.get_dense(z, U), ctx.get_dense(z, RSD)) for z in range(zones.count)]
    ctx.tune(200)
    assert "sweep=" in ctx.describe()
    for z, (u, r) in enumerate(before):
        assert np.array_equal(ctx.get_dense(z, U), u)
        assert np.array_equal(ctx.get_dense(z, RSD), r)
    ctx.close()
