"""BT-MZ on the GPU against the CPU oracle, through the C ABI.

Strict-FP kernels (no FMA contraction) must reproduce the oracle bit for bit:
initial state, every phase, whole runs.  The FMA kernels (the product's
default) must keep the norms within 1e-12 per step and 1e-10 at the end of a
run, and every run must pass the reference's verification table.
"""
import numpy as np
import pytest

from oracle_lib import OracleBT

pytestmark = pytest.mark.gpu


def _ctx(P, cls, strict, pair=None):
    """pair: bit d set solves direction d with the builder/eliminator wave
    pair instead of one wave per 64 lines (None keeps the default)."""
    zones, p = P.zones_for("BT", cls)
    ctx = P.Context("BT", zones, p.dt, strict=strict)
    if pair is not None:
        # the one-wave and wave-pair kernels themselves, not the rows kernel
        ctx.set_schedule(P.native.SCHED_SOLVE_ROWS, 0)
        ctx.set_schedule(P.native.SCHED_SOLVE_PAIR, pair)
    return ctx, zones, p


def _load_oracle_state(ctx, o, zones):
    for z in range(zones.count):
        ctx.put_dense(z, 0, o.get(z, OracleBT.U))
        ctx.put_dense(z, 1, o.get(z, OracleBT.FORCING))


def _interior(a):
    return a[1:-1, 1:-1, 1:-1]


@pytest.mark.parametrize("cls", ["S", "W", "A"])
def test_init_exact_is_bit_exact(gpu, cls):
    """initialize + exact_rhs on the device equal the host routines."""
    ctx, zones, p = _ctx(gpu, cls, True)
    o = OracleBT.for_class(p)
    ctx.init_exact()
    for z in range(zones.count):
        assert np.array_equal(ctx.get_dense(z, 0), o.get(z, OracleBT.U)), z
        assert np.array_equal(ctx.get_dense(z, 1), o.get(z, OracleBT.FORCING)), z
    ctx.close()


@pytest.mark.parametrize("pair", [0, 7])
@pytest.mark.parametrize("cls", ["S", "W"])
@pytest.mark.parametrize("mode", ["perzone", "batched"])
def test_each_phase_is_bit_exact(gpu, cls, mode, pair):
    """exch_qbc, compute_rhs, x/y/z_solve and add, one at a time, from the
    same state as the oracle (a few steps into the run), with either line
    solve kernel."""
    ctx, zones, p = _ctx(gpu, cls, True, pair)
    o = OracleBT.for_class(p)
    o.step(3)
    _load_oracle_state(ctx, o, zones)
    N = gpu.native
    phases = [N.PHASE_EXCH, N.PHASE_RHS, N.PHASE_XSOLVE, N.PHASE_YSOLVE, N.PHASE_ZSOLVE,
              N.PHASE_ADD]
    for ph in phases:
        o.phase(ph)
        ctx.run_phase(ph, mode)
        for z in range(zones.count):
            if ph in (N.PHASE_EXCH, N.PHASE_ADD):
                assert np.array_equal(ctx.get_dense(z, 0), o.get(z, OracleBT.U)), (ph, z)
            else:
                g = _interior(ctx.get_dense(z, 2))
                r = _interior(o.get(z, OracleBT.RHS))
                assert np.array_equal(g, r), (ph, z, np.abs(g - r).max())
    ctx.close()


@pytest.mark.parametrize("pair,twist", [(0, 0), (7, 0), (0, 7)])
def test_full_run_strict_is_bit_exact_class_S(gpu, pair, twist):
    """60 timesteps of class S: u identical to the oracle's in every zone;
    norms agree to the last bits of a differently ordered sum.  Strict
    arithmetic ignores the two-sided solve (it reorders the elimination)."""
    ctx, zones, p = _ctx(gpu, "S", True, pair)
    ctx.set_schedule(gpu.native.SCHED_SOLVE_TWIST, twist)
    o = OracleBT.for_class(p)
    ctx.init_exact()
    ctx.step(p.niter, "graph")
    o.step(p.niter)
    for z in range(zones.count):
        assert np.array_equal(ctx.get_dense(z, 0), o.get(z, OracleBT.U)), z
    gxcr, gxce, _ = ctx.norms()
    oxcr, oxce = o.norms()
    assert np.abs(gxcr / oxcr - 1).max() < 1e-13
    assert np.abs(gxce / oxce - 1).max() < 1e-13
    ctx.close()


def test_modes_agree_bitwise(gpu):
    """Per-zone launches, batched launches and the hipGraph replays run the
    same kernels on the same data: identical results.  19 steps in two calls
    replay the eight-step graph twice and the one-step graph three times."""
    res = []
    for mode in ("perzone", "batched", "graph"):
        ctx, zones, p = _ctx(gpu, "W", False)
        ctx.init_exact()
        ctx.step(10, mode)
        ctx.step(9, mode)
        res.append([ctx.get_dense(z, 0) for z in range(zones.count)])
        ctx.close()
    for r in res[1:]:
        for a, b in zip(res[0], r):
            assert np.array_equal(a, b)


@pytest.mark.parametrize("cls", ["W", "A"])
def test_line_solve_kernels_agree(gpu, cls):
    """Strict arithmetic: the wave-pair and the one-wave line solves give the
    same bits over 5 steps; FMA arithmetic: each kernel contracts its own
    products, so they agree to 1e-12 relative."""
    for strict, tol in ((True, 0.0), (False, 1e-12)):
        out = []
        for pair in (0, 7):
            ctx, zones, p = _ctx(gpu, cls, strict, pair)
            ctx.init_exact()
            ctx.step(5, "graph")
            out.append(ctx.get_all(0))
            ctx.close()
        err = np.abs(out[0] - out[1]).max() / np.abs(out[0]).max()
        assert err <= tol, (strict, err)


@pytest.mark.parametrize("cls", ["S", "W"])
def test_fma_kernels_track_oracle_per_step(gpu, cls):
    """FMA kernels: rsdnrm/errnorm within 1e-12 relative of the oracle after
    every one of the first 10 steps."""
    ctx, zones, p = _ctx(gpu, cls, False)
    o = OracleBT.for_class(p)
    ctx.init_exact()
    for s in range(10):
        ctx.step(1, "batched")
        o.step(1)
        gxcr, gxce, _ = ctx.norms()
        oxcr, oxce = o.norms()
        assert np.abs(gxcr / oxcr - 1).max() < 1e-12, (s, gxcr, oxcr)
        assert np.abs(gxce / oxce - 1).max() < 1e-12, (s, gxce, oxce)
    ctx.close()


@pytest.mark.parametrize("cls", ["S", "W", "A"])
def test_two_sided_solve_tracks_oracle(gpu, cls):
    """The two-sided elimination in all three directions (FMA arithmetic):
    rsdnrm/errnorm within 1e-12 of the oracle after each of the first 10
    steps; the whole class run passes the reference's verification and ends
    within 1e-10 of the oracle; and 5 steps agree with the one-wave solve to
    1e-12."""
    N = gpu.native
    ctx, zones, p = _ctx(gpu, cls, False)
    ctx.set_schedule(N.SCHED_SOLVE_ROWS, 0)
    ctx.set_schedule(N.SCHED_SOLVE_TWIST, 7)
    o = OracleBT.for_class(p, threads=16)
    ctx.init_exact()
    for s in range(10):
        ctx.step(1, "graph")
        o.step(1)
        gxcr, gxce, _ = ctx.norms()
        oxcr, oxce = o.norms()
        assert np.abs(gxcr / oxcr - 1).max() < 1e-12, (s, gxcr, oxcr)
        assert np.abs(gxce / oxce - 1).max() < 1e-12, (s, gxce, oxce)
    ctx.step(p.niter - 10, "graph")
    o.step(p.niter - 10)
    gxcr, gxce, _ = ctx.norms()
    oxcr, oxce = o.norms()
    ok, rel = gpu.verify.verify("BT", cls, p.dt, p.niter, gxcr, gxce)
    assert ok, rel
    assert np.abs(gxcr / oxcr - 1).max() < 1e-10
    assert np.abs(gxce / oxce - 1).max() < 1e-10
    out = []
    for twist in (0, 7):
        ctx.set_schedule(N.SCHED_SOLVE_TWIST, twist)
        ctx.init_exact()
        ctx.step(5, "graph")
        out.append(ctx.get_all(0))
    assert np.abs(out[0] - out[1]).max() / np.abs(out[0]).max() < 1e-12
    ctx.close()


@pytest.mark.parametrize("cls", ["S", "W", "A"])
def test_fma_run_verifies_and_tracks_oracle(gpu, cls):
    """A whole class run with the FMA kernels: the reference's verification
    passes and the end-of-run norms are within 1e-10 of the oracle."""
    r = gpu.go("BT", cls, mode="graph", verbose=False)
    assert r["verified"], r["rel"]
    zones, p = gpu.zones_for("BT", cls)
    o = OracleBT.for_class(p)
    o.step(p.niter)
    oxcr, oxce = o.norms()
    assert np.abs(np.array(r["xcr"]) / oxcr - 1).max() < 1e-10
    assert np.abs(np.array(r["xce"]) / oxce - 1).max() < 1e-10


def _host_threads():
    """Host cores for the oracle: this process's CPU affinity, within the
    thread budget the environment grants (OMP_NUM_THREADS)."""
    import os
    n = len(os.sched_getaffinity(0))
    cap = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return max(1, min(n, cap) if cap > 0 else n)


def test_class_B_headline_config(gpu):
    """BT-MZ 304x208x17 in 8x8 uneven zones (the benchmarked configuration):
    - strict kernels: two steps bit-identical to the oracle;
    - the product path (FMA kernels, the default schedule, hipGraph replays):
      rsdnrm/errnorm within 1e-12 of the oracle after each of the first 5
      steps, and within 1e-10 after all 200, which also pass the reference's
      class B table (src/BT/verify.jl:213-233)."""
    ctx, zones, p = _ctx(gpu, "B", True)
    o = OracleBT.for_class(p, threads=_host_threads())
    ctx.init_exact()
    ctx.step(2, "graph")
    o.step(2)
    for z in range(zones.count):
        assert np.array_equal(ctx.get_dense(z, 0), o.get(z, OracleBT.U)), z
    ctx.close()
    o.close()

    ctx, zones, p = _ctx(gpu, "B", False)
    o = OracleBT.for_class(p, threads=_host_threads())
    ctx.init_exact()
    for s in range(5):
        ctx.step(1, "graph")
        o.step(1)
        gxcr, gxce, _ = ctx.norms()
        oxcr, oxce = o.norms()
        assert np.abs(gxcr / oxcr - 1).max() < 1e-12, (s, gxcr, oxcr)
        assert np.abs(gxce / oxce - 1).max() < 1e-12, (s, gxce, oxce)
    ctx.step(p.niter - 5, "graph")
    o.step(p.niter - 5)
    gxcr, gxce, _ = ctx.norms()
    oxcr, oxce = o.norms()
    ok, rel = gpu.verify.verify("BT", "B", p.dt, p.niter, gxcr, gxce)
    assert ok, rel
    assert np.abs(gxcr / oxcr - 1).max() < 1e-10, (gxcr, oxcr)
    assert np.abs(gxce / oxce - 1).max() < 1e-10, (gxce, oxce)
    ctx.close()
    o.close()


def test_class_B_perzone_matches_graph_bitwise(gpu):
    """At the benchmarked configuration, the per-zone launches (the baseline
    the bench compares with) and the hipGraph replays give identical u after
    4 steps: same kernels, same data."""
    out = []
    for mode in ("perzone", "graph"):
        ctx, zones, p = _ctx(gpu, "B", False)
        ctx.init_exact()
        ctx.step(4, mode)
        out.append(ctx.get_all(0))
        ctx.close()
    assert np.array_equal(out[0], out[1])


def test_solve_grid_out_of_range_is_rejected_and_harmless(gpu):
    """A persistent solve grid beyond the accepted range fails with EINVAL
    before touching the context: the schedule is unchanged and a class S
    run still verifies; a valid grid still works afterwards."""
    N = gpu.native
    ctx, zones, p = _ctx(gpu, "S", False)
    before = ctx.describe()
    with pytest.raises(gpu.NpbmzError, match="solve grid"):
        ctx.set_schedule(N.SCHED_SOLVE_GRID, 1 << 30)
    with pytest.raises(gpu.NpbmzError, match="solve grid"):
        ctx.set_schedule(N.SCHED_SOLVE_GRID, -1)
    assert ctx.describe() == before
    ctx.init_exact()
    ctx.step(p.niter, "graph")
    xcr, xce, _ = ctx.norms()
    assert gpu.verify.verify("BT", "S", p.dt, p.niter, xcr, xce)[0]
    ctx.set_schedule(N.SCHED_SOLVE_GRID, 65536)
    ctx.init_exact()
    ctx.step(p.niter, "graph")
    xcr2, xce2, _ = ctx.norms()
    assert np.array_equal(xcr, xcr2) and np.array_equal(xce, xce2)
    ctx.close()


def test_put_get_julia_layout(gpu):
    """put/get through the reference's own OffsetArray layout of BT u
    (1:5, -2:nx+2, -2:ny+2, -2:nz+2), touching only the grid points."""
    ctx, zones, p = _ctx(gpu, "S", False)
    z = 1
    nx, ny, nz = zones.nx[z], zones.ny[z], zones.nz[z]
    lo, hi = (1, -2, -2, -2), (5, nx + 2, ny + 2, nz + 2)
    rng = np.random.default_rng(3)
    host = rng.standard_normal((nz + 5, ny + 5, nx + 5, 5))
    ctx.put_field(z, 0, host, lo, hi)
    back = np.full_like(host, -7.0)
    ctx.get_field(z, 0, back, lo, hi)
    grid = (slice(2, nz + 2), slice(2, ny + 2), slice(2, nx + 2))
    assert np.array_equal(back[grid], host[grid])
    mask = np.ones(host.shape[:3], bool)
    mask[grid] = False
    assert np.all(back[mask] == -7.0)
    assert np.array_equal(ctx.get_dense(z, 0), host[grid])
    ctx.close()


def test_errors_are_reported_not_raised_across_the_abi(gpu):
    ctx, zones, p = _ctx(gpu, "S", False)
    with pytest.raises(gpu.NpbmzError, match="zone out of range"):
        ctx.put_field(99, 0, np.zeros(40), (1, 0, 0, 0), (5, 1, 1, 1))
    with pytest.raises(gpu.NpbmzError, match="bounds describe"):
        ctx.put_field(0, 0, np.zeros(10), (1, 0, 0, 0), (5, 1, 1, 1))
    with pytest.raises(gpu.NpbmzError, match="does not cover"):
        ctx.put_field(0, 0, np.zeros((2, 2, 2, 5)), (1, 0, 0, 0), (5, 1, 1, 1))
    with pytest.raises(gpu.NpbmzError, match="unknown mode"):
        ctx.step(1, 7)
    N = gpu.native
    with pytest.raises(gpu.NpbmzError, match="LU sweep"):
        ctx.set_schedule(N.SCHED_LU_SWEEP, 1)
    with pytest.raises(gpu.NpbmzError, match="two-sided"):
        ctx.set_schedule(N.SCHED_SOLVE_TWIST, 9)
    with pytest.raises(gpu.NpbmzError, match="forwarding"):
        ctx.set_schedule(N.SCHED_LU_FORWARD, 1)
    with pytest.raises(gpu.NpbmzError, match="unknown phase"):
        ctx.run_phase(17)
    with pytest.raises(gpu.NpbmzError, match="nsteps"):
        ctx.phase_times(0)
    with pytest.raises(gpu.NpbmzError, match="nsteps and ckpt_every"):
        ctx.adjoint(0, 20, np.zeros(ctx.size))
    with pytest.raises(gpu.NpbmzError, match="nsteps and ckpt_every"):
        ctx.adjoint(4, 0, np.zeros(ctx.size))
    # the context is still usable: a class S run from here verifies
    ctx.init_exact()
    ctx.step(p.niter, "graph")
    xcr, xce, _ = ctx.norms()
    assert gpu.verify.verify("BT", "S", p.dt, p.niter, xcr, xce)[0]
    ctx.close()
