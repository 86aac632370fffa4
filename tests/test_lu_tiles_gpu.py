"""LU's pipelined tile sweep (NPBMZ_SCHED_LU_SWEEP = 2) against the other
sweep schedules, its two hand-offs against each other, and the CPU oracle.

Each zone's padded hyperplane is split by k rows over W workgroups that run
the zone's hyperplanes as a pipeline, every tile handing its edge row to the
next through global memory and publishing its progress every KB
hyperplanes.  The point arithmetic is the same as the padded-halo and
hyperplane-launch sweeps', on the same neighbour values, so results agree bit
for bit with them; runs track the oracle per iteration and verify.  Every
wait between tiles is bounded: with the bound forced to fail, the step
reports an error and the context stays usable.
"""
import numpy as np
import pytest

from oracle_lib import OracleLU

pytestmark = pytest.mark.gpu


def _ctx(P, cls, sweep=2, tiles=None, batch=None):
    N = P.native
    zones, p = P.zones_for("LU", cls)
    ctx = P.Context("LU", zones, p.dt, omega=1.2, npb_exchange=True)
    ctx.set_schedule(N.SCHED_LU_SWEEP, sweep)
    if tiles is not None:
        ctx.set_schedule(N.SCHED_LU_TILES, tiles)
    if batch is not None:
        ctx.set_schedule(N.SCHED_LU_TILE_BATCH, batch)
    return ctx, zones, p


def _run(P, cls, n, mode="graph", **kw):
    ctx, zones, p = _ctx(P, cls, **kw)
    desc = ctx.describe()
    ctx.init_exact()
    ctx.step(n, mode)
    out = (ctx.get_all(0), ctx.get_all(2))
    ctx.close()
    return out, desc


@pytest.mark.parametrize("cls,ref_sweep", [("W", 1), ("A", 1), ("B", 0)])
def test_tiles_match_other_sweeps(gpu, cls, ref_sweep):
    """Several iterations with the tile sweep (the class's default split,
    batch 4) against the
    workgroup-per-zone sweep (W: from global memory, A: the padded-halo form)
    and the hyperplane launches (B).  The point code and its neighbour values
    are the same; each kernel's FMA contraction is the compiler's, so they
    agree to 1e-13 relative (not always bit for bit)."""
    tol = 1e-13
    (u_t, r_t), desc = _run(gpu, cls, 6)
    assert "pipelined-tiles" in desc, desc
    (u_r, r_r), _ = _run(gpu, cls, 6, sweep=ref_sweep)
    eu = np.abs(u_t - u_r).max() / np.abs(u_r).max()
    er = np.abs(r_t - r_r).max() / np.abs(r_r).max()
    assert eu <= tol and er <= tol, (eu, er)


@pytest.mark.parametrize("tiles,batch", [(2, 1), (3, 5), (8, 8), (4, 64)])
def test_tile_counts_and_batches_agree_bitwise(gpu, tiles, batch):
    """Any split and publishing batch gives the same bits as the default
    split (2 tiles, batch 4) at class A: only which workgroup computes a
    point, and when its neighbours arrive, changes."""
    (u_ref, r_ref), _ = _run(gpu, "A", 4)
    (u, r), desc = _run(gpu, "A", 4, tiles=tiles, batch=batch)
    assert "pipelined-tiles" in desc
    assert np.array_equal(u, u_ref) and np.array_equal(r, r_ref), (tiles, batch)


def test_tiles_modes_agree_bitwise(gpu):
    """Per-zone launches (one zone's tiles per launch), batched launches and
    graph replays agree."""
    res = [_run(gpu, "A", 9, mode=m)[0] for m in ("perzone", "batched", "graph")]
    for u, r in res[1:]:
        assert np.array_equal(u, res[0][0]) and np.array_equal(r, res[0][1])


@pytest.mark.parametrize("cls", ["A", "B"])
def test_tiles_track_oracle_and_verify(gpu, cls):
    """rsdnrm/errnorm within 1e-12 of the oracle after each of the first 3
    iterations; the whole class run passes the reference's verification
    table (src/LU/verify.jl), and at class A ends within 1e-10 of the
    oracle's own whole run."""
    ctx, zones, p = _ctx(gpu, cls)
    o = OracleLU.for_class(p, flags=OracleLU.NPB_EXCH, threads=16)
    ctx.init_exact()
    for s in range(3):
        ctx.step(1, "graph")
        o.step(1)
        g = ctx.norms()
        r = o.norms()
        for a, b in ((g[0], r[0]), (g[1], r[1])):
            assert np.abs(np.asarray(a) / b - 1).max() < 1e-12, (s, a, b)
    ctx.step(p.niter - 3, "graph")
    g = ctx.norms()
    ok, rel = gpu.verify.verify("LU", cls, p.dt, p.niter, g[0], g[1], g[2])
    assert ok, rel
    if cls == "A":
        o.step(p.niter - 3)
        r = o.norms()
        assert np.abs(np.asarray(g[0]) / r[0] - 1).max() < 1e-10
        assert np.abs(np.asarray(g[1]) / r[1] - 1).max() < 1e-10
        assert abs(g[2] / r[2] - 1) < 1e-10
    ctx.close()
    o.close()


def test_tile_wait_bound_reports_an_error_and_recovers(gpu):
    """A wait bound of -1 fails every wait between tiles at once: the step
    returns an error naming the pipelined sweep (every workgroup still exits,
    and the counters are left reset); with the bound restored the context
    runs a verified class W run."""
    N = gpu.native
    ctx, zones, p = _ctx(gpu, "W")
    ctx.set_schedule(N.SCHED_LU_SPIN, -1)
    ctx.init_exact()
    with pytest.raises(gpu.NpbmzError, match="pipelined sweep"):
        ctx.step(2, "batched")
    ctx.set_schedule(N.SCHED_LU_SPIN, 1 << 20)
    ctx.init_exact()
    ctx.step(p.niter, "graph")
    xcr, xce, xci = ctx.norms()
    assert gpu.verify.verify("LU", "W", p.dt, p.niter, xcr, xce, xci)[0]
    ctx.close()


@pytest.mark.parametrize("cls,tiles,batch", [("A", 4, 4), ("A", 2, 8), ("B", 4, 4), ("B", 8, 2)])
def test_handoffs_agree_bitwise(gpu, cls, tiles, batch):
    """The XCD-local hand-off (edge rows through the shared L2 of the zone's
    XCD, no cache maintenance) and the agent-scope release/acquire give the
    same bits: the rows they hand on are the same values."""
    N = gpu.native
    out = []
    for ho in (0, 1):
        ctx, zones, p = _ctx(gpu, cls, tiles=tiles, batch=batch)
        ctx.set_schedule(N.SCHED_LU_HANDOFF, ho)
        desc = ctx.describe()
        assert ("handoff=xcd-local" if ho else "handoff=agent") in desc, desc
        ctx.init_exact()
        ctx.step(5, "graph")
        out.append((ctx.get_all(0), ctx.get_all(2)))
        ctx.close()
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


def test_handoff_placement_check_reports_an_error_and_recovers(gpu):
    """Hand-off 2 runs the XCD-local kernel with a placement map that no
    tile matches: every tile stops at its check, the step reports the
    placement error, and with hand-off 1 the context runs a verified class W
    run."""
    N = gpu.native
    ctx, zones, p = _ctx(gpu, "W")
    ctx.set_schedule(N.SCHED_LU_HANDOFF, 2)
    ctx.init_exact()
    with pytest.raises(gpu.NpbmzError, match="XCD"):
        ctx.step(2, "batched")
    ctx.set_schedule(N.SCHED_LU_HANDOFF, 1)
    ctx.init_exact()
    ctx.step(p.niter, "graph")
    xcr, xce, xci = ctx.norms()
    assert gpu.verify.verify("LU", "W", p.dt, p.niter, xcr, xce, xci)[0]
    ctx.close()


def test_tile_knobs_are_validated(gpu):
    N = gpu.native
    ctx, zones, p = _ctx(gpu, "S")
    for which, bad in ((N.SCHED_LU_SWEEP, 3), (N.SCHED_LU_TILES, 0), (N.SCHED_LU_TILES, 9),
                       (N.SCHED_LU_TILE_BATCH, 0), (N.SCHED_LU_SPIN, -2),
                       (N.SCHED_LU_HANDOFF, -1), (N.SCHED_LU_HANDOFF, 3)):
        with pytest.raises(gpu.NpbmzError):
            ctx.set_schedule(which, bad)
    ctx.close()


def test_default_sweeps_by_zone_size(gpu):
    """The library's defaults: zones whose padded hyperplane exceeds 256
    slots run pipelined tiles of at most 256 slots (class A, 31x15 slots: 2
    tiles; class B, 51x16: 4), smaller ones one workgroup per zone (class
    W); strict arithmetic (no tile kernel) keeps the padded-halo-free zone
    workgroups at class A and hyperplane launches at class B."""
    for cls, strict, want in (("W", False, "zone-workgroup"),
                              ("A", False, "pipelined-tiles tiles=2 batch=4 nt=256 handoff=xcd-local"),
                              ("A", True, "zone-workgroup"),
                              ("B", False, "pipelined-tiles tiles=4 batch=4 nt=256 handoff=xcd-local"),
                              ("B", True, "hyperplane-launch")):
        zones, p = gpu.zones_for("LU", cls)
        ctx = gpu.Context("LU", zones, p.dt, omega=1.2, npb_exchange=True, strict=strict)
        assert want in ctx.describe(), (cls, strict, ctx.describe())
        ctx.close()


@pytest.mark.parametrize("cls,sweep", [("W", 1), ("A", 2), ("B", 2)])
def test_update_in_upper_sweep_matches_separate_update(gpu, cls, sweep):
    """A timestep's upper sweep takes ssor's update (u += rsd/(omega(2 -
    omega))) itself in the padded-halo zone sweep (class W) and the
    pipelined tiles (A, B); phase by phase, the upper sweep stores rsd and
    the update runs as its own launch.  Both give the same bits in u.
    (Class W runs the padded-halo form here; its default forwards nothing.)"""
    N = gpu.native
    out = []
    for phases in (False, True):
        ctx, zones, p = _ctx(gpu, cls, sweep=sweep)
        if sweep == 1:
            ctx.set_schedule(N.SCHED_LU_FORWARD, 3)  # the padded-halo form
        assert ("pipelined-tiles" if sweep == 2 else "lds-halo") in ctx.describe()
        ctx.init_exact()
        for _ in range(3):
            if phases:
                for ph in (N.PHASE_EXCH, N.PHASE_SSOR_LOWER, N.PHASE_SSOR_UPPER,
                           N.PHASE_SSOR_UPDATE, N.PHASE_RHS):
                    ctx.run_phase(ph)
            else:
                ctx.step(1, "batched")
        out.append(ctx.get_all(0))
        ctx.close()
    assert np.array_equal(out[0], out[1])
