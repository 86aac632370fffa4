"""The rows-over-lanes BT line solve (NPBMZ_SCHED_SOLVE_ROWS) against the CPU
oracle.

Each line's five block rows are solved on five lanes, two-sided, with the
last reduced rows of each half in registers and older rows spilled (class A's
and B's longer x and y lines spill; S and W fit in registers).  FMA
arithmetic only: rsdnrm/errnorm within 1e-12 of the oracle after each of the
first steps and within 1e-10 at the end of the class run, which must also
pass the reference's verification table; and per-zone, batched and graph
launches agree bit for bit.
"""
import numpy as np
import pytest

from oracle_lib import OracleBT

pytestmark = pytest.mark.gpu


def _threads():
    import os
    n = len(os.sched_getaffinity(0))
    cap = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return max(1, min(n, cap) if cap > 0 else n)


def _ctx(P, cls, rows=7):
    zones, p = P.zones_for("BT", cls)
    ctx = P.Context("BT", zones, p.dt)
    ctx.set_schedule(P.native.SCHED_SOLVE_ROWS, rows)
    return ctx, zones, p


@pytest.mark.parametrize("cls,per_step", [("S", 10), ("W", 10), ("A", 5), ("B", 3)])
def test_rows_solve_tracks_oracle(gpu, cls, per_step):
    ctx, zones, p = _ctx(gpu, cls)
    assert "rows=7" in ctx.describe()
    o = OracleBT.for_class(p, threads=_threads())
    ctx.init_exact()
    for s in range(per_step):
        ctx.step(1, "graph")
        o.step(1)
        gxcr, gxce, _ = ctx.norms()
        oxcr, oxce = o.norms()
        assert np.abs(gxcr / oxcr - 1).max() < 1e-12, (s, gxcr, oxcr)
        assert np.abs(gxce / oxce - 1).max() < 1e-12, (s, gxce, oxce)
    ctx.step(p.niter - per_step, "graph")
    o.step(p.niter - per_step)
    gxcr, gxce, _ = ctx.norms()
    oxcr, oxce = o.norms()
    ok, rel = gpu.verify.verify("BT", cls, p.dt, p.niter, gxcr, gxce)
    assert ok, rel
    assert np.abs(gxcr / oxcr - 1).max() < 1e-10
    assert np.abs(gxce / oxce - 1).max() < 1e-10
    ctx.close()
    o.close()


@pytest.mark.parametrize("cls", ["W", "A"])
def test_rows_solve_each_direction_matches_two_sided(gpu, cls):
    """One direction at a time on the rows kernel (the others two-sided): the
    solve of that phase agrees with the two-sided kernel's to 1e-12 on the
    same input, interior points of every zone."""
    N = gpu.native
    zones, p = gpu.zones_for("BT", cls)
    o = OracleBT.for_class(p)
    o.step(3)
    for d, ph in enumerate((N.PHASE_XSOLVE, N.PHASE_YSOLVE, N.PHASE_ZSOLVE)):
        out = []
        for rows in (0, 1 << d):
            ctx = gpu.Context("BT", zones, p.dt)
            ctx.set_schedule(N.SCHED_SOLVE_ROWS, rows)
            for z in range(zones.count):
                ctx.put_dense(z, 0, o.get(z, OracleBT.U))
                ctx.put_dense(z, 1, o.get(z, OracleBT.FORCING))
            ctx.run_phase(N.PHASE_RHS)
            ctx.run_phase(ph)
            out.append(np.concatenate([ctx.get_dense(z, 2)[1:-1, 1:-1, 1:-1].ravel()
                                       for z in range(zones.count)]))
            ctx.close()
        err = np.abs(out[0] - out[1]).max() / np.abs(out[0]).max()
        assert err < 1e-12, (d, err)
    o.close()


def test_rows_solve_modes_agree_bitwise(gpu):
    """Per-zone launches, batched launches and graph replays of the rows
    kernel give identical u (class A: x and y lines spill)."""
    res = []
    for mode in ("perzone", "batched", "graph"):
        ctx, zones, p = _ctx(gpu, "A")
        ctx.init_exact()
        ctx.step(9, mode)
        res.append(ctx.get_all(0))
        ctx.close()
    assert np.array_equal(res[0], res[1]) and np.array_equal(res[0], res[2])


def test_rows_solve_persistent_grid_bitwise(gpu):
    """A persistent grid smaller than the task count (each workgroup strides
    over several tasks, reusing its LDS and spill slice) gives the same bits
    as one workgroup per task."""
    res = []
    for grid in (0, 96):
        ctx, zones, p = _ctx(gpu, "A")
        ctx.set_schedule(gpu.native.SCHED_SOLVE_GRID, grid)
        ctx.init_exact()
        ctx.step(4, "graph")
        res.append(ctx.get_all(0))
        ctx.close()
    assert np.array_equal(res[0], res[1])


def test_ring_form_follows_the_task_count(gpu):
    """Directions whose 12-line tasks exceed one round of two waves per SIMD
    and whose half lines exceed 8 rows run the long-ring form (8 reduced rows
    in registers, block rows read as applied), the others the short ring (6):
    class A all short (x and y have few tasks, z short lines), class B long
    in x and y on a 256-CU device."""
    for cls, want in (("A", "ring=6/6/6"), ("B", "ring=8/8/6")):
        zones, p = gpu.zones_for("BT", cls)
        ctx = gpu.Context("BT", zones, p.dt)
        assert want in ctx.describe(), (cls, ctx.describe())
        ctx.close()


def test_rows_solve_rejects_bad_masks_and_strict_ignores_it(gpu):
    N = gpu.native
    ctx, zones, p = _ctx(gpu, "S", rows=0)
    with pytest.raises(gpu.NpbmzError, match="rows-over-lanes"):
        ctx.set_schedule(N.SCHED_SOLVE_ROWS, 8)
    ctx.close()
    # strict arithmetic keeps the reference's elimination: still bit-exact
    zones, p = gpu.zones_for("BT", "S")
    ctx = gpu.Context("BT", zones, p.dt, strict=True)
    ctx.set_schedule(N.SCHED_SOLVE_ROWS, 7)
    o = OracleBT.for_class(p)
    ctx.init_exact()
    ctx.step(5, "graph")
    o.step(5)
    for z in range(zones.count):
        assert np.array_equal(ctx.get_dense(z, 0), o.get(z, OracleBT.U))
    ctx.close()
    o.close()


@pytest.mark.parametrize("bench,cls", [("BT", "W"), ("BT", "A"), ("SP", "W"), ("SP", "A")])
def test_exchange_in_rhs_launch_matches_separate_exchange(gpu, bench, cls):
    """Batched and graph steps run the zone-face exchange in compute_rhs's
    launch by default (the rhs reads each face neighbour from the cell the
    exchange copies; the exchange's blocks fill the halos for the solves);
    NPBMZ_SCHED_RHS_EXCH = 0 gives the exchange a launch of its own.  Both
    give the same bits in u, halos included, after a group of eight graph
    steps and a single one."""
    N = gpu.native
    zones, p = gpu.zones_for(bench, cls)
    out = []
    for fx in (1, 0):
        ctx = gpu.Context(bench, zones, p.dt)
        ctx.set_schedule(N.SCHED_RHS_EXCH, fx)
        ctx.init_exact()
        ctx.step(9, "graph")
        ctx.step(2, "batched")
        out.append(ctx.get_all(0))
        ctx.close()
    assert np.array_equal(out[0], out[1])


def test_rhs_exch_knob_is_validated(gpu):
    N = gpu.native
    zones, p = gpu.zones_for("BT", "S")
    ctx = gpu.Context("BT", zones, p.dt)
    for bad in (-1, 2):
        with pytest.raises(gpu.NpbmzError):
            ctx.set_schedule(N.SCHED_RHS_EXCH, bad)
    ctx.close()
    zones, p = gpu.zones_for("LU", "S")
    ctx = gpu.Context("LU", zones, p.dt, omega=1.2)
    with pytest.raises(gpu.NpbmzError):
        ctx.set_schedule(N.SCHED_RHS_EXCH, 1)
    ctx.close()
