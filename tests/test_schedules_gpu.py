"""Schedules that change only which workgroup runs which task give the same
bits: the persistent solve grid (a few workgroups looping over the line
tasks, NPBMZ_SCHED_SOLVE_GRID) against one workgroup per task, for the BT and
SP steps (one-sided and two-sided solves) and their adjoints.  No counterpart
in the reference, which runs one Julia thread per zone."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run(gpu, bench, cls, strict, grid, twist, nsteps=4):
    N = gpu.native
    zones, p = gpu.zones_for(bench, cls)
    ctx = gpu.Context(bench, zones, p.dt, strict=strict)
    if twist is not None:
        if bench == "BT" and not strict:
            ctx.set_schedule(N.SCHED_SOLVE_ROWS, 0)  # the two-sided kernel itself
        ctx.set_schedule(N.SCHED_SOLVE_TWIST, twist)
    if grid is not None:
        ctx.set_schedule(N.SCHED_SOLVE_GRID, grid)
    ctx.init_exact()
    ctx.step(nsteps, "batched")
    u = ctx.get_all(0)
    ctx.close()
    return u


@pytest.mark.parametrize("bench,cls,strict,twist", [("BT", "W", True, None), ("BT", "W", False, 7),
                                                    ("SP", "W", True, None), ("SP", "W", False, 7),
                                                    ("SP", "A", False, 0)])
def test_persistent_solve_grid_gives_the_same_bits(gpu, bench, cls, strict, twist):
    print("grid", bench, cls, strict, twist, flush=True)
    ref = _run(gpu, bench, cls, strict, None, twist)
    for grid in (7, 64):
        got = _run(gpu, bench, cls, strict, grid, twist)
        assert np.array_equal(got, ref), (grid, np.abs(got - ref).max())


@pytest.mark.parametrize("bench", ["BT", "SP"])
def test_persistent_solve_grid_adjoint_gives_the_same_bits(gpu, bench):
    N = gpu.native
    zones, p = gpu.zones_for(bench, "S")
    res = []
    for grid in (None, 5):
        ctx = gpu.Context(bench, zones, p.dt)
        if grid is not None:
            ctx.set_schedule(N.SCHED_SOLVE_GRID, grid)
        ctx.init_exact()
        w = np.random.Generator(np.random.Philox(7)).random(ctx.size)
        r = ctx.adjoint(6, 3, w)
        res.append(r)
        ctx.close()
    assert res[0]["loss"] == res[1]["loss"]
    assert np.array_equal(res[0]["grad_u0"], res[1]["grad_u0"])
    assert np.array_equal(res[0]["grad_forcing"], res[1]["grad_forcing"])


@pytest.mark.parametrize("bench", ["BT", "SP"])
def test_tune_keeps_the_fields_and_the_tuned_run_verifies(gpu, bench):
    """npbmz_tune times graph replays of candidate schedules on the live
    context: the fields are as before afterwards, and a whole class W run on
    the schedule it picked verifies against the class table."""
    zones, p = gpu.zones_for(bench, "W")
    ctx = gpu.Context(bench, zones, p.dt)
    ctx.init_exact()
    ctx.step(2, "batched")
    before = [ctx.get_all(f) for f in (0, 1)]
    ctx.tune(200)
    after = [ctx.get_all(f) for f in (0, 1)]
    for a, b in zip(before, after):
        assert np.array_equal(a, b)
    sched = ctx.describe()
    ctx.init_exact()
    ctx.step(p.niter, "graph")
    xcr, xce, _ = ctx.norms()
    ok, rel = gpu.verify.verify(bench, "W", p.dt, p.niter, xcr, xce)
    assert ok, (sched, rel)
    ctx.close()


@pytest.mark.parametrize("bench", ["BT", "SP"])
def test_schedule_change_discards_both_step_graphs(gpu, bench):
    """Graph mode replays a captured eight-step graph and a one-step graph;
    a schedule change between calls must discard both.  The two-sided and
    the one-wave solves round differently with FMA arithmetic, so a stale
    graph would show as different bits from the same sequence launched
    batched."""
    N = gpu.native
    zones, p = gpu.zones_for(bench, "W")
    res = []
    for mode in ("batched", "graph"):
        ctx = gpu.Context(bench, zones, p.dt)
        # BT: the rows kernel, then the one-wave solves; SP: two-sided, then one-wave
        knob = N.SCHED_SOLVE_ROWS if bench == "BT" else N.SCHED_SOLVE_TWIST
        if bench == "BT":
            ctx.set_schedule(N.SCHED_SOLVE_TWIST, 0)
        ctx.set_schedule(knob, 7)
        ctx.init_exact()
        ctx.step(10, mode)
        ctx.set_schedule(knob, 0)
        ctx.step(9, mode)
        res.append(ctx.get_all(0))
        ctx.close()
    assert np.array_equal(res[0], res[1])
