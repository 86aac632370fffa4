"""The LU adjoint on the GPU (no counterpart in the reference).

The loss is sum(w*u_N**2) over every grid point after N SSOR iterations, w a
seeded uniform [0,1) field, with rsd recomputed as rhs(u0) first (what it
holds at an iteration's start).  Its gradients w.r.t. the initial u and frct
(every point: frct's boundary values enter through rsd = -frct there) are
pinned against forward-mode dual numbers pushed through the CPU oracle (exact
derivatives, classes S and W, both exchange plane ranges) and against
central differences of the GPU forward run along seeded random directions
(class A).  Directions and w come from numpy's Philox generator with the
seeds written here.
"""
import numpy as np
import pytest

from oracle_lib import OracleLU

pytestmark = pytest.mark.gpu


def _rng(seed):
    return np.random.Generator(np.random.Philox(seed))


@pytest.mark.parametrize("cls,nsteps,ckpt,npb,sweep", [("S", 10, 3, True, None),
                                                       ("S", 6, 20, False, None),
                                                       ("W", 6, 4, True, None),
                                                       ("S", 5, 2, True, (0, 256)),
                                                       ("S", 5, 2, True, (1, 128)),
                                                       ("S", 5, 2, True, (1, 512, 0)),
                                                       ("S", 5, 2, True, (1, 512, 3)),
                                                       ("W", 4, 2, True, (1, 512, 3))])
def test_adjoint_matches_dual_numbers(gpu, cls, nsteps, ckpt, npb, sweep):
    """Reverse mode on the GPU against forward mode on the CPU oracle: the
    directional derivatives along 8 random (u0, frct) directions agree to
    1e-9 and the loss to 1e-12; u is back at u0 afterwards.  sweep: the LU
    sweep schedule (hyperplane launches, or a workgroup per zone of the given
    size, and optionally its forwarding level, 3 handing hyperplanes on
    through LDS and loading a hyperplane ahead) that both the forward and the
    transposed sweeps follow; None keeps the default (at classes S and W a
    workgroup per zone reading global memory)."""
    N = gpu.native
    zones, p = gpu.zones_for("LU", cls)
    ctx = gpu.Context("LU", zones, p.dt, omega=1.2, npb_exchange=npb)
    if sweep is not None:
        ctx.set_schedule(N.SCHED_LU_SWEEP, sweep[0])
        ctx.set_schedule(N.SCHED_LU_THREADS, sweep[1])
        if len(sweep) > 2:
            ctx.set_schedule(N.SCHED_LU_FORWARD, sweep[2])
    ctx.init_exact()
    n = ctx.size
    r = _rng(2026)
    w = r.random(n)
    vu = r.standard_normal((8, n)) * 1e-2
    vf = r.standard_normal((8, n))
    u0 = ctx.get_all(0)
    res = ctx.adjoint(nsteps, ckpt, w)
    assert np.array_equal(ctx.get_all(0), u0)
    o = OracleLU.for_class(p, flags=OracleLU.NPB_EXCH if npb else 0, threads=16)
    L, d = o.dual_gradient(nsteps, w, vu, vf)
    assert abs(res["loss"] / L - 1) < 1e-12
    got = vu @ res["grad_u0"] + vf @ res["grad_forcing"]
    assert np.abs(got / d - 1).max() < 1e-9, np.abs(got / d - 1)
    ctx.close()


def test_adjoint_leaves_the_iteration_unchanged(gpu):
    """After the adjoint the context steps on exactly as one that never ran
    it: u and rsd hold their values from before the call."""
    zones, p = gpu.zones_for("LU", "S")
    a = gpu.Context("LU", zones, p.dt, omega=1.2, npb_exchange=True)
    b = gpu.Context("LU", zones, p.dt, omega=1.2, npb_exchange=True)
    a.init_exact()
    b.init_exact()
    a.adjoint(4, 2, np.ones(a.size))
    a.step(5, "batched")
    b.step(5, "batched")
    assert np.array_equal(a.get_all(0), b.get_all(0))
    assert np.array_equal(a.get_all(2), b.get_all(2))
    a.close()
    b.close()


def test_adjoint_central_differences_class_A(gpu):
    """LU class A, 40 iterations with checkpoints every 10: central
    differences (h = 1e-4 on frct, 1e-6 on u, the loss difference summed
    point by point) along 8 seeded directions against the adjoint's
    directional derivatives, within 1e-7 of the largest derivative among the
    directions."""
    zones, p = gpu.zones_for("LU", "A")
    ctx = gpu.Context("LU", zones, p.dt, omega=1.2, npb_exchange=True)
    ctx.init_exact()
    n = ctx.size
    r = _rng(22)
    w = r.random(n)
    nsteps = 40
    res = ctx.adjoint(nsteps, 10, w)
    u0, f0 = ctx.get_all(0), ctx.get_all(1)

    def final_u(du, df):
        ctx.put_all(0, u0 + du)
        ctx.put_all(1, f0 + df)
        ctx.run_phase(gpu.native.PHASE_RHS)  # rsd = rhs(u)
        ctx.step(nsteps, "batched")
        return ctx.get_all(0)

    h = 1e-4
    rows = []
    for _ in range(8):
        vu = r.standard_normal(n) * 1e-2
        vf = r.standard_normal(n)
        ad = vu @ res["grad_u0"] + vf @ res["grad_forcing"]
        up, um = final_u(h * vu, h * vf), final_u(-h * vu, -h * vf)
        fd = np.sum(w * (up - um) * (up + um)) / (2 * h)
        rows.append((ad, fd))
    rows = np.array(rows)
    scale = np.abs(rows[:, 0]).max()
    assert np.all(np.abs(rows[:, 1] - rows[:, 0]) <= 1e-7 * scale), rows
    ctx.close()
