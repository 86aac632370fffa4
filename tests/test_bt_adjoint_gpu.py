"""The BT adjoint on the GPU (no counterpart in the reference).

The loss is sum(w*u_N**2) over every grid point after N steps, w a seeded
uniform [0,1) field.  Its gradients w.r.t. the initial u and the forcing are
pinned two independent ways: against forward-mode dual numbers pushed through
the CPU oracle (exact derivatives, small class), and against central finite
differences of the GPU forward run along seeded random directions (the
benchmark's adjoint case: BT class A, 4x4 uneven zones, 200 steps,
checkpoints every 20 steps).  Directions and w come from numpy's Philox
counter-based generator with the seeds written here.
"""
import numpy as np
import pytest

from oracle_lib import OracleBT

pytestmark = pytest.mark.gpu


def _rng(seed):
    return np.random.Generator(np.random.Philox(seed))


def _interior_mask(zones):
    parts = []
    for z in range(zones.count):
        nx, ny, nz = zones.nx[z], zones.ny[z], zones.nz[z]
        m = np.zeros((nz, ny, nx, 5))
        m[1:-1, 1:-1, 1:-1] = 1.0
        parts.append(m.ravel())
    return np.concatenate(parts)


@pytest.mark.parametrize("nsteps,ckpt,twist", [(10, 3, None), (6, 20, None), (6, 4, 0)])
def test_adjoint_matches_dual_numbers_class_S(gpu, nsteps, ckpt, twist):
    """Reverse mode on the GPU against forward mode on the CPU oracle: the
    directional derivatives along 8 random (u0, forcing) directions agree to
    1e-9 (both are exact up to rounding; the GPU runs the FMA kernels).
    twist = 0 runs the forward and the transposed solves one wave per 64
    lines instead of the default two-sided ones."""
    zones, p = gpu.zones_for("BT", "S")
    ctx = gpu.Context("BT", zones, p.dt)
    if twist is not None:
        ctx.set_schedule(gpu.native.SCHED_SOLVE_ROWS, 0)
        ctx.set_schedule(gpu.native.SCHED_SOLVE_TWIST, twist)
    ctx.init_exact()
    n = ctx.size
    r = _rng(2024)
    w = r.random(n)
    vu = r.standard_normal((8, n))
    vf = r.standard_normal((8, n)) * _interior_mask(zones)
    u0 = ctx.get_all(0)
    res = ctx.adjoint(nsteps, ckpt, w)
    assert np.array_equal(ctx.get_all(0), u0)  # u is back at u0
    o = OracleBT.for_class(p)
    L, d = o.dual_gradient(nsteps, w, vu, vf)
    assert abs(res["loss"] / L - 1) < 1e-12
    got = vu @ res["grad_u0"] + vf @ res["grad_forcing"]
    assert np.abs(got / d - 1).max() < 1e-9, np.abs(got / d - 1)
    # forcing enters only at interior points
    assert np.all(res["grad_forcing"][_interior_mask(zones) == 0] == 0)
    ctx.close()


def test_adjoint_central_differences_class_A(gpu):
    """The benchmark's adjoint case: central differences along 16 seeded
    random directions in (u0, forcing) against the adjoint's directional
    derivatives.  The loss difference is summed point by point,
    sum(w (u+ - u-)(u+ + u-)) / 2h, i.e. (L+ - L-)/2h without subtracting two
    rounded sums of 1.3 million terms (the loss is about 3e10).

    At h = 1e-6 what remains is the rounding of the two 200-step runs: the
    same quotient formed from the FMA and the strict kernels at the same input
    measures it (about 1e-8 to 1e-6 of the derivative), and the differences
    stay within a small multiple of it.  At h = 1e-4 that noise is 100 times
    smaller and the agreement is better than 1e-7."""
    zones, p = gpu.zones_for("BT", "A")
    ctx = gpu.Context("BT", zones, p.dt)
    ctx.init_exact()
    cs = gpu.Context("BT", zones, p.dt, strict=True)
    n = ctx.size
    r = _rng(20)
    w = r.random(n)
    res = ctx.adjoint(p.niter, 20, w)
    u0, f0 = ctx.get_all(0), ctx.get_all(1)
    mask = _interior_mask(zones)

    def final_u(c, du, df):
        c.put_all(0, u0 + du)
        c.put_all(1, f0 + df)
        c.step(p.niter, "batched")
        return c.get_all(0)

    def quotient(a, b, h):
        return np.sum(w * (a - b) * (a + b)) / (2 * h)

    rows = []
    for _ in range(16):
        vu = r.standard_normal(n)
        vf = r.standard_normal(n) * mask
        ad = vu @ res["grad_u0"] + vf @ res["grad_forcing"]
        h = 1e-6
        up, um = final_u(ctx, h * vu, h * vf), final_u(ctx, -h * vu, -h * vf)
        noise = abs(quotient(up, final_u(cs, h * vu, h * vf), h))
        e6 = abs(quotient(up, um, h) - ad)
        h = 1e-4
        e4 = abs(quotient(final_u(ctx, h * vu, h * vf), final_u(ctx, -h * vu, -h * vf), h) - ad)
        rows.append((abs(ad), e6, noise, e4))
    rows = np.array(rows)
    print("h=1e-6 rel err", ["%.1e" % x for x in rows[:, 1] / rows[:, 0]])
    print("rounding noise ", ["%.1e" % x for x in rows[:, 2] / rows[:, 0]])
    print("h=1e-4 rel err", ["%.1e" % x for x in rows[:, 3] / rows[:, 0]],
          "adjoint/forward time %.2f" % (res["adj_s"] / res["fwd_s"]))
    assert np.all(rows[:, 1] <= 1e-8 * rows[:, 0] + 8.0 * rows[:, 2].max())
    assert np.all(rows[:, 1] <= 1e-5 * rows[:, 0])
    assert np.all(rows[:, 3] <= 1e-7 * rows[:, 0])
    ctx.close()
    cs.close()


def test_adjoint_matches_dual_numbers_class_A(gpu):
    """The same dual-number check on the 4x4 uneven zones of class A, over
    12 steps with checkpoints every 5."""
    zones, p = gpu.zones_for("BT", "A")
    ctx = gpu.Context("BT", zones, p.dt)
    ctx.init_exact()
    n = ctx.size
    r = _rng(7)
    w = r.random(n)
    vu = r.standard_normal((4, n))
    vf = r.standard_normal((4, n)) * _interior_mask(zones)
    res = ctx.adjoint(12, 5, w)
    o = OracleBT.for_class(p, threads=16)
    L, d = o.dual_gradient(12, w, vu, vf)
    assert abs(res["loss"] / L - 1) < 1e-12
    got = vu @ res["grad_u0"] + vf @ res["grad_forcing"]
    assert np.abs(got / d - 1).max() < 1e-9, np.abs(got / d - 1)
    ctx.close()

