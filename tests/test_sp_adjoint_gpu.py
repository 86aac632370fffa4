"""The SP adjoint on the GPU (no counterpart in the reference).

The loss is sum(w*u_N**2) over every grid point after N SP steps, w a seeded
uniform [0,1) field.  Its gradients w.r.t. the initial u and the forcing are
pinned against forward-mode dual numbers pushed through the CPU oracle (exact
derivatives, classes S and W) and against central differences of the GPU
forward run along seeded random directions (class A, 4x4 equal zones).
Directions and w come from numpy's Philox generator with the seeds written
here.
"""
import numpy as np
import pytest

from oracle_lib import OracleSP

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


@pytest.mark.parametrize(
    "cls,nsteps,ckpt,twist",
    [("S", 10, 3, None), ("S", 6, 20, None), ("W", 8, 4, None), ("S", 6, 4, 0), ("W", 4, 2, 0)],
)
def test_adjoint_matches_dual_numbers(gpu, cls, nsteps, ckpt, twist):
    """Reverse mode on the GPU against forward mode on the CPU oracle: the
    directional derivatives along 8 random (u0, forcing) directions agree to
    1e-9 and the loss to 1e-12; the forcing gradient vanishes off the
    interior; u is back at u0.  The default schedule runs the two-sided
    forward and transposed solves in every direction at these sizes; twist =
    0 runs both one wave per 64 lines."""
    zones, p = gpu.zones_for("SP", cls)
    ctx = gpu.Context("SP", zones, p.dt)
    if twist is not None:
        ctx.set_schedule(gpu.native.SCHED_SOLVE_TWIST, twist)
    ctx.init_exact()
    n = ctx.size
    r = _rng(2025)
    w = r.random(n)
    vu = r.standard_normal((8, n))
    vf = r.standard_normal((8, n)) * _interior_mask(zones)
    u0 = ctx.get_all(0)
    res = ctx.adjoint(nsteps, ckpt, w)
    assert np.array_equal(ctx.get_all(0), u0)
    o = OracleSP.for_class(p, threads=16)
    L, d = o.dual_gradient(nsteps, w, vu, vf)
    assert abs(res["loss"] / L - 1) < 1e-12
    got = vu @ res["grad_u0"] + vf @ res["grad_forcing"]
    assert np.abs(got / d - 1).max() < 1e-9, np.abs(got / d - 1)
    assert np.all(res["grad_forcing"][_interior_mask(zones) == 0] == 0)
    ctx.close()


def test_adjoint_central_differences_class_A(gpu):
    """SP class A, 50 steps with checkpoints every 10: central differences
    (h = 1e-4, the loss difference summed point by point) along 8 seeded
    directions in (u0, forcing) against the adjoint's directional
    derivatives.  The differences stay within 1e-8 of the largest derivative
    among the directions (measured: about 1e-9; truncation grows as h^2 and
    rounding as 1/h away from h = 1e-4); an individual derivative can be much
    smaller than the others, so the bound is not taken relative to each."""
    zones, p = gpu.zones_for("SP", "A")
    ctx = gpu.Context("SP", zones, p.dt)
    ctx.init_exact()
    n = ctx.size
    r = _rng(21)
    w = r.random(n)
    nsteps = 50
    res = ctx.adjoint(nsteps, 10, w)
    u0, f0 = ctx.get_all(0), ctx.get_all(1)
    mask = _interior_mask(zones)

    def final_u(du, df):
        ctx.put_all(0, u0 + du)
        ctx.put_all(1, f0 + df)
        ctx.step(nsteps, "batched")
        return ctx.get_all(0)

    h = 1e-4
    rows = []
    for _ in range(8):
        vu = r.standard_normal(n)
        vf = r.standard_normal(n) * mask
        ad = vu @ res["grad_u0"] + vf @ res["grad_forcing"]
        up, um = final_u(h * vu, h * vf), final_u(-h * vu, -h * vf)
        fd = np.sum(w * (up - um) * (up + um)) / (2 * h)
        rows.append((ad, fd))
    rows = np.array(rows)
    scale = np.abs(rows[:, 0]).max()
    assert np.all(np.abs(rows[:, 1] - rows[:, 0]) <= 1e-8 * scale), rows
    ctx.close()
