"""Pins the SP CPU oracle: the reference's verification tables, a dense
solve of each assembled pentadiagonal line system, and the identity of SP's
compute_rhs with BT's (the reference's two copies agree textually)."""
import numpy as np
import pytest

import npbapps_amd as P
from npbapps_amd.verify import TABLES
from oracle_lib import OracleBT, OracleSP


@pytest.mark.parametrize("cls", ["S", "W"])
def test_oracle_reproduces_verification_table(cls):
    """src/SP/verify.jl:137-186: known-answer norms to table precision."""
    zones, p = P.zones_for("SP", cls)
    o = OracleSP.for_class(p)
    o.step(p.niter)
    xcr, xce = o.norms()
    o.close()
    _, _, xcrref, xceref, _ = TABLES["SP"][cls]
    assert np.abs((xcr - xcrref) / xcrref).max() < 5e-12
    assert np.abs((xce - xceref) / xceref).max() < 5e-12


def _penta_solve(L, rhs, o):
    nl = L.shape[0]
    M = np.zeros((nl, nl))
    for l in range(nl):
        for d, c in zip(range(-2, 3), range(5)):
            if 0 <= l + d < nl:
                M[l, l + d] = L[l, o + c]
    return np.linalg.solve(M, rhs)


def _line(arr, d, a, b):
    if d == 0:
        return arr[a, b, :, :]
    if d == 1:
        return arr[a, :, b, :]
    return arr[:, a, b, :]


def _ninvr(r, bt):
    r1, r2, r3, r4, r5 = r.T
    t1, t2 = bt * r3, 0.5 * (r4 + r5)
    return np.stack([-r2, r1, bt * (r4 - r5), -t1 + t2, t1 + t2], axis=1)


def _pinvr(r, bt):
    r1, r2, r3, r4, r5 = r.T
    t1, t2 = bt * r1, 0.5 * (r4 + r5)
    return np.stack([bt * (r4 - r5), -r3, r2, -t1 + t2, t1 + t2], axis=1)


@pytest.mark.parametrize("d", [0, 1])
def test_line_sweeps_are_dense_pentadiagonal_solves(d):
    """x_solve / y_solve: three scalar pentadiagonal systems per line (the
    first shared by components 1..3), solved by forward elimination and back
    substitution, then ninvr / pinvr.  Equal to a dense solve of each
    assembled system followed by the same transform."""
    zones, p = P.zones_for("SP", "W")
    o = OracleSP.for_class(p, threads=1)
    o.step(4)
    o.phase(0)
    o.phase(1)
    if d == 1:
        o.phase(2)
    z = 5
    nx, ny, nz = o.dims(z)
    ab = [(2, 3), (nz - 2, (nx if d == 1 else ny) - 3)]
    before = o.get(z, OracleSP.RHS)
    systems = {x: o.line_lhs(z, d, *x) for x in ab}
    o.phase(2 + d)
    after = o.get(z, OracleSP.RHS)
    bt = np.sqrt(0.5)
    for x, L in systems.items():
        r = _line(before, d, *x)
        sol = np.zeros_like(r)
        for m in range(5):
            sys = 0 if m < 3 else m - 2
            sol[:, m] = _penta_solve(L, r[:, m], 5 * sys)
        want = (_ninvr if d == 0 else _pinvr)(sol, bt)[1:-1]
        got = _line(after, d, *x)[1:-1]
        assert np.abs(got - want).max() <= 1e-11 * np.abs(want).max(), (d, x)
    o.close()


def test_sp_compute_rhs_equals_bt_compute_rhs():
    """From the same u, SP's compute_rhs (src/SP/rhs.jl) and BT's
    (src/BT/rhs.jl) give identical rhs: phase 1 of SP adds txinvr, so compare
    through the verification norms, which use plain compute_rhs."""
    zones, p = P.zones_for("SP", "W")
    s = OracleSP.for_class(p)
    b = OracleBT(p.x_zones, p.y_zones, p.gx, p.gy, p.gz, p.ratio, p.dt)
    s.step(3)
    for z in range(zones.count):
        b.set(z, 0, s.get(z, 0))
        b.set(z, 1, s.get(z, 1))
    assert all(np.array_equal(a, c) for a, c in zip(s.norms(), b.norms()))
    s.close()
    b.close()


def test_dual_numbers_match_central_differences():
    """The SP oracle's forward-mode dual run (the SP adjoint's reference
    gradient) against central differences of its own double-precision run:
    class S, 5 steps, two directions, the loss difference summed point by
    point at h = 1e-4.  rhon's max is differentiated along the branch that
    attains it; the seeded directions cross no tie."""
    zones, p = P.zones_for("SP", "S")
    o = OracleSP.for_class(p)
    r = np.random.Generator(np.random.Philox(12))
    shapes = [(zones.nz[z], zones.ny[z], zones.nx[z], 5) for z in range(zones.count)]
    n = sum(int(np.prod(s)) for s in shapes)
    w = r.random(n)
    mask = np.concatenate([np.pad(np.ones((a - 2, b - 2, c - 2, 5)),
                                  ((1, 1), (1, 1), (1, 1), (0, 0))).ravel() for a, b, c, _ in shapes])
    vu = r.standard_normal((2, n))
    vf = r.standard_normal((2, n)) * mask
    L, d = o.dual_gradient(5, w, vu, vf)
    u0 = [o.get(z, 0) for z in range(zones.count)]
    f0 = [o.get(z, 1) for z in range(zones.count)]

    def final(du, df):
        off = 0
        for z, s in enumerate(shapes):
            k = int(np.prod(s))
            o.set(z, 0, u0[z] + du[off:off + k].reshape(s))
            o.set(z, 1, f0[z] + df[off:off + k].reshape(s))
            off += k
        o.step(5)
        return np.concatenate([o.get(z, 0).ravel() for z in range(zones.count)])

    base = final(0 * vu[0], 0 * vf[0])
    assert abs(L / np.sum(w * base * base) - 1) < 1e-13
    h = 1e-4
    for q in range(2):
        up, um = final(h * vu[q], h * vf[q]), final(-h * vu[q], -h * vf[q])
        fd = np.sum(w * (up - um) * (up + um)) / (2 * h)
        assert abs(fd / d[q] - 1) < 1e-7, (fd, d[q])
    o.close()
