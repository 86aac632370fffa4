"""Pins the BT CPU oracle before it is trusted as the GPU's checker.

Checks: the reference's own known answers (NPB-MZ verification tables), the
closed-form zone split, textbook reductions of the block solver (Gauss-Jordan
= dense solve; a line sweep = dense block-tridiagonal solve), and the
consistency of compute_rhs with exact_rhs on the exact solution.
"""
import numpy as np
import pytest

import npbapps_amd as P
from oracle_lib import OracleBT, binvcrhs
from npbapps_amd.verify import TABLES


def _run_class(cls):
    zones, p = P.zones_for("BT", cls)
    o = OracleBT.for_class(p)
    o.step(p.niter)
    xcr, xce = o.norms()
    o.close()
    return p, xcr, xce


@pytest.mark.parametrize("cls", ["S", "W"])
def test_oracle_reproduces_verification_table(cls):
    """src/BT/verify.jl:136-186: the reference's known-answer norms.  The
    tables carry 13 significant digits; the oracle agrees to that precision
    (far inside the reference's own epsilon of 1e-8)."""
    p, xcr, xce = _run_class(cls)
    _, _, xcrref, xceref, _ = TABLES["BT"][cls]
    rr = np.abs((xcr - xcrref) / xcrref)
    re = np.abs((xce - xceref) / xceref)
    assert rr.max() < 5e-12, rr
    assert re.max() < 5e-12, re


def test_zone_widths_closed_form():
    """Zone widths of the geometric split (src/BT/zone_setup.jl:3-48) at the
    benchmark classes, as computed independently from the formula."""
    z, _ = P.zones_for("BT", "B")
    assert sorted(set(z.nx)) == [16, 20, 24, 30, 38, 47, 57, 72]
    assert sorted(set(z.ny)) == [11, 13, 17, 21, 26, 31, 40, 49]
    assert sum(z.nx[:8]) == 304 and sum(z.ny[::8]) == 208
    z, _ = P.zones_for("BT", "A")
    assert z.nx[:4] == [13, 21, 36, 58] and z.ny[::4] == [13, 21, 36, 58]
    z, _ = P.zones_for("BT", "S")
    assert z.nx[:2] == [6, 18]
    z, _ = P.zones_for("SP", "A")
    assert set(z.nx) == {32} and set(z.ny) == {32} and set(z.nz) == {16}
    z, _ = P.zones_for("LU", "A")
    assert set(z.nx) == {32} and z.count == 16


def test_zone_setup_matches_oracle_and_neighbours_are_periodic():
    for cls in ("S", "A", "B"):
        z, p = P.zones_for("BT", cls)
        o = OracleBT.for_class(p)
        for q in range(z.count):
            assert o.zones[q] == (z.nx[q], z.ny[q], z.nz[q], z.west[q], z.east[q], z.south[q],
                                  z.north[q])
        o.close()
        for q in range(z.count):
            assert z.east[z.west[q]] == q and z.north[z.south[q]] == q


def test_binvcrhs_is_a_dense_solve():
    """binvcrhs (solve_subs.jl:468-723) computes B^-1 [C | r]."""
    rng = np.random.default_rng(7)
    for _ in range(20):
        B = rng.standard_normal((5, 5)) + 6.0 * np.eye(5)
        C = rng.standard_normal((5, 5))
        r = rng.standard_normal(5)
        _, C2, r2 = binvcrhs(B, C, r, True)
        ref = np.linalg.solve(B, np.column_stack([C, r]))
        assert np.allclose(C2, ref[:, :5], rtol=1e-12, atol=1e-13)
        assert np.allclose(r2, ref[:, 5], rtol=1e-12, atol=1e-13)
        _, C3, r3 = binvcrhs(B, C, r, False)
        assert np.array_equal(C3, C) and np.array_equal(r3, r2)


def _dense_line_solve(A, B, C, rhs):
    nl = A.shape[0]
    M = np.zeros((5 * nl, 5 * nl))
    for l in range(nl):
        M[5 * l:5 * l + 5, 5 * l:5 * l + 5] = B[l]
        if l > 0:
            M[5 * l:5 * l + 5, 5 * (l - 1):5 * l] = A[l]
        if l < nl - 1:
            M[5 * l:5 * l + 5, 5 * (l + 1):5 * l + 10] = C[l]
    return np.linalg.solve(M, rhs.reshape(-1)).reshape(nl, 5)


def _line(arr, d, a, b):
    """Values along line (a, b) of direction d from a (nz, ny, nx, 5) array."""
    if d == 0:   # a = k, b = j
        return arr[a, b, :, :]
    if d == 1:   # a = k, b = i
        return arr[a, :, b, :]
    return arr[:, a, b, :]  # a = j, b = i


@pytest.mark.parametrize("cls", ["S", "W"])
def test_line_sweeps_are_dense_block_tridiagonal_solves(cls):
    """x/y/z_solve (Thomas forward elimination with binvcrhs, then back
    substitution) equal a dense solve of the assembled block-tridiagonal
    system of each line."""
    zones, p = P.zones_for("BT", cls)
    o = OracleBT.for_class(p, threads=1)
    o.step(3)
    o.phase(0)
    o.phase(1)
    z = zones.count - 1
    nx, ny, nz = o.dims(z)
    lines = {0: [(1, 1), (nz // 2, ny - 2)], 1: [(1, 1), (nz - 2, nx // 2)],
             2: [(1, 1), (ny // 2, nx - 2)]}
    for d in range(3):
        before = o.get(z, OracleBT.RHS)
        blocks = {ab: o.line_blocks(z, d, *ab) for ab in lines[d]}
        o.phase(2 + d)
        after = o.get(z, OracleBT.RHS)
        for ab, (A, B, C) in blocks.items():
            x = _dense_line_solve(A, B, C, _line(before, d, *ab))
            got = _line(after, d, *ab)
            scale = np.abs(x).max()
            assert np.abs(got - x).max() <= 1e-11 * scale, (d, ab)
    o.close()


def _exact_field(nx, ny, nz, dn):
    """u_exact at every grid point: the ce polynomial of
    src/BT/exact_solution.jl evaluated with numpy's polyval."""
    ce = np.array([
        [2.0, 0.0, 0.0, 4.0, 5.0, 3.0, 0.5, 0.02, 0.01, 0.03, 0.5, 0.4, 0.3],
        [1.0, 0.0, 0.0, 0.0, 1.0, 2.0, 3.0, 0.01, 0.03, 0.02, 0.4, 0.3, 0.5],
        [2.0, 2.0, 0.0, 0.0, 0.0, 2.0, 3.0, 0.04, 0.03, 0.05, 0.3, 0.5, 0.4],
        [2.0, 2.0, 0.0, 0.0, 0.0, 2.0, 3.0, 0.03, 0.05, 0.04, 0.2, 0.1, 0.3],
        [5.0, 4.0, 3.0, 2.0, 0.1, 0.4, 0.3, 0.05, 0.04, 0.03, 0.1, 0.3, 0.2]])
    xi = np.arange(nx) * dn[0]
    eta = np.arange(ny) * dn[1]
    zeta = np.arange(nz) * dn[2]
    Z, Y, X = np.meshgrid(zeta, eta, xi, indexing="ij")
    out = np.zeros((nz, ny, nx, 5))
    for m in range(5):
        c = ce[m]
        px = np.polynomial.polynomial.polyval(X, [0.0, c[1], c[4], c[7], c[10]])
        py = np.polynomial.polynomial.polyval(Y, [0.0, c[2], c[5], c[8], c[11]])
        pz = np.polynomial.polynomial.polyval(Z, [0.0, c[3], c[6], c[9], c[12]])
        out[..., m] = c[0] + px + py + pz
    return out


def test_compute_rhs_vanishes_on_exact_solution():
    """exact_rhs builds forcing as minus the discrete operator applied to the
    exact solution (src/BT/exact_rhs.jl), with its own formula set; so
    compute_rhs on u = exact solution leaves a rounding-level residual."""
    zones, p = P.zones_for("BT", "W")
    o = OracleBT.for_class(p, threads=1)
    dn = (1.0 / ((p.gx - 1) / p.x_zones), 1.0 / ((p.gy - 1) / p.y_zones), 1.0 / (p.gz - 1))
    worst = 0.0
    for z in range(zones.count):
        nx, ny, nz = o.dims(z)
        o.set(z, OracleBT.U, _exact_field(nx, ny, nz, dn))
    o.phase(1)
    for z in range(zones.count):
        rhs = o.get(z, OracleBT.RHS)[1:-1, 1:-1, 1:-1]
        frc = o.get(z, OracleBT.FORCING)[1:-1, 1:-1, 1:-1]
        scale = np.abs(frc).max() * p.dt
        worst = max(worst, np.abs(rhs).max() / scale)
    o.close()
    # measured 2.4e-15; a 1e-12 relative perturbation of u already gives 1e-12
    assert worst < 5e-13, worst


def test_repo_flop_counts():
    """Mflop per step from the repo's formulas at the benchmark configs
    (src/BT/bt-driver.jl:131-140, src/LU/lu-driver.jl:142-151 with nx)."""
    z, _ = P.zones_for("BT", "B")
    assert abs(P.mflops_per_step("BT", z) - 758.383) < 5e-3
    z, _ = P.zones_for("BT", "A")
    assert abs(P.mflops_per_step("BT", z) - 184.425) < 5e-3
    z, _ = P.zones_for("BT", "S")
    assert abs(P.mflops_per_step("BT", z) - 1.629) < 5e-3
    z, _ = P.zones_for("SP", "A")
    assert abs(P.mflops_per_step("SP", z) - 184.425) < 5e-3
    z, _ = P.zones_for("LU", "A")
    assert abs(P.mflops_per_step("LU", z) - 410.53) < 5e-2


def test_dual_numbers_match_central_differences():
    """The oracle's forward-mode dual run (the adjoint's reference gradient)
    against central differences of its own double-precision run: class S,
    5 steps, two directions, the loss difference summed point by point.
    h = 1e-4: rounding noise (which grows as 1/h) and truncation (as h^2)
    are both below 1e-8 there."""
    zones, p = P.zones_for("BT", "S")
    o = OracleBT.for_class(p)
    r = np.random.Generator(np.random.Philox(11))
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

    h = 1e-4
    for q in range(2):
        up, um = final(h * vu[q], h * vf[q]), final(-h * vu[q], -h * vf[q])
        fd = np.sum(w * (up - um) * (up + um)) / (2 * h)
        assert abs(fd / d[q] - 1) < 1e-7, (fd, d[q])
    o.close()
