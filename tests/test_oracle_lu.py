"""Pins the LU-MZ CPU oracle before it is trusted as the GPU's checker.

Checks: the reference's known-answer tables (src/LU/verify.jl:89-170), the
closed-form initial field (setbv/setiv), the exchange's index algebra, the
textbook reductions of the two SSOR sweeps (blts = a dense block
lower-triangular solve, buts = a dense block upper-triangular solve), the
norms recomputed with numpy (RMS, and pintgr as a trapezoid rule), and the
two reference quirks: the exchange plane range (it decides verification) and
buts' sign slips (they change nothing).
"""
import numpy as np
import pytest

import npbapps_amd as P
from npbapps_amd.verify import TABLES
from oracle_lib import OracleLU

NPB = OracleLU.NPB_EXCH

# ce of the exact solution, src/LU/setcoeff.jl:36-40.
CE = np.array([
    [2.0, 0.0, 0.0, 4.0, 5.0, 3.0, 0.5, 0.02, 0.01, 0.03, 0.5, 0.4, 0.3],
    [1.0, 0.0, 0.0, 0.0, 1.0, 2.0, 3.0, 0.01, 0.03, 0.02, 0.4, 0.3, 0.5],
    [2.0, 2.0, 0.0, 0.0, 0.0, 2.0, 3.0, 0.04, 0.03, 0.05, 0.3, 0.5, 0.4],
    [2.0, 2.0, 0.0, 0.0, 0.0, 2.0, 3.0, 0.03, 0.05, 0.04, 0.2, 0.1, 0.3],
    [5.0, 4.0, 3.0, 2.0, 0.1, 0.4, 0.3, 0.05, 0.04, 0.03, 0.1, 0.3, 0.2]])


def exact_grid(nx, ny, nz):
    """The exact polynomial at every point of a zone, [k][j][i][m]."""
    xi = np.arange(nx) / (nx - 1)
    eta = np.arange(ny) / (ny - 1)
    zeta = np.arange(nz) / (nz - 1)
    Z, Y, X = np.meshgrid(zeta, eta, xi, indexing="ij")
    out = np.empty((nz, ny, nx, 5))
    for m in range(5):
        c = CE[m]
        out[..., m] = (c[0] + X * (c[1] + X * (c[4] + X * (c[7] + X * c[10])))
                       + Y * (c[2] + Y * (c[5] + Y * (c[8] + Y * c[11])))
                       + Z * (c[3] + Z * (c[6] + Z * (c[9] + Z * c[12]))))
    return out


def _oracle(cls, flags=NPB, threads=None):
    zones, p = P.zones_for("LU", cls)
    return OracleLU.for_class(p, flags=flags, threads=threads), zones, p


@pytest.mark.parametrize("cls", ["S", "W", "A"])
def test_oracle_reproduces_verification_table(cls):
    """src/LU/verify.jl:89-170: residual norms, error norms and the surface
    integral, to the tables' 13 digits, with NPB-MZ's exchange planes."""
    o, _, p = _oracle(cls)
    o.step(p.niter)
    xcr, xce, xci = o.norms()
    o.close()
    _, _, xcrref, xceref, xciref = TABLES["LU"][cls]
    assert np.abs((xcr - xcrref) / xcrref).max() < 5e-12
    assert np.abs((xce - xceref) / xceref).max() < 5e-12
    assert abs((xci - xciref) / xciref) < 5e-12


def test_reference_exchange_planes_do_not_verify():
    """The reference exchanges planes k = 1..nz-2 (src/LU/exchange_qbc.jl:85,
    109,135,160,205,226,247,268): the bottom boundary plane is overwritten and
    the top interior plane's halo is never refreshed.  Run as written, LU-MZ
    misses its own class S table by far more than the 1e-8 epsilon."""
    o, _, p = _oracle("S", flags=0)
    o.step(p.niter)
    xcr, xce, xci = o.norms()
    _, _, xcrref, xceref, xciref = TABLES["LU"]["S"]
    assert np.abs((xcr - xcrref) / xcrref).max() > 1e-2
    assert abs((xci - xciref) / xciref) > 1e-3


def test_buts_sign_slips_change_nothing():
    """src/LU/buts.jl:97,126,129,132 flip signs of products with entries of
    the eliminated diagonal block's upper triangle.  That triangle is
    structurally zero (jacld/jacu's d has nonzeros only on the diagonal, the
    first column and the last row, and elimination keeps it so), so the
    reference's buts and NPB's give identical fields."""
    runs = []
    for flags in (NPB, NPB | OracleLU.NPB_BUTS):
        o, zones, _ = _oracle("W", flags=flags)
        o.step(12)
        runs.append([(o.get(z, 0), o.get(z, 2)) for z in range(zones.count)])
        o.close()
    for (ua, ra), (ub, rb) in zip(*runs):
        assert np.array_equal(ua, ub) and np.array_equal(ra, rb)


def test_zones_are_uniform_and_neighbours_periodic():
    """src/LU/zone_setup.jl:22-88: equal zones; north/south are the x-minus/
    x-plus neighbours and west/east the y-minus/y-plus ones, periodic."""
    for cls, n in (("S", 6), ("A", 32)):
        o, zones, p = _oracle(cls)
        for z, (nx, ny, nz, no, so, we, ea) in enumerate(o.zones):
            assert (nx, ny, nz) == (n, n, p.gz) == (zones.nx[z], zones.ny[z], zones.nz[z])
            i, j = z % p.x_zones, z // p.x_zones
            assert no == (i - 1) % p.x_zones + j * p.x_zones
            assert so == (i + 1) % p.x_zones + j * p.x_zones
            assert we == i + ((j - 1) % p.y_zones) * p.x_zones
            assert ea == i + ((j + 1) % p.y_zones) * p.x_zones
            # the geometric neighbour table the library is given
            assert (no, so, we, ea) == (zones.west[z], zones.east[z], zones.south[z],
                                        zones.north[z])


def test_initial_field_is_boundary_values_plus_transfinite_interpolation():
    """setbv (src/LU/setbv.jl) puts the exact solution on all six faces;
    setiv (src/LU/setiv.jl:15-47) fills the interior with the transfinite
    interpolation of the face values."""
    o, zones, _ = _oracle("S")
    for z in range(zones.count):
        nx, ny, nz = o.dims(z)
        ue = exact_grid(nx, ny, nz)
        u = o.get(z, 0)
        bnd = np.ones((nz, ny, nx), bool)
        bnd[1:-1, 1:-1, 1:-1] = False
        assert np.allclose(u[bnd], ue[bnd], rtol=1e-14, atol=0)
        xi = (np.arange(nx) / (nx - 1))[None, None, :, None]
        eta = (np.arange(ny) / (ny - 1))[None, :, None, None]
        zeta = (np.arange(nz) / (nz - 1))[:, None, None, None]
        pxi = (1 - xi) * ue[:, :, :1] + xi * ue[:, :, -1:]
        peta = (1 - eta) * ue[:, :1] + eta * ue[:, -1:]
        pzeta = (1 - zeta) * ue[:1] + zeta * ue[-1:]
        ti = pxi + peta + pzeta - pxi * peta - peta * pzeta - pzeta * pxi + pxi * peta * pzeta
        assert np.allclose(u[1:-1, 1:-1, 1:-1], ti[1:-1, 1:-1, 1:-1], rtol=1e-13, atol=0)


@pytest.mark.parametrize("flags", [0, NPB])
def test_exchange_index_algebra(flags):
    """exch_qbc: i = 1 takes the north zone's i = nx-1 plane, i = nx the
    south zone's i = 2, then j = 1 the west zone's j = ny-1 and j = ny the east
    zone's j = 2 over the full i range (so j faces win at the corners), all
    from the values before the exchange, on planes k = 1..nz-2 (reference) or
    2..nz-1 (NPB)."""
    o, zones, _ = _oracle("S", flags=flags)
    o.step(2)
    pre = [o.get(z, 0) for z in range(zones.count)]
    o.phase(0)
    for z, (nx, ny, nz, no, so, we, ea) in enumerate(o.zones):
        ks = slice(0, nz - 2) if flags == 0 else slice(1, nz - 1)
        want = pre[z].copy()
        want[ks, :, nx - 1] = pre[so][ks, :, 1]
        want[ks, :, 0] = pre[no][ks, :, nx - 2]
        want[ks, 0, :] = pre[we][ks, ny - 2, :]
        want[ks, ny - 1, :] = pre[ea][ks, 1, :]
        assert np.array_equal(o.get(z, 0), want), z


def test_rhs_of_the_exact_solution_is_rounding_level():
    """frct is minus the discrete operator applied to the exact polynomial
    (src/LU/erhs.jl), and rhs (src/LU/rhs.jl) adds the same operator applied
    to u.  With u exact everywhere, the interior residual cancels to
    rounding level (the two evaluate the polynomial and the sums in different
    orders, so not to zero)."""
    o, zones, _ = _oracle("W")
    for z in range(zones.count):
        o.set_exact(z)
    o.phase(5)
    for z in range(zones.count):
        r = o.get(z, 2)[1:-1, 1:-1, 1:-1]
        f = o.get(z, 1)[1:-1, 1:-1, 1:-1]
        assert np.abs(r).max() < 1e-12 * np.abs(f).max(), z


def _interior_points(nx, ny, nz):
    return [(i, j, k) for k in range(2, nz) for j in range(2, ny) for i in range(2, nx)]


def test_sweeps_are_dense_block_triangular_solves():
    """jacld + blts (src/LU/blts.jl:9-152) solve (D + w L) x = dt*rsd for the
    block lower-triangular L of the k-1, j-1 and i-1 neighbours; jacu + buts
    (src/LU/buts.jl:9-148) then solve (D + w U) y = D x.  Both equal dense
    solves of the assembled matrices on a class S zone (320 unknowns).  Also
    checks the structure that makes buts' sign slips inert: d is the same in
    both sweeps and its strict upper triangle is zero."""
    o, zones, p = _oracle("S")
    o.step(3)
    o.phase(0)
    o.phase(1)
    z = 5
    nx, ny, nz = o.dims(z)
    pts = _interior_points(nx, ny, nz)
    idx = {q: n for n, q in enumerate(pts)}
    N = len(pts) * 5
    w = o.omega

    def assemble(upper):
        M = np.zeros((N, N))
        D = np.zeros((N, N))
        s = 1 if upper else -1
        for (i, j, k), n in idx.items():
            blk = o.point_blocks(z, upper, i, j, k)
            assert np.all(blk[0][np.triu_indices(5, 1)] == 0)
            M[5 * n:5 * n + 5, 5 * n:5 * n + 5] = blk[0]
            D[5 * n:5 * n + 5, 5 * n:5 * n + 5] = blk[0]
            for b, q in ((1, (i, j, k + s)), (2, (i, j + s, k)), (3, (i + s, j, k))):
                if q in idx:
                    m = idx[q]
                    M[5 * n:5 * n + 5, 5 * m:5 * m + 5] = w * blk[b]
        return M, D

    def interior(a):
        return np.array([a[k - 1, j - 1, i - 1] for (i, j, k) in pts]).ravel()

    v0 = interior(o.get(z, 2))
    ML, DL = assemble(False)
    o.phase(2)
    x = interior(o.get(z, 2))
    want = np.linalg.solve(ML, v0)
    assert np.abs(x - want).max() <= 1e-11 * np.abs(want).max()
    MU, DU = assemble(True)
    assert np.array_equal(DL, DU)
    o.phase(3)
    y = interior(o.get(z, 2))
    want = np.linalg.solve(MU, DU @ x)
    assert np.abs(y - want).max() <= 1e-11 * np.abs(want).max()


def _pressure(u):
    return 0.4 * (u[..., 4] - 0.5 * (u[..., 1] ** 2 + u[..., 2] ** 2 + u[..., 3] ** 2) / u[..., 0])


def test_norms_recomputed_with_numpy():
    """ERROR (src/LU/error.jl) and l2norm (src/LU/l2norm.jl) are RMS values
    over the zone interior; pintgr (src/LU/pintgr.jl) is the trapezoid rule
    for the pressure integral over the face pairs i = 2 / nx-1, j = 2 / ny-2
    and k = 3 / nz-1, on the box i in [2, nx-1], j in [2, ny-2], k in [3, nz-1]
    (1-based).  numpy's trapezoid gives the same values."""
    o, zones, p = _oracle("W")
    o.step(7)
    xcr, xce, xci, per = o.norms(per_zone=True)
    trap = getattr(np, "trapezoid", None) or np.trapz
    for z in range(zones.count):
        nx, ny, nz = o.dims(z)
        u, r = o.get(z, 0), o.get(z, 2)
        sl = (slice(1, -1),) * 3
        n = (nx - 2) * (ny - 2) * (nz - 2)
        errnm = np.sqrt(((exact_grid(nx, ny, nz)[sl] - u[sl]) ** 2).sum(axis=(0, 1, 2)) / n)
        rsdnm = np.sqrt((r[sl] ** 2).sum(axis=(0, 1, 2)) / n)
        assert np.allclose(per[z, :5], errnm, rtol=1e-12, atol=0)
        assert np.allclose(per[z, 5:10], rsdnm, rtol=1e-12, atol=0)
        dxi, deta, dzeta = 1.0 / (nx - 1), 1.0 / (ny - 1), 1.0 / (nz - 1)
        pr = _pressure(u)                      # [k][j][i], 0-based
        ib, ie, jb, je, kb, ke = 1, nx - 2, 1, ny - 3, 2, nz - 2
        t2 = lambda a, da, db: trap(trap(a, dx=da, axis=1), dx=db)
        f1 = sum(t2(pr[k, jb:je + 1, ib:ie + 1], dxi, deta) for k in (kb, ke))
        f2 = sum(t2(pr[kb:ke + 1, j, ib:ie + 1], dxi, dzeta) for j in (jb, je))
        f3 = sum(t2(pr[kb:ke + 1, jb:je + 1, i], deta, dzeta) for i in (ib, ie))
        assert abs(per[z, 10] - (f1 + f2 + f3)) <= 1e-12 * abs(f1 + f2 + f3)
    assert np.allclose(xcr, per[:, 5:10].sum(axis=0), rtol=1e-14)
    assert np.allclose(xce, per[:, :5].sum(axis=0), rtol=1e-14)


def test_dual_numbers_match_central_differences():
    """The LU oracle's forward-mode dual run (the LU adjoint's reference
    gradient) against central differences of its own double-precision run:
    class S, 5 iterations with NPB's exchange planes, two directions on u and
    frct over every point (frct's boundary values enter through rsd = -frct
    there), the loss difference summed point by point at h = 1e-4."""
    zones, p = P.zones_for("LU", "S")
    o = OracleLU.for_class(p, flags=NPB)
    r = np.random.Generator(np.random.Philox(13))
    shapes = [(zones.nz[z], zones.ny[z], zones.nx[z], 5) for z in range(zones.count)]
    n = sum(int(np.prod(s)) for s in shapes)
    w = r.random(n)
    vu = r.standard_normal((2, n)) * 1e-2
    vf = r.standard_normal((2, n))
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
        o.phase(5)  # rsd = rhs(u)
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
