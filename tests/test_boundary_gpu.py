"""The drop-in boundary with the reference's own array layouts.

npbmz_put_field / npbmz_get_field take the parent array of a Julia
OffsetArray plus its axis bounds.  These tests use the exact bounds the
reference allocates with one rank (maxcells = 1, so IMAX = nx+1):
- BT (src/BT/bt_data.jl:138-149): u (1:5, -2:IMAX+1, -2:JMAX+1, -2:KMAX+1,
  1:1), forcing (1:5, 0:IMAX-1, ..., 1:1), rhs (1:5, -1:IMAX-1, ..., 1:1);
  the trailing cell axis has length 1, so the parent array is the 4-D one the
  ABI takes;
- SP (src/SP/sp_data.jl:115-171), index order [i,j,k,m]: u (-2:IMAXP+1,
  -2:JMAXP+1, -2:KMAX+1, 1:5, 1:1) with IMAXP = div(IMAX,2)*2+1, forcing
  and rhs (0:IMAXP-1, 0:JMAXP-1, 0:KMAX-1, 1:5, 1:1);
- LU (src/LU/lu_data.jl:77-84): u, rsd, frct (1:5, -1:nx+2, -1:ny+2, 1:nz).

Julia arrays are column-major, so a numpy array in C order holds them with
the axes reversed.  Each test checks that only the zone's grid points move,
that padding is left untouched, and that the values land where the kernels
expect them: an oracle state loaded through these layouts and stepped by the
strict kernels stays bit-identical to the oracle.
"""
import numpy as np
import pytest

from oracle_lib import OracleBT, OracleLU, OracleSP

pytestmark = pytest.mark.gpu
SENTINEL = -7.25


def _bt_bounds(nx, ny, nz):
    imax, jmax, kmax = nx + 1, ny + 1, nz + 1
    return {0: ((1, -2, -2, -2), (5, imax + 1, jmax + 1, kmax + 1)),
            1: ((1, 0, 0, 0), (5, imax - 1, jmax - 1, kmax - 1)),
            2: ((1, -1, -1, -1), (5, imax - 1, jmax - 1, kmax - 1))}


def _sp_bounds(nx, ny, nz):
    imax, jmax, kmax = nx + 1, ny + 1, nz + 1
    imaxp, jmaxp = imax // 2 * 2 + 1, jmax // 2 * 2 + 1
    return {0: ((-2, -2, -2, 1), (imaxp + 1, jmaxp + 1, kmax + 1, 5)),
            1: ((0, 0, 0, 1), (imaxp - 1, jmaxp - 1, kmax - 1, 5)),
            2: ((0, 0, 0, 1), (imaxp - 1, jmaxp - 1, kmax - 1, 5))}


def _lu_bounds(nx, ny, nz):
    b = ((1, -1, -1, 1), (5, nx + 2, ny + 2, nz))
    return {0: b, 1: b, 2: b}


def _shape(lo, hi):
    """numpy C-order shape of a Julia array with these axis bounds."""
    return tuple(int(h) - int(l) + 1 for l, h in zip(lo, hi))[::-1]


def _grid_view(a, lo, order, nx, ny, nz, origin):
    """The zone's grid points of host array a as (nz, ny, nx, 5)."""
    if order == 0:  # [m,i,j,k]: numpy axes (k, j, i, m)
        m0, i0, j0, k0 = lo
        return a[origin - k0:origin - k0 + nz, origin - j0:origin - j0 + ny,
                 origin - i0:origin - i0 + nx, 1 - m0:1 - m0 + 5]
    i0, j0, k0, m0 = lo  # [i,j,k,m]: numpy axes (m, k, j, i)
    v = a[1 - m0:1 - m0 + 5, origin - k0:origin - k0 + nz, origin - j0:origin - j0 + ny,
          origin - i0:origin - i0 + nx]
    return np.moveaxis(v, 0, -1)


CASES = {
    "BT": (_bt_bounds, 0, 0),
    "SP": (_sp_bounds, 1, 0),
    "LU": (_lu_bounds, 0, 1),
}


@pytest.mark.parametrize("bench", ["BT", "SP", "LU"])
def test_reference_layouts_round_trip(gpu, bench):
    """Random values put through the reference's bounds come back at the grid
    points only (padding keeps its sentinel) and agree with the dense view."""
    bounds_of, order, origin = CASES[bench]
    zones, p = gpu.zones_for(bench, "S")
    ctx = gpu.Context(bench, zones, p.dt, omega=1.2)
    rng = np.random.default_rng(11)
    for z in (0, zones.count - 1):
        nx, ny, nz = zones.nx[z], zones.ny[z], zones.nz[z]
        for field, (lo, hi) in bounds_of(nx, ny, nz).items():
            host = rng.standard_normal(_shape(lo, hi))
            if bench != "LU":
                # the reference's trailing cell axis (1:maxcells, one rank):
                # the same memory with one more axis of length 1
                host5 = host.reshape((1,) + host.shape)
                assert np.shares_memory(host5, host)
            ctx.put_field(z, field, host, lo, hi, order)
            back = np.full_like(host, SENTINEL)
            ctx.get_field(z, field, back, lo, hi, order)
            g_in = _grid_view(host, lo, order, nx, ny, nz, origin)
            g_out = _grid_view(back, lo, order, nx, ny, nz, origin)
            assert np.array_equal(g_out, g_in), (z, field)
            assert np.array_equal(ctx.get_dense(z, field), g_in), (z, field)
            mask = np.ones(back.shape, bool)
            _grid_view(mask, lo, order, nx, ny, nz, origin)[...] = False
            assert np.all(back[mask] == SENTINEL), (z, field)
    ctx.close()


def _oracle(bench, p):
    if bench == "LU":
        return OracleLU.for_class(p, flags=OracleLU.NPB_EXCH)
    return {"BT": OracleBT, "SP": OracleSP}[bench].for_class(p)


@pytest.mark.parametrize("bench", ["BT", "SP", "LU"])
def test_reference_layouts_feed_the_kernels(gpu, bench):
    """The oracle's state after 3 steps, written into host arrays with the
    reference's layout and bounds, put through the ABI, stepped once more by
    the strict kernels and read back through the same layout: bit-identical
    to the oracle's next step.  A transposed or shifted axis in the mapping
    would scramble the stencils and fail this."""
    bounds_of, order, origin = CASES[bench]
    zones, p = gpu.zones_for(bench, "S")
    ctx = gpu.Context(bench, zones, p.dt, omega=1.2, strict=True, npb_exchange=bench == "LU")
    o = _oracle(bench, p)
    o.step(3)
    fields = (0, 1, 2) if bench == "LU" else (0, 1)
    for z in range(zones.count):
        nx, ny, nz = zones.nx[z], zones.ny[z], zones.nz[z]
        bnd = bounds_of(nx, ny, nz)
        for f in fields:
            lo, hi = bnd[f]
            host = np.full(_shape(lo, hi), SENTINEL)
            _grid_view(host, lo, order, nx, ny, nz, origin)[...] = o.get(z, f)
            ctx.put_field(z, f, host, lo, hi, order)
    ctx.step(1, "batched")
    o.step(1)
    for z in range(zones.count):
        nx, ny, nz = zones.nx[z], zones.ny[z], zones.nz[z]
        lo, hi = bounds_of(nx, ny, nz)[0]
        back = np.full(_shape(lo, hi), SENTINEL)
        ctx.get_field(z, 0, back, lo, hi, order)
        got = _grid_view(back, lo, order, nx, ny, nz, origin)
        assert np.array_equal(got, o.get(z, 0)), z
    ctx.close()
    o.close()
