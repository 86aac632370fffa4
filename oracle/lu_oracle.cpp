// Test infrastructure only: the CPU oracle for the LU-MZ SSOR timestep.
//
// Nothing under oracle/ is part of the product (see npb_common.h).  This file
// restates the Julia LU-MZ path zone by zone, with one MPI rank per zone as
// the reference runs it on one node: every zone is its own subdomain, so the
// intra-zone neighbours of src/LU/neighbors.jl are all -1, ist = jst = 2,
// iend = nx-1, jend = ny-1 (src/LU/subdomain.jl), ipt = jpt = 0, and the
// intra-zone exchanges (exchange_1/3/4/5/6) do nothing.
//
// Indices are 1-based like the reference's, (m, i, j, k) with m = 1..5,
// i = 1..nx, j = 1..ny, k = 1..nz; storage is m fastest, then i, j, k, which is
// also the dense layout of oracle_lu_get / oracle_lu_set.
//
// Every expression keeps the reference's operand order and grouping: C++ and
// Julia share precedence and left associativity for + - * /, unary minus binds
// tighter than * in both, x^2 is written sq(x) = x*x (Julia lowers the literal
// power to one multiply), and the build uses -ffp-contract=off.
//
// Two choices are selectable so that tests can show what they change:
//   LU_NPB_EXCH  exchange planes k = 2..nz-1 (NPB-MZ) instead of the
//                reference's k = 1..nz-2 (src/LU/exchange_qbc.jl:85,109,...).
//   LU_NPB_BUTS  back substitution of buts with NPB's signs instead of the
//                reference's (src/LU/buts.jl:97,126,129,132).
// The default (flags 0) is the reference as written.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "dual.h"
#include "npb_common.h"

namespace npb {
namespace lu {

enum { LU_NPB_EXCH = 1, LU_NPB_BUTS = 2 };

// Module constants, src/LU/setcoeff.jl:1-41 and src/LU/lu_data.jl:13-17.
static const double c1 = 1.40, c2 = 0.40, c3 = 1.00e-01, c4 = 1.00, c5 = 1.40;
static const double dx1 = 0.75, dx2 = 0.75, dx3 = 0.75, dx4 = 0.75, dx5 = 0.75;
static const double dy1 = 0.75, dy2 = 0.75, dy3 = 0.75, dy4 = 0.75, dy5 = 0.75;
static const double dz1 = 1.00, dz2 = 1.00, dz3 = 1.00, dz4 = 1.00, dz5 = 1.00;
static const double dssp = std::max(std::max(dx1, dy1), dz1) / 4.0;
static const int ii1 = 2, ji1 = 2, ki1 = 3;

template <class T>
static inline T sq(const T& x) { return x * x; }

// setcoeff, src/LU/setcoeff.jl:43-70.  The reference keeps these in module
// globals set per zone in the setup loop, so every later call sees the last
// zone's values; LU zones are all the same size, so per-zone values agree.
struct Coeff {
  double dxi, deta, dzeta, tx1, tx2, tx3, ty1, ty2, ty3, tz1, tz2, tz3;
  int ii2, ji2, ki2;
};

static Coeff setcoeff(int nx0, int ny0, int nz0) {
  Coeff c;
  c.dxi = 1.0e+00 / (nx0 - 1);
  c.deta = 1.0e+00 / (ny0 - 1);
  c.dzeta = 1.0e+00 / (nz0 - 1);
  c.tx1 = 1.0e+00 / (c.dxi * c.dxi);
  c.tx2 = 1.0e+00 / (2.0e+00 * c.dxi);
  c.tx3 = 1.0e+00 / c.dxi;
  c.ty1 = 1.0e+00 / (c.deta * c.deta);
  c.ty2 = 1.0e+00 / (2.0e+00 * c.deta);
  c.ty3 = 1.0e+00 / c.deta;
  c.tz1 = 1.0e+00 / (c.dzeta * c.dzeta);
  c.tz2 = 1.0e+00 / (2.0e+00 * c.dzeta);
  c.tz3 = 1.0e+00 / c.dzeta;
  c.ii2 = nx0 - 1;
  c.ji2 = ny0 - 2;
  c.ki2 = nz0 - 1;
  return c;
}

// A 5 x nx x ny x nz field, 1-based, of scalars T (double, or the dual
// numbers of oracle_lu_dual).
template <class T>
struct FieldT {
  int nx = 0, ny = 0, nz = 0;
  std::vector<T> v;
  void alloc(int a, int b, int c) {
    nx = a; ny = b; nz = c;
    v.assign((size_t)5 * a * b * c, T(0.0));
  }
  T& operator()(int m, int i, int j, int k) {
    return v[(size_t)(m - 1) + 5 * ((size_t)(i - 1) + (size_t)nx * ((size_t)(j - 1) + (size_t)ny * (k - 1)))];
  }
  T* at(int i, int j, int k) { return &(*this)(1, i, j, k); }
};
using Field = FieldT<double>;

// Per-row 5 x 5 x nx block workspace of jacld/jacu, src/LU/lu_data.jl:89-92.
template <class T>
struct BlocksT {
  std::vector<T> v;
  void alloc(int nx) { v.assign((size_t)25 * (nx + 2), T(0.0)); }
  T& operator()(int m, int n, int i) { return v[(size_t)(m - 1) + 5 * ((n - 1) + 5 * (size_t)i)]; }
};

template <class T>
struct ZoneT {
  int nx, ny, nz;
  FieldT<T> u, rsd, frct, flux;
  BlocksT<T> a, b, c, d;
  std::vector<T> tv;  // 5 x nx, src/LU/ssor.jl:49
  void alloc() {
    u.alloc(nx, ny, nz); rsd.alloc(nx, ny, nz); frct.alloc(nx, ny, nz); flux.alloc(nx, ny, nz);
    a.alloc(nx); b.alloc(nx); c.alloc(nx); d.alloc(nx);
    tv.assign((size_t)5 * (nx + 2), T(0.0));
  }
  T& TV(int m, int i) { return tv[(size_t)(m - 1) + 5 * (size_t)i]; }
};
using Zone = ZoneT<double>;

// exact, src/LU/exact.jl:7-22 (nested products; the same values as BT's
// exact_solution, whose products only have their operands swapped).
static void exact(int i, int j, int k, double* u, int nx0, int ny0, int nz0) {
  double xi = double(i - 1) / (nx0 - 1);
  double eta = double(j - 1) / (ny0 - 1);
  double zeta = double(k - 1) / (nz0 - 1);
  for (int m = 0; m < 5; ++m) {
    const double* ce = CE[m];
    u[m] = ce[0] + (ce[1] + (ce[4] + (ce[7] + ce[10] * xi) * xi) * xi) * xi +
           (ce[2] + (ce[5] + (ce[8] + ce[11] * eta) * eta) * eta) * eta +
           (ce[3] + (ce[6] + (ce[9] + ce[12] * zeta) * zeta) * zeta) * zeta;
  }
}

// setbv, src/LU/setbv.jl:5-62 (every face is a physical boundary of the
// subdomain, so all six faces are set).
static void setbv(Zone& z) {
  const int nx = z.nx, ny = z.ny, nz = z.nz;
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) {
      exact(i, j, 1, z.u.at(i, j, 1), nx, ny, nz);
      exact(i, j, nz, z.u.at(i, j, nz), nx, ny, nz);
    }
  for (int k = 1; k <= nz; ++k)
    for (int i = 1; i <= nx; ++i) exact(i, 1, k, z.u.at(i, 1, k), nx, ny, nz);
  for (int k = 1; k <= nz; ++k)
    for (int i = 1; i <= nx; ++i) exact(i, ny, k, z.u.at(i, ny, k), nx, ny, nz);
  for (int k = 1; k <= nz; ++k)
    for (int j = 1; j <= ny; ++j) exact(1, j, k, z.u.at(1, j, k), nx, ny, nz);
  for (int k = 1; k <= nz; ++k)
    for (int j = 1; j <= ny; ++j) exact(nx, j, k, z.u.at(nx, j, k), nx, ny, nz);
}

// setiv, src/LU/setiv.jl:15-47: transfinite interpolation of the boundary
// values into the interior.
static void setiv(Zone& z) {
  const int nx0 = z.nx, ny0 = z.ny, nz0 = z.nz, nx = z.nx, ny = z.ny, nz = z.nz;
  double ue_1jk[5], ue_nx0jk[5], ue_i1k[5], ue_iny0k[5], ue_ij1[5], ue_ijnz[5];
  for (int k = 2; k <= nz - 1; ++k) {
    double zeta = double(k - 1) / (nz - 1);
    for (int j = 1; j <= ny; ++j) {
      if (j == 1 || j == ny0) continue;
      double eta = double(j - 1) / (ny0 - 1);
      for (int i = 1; i <= nx; ++i) {
        if (i == 1 || i == nx0) continue;
        double xi = double(i - 1) / (nx0 - 1);
        exact(1, j, k, ue_1jk, nx0, ny0, nz0);
        exact(nx0, j, k, ue_nx0jk, nx0, ny0, nz0);
        exact(i, 1, k, ue_i1k, nx0, ny0, nz0);
        exact(i, ny0, k, ue_iny0k, nx0, ny0, nz0);
        exact(i, j, 1, ue_ij1, nx0, ny0, nz0);
        exact(i, j, nz, ue_ijnz, nx0, ny0, nz0);
        for (int m = 1; m <= 5; ++m) {
          double pxi = (1.0e+00 - xi) * ue_1jk[m - 1] + xi * ue_nx0jk[m - 1];
          double peta = (1.0e+00 - eta) * ue_i1k[m - 1] + eta * ue_iny0k[m - 1];
          double pzeta = (1.0e+00 - zeta) * ue_ij1[m - 1] + zeta * ue_ijnz[m - 1];
          z.u(m, i, j, k) = pxi + peta + pzeta - pxi * peta - peta * pzeta - pzeta * pxi +
                            pxi * peta * pzeta;
        }
      }
    }
  }
}

// erhs, src/LU/erhs.jl:5-524: frct = minus the discrete operator applied to
// the exact polynomial (written term by term, not nested), staged in rsd.
static void erhs(Zone& z, const Coeff& C) {
  const int nx = z.nx, ny = z.ny, nz = z.nz, nx0 = nx, ny0 = ny;
  const int ist = 2, iend = nx - 1, jst = 2, jend = ny - 1;
  Field& frct = z.frct;
  Field& rsd = z.rsd;
  Field& flux = z.flux;
  const double dsspm = dssp;
  const double tx1 = C.tx1, tx2 = C.tx2, tx3 = C.tx3, ty1 = C.ty1, ty2 = C.ty2, ty3 = C.ty3;
  const double tz1 = C.tz1, tz2 = C.tz2, tz3 = C.tz3;
  std::fill(frct.v.begin(), frct.v.end(), 0.0);
  for (int k = 1; k <= nz; ++k) {
    double zeta = double(k - 1) / (nz - 1);
    for (int j = 1; j <= ny; ++j) {
      double eta = double(j - 1) / (ny0 - 1);
      for (int i = 1; i <= nx; ++i) {
        double xi = double(i - 1) / (nx0 - 1);
        for (int m = 1; m <= 5; ++m) {
          const double* ce = CE[m - 1];
          rsd(m, i, j, k) = ce[0] + ce[1] * xi + ce[2] * eta + ce[3] * zeta + ce[4] * xi * xi +
                            ce[5] * eta * eta + ce[6] * zeta * zeta + ce[7] * xi * xi * xi +
                            ce[8] * eta * eta * eta + ce[9] * zeta * zeta * zeta +
                            ce[10] * xi * xi * xi * xi + ce[11] * eta * eta * eta * eta +
                            ce[12] * zeta * zeta * zeta * zeta;
        }
      }
    }
  }
  // xi-direction flux differences, src/LU/erhs.jl:73-205.
  const int L1 = 1, L2 = nx, ist1 = 4, iend1 = nx - 3;
  for (int k = 2; k <= nz - 1; ++k)
    for (int j = jst; j <= jend; ++j)
      for (int i = L1; i <= L2; ++i) {
        flux(1, i, j, k) = rsd(2, i, j, k);
        double u21 = rsd(2, i, j, k) / rsd(1, i, j, k);
        double q = 0.50e+00 *
                   (rsd(2, i, j, k) * rsd(2, i, j, k) + rsd(3, i, j, k) * rsd(3, i, j, k) +
                    rsd(4, i, j, k) * rsd(4, i, j, k)) /
                   rsd(1, i, j, k);
        flux(2, i, j, k) = rsd(2, i, j, k) * u21 + c2 * (rsd(5, i, j, k) - q);
        flux(3, i, j, k) = rsd(3, i, j, k) * u21;
        flux(4, i, j, k) = rsd(4, i, j, k) * u21;
        flux(5, i, j, k) = (c1 * rsd(5, i, j, k) - c2 * q) * u21;
      }
  for (int k = 2; k <= nz - 1; ++k)
    for (int j = jst; j <= jend; ++j) {
      for (int i = ist; i <= iend; ++i)
        for (int m = 1; m <= 5; ++m)
          frct(m, i, j, k) = frct(m, i, j, k) - tx2 * (flux(m, i + 1, j, k) - flux(m, i - 1, j, k));
      for (int i = ist; i <= L2; ++i) {
        double tmp = 1.0e+00 / rsd(1, i, j, k);
        double u21i = tmp * rsd(2, i, j, k), u31i = tmp * rsd(3, i, j, k);
        double u41i = tmp * rsd(4, i, j, k), u51i = tmp * rsd(5, i, j, k);
        tmp = 1.0e+00 / rsd(1, i - 1, j, k);
        double u21im1 = tmp * rsd(2, i - 1, j, k), u31im1 = tmp * rsd(3, i - 1, j, k);
        double u41im1 = tmp * rsd(4, i - 1, j, k), u51im1 = tmp * rsd(5, i - 1, j, k);
        flux(2, i, j, k) = (4.0e+00 / 3.0e+00) * tx3 * (u21i - u21im1);
        flux(3, i, j, k) = tx3 * (u31i - u31im1);
        flux(4, i, j, k) = tx3 * (u41i - u41im1);
        flux(5, i, j, k) = 0.50e+00 * (1.0e+00 - c1 * c5) * tx3 *
                               ((sq(u21i) + sq(u31i) + sq(u41i)) - (sq(u21im1) + sq(u31im1) + sq(u41im1))) +
                           (1.0e+00 / 6.0e+00) * tx3 * (sq(u21i) - sq(u21im1)) +
                           c1 * c5 * tx3 * (u51i - u51im1);
      }
      for (int i = ist; i <= iend; ++i) {
        frct(1, i, j, k) = frct(1, i, j, k) +
                           dx1 * tx1 * (rsd(1, i - 1, j, k) - 2.0e+00 * rsd(1, i, j, k) + rsd(1, i + 1, j, k));
        frct(2, i, j, k) = frct(2, i, j, k) + tx3 * c3 * c4 * (flux(2, i + 1, j, k) - flux(2, i, j, k)) +
                           dx2 * tx1 * (rsd(2, i - 1, j, k) - 2.0e+00 * rsd(2, i, j, k) + rsd(2, i + 1, j, k));
        frct(3, i, j, k) = frct(3, i, j, k) + tx3 * c3 * c4 * (flux(3, i + 1, j, k) - flux(3, i, j, k)) +
                           dx3 * tx1 * (rsd(3, i - 1, j, k) - 2.0e+00 * rsd(3, i, j, k) + rsd(3, i + 1, j, k));
        frct(4, i, j, k) = frct(4, i, j, k) + tx3 * c3 * c4 * (flux(4, i + 1, j, k) - flux(4, i, j, k)) +
                           dx4 * tx1 * (rsd(4, i - 1, j, k) - 2.0e+00 * rsd(4, i, j, k) + rsd(4, i + 1, j, k));
        frct(5, i, j, k) = frct(5, i, j, k) + tx3 * c3 * c4 * (flux(5, i + 1, j, k) - flux(5, i, j, k)) +
                           dx5 * tx1 * (rsd(5, i - 1, j, k) - 2.0e+00 * rsd(5, i, j, k) + rsd(5, i + 1, j, k));
      }
      for (int m = 1; m <= 5; ++m) {
        frct(m, 2, j, k) = frct(m, 2, j, k) -
                           dsspm * (+5.0e+00 * rsd(m, 2, j, k) - 4.0e+00 * rsd(m, 3, j, k) + rsd(m, 4, j, k));
        frct(m, 3, j, k) = frct(m, 3, j, k) - dsspm * (-4.0e+00 * rsd(m, 2, j, k) + 6.0e+00 * rsd(m, 3, j, k) -
                                                        4.0e+00 * rsd(m, 4, j, k) + rsd(m, 5, j, k));
      }
      for (int i = ist1; i <= iend1; ++i)
        for (int m = 1; m <= 5; ++m)
          frct(m, i, j, k) = frct(m, i, j, k) -
                             dsspm * (rsd(m, i - 2, j, k) - 4.0e+00 * rsd(m, i - 1, j, k) + 6.0e+00 * rsd(m, i, j, k) -
                                      4.0e+00 * rsd(m, i + 1, j, k) + rsd(m, i + 2, j, k));
      for (int m = 1; m <= 5; ++m) {
        frct(m, nx - 2, j, k) = frct(m, nx - 2, j, k) -
                                dsspm * (rsd(m, nx - 4, j, k) - 4.0e+00 * rsd(m, nx - 3, j, k) +
                                         6.0e+00 * rsd(m, nx - 2, j, k) - 4.0e+00 * rsd(m, nx - 1, j, k));
        frct(m, nx - 1, j, k) = frct(m, nx - 1, j, k) -
                                dsspm * (rsd(m, nx - 3, j, k) - 4.0e+00 * rsd(m, nx - 2, j, k) +
                                         5.0e+00 * rsd(m, nx - 1, j, k));
      }
    }
  // eta-direction flux differences, src/LU/erhs.jl:233-377.
  const int M1 = 1, M2 = ny, jst1 = 4, jend1 = ny - 3;
  for (int k = 2; k <= nz - 1; ++k) {
    for (int j = M1; j <= M2; ++j)
      for (int i = ist; i <= iend; ++i) {
        flux(1, i, j, k) = rsd(3, i, j, k);
        double u31 = rsd(3, i, j, k) / rsd(1, i, j, k);
        double q = 0.50e+00 *
                   (rsd(2, i, j, k) * rsd(2, i, j, k) + rsd(3, i, j, k) * rsd(3, i, j, k) +
                    rsd(4, i, j, k) * rsd(4, i, j, k)) /
                   rsd(1, i, j, k);
        flux(2, i, j, k) = rsd(2, i, j, k) * u31;
        flux(3, i, j, k) = rsd(3, i, j, k) * u31 + c2 * (rsd(5, i, j, k) - q);
        flux(4, i, j, k) = rsd(4, i, j, k) * u31;
        flux(5, i, j, k) = (c1 * rsd(5, i, j, k) - c2 * q) * u31;
      }
    for (int i = ist; i <= iend; ++i)
      for (int j = jst; j <= jend; ++j)
        for (int m = 1; m <= 5; ++m)
          frct(m, i, j, k) = frct(m, i, j, k) - ty2 * (flux(m, i, j + 1, k) - flux(m, i, j - 1, k));
    for (int j = jst; j <= M2; ++j)
      for (int i = ist; i <= iend; ++i) {
        double tmp = 1.0e+00 / rsd(1, i, j, k);
        double u21j = tmp * rsd(2, i, j, k), u31j = tmp * rsd(3, i, j, k);
        double u41j = tmp * rsd(4, i, j, k), u51j = tmp * rsd(5, i, j, k);
        tmp = 1.0e+00 / rsd(1, i, j - 1, k);
        double u21jm1 = tmp * rsd(2, i, j - 1, k), u31jm1 = tmp * rsd(3, i, j - 1, k);
        double u41jm1 = tmp * rsd(4, i, j - 1, k), u51jm1 = tmp * rsd(5, i, j - 1, k);
        flux(2, i, j, k) = ty3 * (u21j - u21jm1);
        flux(3, i, j, k) = (4.0e+00 / 3.0e+00) * ty3 * (u31j - u31jm1);
        flux(4, i, j, k) = ty3 * (u41j - u41jm1);
        flux(5, i, j, k) = 0.50e+00 * (1.0e+00 - c1 * c5) * ty3 *
                               ((sq(u21j) + sq(u31j) + sq(u41j)) - (sq(u21jm1) + sq(u31jm1) + sq(u41jm1))) +
                           (1.0e+00 / 6.0e+00) * ty3 * (sq(u31j) - sq(u31jm1)) +
                           c1 * c5 * ty3 * (u51j - u51jm1);
      }
    for (int j = jst; j <= jend; ++j)
      for (int i = ist; i <= iend; ++i) {
        frct(1, i, j, k) = frct(1, i, j, k) +
                           dy1 * ty1 * (rsd(1, i, j - 1, k) - 2.0e+00 * rsd(1, i, j, k) + rsd(1, i, j + 1, k));
        frct(2, i, j, k) = frct(2, i, j, k) + ty3 * c3 * c4 * (flux(2, i, j + 1, k) - flux(2, i, j, k)) +
                           dy2 * ty1 * (rsd(2, i, j - 1, k) - 2.0e+00 * rsd(2, i, j, k) + rsd(2, i, j + 1, k));
        frct(3, i, j, k) = frct(3, i, j, k) + ty3 * c3 * c4 * (flux(3, i, j + 1, k) - flux(3, i, j, k)) +
                           dy3 * ty1 * (rsd(3, i, j - 1, k) - 2.0e+00 * rsd(3, i, j, k) + rsd(3, i, j + 1, k));
        frct(4, i, j, k) = frct(4, i, j, k) + ty3 * c3 * c4 * (flux(4, i, j + 1, k) - flux(4, i, j, k)) +
                           dy4 * ty1 * (rsd(4, i, j - 1, k) - 2.0e+00 * rsd(4, i, j, k) + rsd(4, i, j + 1, k));
        frct(5, i, j, k) = frct(5, i, j, k) + ty3 * c3 * c4 * (flux(5, i, j + 1, k) - flux(5, i, j, k)) +
                           dy5 * ty1 * (rsd(5, i, j - 1, k) - 2.0e+00 * rsd(5, i, j, k) + rsd(5, i, j + 1, k));
      }
    for (int i = ist; i <= iend; ++i)
      for (int m = 1; m <= 5; ++m) {
        frct(m, i, 2, k) = frct(m, i, 2, k) -
                           dsspm * (+5.0e+00 * rsd(m, i, 2, k) - 4.0e+00 * rsd(m, i, 3, k) + rsd(m, i, 4, k));
        frct(m, i, 3, k) = frct(m, i, 3, k) - dsspm * (-4.0e+00 * rsd(m, i, 2, k) + 6.0e+00 * rsd(m, i, 3, k) -
                                                        4.0e+00 * rsd(m, i, 4, k) + rsd(m, i, 5, k));
      }
    for (int j = jst1; j <= jend1; ++j)
      for (int i = ist; i <= iend; ++i)
        for (int m = 1; m <= 5; ++m)
          frct(m, i, j, k) = frct(m, i, j, k) -
                             dsspm * (rsd(m, i, j - 2, k) - 4.0e+00 * rsd(m, i, j - 1, k) + 6.0e+00 * rsd(m, i, j, k) -
                                      4.0e+00 * rsd(m, i, j + 1, k) + rsd(m, i, j + 2, k));
    for (int i = ist; i <= iend; ++i)
      for (int m = 1; m <= 5; ++m) {
        frct(m, i, ny - 2, k) = frct(m, i, ny - 2, k) -
                                dsspm * (rsd(m, i, ny - 4, k) - 4.0e+00 * rsd(m, i, ny - 3, k) +
                                         6.0e+00 * rsd(m, i, ny - 2, k) - 4.0e+00 * rsd(m, i, ny - 1, k));
        frct(m, i, ny - 1, k) = frct(m, i, ny - 1, k) -
                                dsspm * (rsd(m, i, ny - 3, k) - 4.0e+00 * rsd(m, i, ny - 2, k) +
                                         5.0e+00 * rsd(m, i, ny - 1, k));
      }
  }
  // zeta-direction flux differences, src/LU/erhs.jl:382-520.
  for (int k = 1; k <= nz; ++k)
    for (int j = jst; j <= jend; ++j)
      for (int i = ist; i <= iend; ++i) {
        flux(1, i, j, k) = rsd(4, i, j, k);
        double u41 = rsd(4, i, j, k) / rsd(1, i, j, k);
        double q = 0.50e+00 *
                   (rsd(2, i, j, k) * rsd(2, i, j, k) + rsd(3, i, j, k) * rsd(3, i, j, k) +
                    rsd(4, i, j, k) * rsd(4, i, j, k)) /
                   rsd(1, i, j, k);
        flux(2, i, j, k) = rsd(2, i, j, k) * u41;
        flux(3, i, j, k) = rsd(3, i, j, k) * u41;
        flux(4, i, j, k) = rsd(4, i, j, k) * u41 + c2 * (rsd(5, i, j, k) - q);
        flux(5, i, j, k) = (c1 * rsd(5, i, j, k) - c2 * q) * u41;
      }
  for (int k = 2; k <= nz - 1; ++k)
    for (int j = jst; j <= jend; ++j)
      for (int i = ist; i <= iend; ++i)
        for (int m = 1; m <= 5; ++m)
          frct(m, i, j, k) = frct(m, i, j, k) - tz2 * (flux(m, i, j, k + 1) - flux(m, i, j, k - 1));
  for (int k = 2; k <= nz; ++k)
    for (int j = jst; j <= jend; ++j)
      for (int i = ist; i <= iend; ++i) {
        double tmp = 1.0e+00 / rsd(1, i, j, k);
        double u21k = tmp * rsd(2, i, j, k), u31k = tmp * rsd(3, i, j, k);
        double u41k = tmp * rsd(4, i, j, k), u51k = tmp * rsd(5, i, j, k);
        tmp = 1.0e+00 / rsd(1, i, j, k - 1);
        double u21km1 = tmp * rsd(2, i, j, k - 1), u31km1 = tmp * rsd(3, i, j, k - 1);
        double u41km1 = tmp * rsd(4, i, j, k - 1), u51km1 = tmp * rsd(5, i, j, k - 1);
        flux(2, i, j, k) = tz3 * (u21k - u21km1);
        flux(3, i, j, k) = tz3 * (u31k - u31km1);
        flux(4, i, j, k) = (4.0e+00 / 3.0e+00) * tz3 * (u41k - u41km1);
        flux(5, i, j, k) = 0.50e+00 * (1.0e+00 - c1 * c5) * tz3 *
                               ((sq(u21k) + sq(u31k) + sq(u41k)) - (sq(u21km1) + sq(u31km1) + sq(u41km1))) +
                           (1.0e+00 / 6.0e+00) * tz3 * (sq(u41k) - sq(u41km1)) +
                           c1 * c5 * tz3 * (u51k - u51km1);
      }
  for (int k = 2; k <= nz - 1; ++k)
    for (int j = jst; j <= jend; ++j)
      for (int i = ist; i <= iend; ++i) {
        frct(1, i, j, k) = frct(1, i, j, k) +
                           dz1 * tz1 * (rsd(1, i, j, k + 1) - 2.0e+00 * rsd(1, i, j, k) + rsd(1, i, j, k - 1));
        frct(2, i, j, k) = frct(2, i, j, k) + tz3 * c3 * c4 * (flux(2, i, j, k + 1) - flux(2, i, j, k)) +
                           dz2 * tz1 * (rsd(2, i, j, k + 1) - 2.0e+00 * rsd(2, i, j, k) + rsd(2, i, j, k - 1));
        frct(3, i, j, k) = frct(3, i, j, k) + tz3 * c3 * c4 * (flux(3, i, j, k + 1) - flux(3, i, j, k)) +
                           dz3 * tz1 * (rsd(3, i, j, k + 1) - 2.0e+00 * rsd(3, i, j, k) + rsd(3, i, j, k - 1));
        frct(4, i, j, k) = frct(4, i, j, k) + tz3 * c3 * c4 * (flux(4, i, j, k + 1) - flux(4, i, j, k)) +
                           dz4 * tz1 * (rsd(4, i, j, k + 1) - 2.0e+00 * rsd(4, i, j, k) + rsd(4, i, j, k - 1));
        frct(5, i, j, k) = frct(5, i, j, k) + tz3 * c3 * c4 * (flux(5, i, j, k + 1) - flux(5, i, j, k)) +
                           dz5 * tz1 * (rsd(5, i, j, k + 1) - 2.0e+00 * rsd(5, i, j, k) + rsd(5, i, j, k - 1));
      }
  for (int j = jst; j <= jend; ++j)
    for (int i = ist; i <= iend; ++i)
      for (int m = 1; m <= 5; ++m) {
        frct(m, i, j, 2) = frct(m, i, j, 2) -
                           dsspm * (+5.0e+00 * rsd(m, i, j, 2) - 4.0e+00 * rsd(m, i, j, 3) + rsd(m, i, j, 4));
        frct(m, i, j, 3) = frct(m, i, j, 3) - dsspm * (-4.0e+00 * rsd(m, i, j, 2) + 6.0e+00 * rsd(m, i, j, 3) -
                                                        4.0e+00 * rsd(m, i, j, 4) + rsd(m, i, j, 5));
      }
  for (int k = 4; k <= nz - 3; ++k)
    for (int j = jst; j <= jend; ++j)
      for (int i = ist; i <= iend; ++i)
        for (int m = 1; m <= 5; ++m)
          frct(m, i, j, k) = frct(m, i, j, k) -
                             dsspm * (rsd(m, i, j, k - 2) - 4.0e+00 * rsd(m, i, j, k - 1) + 6.0e+00 * rsd(m, i, j, k) -
                                      4.0e+00 * rsd(m, i, j, k + 1) + rsd(m, i, j, k + 2));
  for (int j = jst; j <= jend; ++j)
    for (int i = ist; i <= iend; ++i)
      for (int m = 1; m <= 5; ++m) {
        frct(m, i, j, nz - 2) = frct(m, i, j, nz - 2) -
                                dsspm * (rsd(m, i, j, nz - 4) - 4.0e+00 * rsd(m, i, j, nz - 3) +
                                         6.0e+00 * rsd(m, i, j, nz - 2) - 4.0e+00 * rsd(m, i, j, nz - 1));
        frct(m, i, j, nz - 1) = frct(m, i, j, nz - 1) -
                                dsspm * (rsd(m, i, j, nz - 3) - 4.0e+00 * rsd(m, i, j, nz - 2) +
                                         5.0e+00 * rsd(m, i, j, nz - 1));
      }
}

// rhs, src/LU/rhs.jl:5-519: rsd = -frct everywhere, then the flux
// differences, viscous terms and fourth-order dissipation on the interior.
template <class T>
static void rhs(ZoneT<T>& z, const Coeff& C) {
  const int nx = z.nx, ny = z.ny, nz = z.nz;
  const int ist = 2, iend = nx - 1, jst = 2, jend = ny - 1;
  FieldT<T>& u = z.u;
  FieldT<T>& rsd = z.rsd;
  FieldT<T>& frct = z.frct;
  FieldT<T>& flux = z.flux;
  const double tx1 = C.tx1, tx2 = C.tx2, tx3 = C.tx3, ty1 = C.ty1, ty2 = C.ty2, ty3 = C.ty3;
  const double tz1 = C.tz1, tz2 = C.tz2, tz3 = C.tz3;
  for (size_t p = 0; p < rsd.v.size(); ++p) rsd.v[p] = -frct.v[p];
  // xi-direction, src/LU/rhs.jl:76-206.
  const int L1 = 1, L2 = nx, ist1 = 4, iend1 = nx - 3;
  for (int k = 2; k <= nz - 1; ++k)
    for (int j = jst; j <= jend; ++j) {
      for (int i = L1; i <= L2; ++i) {
        flux(1, i, j, k) = u(2, i, j, k);
        T u21 = u(2, i, j, k) / u(1, i, j, k);
        T q = 0.50e+00 * (u(2, i, j, k) * u(2, i, j, k) + u(3, i, j, k) * u(3, i, j, k) +
                               u(4, i, j, k) * u(4, i, j, k)) /
                   u(1, i, j, k);
        flux(2, i, j, k) = u(2, i, j, k) * u21 + c2 * (u(5, i, j, k) - q);
        flux(3, i, j, k) = u(3, i, j, k) * u21;
        flux(4, i, j, k) = u(4, i, j, k) * u21;
        flux(5, i, j, k) = (c1 * u(5, i, j, k) - c2 * q) * u21;
      }
      for (int i = ist; i <= iend; ++i)
        for (int m = 1; m <= 5; ++m)
          rsd(m, i, j, k) -= tx2 * (flux(m, i + 1, j, k) - flux(m, i - 1, j, k));
      for (int i = ist; i <= L2; ++i) {
        T tmp = 1.0e+00 / u(1, i, j, k);
        T u21i = tmp * u(2, i, j, k), u31i = tmp * u(3, i, j, k);
        T u41i = tmp * u(4, i, j, k), u51i = tmp * u(5, i, j, k);
        tmp = 1.0e+00 / u(1, i - 1, j, k);
        T u21im1 = tmp * u(2, i - 1, j, k), u31im1 = tmp * u(3, i - 1, j, k);
        T u41im1 = tmp * u(4, i - 1, j, k), u51im1 = tmp * u(5, i - 1, j, k);
        flux(2, i, j, k) = (4.0e+00 / 3.0e+00) * tx3 * (u21i - u21im1);
        flux(3, i, j, k) = tx3 * (u31i - u31im1);
        flux(4, i, j, k) = tx3 * (u41i - u41im1);
        flux(5, i, j, k) = 0.50e+00 * (1.0e+00 - c1 * c5) * tx3 *
                               ((sq(u21i) + sq(u31i) + sq(u41i)) - (sq(u21im1) + sq(u31im1) + sq(u41im1))) +
                           (1.0e+00 / 6.0e+00) * tx3 * (sq(u21i) - sq(u21im1)) +
                           c1 * c5 * tx3 * (u51i - u51im1);
      }
      for (int i = ist; i <= iend; ++i) {
        rsd(1, i, j, k) += dx1 * tx1 * (u(1, i - 1, j, k) - 2.0e+00 * u(1, i, j, k) + u(1, i + 1, j, k));
        rsd(2, i, j, k) += tx3 * c3 * c4 * (flux(2, i + 1, j, k) - flux(2, i, j, k)) +
                           dx2 * tx1 * (u(2, i - 1, j, k) - 2.0e+00 * u(2, i, j, k) + u(2, i + 1, j, k));
        rsd(3, i, j, k) += tx3 * c3 * c4 * (flux(3, i + 1, j, k) - flux(3, i, j, k)) +
                           dx3 * tx1 * (u(3, i - 1, j, k) - 2.0e+00 * u(3, i, j, k) + u(3, i + 1, j, k));
        rsd(4, i, j, k) += tx3 * c3 * c4 * (flux(4, i + 1, j, k) - flux(4, i, j, k)) +
                           dx4 * tx1 * (u(4, i - 1, j, k) - 2.0e+00 * u(4, i, j, k) + u(4, i + 1, j, k));
        rsd(5, i, j, k) += tx3 * c3 * c4 * (flux(5, i + 1, j, k) - flux(5, i, j, k)) +
                           dx5 * tx1 * (u(5, i - 1, j, k) - 2.0e+00 * u(5, i, j, k) + u(5, i + 1, j, k));
      }
      for (int m = 1; m <= 5; ++m) {
        rsd(m, 2, j, k) -= dssp * (+5.0e+00 * u(m, 2, j, k) - 4.0e+00 * u(m, 3, j, k) + u(m, 4, j, k));
        rsd(m, 3, j, k) -= dssp * (-4.0e+00 * u(m, 2, j, k) + 6.0e+00 * u(m, 3, j, k) -
                                   4.0e+00 * u(m, 4, j, k) + u(m, 5, j, k));
      }
      for (int i = ist1; i <= iend1; ++i)
        for (int m = 1; m <= 5; ++m)
          rsd(m, i, j, k) -= dssp * (u(m, i - 2, j, k) - 4.0e+00 * u(m, i - 1, j, k) + 6.0e+00 * u(m, i, j, k) -
                                     4.0e+00 * u(m, i + 1, j, k) + u(m, i + 2, j, k));
      for (int m = 1; m <= 5; ++m) {
        rsd(m, nx - 2, j, k) -= dssp * (u(m, nx - 4, j, k) - 4.0e+00 * u(m, nx - 3, j, k) +
                                        6.0e+00 * u(m, nx - 2, j, k) - 4.0e+00 * u(m, nx - 1, j, k));
        rsd(m, nx - 1, j, k) -= dssp * (u(m, nx - 3, j, k) - 4.0e+00 * u(m, nx - 2, j, k) +
                                        5.0e+00 * u(m, nx - 1, j, k));
      }
    }
  // eta-direction, src/LU/rhs.jl:224-370.
  const int M1 = 1, M2 = ny, jst1 = 4, jend1 = ny - 3;
  for (int k = 2; k <= nz - 1; ++k) {
    for (int j = M1; j <= M2; ++j)
      for (int i = ist; i <= iend; ++i) {
        flux(1, i, j, k) = u(3, i, j, k);
        T u31 = u(3, i, j, k) / u(1, i, j, k);
        T q = 0.50e+00 * (u(2, i, j, k) * u(2, i, j, k) + u(3, i, j, k) * u(3, i, j, k) +
                               u(4, i, j, k) * u(4, i, j, k)) /
                   u(1, i, j, k);
        flux(2, i, j, k) = u(2, i, j, k) * u31;
        flux(3, i, j, k) = u(3, i, j, k) * u31 + c2 * (u(5, i, j, k) - q);
        flux(4, i, j, k) = u(4, i, j, k) * u31;
        flux(5, i, j, k) = (c1 * u(5, i, j, k) - c2 * q) * u31;
      }
    for (int j = jst; j <= jend; ++j)
      for (int i = ist; i <= iend; ++i)
        for (int m = 1; m <= 5; ++m)
          rsd(m, i, j, k) -= ty2 * (flux(m, i, j + 1, k) - flux(m, i, j - 1, k));
    for (int j = jst; j <= M2; ++j)
      for (int i = ist; i <= iend; ++i) {
        T tmp = 1.0e+00 / u(1, i, j, k);
        T u21j = tmp * u(2, i, j, k), u31j = tmp * u(3, i, j, k);
        T u41j = tmp * u(4, i, j, k), u51j = tmp * u(5, i, j, k);
        tmp = 1.0e+00 / u(1, i, j - 1, k);
        T u21jm1 = tmp * u(2, i, j - 1, k), u31jm1 = tmp * u(3, i, j - 1, k);
        T u41jm1 = tmp * u(4, i, j - 1, k), u51jm1 = tmp * u(5, i, j - 1, k);
        flux(2, i, j, k) = ty3 * (u21j - u21jm1);
        flux(3, i, j, k) = (4.0e+00 / 3.0e+00) * ty3 * (u31j - u31jm1);
        flux(4, i, j, k) = ty3 * (u41j - u41jm1);
        flux(5, i, j, k) = 0.50e+00 * (1.0e+00 - c1 * c5) * ty3 *
                               ((sq(u21j) + sq(u31j) + sq(u41j)) - (sq(u21jm1) + sq(u31jm1) + sq(u41jm1))) +
                           (1.0e+00 / 6.0e+00) * ty3 * (sq(u31j) - sq(u31jm1)) +
                           c1 * c5 * ty3 * (u51j - u51jm1);
      }
    for (int j = jst; j <= jend; ++j)
      for (int i = ist; i <= iend; ++i) {
        rsd(1, i, j, k) += dy1 * ty1 * (u(1, i, j - 1, k) - 2.0e+00 * u(1, i, j, k) + u(1, i, j + 1, k));
        rsd(2, i, j, k) += ty3 * c3 * c4 * (flux(2, i, j + 1, k) - flux(2, i, j, k)) +
                           dy2 * ty1 * (u(2, i, j - 1, k) - 2.0e+00 * u(2, i, j, k) + u(2, i, j + 1, k));
        rsd(3, i, j, k) += ty3 * c3 * c4 * (flux(3, i, j + 1, k) - flux(3, i, j, k)) +
                           dy3 * ty1 * (u(3, i, j - 1, k) - 2.0e+00 * u(3, i, j, k) + u(3, i, j + 1, k));
        rsd(4, i, j, k) += ty3 * c3 * c4 * (flux(4, i, j + 1, k) - flux(4, i, j, k)) +
                           dy4 * ty1 * (u(4, i, j - 1, k) - 2.0e+00 * u(4, i, j, k) + u(4, i, j + 1, k));
        rsd(5, i, j, k) += ty3 * c3 * c4 * (flux(5, i, j + 1, k) - flux(5, i, j, k)) +
                           dy5 * ty1 * (u(5, i, j - 1, k) - 2.0e+00 * u(5, i, j, k) + u(5, i, j + 1, k));
      }
    for (int i = ist; i <= iend; ++i)
      for (int m = 1; m <= 5; ++m) {
        rsd(m, i, 2, k) -= dssp * (+5.0e+00 * u(m, i, 2, k) - 4.0e+00 * u(m, i, 3, k) + u(m, i, 4, k));
        rsd(m, i, 3, k) -= dssp * (-4.0e+00 * u(m, i, 2, k) + 6.0e+00 * u(m, i, 3, k) -
                                   4.0e+00 * u(m, i, 4, k) + u(m, i, 5, k));
      }
    for (int j = jst1; j <= jend1; ++j)
      for (int i = ist; i <= iend; ++i)
        for (int m = 1; m <= 5; ++m)
          rsd(m, i, j, k) -= dssp * (u(m, i, j - 2, k) - 4.0e+00 * u(m, i, j - 1, k) + 6.0e+00 * u(m, i, j, k) -
                                     4.0e+00 * u(m, i, j + 1, k) + u(m, i, j + 2, k));
    for (int i = ist; i <= iend; ++i)
      for (int m = 1; m <= 5; ++m) {
        rsd(m, i, ny - 2, k) -= dssp * (u(m, i, ny - 4, k) - 4.0e+00 * u(m, i, ny - 3, k) +
                                        6.0e+00 * u(m, i, ny - 2, k) - 4.0e+00 * u(m, i, ny - 1, k));
        rsd(m, i, ny - 1, k) -= dssp * (u(m, i, ny - 3, k) - 4.0e+00 * u(m, i, ny - 2, k) +
                                        5.0e+00 * u(m, i, ny - 1, k));
      }
  }
  // zeta-direction, src/LU/rhs.jl:375-517.
  for (int k = 1; k <= nz; ++k)
    for (int j = jst; j <= jend; ++j)
      for (int i = ist; i <= iend; ++i) {
        flux(1, i, j, k) = u(4, i, j, k);
        T u41 = u(4, i, j, k) / u(1, i, j, k);
        T q = 0.50e+00 * (u(2, i, j, k) * u(2, i, j, k) + u(3, i, j, k) * u(3, i, j, k) +
                               u(4, i, j, k) * u(4, i, j, k)) /
                   u(1, i, j, k);
        flux(2, i, j, k) = u(2, i, j, k) * u41;
        flux(3, i, j, k) = u(3, i, j, k) * u41;
        flux(4, i, j, k) = u(4, i, j, k) * u41 + c2 * (u(5, i, j, k) - q);
        flux(5, i, j, k) = (c1 * u(5, i, j, k) - c2 * q) * u41;
      }
  for (int k = 2; k <= nz - 1; ++k)
    for (int j = jst; j <= jend; ++j)
      for (int i = ist; i <= iend; ++i)
        for (int m = 1; m <= 5; ++m)
          rsd(m, i, j, k) -= tz2 * (flux(m, i, j, k + 1) - flux(m, i, j, k - 1));
  for (int k = 2; k <= nz; ++k)
    for (int j = jst; j <= jend; ++j)
      for (int i = ist; i <= iend; ++i) {
        T tmp = 1.0e+00 / u(1, i, j, k);
        T u21k = tmp * u(2, i, j, k), u31k = tmp * u(3, i, j, k);
        T u41k = tmp * u(4, i, j, k), u51k = tmp * u(5, i, j, k);
        tmp = 1.0e+00 / u(1, i, j, k - 1);
        T u21km1 = tmp * u(2, i, j, k - 1), u31km1 = tmp * u(3, i, j, k - 1);
        T u41km1 = tmp * u(4, i, j, k - 1), u51km1 = tmp * u(5, i, j, k - 1);
        flux(2, i, j, k) = tz3 * (u21k - u21km1);
        flux(3, i, j, k) = tz3 * (u31k - u31km1);
        flux(4, i, j, k) = (4.0e+00 / 3.0e+00) * tz3 * (u41k - u41km1);
        flux(5, i, j, k) = 0.50e+00 * (1.0e+00 - c1 * c5) * tz3 *
                               ((sq(u21k) + sq(u31k) + sq(u41k)) - (sq(u21km1) + sq(u31km1) + sq(u41km1))) +
                           (1.0e+00 / 6.0e+00) * tz3 * (sq(u41k) - sq(u41km1)) +
                           c1 * c5 * tz3 * (u51k - u51km1);
      }
  for (int k = 2; k <= nz - 1; ++k)
    for (int j = jst; j <= jend; ++j)
      for (int i = ist; i <= iend; ++i) {
        rsd(1, i, j, k) += dz1 * tz1 * (u(1, i, j, k - 1) - 2.0e+00 * u(1, i, j, k) + u(1, i, j, k + 1));
        rsd(2, i, j, k) += tz3 * c3 * c4 * (flux(2, i, j, k + 1) - flux(2, i, j, k)) +
                           dz2 * tz1 * (u(2, i, j, k - 1) - 2.0e+00 * u(2, i, j, k) + u(2, i, j, k + 1));
        rsd(3, i, j, k) += tz3 * c3 * c4 * (flux(3, i, j, k + 1) - flux(3, i, j, k)) +
                           dz3 * tz1 * (u(3, i, j, k - 1) - 2.0e+00 * u(3, i, j, k) + u(3, i, j, k + 1));
        rsd(4, i, j, k) += tz3 * c3 * c4 * (flux(4, i, j, k + 1) - flux(4, i, j, k)) +
                           dz4 * tz1 * (u(4, i, j, k - 1) - 2.0e+00 * u(4, i, j, k) + u(4, i, j, k + 1));
        rsd(5, i, j, k) += tz3 * c3 * c4 * (flux(5, i, j, k + 1) - flux(5, i, j, k)) +
                           dz5 * tz1 * (u(5, i, j, k - 1) - 2.0e+00 * u(5, i, j, k) + u(5, i, j, k + 1));
      }
  for (int j = jst; j <= jend; ++j)
    for (int i = ist; i <= iend; ++i)
      for (int m = 1; m <= 5; ++m) {
        rsd(m, i, j, 2) -= dssp * (+5.0e+00 * u(m, i, j, 2) - 4.0e+00 * u(m, i, j, 3) + u(m, i, j, 4));
        rsd(m, i, j, 3) -= dssp * (-4.0e+00 * u(m, i, j, 2) + 6.0e+00 * u(m, i, j, 3) -
                                   4.0e+00 * u(m, i, j, 4) + u(m, i, j, 5));
      }
  for (int k = 4; k <= nz - 3; ++k)
    for (int j = jst; j <= jend; ++j)
      for (int i = ist; i <= iend; ++i)
        for (int m = 1; m <= 5; ++m)
          rsd(m, i, j, k) -= dssp * (u(m, i, j, k - 2) - 4.0e+00 * u(m, i, j, k - 1) + 6.0e+00 * u(m, i, j, k) -
                                     4.0e+00 * u(m, i, j, k + 1) + u(m, i, j, k + 2));
  for (int j = jst; j <= jend; ++j)
    for (int i = ist; i <= iend; ++i)
      for (int m = 1; m <= 5; ++m) {
        rsd(m, i, j, nz - 2) -= dssp * (u(m, i, j, nz - 4) - 4.0e+00 * u(m, i, j, nz - 3) +
                                        6.0e+00 * u(m, i, j, nz - 2) - 4.0e+00 * u(m, i, j, nz - 1));
        rsd(m, i, j, nz - 1) -= dssp * (u(m, i, j, nz - 3) - 4.0e+00 * u(m, i, j, nz - 2) +
                                        5.0e+00 * u(m, i, j, nz - 1));
      }
}

// The four 5x5 blocks of one point.  jacld (src/LU/jacld.jl:5-370) and jacu
// (src/LU/jacu.jl:5-370) compute the same diagonal block d, and neighbour
// blocks that differ only in the sign of the convective (t?2) part: every
// jacld entry starts "- dt * t?2 * ..." where jacu's starts "dt * t?2 * ...".
// s = -1 gives jacld's blocks, s = +1 jacu's; (s*dt)*t?2 rounds exactly like
// (-dt)*t?2 and dt*t?2, so both are reproduced bit for bit.
struct Jac {
  double dt, tx1, tx2, ty1, ty2, tz1, tz2;
  double r43, c1345, c34;
};

static Jac make_jac(const Coeff& C, double dt) {
  Jac J;
  J.dt = dt; J.tx1 = C.tx1; J.tx2 = C.tx2; J.ty1 = C.ty1; J.ty2 = C.ty2; J.tz1 = C.tz1; J.tz2 = C.tz2;
  J.r43 = (4.0e+00 / 3.0e+00);
  J.c1345 = c1 * c3 * c4 * c5;
  J.c34 = c3 * c4;
  return J;
}

// d block at the point: src/LU/jacld.jl:31-123 (src/LU/jacu.jl the same).
template <class T>
static void block_d(const Jac& J, const T* uu, T D[5][5]) {
  const double dt = J.dt, tx1 = J.tx1, ty1 = J.ty1, tz1 = J.tz1, r43 = J.r43, c1345 = J.c1345, c34 = J.c34;
  const T u2 = uu[1], u3 = uu[2], u4 = uu[3], u5 = uu[4];
  T tmp1 = 1.0e+00 / uu[0];
  T tmp2 = tmp1 * tmp1;
  T tmp3 = tmp1 * tmp2;
  for (int m = 0; m < 5; ++m)
    for (int n = 0; n < 5; ++n) D[m][n] = 0.0;
  D[0][0] = 1.0e+00 + dt * 2.0e+00 * (tx1 * dx1 + ty1 * dy1 + tz1 * dz1);
  D[1][0] = dt * 2.0e+00 *
            (tx1 * (-r43 * c34 * tmp2 * u2) + ty1 * (-c34 * tmp2 * u2) + tz1 * (-c34 * tmp2 * u2));
  D[1][1] = 1.0e+00 + dt * 2.0e+00 * (tx1 * r43 * c34 * tmp1 + ty1 * c34 * tmp1 + tz1 * c34 * tmp1) +
            dt * 2.0e+00 * (tx1 * dx2 + ty1 * dy2 + tz1 * dz2);
  D[2][0] = dt * 2.0e+00 *
            (tx1 * (-c34 * tmp2 * u3) + ty1 * (-r43 * c34 * tmp2 * u3) + tz1 * (-c34 * tmp2 * u3));
  D[2][2] = 1.0e+00 + dt * 2.0e+00 * (tx1 * c34 * tmp1 + ty1 * r43 * c34 * tmp1 + tz1 * c34 * tmp1) +
            dt * 2.0e+00 * (tx1 * dx3 + ty1 * dy3 + tz1 * dz3);
  D[3][0] = dt * 2.0e+00 *
            (tx1 * (-c34 * tmp2 * u4) + ty1 * (-c34 * tmp2 * u4) + tz1 * (-r43 * c34 * tmp2 * u4));
  D[3][3] = 1.0e+00 + dt * 2.0e+00 * (tx1 * c34 * tmp1 + ty1 * c34 * tmp1 + tz1 * r43 * c34 * tmp1) +
            dt * 2.0e+00 * (tx1 * dx4 + ty1 * dy4 + tz1 * dz4);
  D[4][0] = dt * 2.0e+00 *
            (tx1 * (-(r43 * c34 - c1345) * tmp3 * sq(u2) - (c34 - c1345) * tmp3 * sq(u3) -
                    (c34 - c1345) * tmp3 * sq(u4) - (c1345)*tmp2 * u5) +
             ty1 * (-(c34 - c1345) * tmp3 * sq(u2) - (r43 * c34 - c1345) * tmp3 * sq(u3) -
                    (c34 - c1345) * tmp3 * sq(u4) - (c1345)*tmp2 * u5) +
             tz1 * (-(c34 - c1345) * tmp3 * sq(u2) - (c34 - c1345) * tmp3 * sq(u3) -
                    (r43 * c34 - c1345) * tmp3 * sq(u4) - (c1345)*tmp2 * u5));
  D[4][1] = dt * 2.0e+00 *
            (tx1 * (r43 * c34 - c1345) * tmp2 * u2 + ty1 * (c34 - c1345) * tmp2 * u2 +
             tz1 * (c34 - c1345) * tmp2 * u2);
  D[4][2] = dt * 2.0e+00 *
            (tx1 * (c34 - c1345) * tmp2 * u3 + ty1 * (r43 * c34 - c1345) * tmp2 * u3 +
             tz1 * (c34 - c1345) * tmp2 * u3);
  D[4][3] = dt * 2.0e+00 *
            (tx1 * (c34 - c1345) * tmp2 * u4 + ty1 * (c34 - c1345) * tmp2 * u4 +
             tz1 * (r43 * c34 - c1345) * tmp2 * u4);
  D[4][4] = 1.0e+00 + dt * 2.0e+00 * (tx1 * c1345 * tmp1 + ty1 * c1345 * tmp1 + tz1 * c1345 * tmp1) +
            dt * 2.0e+00 * (tx1 * dx5 + ty1 * dy5 + tz1 * dz5);
}

// xi-neighbour block from u at (i-1) for jacld's c (src/LU/jacld.jl:288-364)
// or at (i+1) for jacu's a (src/LU/jacu.jl, the udx block).
template <class T>
static void block_x(const Jac& J, double s, const T* uu, T B[5][5]) {
  const double dt = J.dt, sdt = s * J.dt, tx1 = J.tx1, tx2 = J.tx2, r43 = J.r43, c1345 = J.c1345, c34 = J.c34;
  const T u2 = uu[1], u3 = uu[2], u4 = uu[3], u5 = uu[4];
  T tmp1 = 1.0e+00 / uu[0];
  T tmp2 = tmp1 * tmp1;
  T tmp3 = tmp1 * tmp2;
  for (int m = 0; m < 5; ++m)
    for (int n = 0; n < 5; ++n) B[m][n] = 0.0;
  B[0][0] = -dt * tx1 * dx1;
  B[0][1] = sdt * tx2;
  B[1][0] = sdt * tx2 * (-sq(u2 * tmp1) + c2 * 0.50e+00 * (u2 * u2 + u3 * u3 + u4 * u4) * tmp2) -
            dt * tx1 * (-r43 * c34 * tmp2 * u2);
  B[1][1] = sdt * tx2 * ((2.0e+00 - c2) * (u2 * tmp1)) - dt * tx1 * (r43 * c34 * tmp1) - dt * tx1 * dx2;
  B[1][2] = sdt * tx2 * (-c2 * (u3 * tmp1));
  B[1][3] = sdt * tx2 * (-c2 * (u4 * tmp1));
  B[1][4] = sdt * tx2 * c2;
  B[2][0] = sdt * tx2 * (-(u2 * u3) * tmp2) - dt * tx1 * (-c34 * tmp2 * u3);
  B[2][1] = sdt * tx2 * (u3 * tmp1);
  B[2][2] = sdt * tx2 * (u2 * tmp1) - dt * tx1 * (c34 * tmp1) - dt * tx1 * dx3;
  B[3][0] = sdt * tx2 * (-(u2 * u4) * tmp2) - dt * tx1 * (-c34 * tmp2 * u4);
  B[3][1] = sdt * tx2 * (u4 * tmp1);
  B[3][3] = sdt * tx2 * (u2 * tmp1) - dt * tx1 * (c34 * tmp1) - dt * tx1 * dx4;
  B[4][0] = sdt * tx2 * ((c2 * (u2 * u2 + u3 * u3 + u4 * u4) * tmp2 - c1 * (u5 * tmp1)) * (u2 * tmp1)) -
            dt * tx1 *
                (-(r43 * c34 - c1345) * tmp3 * sq(u2) - (c34 - c1345) * tmp3 * sq(u3) -
                 (c34 - c1345) * tmp3 * sq(u4) - c1345 * tmp2 * u5);
  B[4][1] = sdt * tx2 * (c1 * (u5 * tmp1) - 0.50e+00 * c2 * ((3.0e+00 * u2 * u2 + u3 * u3 + u4 * u4) * tmp2)) -
            dt * tx1 * (r43 * c34 - c1345) * tmp2 * u2;
  B[4][2] = sdt * tx2 * (-c2 * (u3 * u2) * tmp2) - dt * tx1 * (c34 - c1345) * tmp2 * u3;
  B[4][3] = sdt * tx2 * (-c2 * (u4 * u2) * tmp2) - dt * tx1 * (c34 - c1345) * tmp2 * u4;
  B[4][4] = sdt * tx2 * (c1 * (u2 * tmp1)) - dt * tx1 * c1345 * tmp1 - dt * tx1 * dx5;
}

// eta-neighbour block from u at (j-1) for jacld's b (src/LU/jacld.jl:207-283)
// or at (j+1) for jacu's b.
template <class T>
static void block_y(const Jac& J, double s, const T* uu, T B[5][5]) {
  const double dt = J.dt, sdt = s * J.dt, ty1 = J.ty1, ty2 = J.ty2, r43 = J.r43, c1345 = J.c1345, c34 = J.c34;
  const T u2 = uu[1], u3 = uu[2], u4 = uu[3], u5 = uu[4];
  T tmp1 = 1.0e+00 / uu[0];
  T tmp2 = tmp1 * tmp1;
  T tmp3 = tmp1 * tmp2;
  for (int m = 0; m < 5; ++m)
    for (int n = 0; n < 5; ++n) B[m][n] = 0.0;
  B[0][0] = -dt * ty1 * dy1;
  B[0][2] = sdt * ty2;
  B[1][0] = sdt * ty2 * (-(u2 * u3) * tmp2) - dt * ty1 * (-c34 * tmp2 * u2);
  B[1][1] = sdt * ty2 * (u3 * tmp1) - dt * ty1 * (c34 * tmp1) - dt * ty1 * dy2;
  B[1][2] = sdt * ty2 * (u2 * tmp1);
  B[2][0] = sdt * ty2 * (-sq(u3 * tmp1) + 0.50e+00 * c2 * ((u2 * u2 + u3 * u3 + u4 * u4) * tmp2)) -
            dt * ty1 * (-r43 * c34 * tmp2 * u3);
  B[2][1] = sdt * ty2 * (-c2 * (u2 * tmp1));
  B[2][2] = sdt * ty2 * ((2.0e+00 - c2) * (u3 * tmp1)) - dt * ty1 * (r43 * c34 * tmp1) - dt * ty1 * dy3;
  B[2][3] = sdt * ty2 * (-c2 * (u4 * tmp1));
  B[2][4] = sdt * ty2 * c2;
  B[3][0] = sdt * ty2 * (-(u3 * u4) * tmp2) - dt * ty1 * (-c34 * tmp2 * u4);
  B[3][2] = sdt * ty2 * (u4 * tmp1);
  B[3][3] = sdt * ty2 * (u3 * tmp1) - dt * ty1 * (c34 * tmp1) - dt * ty1 * dy4;
  B[4][0] = sdt * ty2 * ((c2 * (u2 * u2 + u3 * u3 + u4 * u4) * tmp2 - c1 * (u5 * tmp1)) * (u3 * tmp1)) -
            dt * ty1 *
                (-(c34 - c1345) * tmp3 * sq(u2) - (r43 * c34 - c1345) * tmp3 * sq(u3) -
                 (c34 - c1345) * tmp3 * sq(u4) - c1345 * tmp2 * u5);
  B[4][1] = sdt * ty2 * (-c2 * (u2 * u3) * tmp2) - dt * ty1 * (c34 - c1345) * tmp2 * u2;
  B[4][2] = sdt * ty2 * (c1 * (u5 * tmp1) - 0.50e+00 * c2 * ((u2 * u2 + 3.0e+00 * u3 * u3 + u4 * u4) * tmp2)) -
            dt * ty1 * (r43 * c34 - c1345) * tmp2 * u3;
  B[4][3] = sdt * ty2 * (-c2 * (u3 * u4) * tmp2) - dt * ty1 * (c34 - c1345) * tmp2 * u4;
  B[4][4] = sdt * ty2 * (c1 * (u3 * tmp1)) - dt * ty1 * c1345 * tmp1 - dt * ty1 * dy5;
}

// zeta-neighbour block from u at (k-1) for jacld's a (src/LU/jacld.jl:128-202)
// or at (k+1) for jacu's c.
template <class T>
static void block_z(const Jac& J, double s, const T* uu, T B[5][5]) {
  const double dt = J.dt, sdt = s * J.dt, tz1 = J.tz1, tz2 = J.tz2, r43 = J.r43, c1345 = J.c1345, c34 = J.c34;
  const T u2 = uu[1], u3 = uu[2], u4 = uu[3], u5 = uu[4];
  T tmp1 = 1.0e+00 / uu[0];
  T tmp2 = tmp1 * tmp1;
  T tmp3 = tmp1 * tmp2;
  for (int m = 0; m < 5; ++m)
    for (int n = 0; n < 5; ++n) B[m][n] = 0.0;
  B[0][0] = -dt * tz1 * dz1;
  B[0][3] = sdt * tz2;
  B[1][0] = sdt * tz2 * (-(u2 * u4) * tmp2) - dt * tz1 * (-c34 * tmp2 * u2);
  B[1][1] = sdt * tz2 * (u4 * tmp1) - dt * tz1 * c34 * tmp1 - dt * tz1 * dz2;
  B[1][3] = sdt * tz2 * (u2 * tmp1);
  B[2][0] = sdt * tz2 * (-(u3 * u4) * tmp2) - dt * tz1 * (-c34 * tmp2 * u3);
  B[2][2] = sdt * tz2 * (u4 * tmp1) - dt * tz1 * (c34 * tmp1) - dt * tz1 * dz3;
  B[2][3] = sdt * tz2 * (u3 * tmp1);
  B[3][0] = sdt * tz2 * (-sq(u4 * tmp1) + 0.50e+00 * c2 * ((u2 * u2 + u3 * u3 + u4 * u4) * tmp2)) -
            dt * tz1 * (-r43 * c34 * tmp2 * u4);
  B[3][1] = sdt * tz2 * (-c2 * (u2 * tmp1));
  B[3][2] = sdt * tz2 * (-c2 * (u3 * tmp1));
  B[3][3] = sdt * tz2 * (2.0e+00 - c2) * (u4 * tmp1) - dt * tz1 * (r43 * c34 * tmp1) - dt * tz1 * dz4;
  B[3][4] = sdt * tz2 * c2;
  B[4][0] = sdt * tz2 * ((c2 * (u2 * u2 + u3 * u3 + u4 * u4) * tmp2 - c1 * (u5 * tmp1)) * (u4 * tmp1)) -
            dt * tz1 *
                (-(c34 - c1345) * tmp3 * sq(u2) - (c34 - c1345) * tmp3 * sq(u3) -
                 (r43 * c34 - c1345) * tmp3 * sq(u4) - c1345 * tmp2 * u5);
  B[4][1] = sdt * tz2 * (-c2 * (u2 * u4) * tmp2) - dt * tz1 * (c34 - c1345) * tmp2 * u2;
  B[4][2] = sdt * tz2 * (-c2 * (u3 * u4) * tmp2) - dt * tz1 * (c34 - c1345) * tmp2 * u3;
  B[4][3] = sdt * tz2 * (c1 * (u5 * tmp1) - 0.50e+00 * c2 * ((u2 * u2 + u3 * u3 + 3.0e+00 * u4 * u4) * tmp2)) -
            dt * tz1 * (r43 * c34 - c1345) * tmp2 * u4;
  B[4][4] = sdt * tz2 * (c1 * (u4 * tmp1)) - dt * tz1 * c1345 * tmp1 - dt * tz1 * dz5;
}

template <class T>
static void store(BlocksT<T>& X, int i, const T B[5][5]) {
  for (int m = 1; m <= 5; ++m)
    for (int n = 1; n <= 5; ++n) X(m, n, i) = B[m - 1][n - 1];
}

// jacld(j, k): d at (i,j,k), a from (i,j,k-1), b from (i,j-1,k), c from
// (i-1,j,k) for i = ist..iend.
template <class T>
static void jacld(ZoneT<T>& z, const Jac& J, int j, int k) {
  T B[5][5];
  for (int i = 2; i <= z.nx - 1; ++i) {
    block_d(J, z.u.at(i, j, k), B); store(z.d, i, B);
    block_z(J, -1.0, z.u.at(i, j, k - 1), B); store(z.a, i, B);
    block_y(J, -1.0, z.u.at(i, j - 1, k), B); store(z.b, i, B);
    block_x(J, -1.0, z.u.at(i - 1, j, k), B); store(z.c, i, B);
  }
}

// jacu(j, k): d at (i,j,k), a from (i+1,j,k), b from (i,j+1,k), c from
// (i,j,k+1).
template <class T>
static void jacu(ZoneT<T>& z, const Jac& J, int j, int k) {
  T B[5][5];
  for (int i = 2; i <= z.nx - 1; ++i) {
    block_d(J, z.u.at(i, j, k), B); store(z.d, i, B);
    block_x(J, 1.0, z.u.at(i + 1, j, k), B); store(z.a, i, B);
    block_y(J, 1.0, z.u.at(i, j + 1, k), B); store(z.b, i, B);
    block_z(J, 1.0, z.u.at(i, j, k + 1), B); store(z.c, i, B);
  }
}

// blts, src/LU/blts.jl:9-152: the lower-triangular sweep of one (j,k) row
// with ldz = a, ldy = b, ldx = c.
template <class T>
static void blts(ZoneT<T>& z, int j, int k, double omega) {
  FieldT<T>& v = z.rsd;
  const int ist = 2, iend = z.nx - 1;
  for (int i = ist; i <= iend; ++i)
    for (int m = 1; m <= 5; ++m)
      v(m, i, j, k) -= omega * (z.a(m, 1, i) * v(1, i, j, k - 1) + z.a(m, 2, i) * v(2, i, j, k - 1) +
                                z.a(m, 3, i) * v(3, i, j, k - 1) + z.a(m, 4, i) * v(4, i, j, k - 1) +
                                z.a(m, 5, i) * v(5, i, j, k - 1));
  T t[6][6];
  for (int i = ist; i <= iend; ++i) {
    for (int m = 1; m <= 5; ++m)
      v(m, i, j, k) -= omega * (z.b(m, 1, i) * v(1, i, j - 1, k) + z.c(m, 1, i) * v(1, i - 1, j, k) +
                                z.b(m, 2, i) * v(2, i, j - 1, k) + z.c(m, 2, i) * v(2, i - 1, j, k) +
                                z.b(m, 3, i) * v(3, i, j - 1, k) + z.c(m, 3, i) * v(3, i - 1, j, k) +
                                z.b(m, 4, i) * v(4, i, j - 1, k) + z.c(m, 4, i) * v(4, i - 1, j, k) +
                                z.b(m, 5, i) * v(5, i, j - 1, k) + z.c(m, 5, i) * v(5, i - 1, j, k));
    for (int m = 1; m <= 5; ++m)
      for (int n = 1; n <= 5; ++n) t[m][n] = z.d(m, n, i);
    T* w = z.rsd.at(i, j, k) - 1;  // w[1..5]
    // Forward elimination without pivoting, then back substitution.
    for (int p = 1; p <= 4; ++p) {
      T tmp1 = 1.0e+00 / t[p][p];
      for (int r = p + 1; r <= 5; ++r) {
        T tmp = tmp1 * t[r][p];
        for (int c = p + 1; c <= 5; ++c) t[r][c] -= tmp * t[p][c];
        w[r] -= w[p] * tmp;
      }
    }
    w[5] /= t[5][5];
    w[4] -= t[4][5] * w[5];
    w[4] /= t[4][4];
    w[3] -= t[3][4] * w[4] + t[3][5] * w[5];
    w[3] /= t[3][3];
    w[2] -= t[2][3] * w[3] + t[2][4] * w[4] + t[2][5] * w[5];
    w[2] /= t[2][2];
    w[1] -= t[1][2] * w[2] + t[1][3] * w[3] + t[1][4] * w[4] + t[1][5] * w[5];
    w[1] /= t[1][1];
  }
}

// buts, src/LU/buts.jl:9-148: the upper-triangular sweep of one (j,k) row
// with udx = a, udy = b, udz = c, processed for i = iend down to ist.  With
// npb false the signs of :97, :126, :129 and :132 are kept as written.
template <class T>
static void buts(ZoneT<T>& z, int j, int k, double omega, bool npb) {
  FieldT<T>& v = z.rsd;
  const int ist = 2, iend = z.nx - 1;
  for (int i = iend; i >= ist; --i)
    for (int m = 1; m <= 5; ++m)
      z.TV(m, i) = omega * (z.c(m, 1, i) * v(1, i, j, k + 1) + z.c(m, 2, i) * v(2, i, j, k + 1) +
                            z.c(m, 3, i) * v(3, i, j, k + 1) + z.c(m, 4, i) * v(4, i, j, k + 1) +
                            z.c(m, 5, i) * v(5, i, j, k + 1));
  T t[6][6];
  for (int i = iend; i >= ist; --i) {
    for (int m = 1; m <= 5; ++m)
      z.TV(m, i) += omega * (z.b(m, 1, i) * v(1, i, j + 1, k) + z.a(m, 1, i) * v(1, i + 1, j, k) +
                             z.b(m, 2, i) * v(2, i, j + 1, k) + z.a(m, 2, i) * v(2, i + 1, j, k) +
                             z.b(m, 3, i) * v(3, i, j + 1, k) + z.a(m, 3, i) * v(3, i + 1, j, k) +
                             z.b(m, 4, i) * v(4, i, j + 1, k) + z.a(m, 4, i) * v(4, i + 1, j, k) +
                             z.b(m, 5, i) * v(5, i, j + 1, k) + z.a(m, 5, i) * v(5, i + 1, j, k));
    for (int m = 1; m <= 5; ++m)
      for (int n = 1; n <= 5; ++n) t[m][n] = z.d(m, n, i);
    T* w = &z.TV(1, i) - 1;  // w[1..5]
    for (int p = 1; p <= 4; ++p) {
      T tmp1 = 1.0e+00 / t[p][p];
      for (int r = p + 1; r <= 5; ++r) {
        T tmp = tmp1 * t[r][p];
        for (int c = p + 1; c <= 5; ++c) {
          if (!npb && p == 2 && r == 5 && c == 5)
            t[r][c] -= -tmp * t[p][c];  // :97 as written
          else
            t[r][c] -= tmp * t[p][c];
        }
        w[r] -= w[p] * tmp;
      }
    }
    w[5] /= t[5][5];
    w[4] -= t[4][5] * w[5];
    w[4] /= t[4][4];
    if (npb) {
      w[3] -= t[3][4] * w[4] + t[3][5] * w[5];
      w[3] /= t[3][3];
      w[2] -= t[2][3] * w[3] + t[2][4] * w[4] + t[2][5] * w[5];
      w[2] /= t[2][2];
      w[1] -= t[1][2] * w[2] + t[1][3] * w[3] + t[1][4] * w[4] + t[1][5] * w[5];
      w[1] /= t[1][1];
    } else {
      w[3] -= t[3][4] * w[4] - t[3][5] * w[5];
      w[3] /= t[3][3];
      w[2] -= t[2][3] * w[3] - t[2][4] * w[4] - t[2][5] * w[5];
      w[2] /= t[2][2];
      w[1] -= t[1][2] * w[2] - t[1][3] * w[3] - t[1][4] * w[4] - t[1][5] * w[5];
      w[1] /= t[1][1];
    }
    for (int m = 1; m <= 5; ++m) v(m, i, j, k) -= z.TV(m, i);
  }
}

// The five parts of ssor (src/LU/ssor.jl:9-254).
template <class T>
static void ssor_scale(ZoneT<T>& z, double dt) {
  for (int k = 2; k <= z.nz - 1; ++k)
    for (int j = 2; j <= z.ny - 1; ++j)
      for (int i = 2; i <= z.nx - 1; ++i)
        for (int m = 1; m <= 5; ++m) z.rsd(m, i, j, k) = dt * z.rsd(m, i, j, k);
}

template <class T>
static void ssor_lower(ZoneT<T>& z, const Jac& J, double omega) {
  for (int k = 2; k <= z.nz - 1; ++k)
    for (int j = 2; j <= z.ny - 1; ++j) {
      jacld(z, J, j, k);
      blts(z, j, k, omega);
    }
}

template <class T>
static void ssor_upper(ZoneT<T>& z, const Jac& J, double omega, bool npb) {
  for (int k = z.nz - 1; k >= 2; --k)
    for (int j = z.ny - 1; j >= 2; --j) {
      jacu(z, J, j, k);
      buts(z, j, k, omega, npb);
    }
}

template <class T>
static void ssor_update(ZoneT<T>& z, double omega) {
  double tmp = 1.0e+00 / (omega * (2.0e+00 - omega));
  for (int k = 2; k <= z.nz - 1; ++k)
    for (int j = 2; j <= z.ny - 1; ++j)
      for (int i = 2; i <= z.nx - 1; ++i)
        for (int m = 1; m <= 5; ++m) z.u(m, i, j, k) = z.u(m, i, j, k) + tmp * z.rsd(m, i, j, k);
}

// Zone neighbours of src/LU/zone_setup.jl:63-88 (1-based there, 0-based
// here).  LU names the x-direction neighbours north (i-1) and south (i+1) and
// the y-direction ones west (j-1) and east (j+1).
struct LuNbr { int north, south, west, east; };

static std::vector<LuNbr> lu_neighbors(int x_zones, int y_zones) {
  std::vector<LuNbr> n((size_t)x_zones * y_zones);
  for (int j = 1; j <= y_zones; ++j)
    for (int i = 1; i <= x_zones; ++i) {
      int zone_no = (i - 1) + (j - 1) * x_zones + 1;
      int row = i, col = j;
      int no = row > 1 ? zone_no - 1 : col * x_zones;
      int so = row < x_zones ? zone_no + 1 : zone_no - x_zones + 1;
      int we = col > 1 ? zone_no - x_zones : zone_no + y_zones * (x_zones - 1);
      int ea = col < y_zones ? zone_no + x_zones : zone_no - y_zones * (x_zones - 1);
      n[(size_t)zone_no - 1] = {no - 1, so - 1, we - 1, ea - 1};
    }
  return n;
}

// exch_qbc, src/LU/exchange_qbc.jl:17-288: every zone packs its four send
// planes first (:69-181), then each zone unpacks south, north, west, east in
// that order (:183-288), so the j faces win at the corners.  Planes are
// k = 1..nz-2 as written, or k = 2..nz-1 with LU_NPB_EXCH.
template <class T>
static void exch_qbc(std::vector<ZoneT<T>>& Z, const std::vector<LuNbr>& N, bool npb) {
  const int nzn = (int)Z.size();
  // Send buffers, indexed by the sending zone: [0] to east (j = ny-1),
  // [1] to west (j = 2), [2] to south (i = nx-1), [3] to north (i = 2).
  std::vector<std::vector<T>> out((size_t)nzn * 4);
  for (int z = 0; z < nzn; ++z) {
    ZoneT<T>& a = Z[z];
    int k0 = npb ? 2 : 1, k1 = npb ? a.nz - 1 : a.nz - 2;
    auto pack_j = [&](std::vector<T>& b, int j) {
      for (int k = k0; k <= k1; ++k)
        for (int i = 1; i <= a.nx; ++i)
          for (int m = 1; m <= 5; ++m) b.push_back(a.u(m, i, j, k));
    };
    auto pack_i = [&](std::vector<T>& b, int i) {
      for (int k = k0; k <= k1; ++k)
        for (int j = 1; j <= a.ny; ++j)
          for (int m = 1; m <= 5; ++m) b.push_back(a.u(m, i, j, k));
    };
    pack_j(out[(size_t)z * 4 + 0], a.ny - 1);
    pack_j(out[(size_t)z * 4 + 1], 2);
    pack_i(out[(size_t)z * 4 + 2], a.nx - 1);
    pack_i(out[(size_t)z * 4 + 3], 2);
  }
  for (int z = 0; z < nzn; ++z) {
    ZoneT<T>& a = Z[z];
    int k0 = npb ? 2 : 1, k1 = npb ? a.nz - 1 : a.nz - 2;
    auto unpack_i = [&](const std::vector<T>& b, int i) {
      size_t p = 0;
      for (int k = k0; k <= k1; ++k)
        for (int j = 1; j <= a.ny; ++j)
          for (int m = 1; m <= 5; ++m) a.u(m, i, j, k) = b[p++];
    };
    auto unpack_j = [&](const std::vector<T>& b, int j) {
      size_t p = 0;
      for (int k = k0; k <= k1; ++k)
        for (int i = 1; i <= a.nx; ++i)
          for (int m = 1; m <= 5; ++m) a.u(m, i, j, k) = b[p++];
    };
    unpack_i(out[(size_t)N[z].south * 4 + 3], a.nx);  // south zone's i = 2
    unpack_i(out[(size_t)N[z].north * 4 + 2], 1);     // north zone's i = nx-1
    unpack_j(out[(size_t)N[z].west * 4 + 0], 1);      // west zone's j = ny-1
    unpack_j(out[(size_t)N[z].east * 4 + 1], a.ny);   // east zone's j = 2
  }
}

// l2norm, src/LU/l2norm.jl:7-36.
static void l2norm(Zone& z, double out[5]) {
  double sum[5] = {0, 0, 0, 0, 0};
  for (int k = 2; k <= z.nz - 1; ++k)
    for (int j = 2; j <= z.ny - 1; ++j)
      for (int i = 2; i <= z.nx - 1; ++i)
        for (int m = 1; m <= 5; ++m) sum[m - 1] += z.rsd(m, i, j, k) * z.rsd(m, i, j, k);
  for (int m = 0; m < 5; ++m)
    out[m] = std::sqrt(sum[m] / (double(z.nx - 2) * (z.ny - 2) * (z.nz - 2)));
}

// ERROR, src/LU/error.jl:7-45.
static void error_norm(Zone& z, double out[5]) {
  double sum[5] = {0, 0, 0, 0, 0}, ue[5];
  for (int k = 2; k <= z.nz - 1; ++k)
    for (int j = 2; j <= z.ny - 1; ++j)
      for (int i = 2; i <= z.nx - 1; ++i) {
        exact(i, j, k, ue, z.nx, z.ny, z.nz);
        for (int m = 1; m <= 5; ++m) {
          double tmp = (ue[m - 1] - z.u(m, i, j, k));
          sum[m - 1] = sum[m - 1] + sq(tmp);
        }
      }
  for (int m = 0; m < 5; ++m)
    out[m] = std::sqrt(sum[m] / (double(z.nx - 2) * (z.ny - 2) * (z.nz - 2)));
}

// pintgr, src/LU/pintgr.jl:4-248, with ipt = jpt = 0: the surface integral of
// the pressure over three faces of the zone's box.
static double pintgr(Zone& z, const Coeff& C) {
  const int nx = z.nx, ny = z.ny, nz = z.nz, ipt = 0, jpt = 0;
  const int ii2 = C.ii2, ji2 = C.ji2, ki2 = C.ki2;
  int ibeg = nx + 1, ifin = 0;
  int iglob1 = ipt + 1, iglob2 = ipt + nx;
  if (iglob1 >= ii1 && iglob2 < ii2 + nx) ibeg = 1;
  if (iglob1 > ii1 - nx && iglob2 <= ii2) ifin = nx;
  if (ii1 >= iglob1 && ii1 <= iglob2) ibeg = ii1 - ipt;
  if (ii2 >= iglob1 && ii2 <= iglob2) ifin = ii2 - ipt;
  int jbeg = ny + 1, jfin = 0;
  int jglob1 = jpt + 1, jglob2 = jpt + ny;
  if (jglob1 >= ji1 && jglob2 < ji2 + ny) jbeg = 1;
  if (jglob1 > ji1 - ny && jglob2 <= ji2) jfin = ny;
  if (ji1 >= jglob1 && ji1 <= jglob2) jbeg = ji1 - jpt;
  if (ji2 >= jglob1 && ji2 <= jglob2) jfin = ji2 - jpt;
  int ifin1 = ifin, jfin1 = jfin;
  if (ipt + ifin1 == ii2) ifin1 = ifin - 1;
  if (jpt + jfin1 == ji2) jfin1 = jfin - 1;
  const int W = std::max(std::max(nx, ny), nz) + 2;
  std::vector<double> p1((size_t)W * W, 0.0), p2((size_t)W * W, 0.0);
  auto phi1 = [&](int a, int b) -> double& { return p1[(size_t)a + (size_t)W * b]; };
  auto phi2 = [&](int a, int b) -> double& { return p2[(size_t)a + (size_t)W * b]; };
  auto pres = [&](int i, int j, int k) {
    return c2 * (z.u(5, i, j, k) -
                 0.50e+00 * (sq(z.u(2, i, j, k)) + sq(z.u(3, i, j, k)) + sq(z.u(4, i, j, k))) / z.u(1, i, j, k));
  };
  for (int j = jbeg; j <= jfin; ++j)
    for (int i = ibeg; i <= ifin; ++i) {
      phi1(i, j) = pres(i, j, ki1);
      phi2(i, j) = pres(i, j, ki2);
    }
  double frc1 = 0.0;
  for (int j = jbeg; j <= jfin1; ++j)
    for (int i = ibeg; i <= ifin1; ++i)
      frc1 += (phi1(i, j) + phi1(i + 1, j) + phi1(i, j + 1) + phi1(i + 1, j + 1) + phi2(i, j) +
               phi2(i + 1, j) + phi2(i, j + 1) + phi2(i + 1, j + 1));
  frc1 = C.dxi * C.deta * frc1;

  std::fill(p1.begin(), p1.end(), 0.0);
  std::fill(p2.begin(), p2.end(), 0.0);
  if (jpt + jbeg == ji1)
    for (int k = ki1; k <= ki2; ++k)
      for (int i = ibeg; i <= ifin; ++i) phi1(i, k) = pres(i, jbeg, k);
  if (jpt + jfin == ji2)
    for (int k = ki1; k <= ki2; ++k)
      for (int i = ibeg; i <= ifin; ++i) phi2(i, k) = pres(i, jfin, k);
  double frc2 = 0.0;
  for (int k = ki1; k <= ki2 - 1; ++k)
    for (int i = ibeg; i <= ifin1; ++i)
      frc2 += (phi1(i, k) + phi1(i + 1, k) + phi1(i, k + 1) + phi1(i + 1, k + 1) + phi2(i, k) +
               phi2(i + 1, k) + phi2(i, k + 1) + phi2(i + 1, k + 1));
  frc2 = C.dxi * C.dzeta * frc2;

  std::fill(p1.begin(), p1.end(), 0.0);
  std::fill(p2.begin(), p2.end(), 0.0);
  if (ipt + ibeg == ii1)
    for (int k = ki1; k <= ki2; ++k)
      for (int j = jbeg; j <= jfin; ++j) phi1(j, k) = pres(ibeg, j, k);
  if (ipt + ifin == ii2)
    for (int k = ki1; k <= ki2; ++k)
      for (int j = jbeg; j <= jfin; ++j) phi2(j, k) = pres(ifin, j, k);
  double frc3 = 0.0;
  for (int k = ki1; k <= ki2 - 1; ++k)
    for (int j = jbeg; j <= jfin1; ++j)
      frc3 += (phi1(j, k) + phi1(j + 1, k) + phi1(j, k + 1) + phi1(j + 1, k + 1) + phi2(j, k) +
               phi2(j + 1, k) + phi2(j, k + 1) + phi2(j + 1, k + 1));
  frc3 = C.deta * C.dzeta * frc3;
  return 0.25e+00 * (frc1 + frc2 + frc3);
}

struct Run {
  int x_zones, y_zones;
  double dt, omega;
  int flags = 0;
  std::vector<Zone> Z;
  std::vector<LuNbr> N;
  std::vector<Coeff> C;
  int nthreads = 1;
};

}  // namespace lu
}  // namespace npb

using namespace npb;
using namespace npb::lu;

extern "C" {

// Setup of src/LU/lu-node.jl:111-196: per zone setcoeff, setbv, setiv, erhs
// and rhs.  The reference then exchanges, runs one SSOR iteration and
// re-runs setbv, setiv and rhs (:203-311); setbv and setiv rewrite every
// point of u and rhs rewrites rsd from u and the unchanged frct, so that
// warm-up leaves no trace and is not repeated here.  Zones must all have the
// same size (the reference's LU zone split, src/LU/zone_setup.jl:22-33).
void* oracle_lu_create(int x_zones, int y_zones, int gx, int gy, int gz, double dt, double omega, int flags) {
  Run* r = new Run();
  r->x_zones = x_zones; r->y_zones = y_zones; r->dt = dt; r->omega = omega; r->flags = flags;
  Zones G = zone_setup(x_zones, y_zones, gx, gy, gz, 1.0, true);
  int n = G.count();
  r->Z.resize(n);
  r->C.resize(n);
  r->N = lu_neighbors(x_zones, y_zones);
  for (int z = 0; z < n; ++z) {
    Zone& a = r->Z[z];
    a.nx = G.nx[z]; a.ny = G.ny[z]; a.nz = G.nz[z];
    a.alloc();
    r->C[z] = setcoeff(a.nx, a.ny, a.nz);
    setbv(a);
    setiv(a);
    erhs(a, r->C[z]);
    rhs(a, r->C[z]);
  }
  return r;
}

void oracle_lu_destroy(void* h) { delete (Run*)h; }
void oracle_lu_set_threads(void* h, int n) { ((Run*)h)->nthreads = n < 1 ? 1 : n; }
int oracle_lu_nzones(void* h) { return (int)((Run*)h)->Z.size(); }

// out7: nx, ny, nz, then the 0-based zones across the x-minus, x-plus,
// y-minus and y-plus faces (LU's north, south, west, east).
void oracle_lu_zone(void* h, int z, int* out7) {
  Run* r = (Run*)h;
  out7[0] = r->Z[z].nx; out7[1] = r->Z[z].ny; out7[2] = r->Z[z].nz;
  out7[3] = r->N[z].north; out7[4] = r->N[z].south; out7[5] = r->N[z].west; out7[6] = r->N[z].east;
}

// Phases: 0 exch_qbc, 1 rsd *= dt, 2 lower sweep (jacld + blts),
// 3 upper sweep (jacu + buts), 4 u += rsd/(omega(2-omega)), 5 rhs.
void oracle_lu_phase(void* h, int phase) {
  Run* r = (Run*)h;
  if (phase == 0) { exch_qbc(r->Z, r->N, r->flags & LU_NPB_EXCH); return; }
  int nzn = (int)r->Z.size();
#pragma omp parallel for schedule(dynamic, 1) num_threads(r->nthreads)
  for (int z = 0; z < nzn; ++z) {
    Zone& a = r->Z[z];
    Jac J = make_jac(r->C[z], r->dt);
    switch (phase) {
      case 1: ssor_scale(a, r->dt); break;
      case 2: ssor_lower(a, J, r->omega); break;
      case 3: ssor_upper(a, J, r->omega, r->flags & LU_NPB_BUTS); break;
      case 4: ssor_update(a, r->omega); break;
      case 5: rhs(a, r->C[z]); break;
    }
  }
}

// nsteps iterations of the timed loop of src/LU/lu-node.jl:338-428:
// exch_qbc, then ssor per zone.
void oracle_lu_step(void* h, int nsteps) {
  Run* r = (Run*)h;
  int nzn = (int)r->Z.size();
  for (int s = 0; s < nsteps; ++s) {
    exch_qbc(r->Z, r->N, r->flags & LU_NPB_EXCH);
#pragma omp parallel for schedule(dynamic, 1) num_threads(r->nthreads)
    for (int z = 0; z < nzn; ++z) {
      Zone& a = r->Z[z];
      Jac J = make_jac(r->C[z], r->dt);
      ssor_scale(a, r->dt);
      ssor_lower(a, J, r->omega);
      ssor_upper(a, J, r->omega, r->flags & LU_NPB_BUTS);
      ssor_update(a, r->omega);
      rhs(a, r->C[z]);
    }
  }
}

// Final verification values of src/LU/lu-node.jl:433-458: per zone ERROR,
// l2norm of rsd and pintgr, summed over zones in zone order.  zone_out, if
// not null, receives 11 values per zone (errnm[5], rsdnm[5], frc).
void oracle_lu_norms(void* h, double* xcr, double* xce, double* xci, double* zone_out) {
  Run* r = (Run*)h;
  int nzn = (int)r->Z.size();
  std::vector<double> v((size_t)nzn * 11);
#pragma omp parallel for schedule(dynamic, 1) num_threads(r->nthreads)
  for (int z = 0; z < nzn; ++z) {
    error_norm(r->Z[z], &v[(size_t)z * 11]);
    l2norm(r->Z[z], &v[(size_t)z * 11 + 5]);
    v[(size_t)z * 11 + 10] = pintgr(r->Z[z], r->C[z]);
  }
  for (int m = 0; m < 5; ++m) { xcr[m] = 0.0; xce[m] = 0.0; }
  *xci = 0.0;
  for (int z = 0; z < nzn; ++z) {
    for (int m = 0; m < 5; ++m) {
      xce[m] = xce[m] + v[(size_t)z * 11 + m];
      xcr[m] = xcr[m] + v[(size_t)z * 11 + 5 + m];
    }
    *xci = *xci + v[(size_t)z * 11 + 10];
  }
  if (zone_out) std::memcpy(zone_out, v.data(), v.size() * sizeof(double));
}

// field: 0 u, 1 frct, 2 rsd; dense [k][j][i][m] over the zone's nx*ny*nz points.
void oracle_lu_get(void* h, int z, int field, double* out) {
  Run* r = (Run*)h;
  Zone& a = r->Z[z];
  const Field& f = field == 0 ? a.u : field == 1 ? a.frct : a.rsd;
  std::memcpy(out, f.v.data(), f.v.size() * sizeof(double));
}

void oracle_lu_set(void* h, int z, int field, const double* in) {
  Run* r = (Run*)h;
  Zone& a = r->Z[z];
  Field& f = field == 0 ? a.u : field == 1 ? a.frct : a.rsd;
  std::memcpy(f.v.data(), in, f.v.size() * sizeof(double));
}

// Test hook: the blocks jacld (upper = 0) or jacu (upper = 1) gives the point
// (i, j, k) (1-based, interior) of zone z from the current u: d, then the
// k-, j- and i-neighbour blocks (jacld's a, b, c; jacu's c, b, a), each
// 25 values row-major [m][n].
void oracle_lu_point_blocks(void* h, int z, int upper, int i, int j, int k, double* out100) {
  Run* r = (Run*)h;
  Zone& a = r->Z[z];
  Jac J = make_jac(r->C[z], r->dt);
  double s = upper ? 1.0 : -1.0;
  int o = upper ? 1 : -1;
  double B[5][5];
  block_d(J, a.u.at(i, j, k), B);
  std::memcpy(out100, B, 25 * sizeof(double));
  block_z(J, s, a.u.at(i, j, k + o), B);
  std::memcpy(out100 + 25, B, 25 * sizeof(double));
  block_y(J, s, a.u.at(i, j + o, k), B);
  std::memcpy(out100 + 50, B, 25 * sizeof(double));
  block_x(J, s, a.u.at(i + o, j, k), B);
  std::memcpy(out100 + 75, B, 25 * sizeof(double));
}

// Test hook: u of zone z set to the exact polynomial at every point (no
// interpolation), for the residual check of rhs against erhs.
void oracle_lu_set_exact(void* h, int z) {
  Run* r = (Run*)h;
  Zone& a = r->Z[z];
  for (int k = 1; k <= a.nz; ++k)
    for (int j = 1; j <= a.ny; ++j)
      for (int i = 1; i <= a.nx; ++i) exact(i, j, k, a.u.at(i, j, k), a.nx, a.ny, a.nz);
}

// Forward-mode dual numbers through nsteps LU iterations from the current u
// and frct: out[0] = loss = sum(w*u_N^2) over every point, out[1+d] its
// derivative along (vu[d] on u, vf[d] on frct), both over every point.  rsd
// is first recomputed as rhs(u) (what it holds at an iteration's start after
// setup or an earlier iteration), so the loss is a function of u and frct
// alone; each iteration is oracle_lu_step's.  Arrays in the whole-run layout
// (every zone's [k][j][i][m], zone after zone); the state is left unchanged.
void oracle_lu_dual(void* h, int nsteps, int ndir, const double* w, const double* vu,
                    const double* vf, double* out) {
  constexpr int ND = 4;
  using D = Dual<ND>;
  Run* r = (Run*)h;
  const int nzn = (int)r->Z.size();
  std::vector<size_t> zoff(nzn + 1, 0);
  for (int z = 0; z < nzn; ++z) zoff[z + 1] = zoff[z] + r->Z[z].u.v.size();
  for (int d0 = 0; d0 < std::max(ndir, 1); d0 += ND) {
    std::vector<ZoneT<D>> Z(nzn);
    for (int z = 0; z < nzn; ++z) {
      Zone& a = r->Z[z];
      ZoneT<D>& b = Z[z];
      b.nx = a.nx; b.ny = a.ny; b.nz = a.nz;
      b.alloc();
      for (size_t q = 0; q < a.u.v.size(); ++q) {
        b.u.v[q] = D(a.u.v[q]);
        b.frct.v[q] = D(a.frct.v[q]);
        for (int e = 0; e < ND && d0 + e < ndir; ++e) {
          const size_t off = (size_t)(d0 + e) * zoff[nzn] + zoff[z] + q;
          b.u.v[q].d[e] = vu[off];
          b.frct.v[q].d[e] = vf[off];
        }
      }
      rhs(b, r->C[z]);
    }
    for (int s = 0; s < nsteps; ++s) {
      exch_qbc(Z, r->N, r->flags & LU_NPB_EXCH);
#pragma omp parallel for schedule(dynamic, 1) num_threads(r->nthreads)
      for (int z = 0; z < nzn; ++z) {
        ZoneT<D>& a = Z[z];
        Jac J = make_jac(r->C[z], r->dt);
        ssor_scale(a, r->dt);
        ssor_lower(a, J, r->omega);
        ssor_upper(a, J, r->omega, r->flags & LU_NPB_BUTS);
        ssor_update(a, r->omega);
        rhs(a, r->C[z]);
      }
    }
    D L(0.0);
    for (int z = 0; z < nzn; ++z)
      for (size_t q = 0; q < Z[z].u.v.size(); ++q) {
        const D& x = Z[z].u.v[q];
        L = L + w[zoff[z] + q] * (x * x);
      }
    out[0] = L.v;
    for (int e = 0; e < ND && d0 + e < ndir; ++e) out[1 + d0 + e] = L.d[e];
  }
}

}  // extern "C"
