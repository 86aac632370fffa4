// Test infrastructure only: CPU oracle for the BT-MZ timestep.
//
// A restatement of the reference's Julia BT path for one MPI rank holding
// every zone (ncells = 1, so cell_start = cell_end = 1 in every direction,
// cell_low = 0 and cell_size = the zone's grid size).  Every loop keeps the
// reference's bounds and every expression keeps its evaluation order so that,
// built with -ffp-contract=off, the results are the Julia CPU path's results
// bit for bit (Julia neither contracts nor reassociates).  Only tests/, the
// smoke check and bench.py's cpu_baseline leg load this library.
//
// The scalar type is a template parameter so that a forward-mode dual number
// can be pushed through the same code.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "dual.h"
#include "zone_ops.h"

namespace npb {
namespace bt {

// ---------------------------------------------------------------- 5x5 helpers
// src/BT/solve_subs.jl.  Blocks are addressed through accessor lambdas with
// (row, col) 0-based; the reference's [m, n] is (m-1, n-1).

// matvec_sub, solve_subs.jl:4-37: b -= (sum of five products).
template <class T, class A, class X, class B>
inline void matvec_sub(A a, X x, B b) {
  for (int m = 0; m < 5; ++m)
    b(m) -= a(m, 0) * x(0) + a(m, 1) * x(1) + a(m, 2) * x(2) + a(m, 3) * x(3) + a(m, 4) * x(4);
}

// matmul_sub2, solve_subs.jl:333-455: c = c - a1b1 - a2b2 - ... (sequential).
template <class T, class A, class Bm, class C>
inline void matmul_sub2(A a, Bm b, C c) {
  for (int n = 0; n < 5; ++n)
    for (int m = 0; m < 5; ++m)
      c(m, n) = c(m, n) - a(m, 0) * b(0, n) - a(m, 1) * b(1, n) - a(m, 2) * b(2, n) -
                a(m, 3) * b(3, n) - a(m, 4) * b(4, n);
}

// binvcrhs, solve_subs.jl:468-723: Gauss-Jordan without pivoting on lhs,
// carrying c (5x5) and r (5).  binvrhs (:730-862) is the same without c.
template <class T, class L, class C, class R>
inline void binvcrhs(L lhs, C c, R r, bool with_c) {
  for (int p = 0; p < 5; ++p) {
    T pivot = 1.00 / lhs(p, p);
    for (int n = p + 1; n < 5; ++n) lhs(p, n) *= pivot;
    if (with_c)
      for (int n = 0; n < 5; ++n) c(p, n) *= pivot;
    r(p) *= pivot;
    for (int q = 0; q < 5; ++q) {
      if (q == p) continue;
      T coeff = lhs(q, p);
      for (int n = p + 1; n < 5; ++n) lhs(q, n) -= coeff * lhs(p, n);
      if (with_c)
        for (int n = 0; n < 5; ++n) c(q, n) -= coeff * c(p, n);
      r(q) -= coeff * r(p);
    }
  }
}

// ---------------------------------------------------------------- jacobians
// fjac/njac at one point for each direction: x_solve.jl:536-609,
// y_solve.jl:543-615, z_solve.jl:537-621.  u1..u5 = u[1..5], ri = 1/u1
// (rho_i in x, recomputed as utmp[1] in y and z: same value), q = qs.
template <class T, class FJ, class NJ>
inline void jacobians(int dir, const Consts& K, T u1, T u2, T u3, T u4, T u5, T ri, T q, FJ fj,
                      NJ nj) {
  (void)u1;
  const double c1 = K.c1, c2 = K.c2, c3c4 = K.c3c4, con43 = K.con43, c1345 = K.c1345;
  T tmp1 = ri;
  T tmp2 = tmp1 * tmp1;
  T tmp3 = tmp1 * tmp2;
  for (int m = 0; m < 5; ++m)
    for (int n = 0; n < 5; ++n) { fj(m, n) = T(0.0); nj(m, n) = T(0.0); }
  if (dir == 0) {
    fj(0, 1) = T(1.0);
    fj(1, 0) = -(u2 * tmp2 * u2) + c2 * q;
    fj(1, 1) = (2.0 - c2) * (u2 * tmp1);
    fj(1, 2) = -c2 * (u3 * tmp1);
    fj(1, 3) = -c2 * (u4 * tmp1);
    fj(1, 4) = T(c2);
    fj(2, 0) = -(u2 * u3) * tmp2;
    fj(2, 1) = u3 * tmp1;
    fj(2, 2) = u2 * tmp1;
    fj(3, 0) = -(u2 * u4) * tmp2;
    fj(3, 1) = u4 * tmp1;
    fj(3, 3) = u2 * tmp1;
    fj(4, 0) = (c2 * 2.0 * q - c1 * (u5 * tmp1)) * (u2 * tmp1);
    fj(4, 1) = c1 * u5 * tmp1 - c2 * (u2 * u2 * tmp2 + q);
    fj(4, 2) = -c2 * (u3 * u2) * tmp2;
    fj(4, 3) = -c2 * (u4 * u2) * tmp2;
    fj(4, 4) = c1 * (u2 * tmp1);
    nj(1, 0) = -con43 * c3c4 * tmp2 * u2;
    nj(1, 1) = con43 * c3c4 * tmp1;
    nj(2, 0) = -c3c4 * tmp2 * u3;
    nj(2, 2) = c3c4 * tmp1;
    nj(3, 0) = -c3c4 * tmp2 * u4;
    nj(3, 3) = c3c4 * tmp1;
    nj(4, 0) = -(con43 * c3c4 - c1345) * tmp3 * (u2 * u2) - (c3c4 - c1345) * tmp3 * (u3 * u3) -
               (c3c4 - c1345) * tmp3 * (u4 * u4) - c1345 * tmp2 * u5;
    nj(4, 1) = (con43 * c3c4 - c1345) * tmp2 * u2;
    nj(4, 2) = (c3c4 - c1345) * tmp2 * u3;
    nj(4, 3) = (c3c4 - c1345) * tmp2 * u4;
    nj(4, 4) = (c1345)*tmp1;
  } else if (dir == 1) {
    fj(0, 2) = T(1.0);
    fj(1, 0) = -(u2 * u3) * tmp2;
    fj(1, 1) = u3 * tmp1;
    fj(1, 2) = u2 * tmp1;
    fj(2, 0) = -(u3 * u3 * tmp2) + c2 * q;
    fj(2, 1) = -c2 * u2 * tmp1;
    fj(2, 2) = (2.0 - c2) * u3 * tmp1;
    fj(2, 3) = -c2 * u4 * tmp1;
    fj(2, 4) = T(c2);
    fj(3, 0) = -(u3 * u4) * tmp2;
    fj(3, 2) = u4 * tmp1;
    fj(3, 3) = u3 * tmp1;
    fj(4, 0) = (c2 * 2.0 * q - c1 * u5 * tmp1) * u3 * tmp1;
    fj(4, 1) = -c2 * u2 * u3 * tmp2;
    fj(4, 2) = c1 * u5 * tmp1 - c2 * (q + u3 * u3 * tmp2);
    fj(4, 3) = -c2 * (u3 * u4) * tmp2;
    fj(4, 4) = c1 * u3 * tmp1;
    nj(1, 0) = -c3c4 * tmp2 * u2;
    nj(1, 1) = c3c4 * tmp1;
    nj(2, 0) = -con43 * c3c4 * tmp2 * u3;
    nj(2, 2) = con43 * c3c4 * tmp1;
    nj(3, 0) = -c3c4 * tmp2 * u4;
    nj(3, 3) = c3c4 * tmp1;
    nj(4, 0) = -(c3c4 - c1345) * tmp3 * (u2 * u2) - (con43 * c3c4 - c1345) * tmp3 * (u3 * u3) -
               (c3c4 - c1345) * tmp3 * (u4 * u4) - c1345 * tmp2 * u5;
    nj(4, 1) = (c3c4 - c1345) * tmp2 * u2;
    nj(4, 2) = (con43 * c3c4 - c1345) * tmp2 * u3;
    nj(4, 3) = (c3c4 - c1345) * tmp2 * u4;
    nj(4, 4) = (c1345)*tmp1;
  } else {
    fj(0, 3) = T(1.0);
    fj(1, 0) = -(u2 * u4) * tmp2;
    fj(1, 1) = u4 * tmp1;
    fj(1, 3) = u2 * tmp1;
    fj(2, 0) = -(u3 * u4) * tmp2;
    fj(2, 2) = u4 * tmp1;
    fj(2, 3) = u3 * tmp1;
    fj(3, 0) = -(u4 * u4 * tmp2) + c2 * q;
    fj(3, 1) = -c2 * u2 * tmp1;
    fj(3, 2) = -c2 * u3 * tmp1;
    fj(3, 3) = (2.0 - c2) * u4 * tmp1;
    fj(3, 4) = T(c2);
    fj(4, 0) = (c2 * 2.0 * q - c1 * (u5 * tmp1)) * (u4 * tmp1);
    fj(4, 1) = -c2 * (u2 * u4) * tmp2;
    fj(4, 2) = -c2 * (u3 * u4) * tmp2;
    fj(4, 3) = c1 * (u5 * tmp1) - c2 * (q + u4 * u4 * tmp2);
    fj(4, 4) = c1 * u4 * tmp1;
    nj(1, 0) = -c3c4 * tmp2 * u2;
    nj(1, 1) = c3c4 * tmp1;
    nj(2, 0) = -c3c4 * tmp2 * u3;
    nj(2, 2) = c3c4 * tmp1;
    nj(3, 0) = -con43 * c3c4 * tmp2 * u4;
    nj(3, 3) = con43 * K.c3 * K.c4 * tmp1;
    nj(4, 0) = -(c3c4 - c1345) * tmp3 * (u2 * u2) - (c3c4 - c1345) * tmp3 * (u3 * u3) -
               (con43 * c3c4 - c1345) * tmp3 * (u4 * u4) - c1345 * tmp2 * u5;
    nj(4, 1) = (c3c4 - c1345) * tmp2 * u2;
    nj(4, 2) = (c3c4 - c1345) * tmp2 * u3;
    nj(4, 3) = (con43 * c3c4 - c1345) * tmp2 * u4;
    nj(4, 4) = (c1345)*tmp1;
  }
}

// ---------------------------------------------------------------- solves
// x_solve_cell/x_backsubstitute (x_solve.jl:444-1745), y_solve_cell
// (y_solve.jl:443-821), z_solve_cell (z_solve.jl:438-864) for FIRST = LAST = 1.
// Assemble the block rows of one line (transverse indices a, b) into the
// scratch s and lhsc: fjac/njac for every point, lhsa/lhsb/lhsc for rows
// 1..size-1 (x_solve.jl:536-708 and its y/z twins).
template <class T>
void assemble_line(Zone<T>& z, const Consts& K, int dir, int a, int b, LineScratch<T>& s) {
  const int n[3] = {z.nx, z.ny, z.nz};
  const int nl = n[dir];
  const int size = nl - 1;
  const double d1 = dir == 0 ? K.tx1 : dir == 1 ? K.ty1 : K.tz1;
  const double d2 = dir == 0 ? K.tx2 : dir == 1 ? K.ty2 : K.tz2;
  const double dd[5] = {dir == 0 ? K.dx1 : dir == 1 ? K.dy1 : K.dz1,
                        dir == 0 ? K.dx2 : dir == 1 ? K.dy2 : K.dz2,
                        dir == 0 ? K.dx3 : dir == 1 ? K.dy3 : K.dz3,
                        dir == 0 ? K.dx4 : dir == 1 ? K.dy4 : K.dz4,
                        dir == 0 ? K.dx5 : dir == 1 ? K.dy5 : K.dz5};
  int da, db;
  if (dir == 0) { da = 2; db = 1; }
  else if (dir == 1) { da = 2; db = 0; }
  else { da = 1; db = 0; }
  int idx[3];
  idx[da] = a; idx[db] = b;
  for (int l = 0; l <= nl - 1; ++l) {
    idx[dir] = l;
    const int i = idx[0], j = idx[1], k = idx[2];
    T u1 = z.U(0, i, j, k);
    T ri = dir == 0 ? z.rho_i[z.didx(i, j, k)] : T(1.0 / u1);
    jacobians<T>(dir, K, u1, z.U(1, i, j, k), z.U(2, i, j, k), z.U(3, i, j, k),
                 z.U(4, i, j, k), ri, z.qs[z.didx(i, j, k)],
                 [&](int m, int nn) -> T& { return s.f(m, nn, l); },
                 [&](int m, int nn) -> T& { return s.nj(m, nn, l); });
  }
  for (int l = 1; l <= size - 1; ++l) {
    idx[dir] = l;
    const int i = idx[0], j = idx[1], k = idx[2];
    T tmp1 = K.dt * d1;
    T tmp2 = K.dt * d2;
    for (int m = 0; m < 5; ++m)
      for (int nn = 0; nn < 5; ++nn) {
        T av = -tmp2 * s.f(m, nn, l - 1) - tmp1 * s.nj(m, nn, l - 1);
        if (m == nn) av = av - tmp1 * dd[m];
        s.a(m, nn, l) = av;
        T bv = tmp1 * 2.0 * s.nj(m, nn, l);
        if (m == nn) bv = 1.0 + tmp1 * 2.0 * s.nj(m, nn, l) + tmp1 * 2.0 * dd[m];
        s.b(m, nn, l) = bv;
        T cv = tmp2 * s.f(m, nn, l + 1) - tmp1 * s.nj(m, nn, l + 1);
        if (m == nn) cv = cv - tmp1 * dd[m];
        z.LC(m, nn, i, j, k) = cv;
      }
  }
}

inline void transverse(int dir, int& da, int& db) {
  if (dir == 0) { da = 2; db = 1; }
  else if (dir == 1) { da = 2; db = 0; }
  else { da = 1; db = 0; }
}

// lhsabinit(lhsa, lhsb, size): initialize.jl:233-246
template <class T>
void lhsabinit(LineScratch<T>& s, int size) {
  for (int i = 0; i <= size; ++i)
    for (int m = 0; m < 5; ++m) {
      for (int nn = 0; nn < 5; ++nn) { s.a(m, nn, i) = T(0.0); s.b(m, nn, i) = T(0.0); }
      s.b(m, m, i) = T(1.0);
    }
}

// x_solve_cell/x_backsubstitute (x_solve.jl:444-1745), y_solve_cell
// (y_solve.jl:443-821), z_solve_cell (z_solve.jl:438-864) for FIRST = LAST = 1.
template <class T>
void solve_dir(Zone<T>& z, const Consts& K, int dir) {
  const int n[3] = {z.nx, z.ny, z.nz};
  const int nl = n[dir];
  const int size = nl - 1;  // isize / jsize / ksize
  LineScratch<T> s(nl);
  lhsabinit(s, size);
  int da, db;
  transverse(dir, da, db);
  for (int a = 1; a <= n[da] - 2; ++a)
    for (int b = 1; b <= n[db] - 2; ++b) {
      int idx[3];
      idx[da] = a; idx[db] = b;
      assemble_line(z, K, dir, a, b, s);
      auto at = [&](int l, int& i, int& j, int& k) {
        idx[dir] = l; i = idx[0]; j = idx[1]; k = idx[2];
      };
      auto C = [&](int l) {
        int i, j, k;
        at(l, i, j, k);
        return [&z, i, j, k](int m, int nn) -> T& { return z.LC(m, nn, i, j, k); };
      };
      auto Rv = [&](int l) {
        int i, j, k;
        at(l, i, j, k);
        return [&z, i, j, k](int m) -> T& { return z.R(m, i, j, k); };
      };
      auto Bm = [&](int l) { return [&s, l](int m, int nn) -> T& { return s.b(m, nn, l); }; };
      auto Am = [&](int l) { return [&s, l](int m, int nn) -> T& { return s.a(m, nn, l); }; };
      binvcrhs<T>(Bm(0), C(0), Rv(0), true);
      for (int l = 1; l <= size - 1; ++l) {
        matvec_sub<T>(Am(l), Rv(l - 1), Rv(l));
        matmul_sub2<T>(Am(l), C(l - 1), Bm(l));
        binvcrhs<T>(Bm(l), C(l), Rv(l), true);
      }
      matvec_sub<T>(Am(size), Rv(size - 1), Rv(size));
      matmul_sub2<T>(Am(size), C(size - 1), Bm(size));
      binvcrhs<T>(Bm(size), C(size), Rv(size), false);
    }
  // backsubstitute: rhs(l) -= C'(l) rhs(l+1), l = size-1 .. 0
  for (int a = 1; a <= n[da] - 2; ++a)
    for (int b = 1; b <= n[db] - 2; ++b) {
      int idx[3];
      idx[da] = a; idx[db] = b;
      for (int l = size - 1; l >= 0; --l) {
        idx[dir] = l;
        int i = idx[0], j = idx[1], k = idx[2];
        idx[dir] = l + 1;
        int i1 = idx[0], j1 = idx[1], k1 = idx[2];
        for (int m = 0; m < 5; ++m)
          for (int nn = 0; nn < 5; ++nn)
            z.R(m, i, j, k) -= z.LC(m, nn, i, j, k) * z.R(nn, i1, j1, k1);
      }
    }
}

// The whole run state.
struct Run {
  Zones G;
  Consts K;
  std::vector<Zone<double>> Z;
  int nthreads = 1;
};

// BT stencil widths dx1..dz5: src/BT/set_constants.jl:16-30.
static const double BT_DXYZ[15] = {0.75, 0.75, 0.75, 0.75, 0.75, 0.75, 0.75, 0.75,
                                   0.75, 0.75, 1.0,  1.0,  1.0,  1.0,  1.0};

}  // namespace bt
}  // namespace npb

// ===================================================================== C API
using namespace npb;
using namespace npb::bt;

extern "C" {

void* oracle_bt_create(int x_zones, int y_zones, int gx, int gy, int gz, double ratio, double dt) {
  Run* r = new Run();
  r->G = zone_setup(x_zones, y_zones, gx, gy, gz, ratio, false);
  r->K = set_constants(dt, gx, gy, gz, x_zones, y_zones, BT_DXYZ);
  r->Z.resize(r->G.count());
  for (int z = 0; z < r->G.count(); ++z) {
    Zone<double>& zz = r->Z[z];
    zz.nx = r->G.nx[z]; zz.ny = r->G.ny[z]; zz.nz = r->G.nz[z];
    zz.alloc();
    initialize(zz, r->K);
    exact_rhs(zz, r->K);
  }
  return r;
}

void oracle_bt_destroy(void* h) { delete (Run*)h; }

void oracle_bt_set_threads(void* h, int n) { ((Run*)h)->nthreads = n < 1 ? 1 : n; }

int oracle_bt_nzones(void* h) { return ((Run*)h)->G.count(); }

void oracle_bt_zone(void* h, int z, int* out7) {
  Run* r = (Run*)h;
  out7[0] = r->G.nx[z]; out7[1] = r->G.ny[z]; out7[2] = r->G.nz[z];
  out7[3] = r->G.west[z]; out7[4] = r->G.east[z]; out7[5] = r->G.south[z];
  out7[6] = r->G.north[z];
}

// Phases: 0 exch_qbc, 1 compute_rhs, 2 x_solve, 3 y_solve, 4 z_solve, 5 add.
void oracle_bt_phase(void* h, int phase) {
  Run* r = (Run*)h;
  if (phase == 0) { exch_qbc(r->Z, r->G); return; }
  int nzn = r->G.count();
#pragma omp parallel for schedule(dynamic, 1) num_threads(r->nthreads)
  for (int z = 0; z < nzn; ++z) {
    Zone<double>& zz = r->Z[z];
    switch (phase) {
      case 1: compute_rhs(zz, r->K); break;
      case 2: solve_dir(zz, r->K, 0); break;
      case 3: solve_dir(zz, r->K, 1); break;
      case 4: solve_dir(zz, r->K, 2); break;
      case 5: add(zz); break;
    }
  }
}

// nsteps iterations of the timestep loop body (bt-node.jl:280-417).
void oracle_bt_step(void* h, int nsteps) {
  Run* r = (Run*)h;
  int nzn = r->G.count();
  for (int s = 0; s < nsteps; ++s) {
    if (nzn > 1) exch_qbc(r->Z, r->G);
#pragma omp parallel for schedule(dynamic, 1) num_threads(r->nthreads)
    for (int z = 0; z < nzn; ++z) {
      Zone<double>& zz = r->Z[z];
      compute_rhs(zz, r->K);
      solve_dir(zz, r->K, 0);
      solve_dir(zz, r->K, 1);
      solve_dir(zz, r->K, 2);
      add(zz);
    }
  }
}

// Node-side verify (src/BT/verify.jl:411-501): per zone error_norm, then
// compute_rhs on the current u without a new exchange, then rhs_norm;
// xcr += rms/dt, xce += rms over zones in zone order.
void oracle_bt_norms(void* h, double* xcr, double* xce) {
  Run* r = (Run*)h;
  for (int m = 0; m < 5; ++m) { xcr[m] = 0.0; xce[m] = 0.0; }
  int nzn = r->G.count();
  std::vector<double> ce_all((size_t)nzn * 5), cr_all((size_t)nzn * 5);
#pragma omp parallel for schedule(dynamic, 1) num_threads(r->nthreads)
  for (int z = 0; z < nzn; ++z) {
    error_norm(r->Z[z], r->K, &ce_all[(size_t)z * 5]);
    compute_rhs(r->Z[z], r->K);
    rhs_norm(r->Z[z], &cr_all[(size_t)z * 5]);
  }
  for (int z = 0; z < nzn; ++z)
    for (int m = 0; m < 5; ++m) {
      xcr[m] = xcr[m] + cr_all[(size_t)z * 5 + m] / r->K.dt;
      xce[m] = xce[m] + ce_all[(size_t)z * 5 + m];
    }
}

// Field access in a dense [k][j][i][m] layout over the zone's grid points
// 0..n-1 (m fastest, i.e. the reference's [m,i,j,k] column-major order).
// field: 0 u, 1 forcing, 2 rhs.
void oracle_bt_get(void* h, int z, int field, double* out) {
  Run* r = (Run*)h;
  Zone<double>& a = r->Z[z];
  size_t p = 0;
  for (int k = 0; k < a.nz; ++k)
    for (int j = 0; j < a.ny; ++j)
      for (int i = 0; i < a.nx; ++i)
        for (int m = 0; m < 5; ++m, ++p)
          out[p] = field == 0 ? a.U(m, i, j, k) : field == 1 ? a.F(m, i, j, k) : a.R(m, i, j, k);
}

void oracle_bt_set(void* h, int z, int field, const double* in) {
  Run* r = (Run*)h;
  Zone<double>& a = r->Z[z];
  size_t p = 0;
  for (int k = 0; k < a.nz; ++k)
    for (int j = 0; j < a.ny; ++j)
      for (int i = 0; i < a.nx; ++i)
        for (int m = 0; m < 5; ++m, ++p) {
          if (field == 0) a.U(m, i, j, k) = in[p];
          else if (field == 1) a.F(m, i, j, k) = in[p];
          else a.R(m, i, j, k) = in[p];
        }
}

// Constants as a flat array in declaration order of npb::Consts.
int oracle_bt_consts(void* h, double* out, int cap) {
  Run* r = (Run*)h;
  int n = int(sizeof(Consts) / sizeof(double));
  const double* src = (const double*)&r->K;
  for (int i = 0; i < n && i < cap; ++i) out[i] = src[i];
  return n;
}

// ---- test hooks -----------------------------------------------------------
// binvcrhs / binvrhs of solve_subs.jl on one block system; B, C row-major.
void oracle_binvcrhs(double* B, double* C, double* r, int with_c) {
  binvcrhs<double>([B](int m, int n) -> double& { return B[m * 5 + n]; },
                   [C](int m, int n) -> double& { return C[m * 5 + n]; },
                   [r](int m) -> double& { return r[m]; }, with_c != 0);
}

// The block rows of line (a, b) of zone z in direction dir as the solver
// assembles them (rows 0..nl-1; rows 0 and nl-1 are the identity rows of
// lhsabinit with zero couplings).  Uses the derived fields of the last
// compute_rhs.  A, B, C: nl x 25 row-major blocks.
void oracle_bt_line_blocks(void* h, int z, int dir, int a, int b, double* A, double* B,
                           double* C) {
  Run* r = (Run*)h;
  Zone<double>& zz = r->Z[z];
  const int n[3] = {zz.nx, zz.ny, zz.nz};
  const int nl = n[dir];
  LineScratch<double> s(nl);
  lhsabinit(s, nl - 1);
  int da, db;
  transverse(dir, da, db);
  // keep lhsc intact: assemble into a copy of the zone
  Zone<double> tmp = zz;
  assemble_line(tmp, r->K, dir, a, b, s);
  int idx[3];
  idx[da] = a; idx[db] = b;
  for (int l = 0; l < nl; ++l) {
    idx[dir] = l;
    for (int m = 0; m < 5; ++m)
      for (int q = 0; q < 5; ++q) {
        A[l * 25 + m * 5 + q] = s.a(m, q, l);
        B[l * 25 + m * 5 + q] = s.b(m, q, l);
        C[l * 25 + m * 5 + q] = (l >= 1 && l <= nl - 2) ? tmp.LC(m, q, idx[0], idx[1], idx[2]) : 0.0;
      }
  }
}

// Reference gradient by forward-mode dual numbers (no counterpart in the
// reference, which has no derivatives).  From the run's current state u0 and
// forcing f, runs nsteps timesteps in Dual<4> arithmetic with u0 seeded
// along vu[d] and f along vf[d] (d = 0..ndir-1, four at a time) and
// evaluates the loss L = sum over every grid point of w*u^2 of the final u.
// w, vu, vf hold every zone's nx*ny*nz*5 values, zone after zone, each in the
// dense [k][j][i][m] order of oracle_bt_get; vf is read at interior points
// only (forcing is used nowhere else).  out[0] = L, out[1 + d] = the
// directional derivative of L along (vu[d], vf[d]).  The run's state is left
// unchanged.
void oracle_bt_dual(void* h, int nsteps, int ndir, const double* w, const double* vu,
                    const double* vf, double* out) {
  constexpr int ND = 4;
  using D = Dual<ND>;
  Run* r = (Run*)h;
  const int nzn = r->G.count();
  std::vector<size_t> zoff(nzn + 1, 0);
  for (int z = 0; z < nzn; ++z)
    zoff[z + 1] = zoff[z] + (size_t)5 * r->Z[z].nx * r->Z[z].ny * r->Z[z].nz;
  for (int d0 = 0; d0 < std::max(ndir, 1); d0 += ND) {
    std::vector<Zone<D>> Z(nzn);
    for (int z = 0; z < nzn; ++z) {
      Zone<double>& a = r->Z[z];
      Zone<D>& b = Z[z];
      b.nx = a.nx; b.ny = a.ny; b.nz = a.nz;
      b.alloc();
      for (size_t q = 0; q < a.u.size(); ++q) b.u[q] = D(a.u[q]);
      for (size_t q = 0; q < a.forcing.size(); ++q) b.forcing[q] = D(a.forcing[q]);
      size_t p = zoff[z];
      for (int k = 0; k < a.nz; ++k)
        for (int j = 0; j < a.ny; ++j)
          for (int i = 0; i < a.nx; ++i)
            for (int m = 0; m < 5; ++m, ++p)
              for (int e = 0; e < ND && d0 + e < ndir; ++e) {
                const size_t off = (size_t)(d0 + e) * zoff[nzn];
                b.U(m, i, j, k).d[e] = vu[off + p];
                const bool in = i > 0 && i < a.nx - 1 && j > 0 && j < a.ny - 1 && k > 0 && k < a.nz - 1;
                if (in) b.F(m, i, j, k).d[e] = vf[off + p];
              }
    }
    for (int s = 0; s < nsteps; ++s) {
      if (nzn > 1) exch_qbc(Z, r->G);
#pragma omp parallel for schedule(dynamic, 1) num_threads(r->nthreads)
      for (int z = 0; z < nzn; ++z) {
        compute_rhs(Z[z], r->K);
        solve_dir(Z[z], r->K, 0);
        solve_dir(Z[z], r->K, 1);
        solve_dir(Z[z], r->K, 2);
        add(Z[z]);
      }
    }
    D L(0.0);
    for (int z = 0; z < nzn; ++z) {
      Zone<D>& b = Z[z];
      size_t p = zoff[z];
      for (int k = 0; k < b.nz; ++k)
        for (int j = 0; j < b.ny; ++j)
          for (int i = 0; i < b.nx; ++i)
            for (int m = 0; m < 5; ++m, ++p) {
              const D& x = b.U(m, i, j, k);
              L = L + w[p] * (x * x);
            }
    }
    out[0] = L.v;
    for (int e = 0; e < ND && d0 + e < ndir; ++e) out[1 + d0 + e] = L.d[e];
  }
}

}  // extern "C"
