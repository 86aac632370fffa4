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

#include "npb_common.h"

namespace npb {
namespace bt {

// One zone's arrays, indexed as the reference's OffsetArrays
// (src/BT/bt_data.jl:138-154 with IMAX = nx+1, JMAX = ny+1, KMAX = nz+1).
template <class T>
struct Zone {
  int nx, ny, nz;
  // u: 1:5 x -2:IMAX+1 x -2:JMAX+1 x -2:KMAX+1
  std::vector<T> u;
  // forcing: 1:5 x 0:IMAX-1 x ...   rhs: 1:5 x -1:IMAX-1 x ...
  std::vector<T> forcing, rhs;
  // derived fields on -1:IMAX x -1:JMAX x -1:KMAX
  std::vector<T> rho_i, us, vs, ws, qs, square;
  // lhsc: 1:5 x 1:5 x -1:IMAX-1 x -1:JMAX-1 x -1:KMAX-1
  std::vector<T> lhsc;

  int ux() const { return nx + 5; }
  int uy() const { return ny + 5; }
  T& U(int m, int i, int j, int k) {
    return u[m + 5 * ((i + 2) + (size_t)ux() * ((j + 2) + (size_t)uy() * (k + 2)))];
  }
  T& F(int m, int i, int j, int k) {
    return forcing[m + 5 * (i + (size_t)(nx + 1) * (j + (size_t)(ny + 1) * k))];
  }
  T& R(int m, int i, int j, int k) {
    return rhs[m + 5 * ((i + 1) + (size_t)(nx + 2) * ((j + 1) + (size_t)(ny + 2) * (k + 1)))];
  }
  size_t didx(int i, int j, int k) const {
    return (i + 1) + (size_t)(nx + 3) * ((j + 1) + (size_t)(ny + 3) * (k + 1));
  }
  T& LC(int m, int n, int i, int j, int k) {
    return lhsc[m + 5 * n +
                25 * ((i + 1) + (size_t)(nx + 2) * ((j + 1) + (size_t)(ny + 2) * (k + 1)))];
  }

  void alloc() {
    u.assign((size_t)5 * (nx + 5) * (ny + 5) * (nz + 5), T(0));
    forcing.assign((size_t)5 * (nx + 1) * (ny + 1) * (nz + 1), T(0));
    rhs.assign((size_t)5 * (nx + 2) * (ny + 2) * (nz + 2), T(0));
    size_t nd = (size_t)(nx + 3) * (ny + 3) * (nz + 3);
    rho_i.assign(nd, T(0)); us.assign(nd, T(0)); vs.assign(nd, T(0));
    ws.assign(nd, T(0)); qs.assign(nd, T(0)); square.assign(nd, T(0));
    lhsc.assign((size_t)25 * (nx + 2) * (ny + 2) * (nz + 2), T(0));
  }
};

// Per-line scratch of one solve: fjac, njac (5x5 x -2:MAX+1), lhsa, lhsb
// (5x5 x 0:MAX) as in src/BT/bt_data.jl:204-236.
template <class T>
struct LineScratch {
  int off = 2;
  std::vector<T> fjac, njac, lhsa, lhsb;
  explicit LineScratch(int maxn) {
    fjac.assign((size_t)25 * (maxn + 6), T(0));
    njac.assign((size_t)25 * (maxn + 6), T(0));
    lhsa.assign((size_t)25 * (maxn + 6), T(0));
    lhsb.assign((size_t)25 * (maxn + 6), T(0));
  }
  T& f(int m, int n, int i) { return fjac[m + 5 * n + 25 * (i + off)]; }
  T& nj(int m, int n, int i) { return njac[m + 5 * n + 25 * (i + off)]; }
  T& a(int m, int n, int i) { return lhsa[m + 5 * n + 25 * (i + off)]; }
  T& b(int m, int n, int i) { return lhsb[m + 5 * n + 25 * (i + off)]; }
};

// ---------------------------------------------------------------- initialize
// src/BT/initialize.jl:26-225.
template <class T>
void initialize(Zone<T>& z, const Consts& K) {
  for (int kk = -1; kk <= z.nz + 1; ++kk)
    for (int jj = -1; jj <= z.ny + 1; ++jj)
      for (int ii = -1; ii <= z.nx + 1; ++ii)
        for (int m = 0; m < 5; ++m) z.U(m, ii, jj, kk) = T(1.0);
  double Pface1[2][5], Pface2[2][5], Pface3[2][5];
  for (int k = 0; k < z.nz; ++k) {
    double zeta = double(k) * K.dnzm1;
    for (int j = 0; j < z.ny; ++j) {
      double eta = double(j) * K.dnym1;
      for (int i = 0; i < z.nx; ++i) {
        double xi = double(i) * K.dnxm1;
        for (int ix = 0; ix < 2; ++ix) exact_solution(double(ix), eta, zeta, Pface1[ix]);
        for (int iy = 0; iy < 2; ++iy) exact_solution(xi, double(iy), zeta, Pface2[iy]);
        for (int iz = 0; iz < 2; ++iz) exact_solution(xi, eta, double(iz), Pface3[iz]);
        for (int m = 0; m < 5; ++m) {
          double Pxi = xi * Pface1[1][m] + (1.0 - xi) * Pface1[0][m];
          double Peta = eta * Pface2[1][m] + (1.0 - eta) * Pface2[0][m];
          double Pzeta = zeta * Pface3[1][m] + (1.0 - zeta) * Pface3[0][m];
          z.U(m, i, j, k) = T(Pxi + Peta + Pzeta - Pxi * Peta - Pxi * Pzeta - Peta * Pzeta +
                              Pxi * Peta * Pzeta);
        }
      }
    }
  }
  double temp[5];
  // west, east, south, north, bottom, top: later faces overwrite edges.
  for (int k = 0; k < z.nz; ++k) {
    double zeta = double(k) * K.dnzm1;
    for (int j = 0; j < z.ny; ++j) {
      double eta = double(j) * K.dnym1;
      exact_solution(0.0, eta, zeta, temp);
      for (int m = 0; m < 5; ++m) z.U(m, 0, j, k) = T(temp[m]);
    }
  }
  for (int k = 0; k < z.nz; ++k) {
    double zeta = double(k) * K.dnzm1;
    for (int j = 0; j < z.ny; ++j) {
      double eta = double(j) * K.dnym1;
      exact_solution(1.0, eta, zeta, temp);
      for (int m = 0; m < 5; ++m) z.U(m, z.nx - 1, j, k) = T(temp[m]);
    }
  }
  for (int k = 0; k < z.nz; ++k) {
    double zeta = double(k) * K.dnzm1;
    for (int i = 0; i < z.nx; ++i) {
      double xi = double(i) * K.dnxm1;
      exact_solution(xi, 0.0, zeta, temp);
      for (int m = 0; m < 5; ++m) z.U(m, i, 0, k) = T(temp[m]);
    }
  }
  for (int k = 0; k < z.nz; ++k) {
    double zeta = double(k) * K.dnzm1;
    for (int i = 0; i < z.nx; ++i) {
      double xi = double(i) * K.dnxm1;
      exact_solution(xi, 1.0, zeta, temp);
      for (int m = 0; m < 5; ++m) z.U(m, i, z.ny - 1, k) = T(temp[m]);
    }
  }
  for (int j = 0; j < z.ny; ++j) {
    double eta = double(j) * K.dnym1;
    for (int i = 0; i < z.nx; ++i) {
      double xi = double(i) * K.dnxm1;
      exact_solution(xi, eta, 0.0, temp);
      for (int m = 0; m < 5; ++m) z.U(m, i, j, 0) = T(temp[m]);
    }
  }
  for (int j = 0; j < z.ny; ++j) {
    double eta = double(j) * K.dnym1;
    for (int i = 0; i < z.nx; ++i) {
      double xi = double(i) * K.dnxm1;
      exact_solution(xi, eta, 1.0, temp);
      for (int m = 0; m < 5; ++m) z.U(m, i, j, z.nz - 1) = T(temp[m]);
    }
  }
}

// ---------------------------------------------------------------- exact_rhs
// src/BT/exact_rhs.jl:5-370.  One direction at a time; `v` is the index
// (1..3) of the velocity component along the line (ue column v+1 in Julia).
// The three direction blocks of the reference are the same code with the
// roles of the momentum components permuted; `oth` lists the other two
// components in the order the reference adds their squares.
template <class T>
void exact_rhs_dir(Zone<T>& z, const Consts& K, int dir) {
  const int n[3] = {z.nx, z.ny, z.nz};
  const int v = dir + 1;  // 0-based component index of the line velocity
  const int oth[3][2] = {{2, 3}, {1, 3}, {1, 2}};
  const double t2 = dir == 0 ? K.tx2 : dir == 1 ? K.ty2 : K.tz2;
  const double dd[5] = {dir == 0 ? K.dx1tx1 : dir == 1 ? K.dy1ty1 : K.dz1tz1,
                        dir == 0 ? K.dx2tx1 : dir == 1 ? K.dy2ty1 : K.dz2tz1,
                        dir == 0 ? K.dx3tx1 : dir == 1 ? K.dy3ty1 : K.dz3tz1,
                        dir == 0 ? K.dx4tx1 : dir == 1 ? K.dy4ty1 : K.dz4tz1,
                        dir == 0 ? K.dx5tx1 : dir == 1 ? K.dy5ty1 : K.dz5tz1};
  const double con1 = dir == 0 ? K.xxcon1 : dir == 1 ? K.yycon1 : K.zzcon1;
  const double con2 = dir == 0 ? K.xxcon2 : dir == 1 ? K.yycon2 : K.zzcon2;
  const double con3 = dir == 0 ? K.xxcon3 : dir == 1 ? K.yycon3 : K.zzcon3;
  const double con4 = dir == 0 ? K.xxcon4 : dir == 1 ? K.yycon4 : K.zzcon4;
  const double con5 = dir == 0 ? K.xxcon5 : dir == 1 ? K.yycon5 : K.zzcon5;
  const double dssp = K.dssp;
  const int nl = n[dir];
  // the two transverse directions, outer loop first as in the reference
  int da, db;
  if (dir == 0) { da = 2; db = 1; }        // k outer, j inner
  else if (dir == 1) { da = 2; db = 0; }   // k outer, i inner
  else { da = 1; db = 0; }                 // j outer, i inner
  std::vector<double> ue((size_t)(nl + 4) * 5), buf((size_t)(nl + 4) * 5), cuf(nl + 4), q(nl + 4);
  auto UE = [&](int i, int m) -> double& { return ue[(i + 2) * 5 + m]; };
  auto BUF = [&](int i, int m) -> double& { return buf[(i + 2) * 5 + m]; };
  auto CUF = [&](int i) -> double& { return cuf[i + 2]; };
  auto Q = [&](int i) -> double& { return q[i + 2]; };
  const double dn[3] = {K.dnxm1, K.dnym1, K.dnzm1};
  for (int a = 1; a <= n[da] - 2; ++a) {
    for (int b = 1; b <= n[db] - 2; ++b) {
      int idx[3];
      idx[da] = a; idx[db] = b;
      for (int i = 0; i <= nl - 1; ++i) {
        idx[dir] = i;
        double xi = double(idx[0]) * dn[0], eta = double(idx[1]) * dn[1],
               zeta = double(idx[2]) * dn[2];
        double dtemp[5];
        exact_solution(xi, eta, zeta, dtemp);
        for (int m = 0; m < 5; ++m) UE(i, m) = dtemp[m];
        double dtpp = 1.0 / dtemp[0];
        for (int m = 1; m < 5; ++m) BUF(i, m) = dtpp * dtemp[m];
        CUF(i) = BUF(i, v) * BUF(i, v);
        BUF(i, 0) = CUF(i) + BUF(i, oth[dir][0]) * BUF(i, oth[dir][0]) +
                    BUF(i, oth[dir][1]) * BUF(i, oth[dir][1]);
        Q(i) = 0.5 * (BUF(i, 1) * UE(i, 1) + BUF(i, 2) * UE(i, 2) + BUF(i, 3) * UE(i, 3));
      }
      for (int i = 1; i <= nl - 2; ++i) {
        idx[dir] = i;
        int im1 = i - 1, ip1 = i + 1;
        T* f[5];
        for (int m = 0; m < 5; ++m) f[m] = &z.F(m, idx[0], idx[1], idx[2]);
        *f[0] = *f[0] - t2 * (UE(ip1, v) - UE(im1, v)) +
                dd[0] * (UE(ip1, 0) - 2.0 * UE(i, 0) + UE(im1, 0));
        for (int m = 1; m <= 3; ++m) {
          if (m == v) {
            *f[m] = *f[m] -
                    t2 * ((UE(ip1, m) * BUF(ip1, v) + K.c2 * (UE(ip1, 4) - Q(ip1))) -
                          (UE(im1, m) * BUF(im1, v) + K.c2 * (UE(im1, 4) - Q(im1)))) +
                    con1 * (BUF(ip1, m) - 2.0 * BUF(i, m) + BUF(im1, m)) +
                    dd[m] * (UE(ip1, m) - 2.0 * UE(i, m) + UE(im1, m));
          } else {
            *f[m] = *f[m] - t2 * (UE(ip1, m) * BUF(ip1, v) - UE(im1, m) * BUF(im1, v)) +
                    con2 * (BUF(ip1, m) - 2.0 * BUF(i, m) + BUF(im1, m)) +
                    dd[m] * (UE(ip1, m) - 2.0 * UE(i, m) + UE(im1, m));
          }
        }
        *f[4] = *f[4] -
                t2 * (BUF(ip1, v) * (K.c1 * UE(ip1, 4) - K.c2 * Q(ip1)) -
                      BUF(im1, v) * (K.c1 * UE(im1, 4) - K.c2 * Q(im1))) +
                0.5 * con3 * (BUF(ip1, 0) - 2.0 * BUF(i, 0) + BUF(im1, 0)) +
                con4 * (CUF(ip1) - 2.0 * CUF(i) + CUF(im1)) +
                con5 * (BUF(ip1, 4) - 2.0 * BUF(i, 4) + BUF(im1, 4)) +
                dd[4] * (UE(ip1, 4) - 2.0 * UE(i, 4) + UE(im1, 4));
      }
      // fourth-order dissipation, same loop sequence as the reference
      auto FF = [&](int m, int i) -> T& {
        idx[dir] = i;
        return z.F(m, idx[0], idx[1], idx[2]);
      };
      for (int m = 0; m < 5; ++m) {
        int i = 1;
        FF(m, i) = FF(m, i) - dssp * (5.0 * UE(i, m) - 4.0 * UE(i + 1, m) + UE(i + 2, m));
        i = 2;
        FF(m, i) = FF(m, i) - dssp * (-4.0 * UE(i - 1, m) + 6.0 * UE(i, m) -
                                      4.0 * UE(i + 1, m) + UE(i + 2, m));
      }
      for (int m = 0; m < 5; ++m)
        for (int i = 3; i <= nl - 4; ++i)
          FF(m, i) = FF(m, i) - dssp * (UE(i - 2, m) - 4.0 * UE(i - 1, m) + 6.0 * UE(i, m) -
                                        4.0 * UE(i + 1, m) + UE(i + 2, m));
      for (int m = 0; m < 5; ++m) {
        int i = nl - 3;
        FF(m, i) = FF(m, i) - dssp * (UE(i - 2, m) - 4.0 * UE(i - 1, m) + 6.0 * UE(i, m) -
                                      4.0 * UE(i + 1, m));
        i = nl - 2;
        FF(m, i) = FF(m, i) - dssp * (UE(i - 2, m) - 4.0 * UE(i - 1, m) + 5.0 * UE(i, m));
      }
    }
  }
}

template <class T>
void exact_rhs(Zone<T>& z, const Consts& K) {
  for (int m = 0; m < 5; ++m)
    for (int k = 0; k < z.nz; ++k)
      for (int j = 0; j < z.ny; ++j)
        for (int i = 0; i < z.nx; ++i) z.F(m, i, j, k) = T(0.0);
  exact_rhs_dir(z, K, 0);
  exact_rhs_dir(z, K, 1);
  exact_rhs_dir(z, K, 2);
  for (int m = 0; m < 5; ++m)
    for (int k = 1; k <= z.nz - 2; ++k)
      for (int j = 1; j <= z.ny - 2; ++j)
        for (int i = 1; i <= z.nx - 2; ++i) z.F(m, i, j, k) = -1.0 * z.F(m, i, j, k);
}

// ---------------------------------------------------------------- compute_rhs
// src/BT/rhs.jl:4-460.
template <class T>
void compute_rhs(Zone<T>& z, const Consts& K) {
  const int nx = z.nx, ny = z.ny, nz = z.nz;
  for (int k = -1; k <= nz; ++k)
    for (int j = -1; j <= ny; ++j)
      for (int i = -1; i <= nx; ++i) {
        size_t d = z.didx(i, j, k);
        T rho_inv = 1.0 / z.U(0, i, j, k);
        z.rho_i[d] = rho_inv;
        z.us[d] = z.U(1, i, j, k) * rho_inv;
        z.vs[d] = z.U(2, i, j, k) * rho_inv;
        z.ws[d] = z.U(3, i, j, k) * rho_inv;
        z.square[d] = 0.5 * (z.U(1, i, j, k) * z.U(1, i, j, k) + z.U(2, i, j, k) * z.U(2, i, j, k) +
                             z.U(3, i, j, k) * z.U(3, i, j, k)) * rho_inv;
        z.qs[d] = z.square[d] * rho_inv;
      }
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i)
        for (int m = 0; m < 5; ++m) z.R(m, i, j, k) = z.F(m, i, j, k);

  // One direction's fluxes + dissipation.  s = stride vector of the line.
  for (int dir = 0; dir < 3; ++dir) {
    const int v = dir + 1;
    const std::vector<T>& vel = dir == 0 ? z.us : dir == 1 ? z.vs : z.ws;
    const std::vector<T>* velm[5] = {nullptr, &z.us, &z.vs, &z.ws, nullptr};
    const double t2 = dir == 0 ? K.tx2 : dir == 1 ? K.ty2 : K.tz2;
    const double dd[5] = {dir == 0 ? K.dx1tx1 : dir == 1 ? K.dy1ty1 : K.dz1tz1,
                          dir == 0 ? K.dx2tx1 : dir == 1 ? K.dy2ty1 : K.dz2tz1,
                          dir == 0 ? K.dx3tx1 : dir == 1 ? K.dy3ty1 : K.dz3tz1,
                          dir == 0 ? K.dx4tx1 : dir == 1 ? K.dy4ty1 : K.dz4tz1,
                          dir == 0 ? K.dx5tx1 : dir == 1 ? K.dy5ty1 : K.dz5tz1};
    const double con2 = dir == 0 ? K.xxcon2 : dir == 1 ? K.yycon2 : K.zzcon2;
    const double con3 = dir == 0 ? K.xxcon3 : dir == 1 ? K.yycon3 : K.zzcon3;
    const double con4 = dir == 0 ? K.xxcon4 : dir == 1 ? K.yycon4 : K.zzcon4;
    const double con5 = dir == 0 ? K.xxcon5 : dir == 1 ? K.yycon5 : K.zzcon5;
    const int di = dir == 0, dj = dir == 1, dk = dir == 2;
    for (int k = 1; k <= nz - 2; ++k)
      for (int j = 1; j <= ny - 2; ++j)
        for (int i = 1; i <= nx - 2; ++i) {
          size_t c0 = z.didx(i, j, k), cp = z.didx(i + di, j + dj, k + dk),
                 cm = z.didx(i - di, j - dj, k - dk);
          T vijk = vel[c0], vp1 = vel[cp], vm1 = vel[cm];
          auto up = [&](int m) -> T { return z.U(m, i + di, j + dj, k + dk); };
          auto u0 = [&](int m) -> T { return z.U(m, i, j, k); };
          auto um = [&](int m) -> T { return z.U(m, i - di, j - dj, k - dk); };
          z.R(0, i, j, k) += dd[0] * (up(0) - 2.0 * u0(0) + um(0)) - t2 * (up(v) - um(v));
          for (int m = 1; m <= 3; ++m) {
            const std::vector<T>& vm = *velm[m];
            if (m == v) {
              z.R(m, i, j, k) +=
                  dd[m] * (up(m) - 2.0 * u0(m) + um(m)) +
                  con2 * K.con43 * (vp1 - 2.0 * vijk + vm1) -
                  t2 * (up(m) * vp1 - um(m) * vm1 +
                        (up(4) - z.square[cp] - um(4) + z.square[cm]) * K.c2);
            } else {
              z.R(m, i, j, k) += dd[m] * (up(m) - 2.0 * u0(m) + um(m)) +
                                 con2 * (vm[cp] - 2.0 * vm[c0] + vm[cm]) -
                                 t2 * (up(m) * vp1 - um(m) * vm1);
            }
          }
          z.R(4, i, j, k) +=
              dd[4] * (up(4) - 2.0 * u0(4) + um(4)) +
              con3 * (z.qs[cp] - 2.0 * z.qs[c0] + z.qs[cm]) +
              con4 * (vp1 * vp1 - 2.0 * vijk * vijk + vm1 * vm1) +
              con5 * (up(4) * z.rho_i[cp] - 2.0 * u0(4) * z.rho_i[c0] + um(4) * z.rho_i[cm]) -
              t2 * ((K.c1 * up(4) - K.c2 * z.square[cp]) * vp1 -
                    (K.c1 * um(4) - K.c2 * z.square[cm]) * vm1);
        }
    // Dissipation along dir (rhs.jl:158-210, 269-324, 384-443).  The
    // reference runs five loops: row 1, row 2, rows 3..n-4, row n-3, row n-2.
    // Each point takes the updates of the loops whose rows contain it, in that
    // loop order (more than one only when n < 6).
    const int nl = dir == 0 ? nx : dir == 1 ? ny : nz;
    for (int k = 1; k <= nz - 2; ++k)
      for (int j = 1; j <= ny - 2; ++j)
        for (int i = 1; i <= nx - 2; ++i) {
          const int p[3] = {i, j, k};
          const int l = p[dir];
          for (int m = 0; m < 5; ++m) {
            auto uu = [&](int o) -> T { return z.U(m, i + di * o, j + dj * o, k + dk * o); };
            T& r = z.R(m, i, j, k);
            if (l == 1) r -= K.dssp * (5.0 * uu(0) - 4.0 * uu(1) + uu(2));
            if (l == 2) r -= K.dssp * (-4.0 * uu(-1) + 6.0 * uu(0) - 4.0 * uu(1) + uu(2));
            if (l >= 3 && l <= nl - 4)
              r -= K.dssp * (uu(-2) - 4.0 * uu(-1) + 6.0 * uu(0) - 4.0 * uu(1) + uu(2));
            if (l == nl - 3) r -= K.dssp * (uu(-2) - 4.0 * uu(-1) + 6.0 * uu(0) - 4.0 * uu(1));
            if (l == nl - 2) r -= K.dssp * (uu(-2) - 4.0 * uu(-1) + 5.0 * uu(0));
          }
        }
  }
  for (int k = 1; k <= nz - 2; ++k)
    for (int j = 1; j <= ny - 2; ++j)
      for (int i = 1; i <= nx - 2; ++i)
        for (int m = 0; m < 5; ++m) z.R(m, i, j, k) *= K.dt;
}

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

// add.jl:5-43
template <class T>
void add(Zone<T>& z) {
  for (int k = 1; k <= z.nz - 2; ++k)
    for (int j = 1; j <= z.ny - 2; ++j)
      for (int i = 1; i <= z.nx - 2; ++i)
        for (int m = 0; m < 5; ++m) z.U(m, i, j, k) += z.R(m, i, j, k);
}

// ---------------------------------------------------------------- exch_qbc
// src/BT/exchange_qbc.jl:1-466 with every zone on one rank: all faces are
// packed first (exch_qbc_send, :145-320), then unpacked (exch_qbc_recv,
// :323-466).  Zone z's i = 1 plane feeds its west neighbour's i = nx-1 plane,
// its i = nx-2 plane its east neighbour's i = 0 plane, and likewise in j.
template <class T>
void exch_qbc(std::vector<Zone<T>>& Z, const Zones& G) {
  const int nzn = G.count();
  std::vector<std::vector<T>> east_in(nzn), west_in(nzn), north_in(nzn), south_in(nzn);
  for (int z = 0; z < nzn; ++z) {
    Zone<T>& a = Z[z];
    // what zone z receives on its west (i=0) face: west neighbour's i=nx-2
    Zone<T>& w = Z[G.west[z]];
    Zone<T>& e = Z[G.east[z]];
    Zone<T>& s = Z[G.south[z]];
    Zone<T>& nn = Z[G.north[z]];
    for (int m = 0; m < 5; ++m)
      for (int k = 1; k <= a.nz - 2; ++k)
        for (int j = 1; j <= a.ny - 2; ++j) {
          west_in[z].push_back(w.U(m, w.nx - 2, j, k));
          east_in[z].push_back(e.U(m, 1, j, k));
        }
    for (int m = 0; m < 5; ++m)
      for (int k = 1; k <= a.nz - 2; ++k)
        for (int i = 1; i <= a.nx - 2; ++i) {
          south_in[z].push_back(s.U(m, i, s.ny - 2, k));
          north_in[z].push_back(nn.U(m, i, 1, k));
        }
  }
  for (int z = 0; z < nzn; ++z) {
    Zone<T>& a = Z[z];
    size_t p = 0;
    for (int m = 0; m < 5; ++m)
      for (int k = 1; k <= a.nz - 2; ++k)
        for (int j = 1; j <= a.ny - 2; ++j, ++p) {
          a.U(m, 0, j, k) = west_in[z][p];
          a.U(m, a.nx - 1, j, k) = east_in[z][p];
        }
    p = 0;
    for (int m = 0; m < 5; ++m)
      for (int k = 1; k <= a.nz - 2; ++k)
        for (int i = 1; i <= a.nx - 2; ++i, ++p) {
          a.U(m, i, 0, k) = south_in[z][p];
          a.U(m, i, a.ny - 1, k) = north_in[z][p];
        }
  }
}

// ---------------------------------------------------------------- norms
// error_norm / rhs_norm: src/BT/error.jl:7-79.
template <class T>
void error_norm(Zone<T>& z, const Consts& K, T rms[5]) {
  for (int m = 0; m < 5; ++m) rms[m] = T(0.0);
  for (int k = 0; k < z.nz; ++k) {
    double zeta = double(k) * K.dnzm1;
    for (int j = 0; j < z.ny; ++j) {
      double eta = double(j) * K.dnym1;
      for (int i = 0; i < z.nx; ++i) {
        double xi = double(i) * K.dnxm1;
        double ue[5];
        exact_solution(xi, eta, zeta, ue);
        for (int m = 0; m < 5; ++m) {
          T add = z.U(m, i, j, k) - ue[m];
          rms[m] = rms[m] + add * add;
        }
      }
    }
  }
  for (int m = 0; m < 5; ++m) {
    rms[m] = rms[m] / (double(z.nx - 2) * double(z.ny - 2) * double(z.nz - 2));
    rms[m] = std::sqrt(rms[m]);
  }
}

template <class T>
void rhs_norm(Zone<T>& z, T rms[5]) {
  for (int m = 0; m < 5; ++m) rms[m] = T(0.0);
  for (int k = 1; k <= z.nz - 2; ++k)
    for (int j = 1; j <= z.ny - 2; ++j)
      for (int i = 1; i <= z.nx - 2; ++i)
        for (int m = 0; m < 5; ++m) {
          T add = z.R(m, i, j, k);
          rms[m] = rms[m] + add * add;
        }
  for (int m = 0; m < 5; ++m) {
    rms[m] = rms[m] / (double(z.nx - 2) * double(z.ny - 2) * double(z.nz - 2));
    rms[m] = std::sqrt(rms[m]);
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

}  // extern "C"
