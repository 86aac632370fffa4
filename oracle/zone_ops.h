// Test infrastructure only: zone-level pieces of the CPU oracle shared by
// BT-MZ and SP-MZ (the reference keeps two copies that agree except where
// noted): the zone arrays, initialize, exact_rhs, compute_rhs, add,
// exch_qbc and the verification norms.  Only tests/, the smoke check and
// bench.py's cpu_baseline leg load the oracle.
#pragma once
#include <cmath>
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
  // SP only: speed of sound and its inverse (src/SP/rhs.jl:85-88)
  std::vector<T> speed, ainv;
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
    speed.assign(nd, T(0)); ainv.assign(nd, T(0));
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
// src/BT/initialize.jl:26-225.  SP fills the halo with (1, 0, 0, 0, 1)
// instead of 1.0 (src/SP/initialize.jl:18-26); the halo only feeds derived
// fields the stencils never read.
template <class T>
void initialize(Zone<T>& z, const Consts& K, bool sp_halo = false) {
  for (int kk = -1; kk <= z.nz + 1; ++kk)
    for (int jj = -1; jj <= z.ny + 1; ++jj)
      for (int ii = -1; ii <= z.nx + 1; ++ii)
        for (int m = 0; m < 5; ++m)
          z.U(m, ii, jj, kk) = T(sp_halo && (m >= 1 && m <= 3) ? 0.0 : 1.0);
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
void compute_rhs(Zone<T>& z, const Consts& K, bool with_speed = false) {
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
        if (with_speed) {  // src/SP/rhs.jl:85-88
          T aux = K.c1c2 * rho_inv * (z.U(4, i, j, k) - z.square[d]);
          using std::sqrt;  // a dual scalar brings its own sqrt
          aux = sqrt(aux);
          z.speed[d] = aux;
          z.ainv[d] = 1.0 / aux;
        }
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
// y_swap: SP's exchange (src/SP/exchange_qbc.jl:247-295, 409-455) fills the
// j = 0 plane from the NORTH neighbour's j = ny-2 plane and the j = ny-1
// plane from the SOUTH neighbour's j = 1 plane, the mirror of BT.
void exch_qbc(std::vector<Zone<T>>& Z, const Zones& G, bool y_swap = false) {
  const int nzn = G.count();
  std::vector<std::vector<T>> east_in(nzn), west_in(nzn), north_in(nzn), south_in(nzn);
  for (int z = 0; z < nzn; ++z) {
    Zone<T>& a = Z[z];
    // what zone z receives on its west (i=0) face: west neighbour's i=nx-2
    Zone<T>& w = Z[G.west[z]];
    Zone<T>& e = Z[G.east[z]];
    Zone<T>& s = Z[y_swap ? G.north[z] : G.south[z]];
    Zone<T>& nn = Z[y_swap ? G.south[z] : G.north[z]];
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

}  // namespace bt
}  // namespace npb
