// Test infrastructure only: the CPU oracle for the multi-zone BT/SP/LU timestep.
//
// Nothing under oracle/ is part of the product.  Only tests/, the smoke check
// and the cpu_baseline leg of bench.py load it, and only as the checker or as
// the reported CPU baseline.  The product (npbapps.jl_amd/) never links it.
//
// This header holds the pieces every benchmark shares: the class table, the
// geometric zone split, the exact-solution polynomial and the stencil
// constants.  Each function cites the Julia file:line it restates.  All
// arithmetic is IEEE binary64 compiled with -ffp-contract=off, so every
// expression is evaluated in the reference's left-to-right order with no
// fused multiply-adds (Julia does not contract either).
#pragma once
#include <cmath>
#include <cstdint>
#include <vector>

namespace npb {

// Zone geometry of one multi-zone run.  zone numbers are 0-based here; the
// reference numbers them from 1 (src/BT/zone_setup.jl:55).
struct Zones {
  int x_zones = 0, y_zones = 0, gx = 0, gy = 0, gz = 0;
  std::vector<int> nx, ny, nz;                          // per zone
  std::vector<int> west, east, south, north;            // 0-based neighbours
  int count() const { return x_zones * y_zones; }
};

// src/BT/zone_setup.jl:1-77 (SP identical, src/SP/zone_setup.jl; LU uses
// round() instead of trunc() at src/LU/zone_setup.jl:15,19, selected by
// use_round).  x_r^i is evaluated with std::pow, which, like Julia's
// Float64^Int, is within one ulp; the +0.45 truncation absorbs that.
inline Zones zone_setup(int x_zones, int y_zones, int gx, int gy, int gz,
                        double ratio, bool use_round = false) {
  Zones z;
  z.x_zones = x_zones; z.y_zones = y_zones; z.gx = gx; z.gy = gy; z.gz = gz;
  std::vector<long> x_end(x_zones + 1), y_end(y_zones + 1);
  if (std::fabs(ratio - 1.0) > 1.0e-10 && x_zones > 1 && y_zones > 1) {
    double x_r = std::exp(std::log(ratio) / (x_zones - 1));
    double y_r = std::exp(std::log(ratio) / (y_zones - 1));
    double x_smallest = double(gx) * (x_r - 1.0) / (std::pow(x_r, x_zones) - 1.0);
    double y_smallest = double(gy) * (y_r - 1.0) / (std::pow(y_r, y_zones) - 1.0);
    for (int i = 1; i <= x_zones; ++i) {
      double v = x_smallest * (std::pow(x_r, i) - 1.0) / (x_r - 1.0) + 0.45;
      x_end[i] = use_round ? long(std::nearbyint(v)) : long(std::trunc(v));
    }
    for (int j = 1; j <= y_zones; ++j) {
      double v = y_smallest * (std::pow(y_r, j) - 1.0) / (y_r - 1.0) + 0.45;
      y_end[j] = use_round ? long(std::nearbyint(v)) : long(std::trunc(v));
    }
  } else {
    // (i*gx_size)/x_zones is a Float64 division stored into an Int array;
    // it is exact for every class, so integer division gives the same value.
    for (int i = 1; i <= x_zones; ++i) x_end[i] = (long(i) * gx) / x_zones;
    for (int j = 1; j <= y_zones; ++j) y_end[j] = (long(j) * gy) / y_zones;
  }
  std::vector<long> x_start(x_zones + 1), y_start(y_zones + 1), x_size(x_zones + 1),
      y_size(y_zones + 1);
  x_start[1] = 1;
  for (int i = 1; i <= x_zones; ++i) {
    if (i != x_zones) x_start[i + 1] = x_end[i] + 1;
    x_size[i] = x_end[i] - x_start[i] + 1;
  }
  y_start[1] = 1;
  for (int j = 1; j <= y_zones; ++j) {
    if (j != y_zones) y_start[j + 1] = y_end[j] + 1;
    y_size[j] = y_end[j] - y_start[j] + 1;
  }
  int n = x_zones * y_zones;
  z.nx.resize(n); z.ny.resize(n); z.nz.resize(n);
  z.west.resize(n); z.east.resize(n); z.south.resize(n); z.north.resize(n);
  for (int j = 1; j <= y_zones; ++j)
    for (int i = 1; i <= x_zones; ++i) {
      int zone = (i - 1) + (j - 1) * x_zones;  // 0-based
      z.nx[zone] = int(x_size[i]);
      z.ny[zone] = int(y_size[j]);
      z.nz[zone] = gz;
      int id_west = (i - 2 + x_zones) % x_zones;
      int id_east = i % x_zones;
      int jd_south = (j - 2 + y_zones) % y_zones;
      int jd_north = j % y_zones;
      z.west[zone] = id_west + (j - 1) * x_zones;
      z.east[zone] = id_east + (j - 1) * x_zones;
      z.south[zone] = (i - 1) + jd_south * x_zones;
      z.north[zone] = (i - 1) + jd_north * x_zones;
    }
  return z;
}

// ce table, src/BT/set_constants.jl:47-51 (identical in SP; LU's table is in
// src/LU/setcoeff.jl and is the same matrix).  ce[m][n] with m = 0..4 and
// n = 0..12 is the reference's ce[m+1, n+1].
static const double CE[5][13] = {
    {2.0, 0.0, 0.0, 4.0, 5.0, 3.0, 0.5, 0.02, 0.01, 0.03, 0.5, 0.4, 0.3},
    {1.0, 0.0, 0.0, 0.0, 1.0, 2.0, 3.0, 0.01, 0.03, 0.02, 0.4, 0.3, 0.5},
    {2.0, 2.0, 0.0, 0.0, 0.0, 2.0, 3.0, 0.04, 0.03, 0.05, 0.3, 0.5, 0.4},
    {2.0, 2.0, 0.0, 0.0, 0.0, 2.0, 3.0, 0.03, 0.05, 0.04, 0.2, 0.1, 0.3},
    {5.0, 4.0, 3.0, 2.0, 0.1, 0.4, 0.3, 0.05, 0.04, 0.03, 0.1, 0.3, 0.2}};

// src/BT/exact_solution.jl:5-18: Horner forms summed left to right.
inline void exact_solution(double xi, double eta, double zeta, double out[5]) {
  for (int m = 0; m < 5; ++m) {
    const double* c = CE[m];
    out[m] = c[0] + xi * (c[1] + xi * (c[4] + xi * (c[7] + xi * c[10]))) +
             eta * (c[2] + eta * (c[5] + eta * (c[8] + eta * c[11]))) +
             zeta * (c[3] + zeta * (c[6] + zeta * (c[9] + zeta * c[12])));
  }
}

// Stencil constants shared by BT and SP: src/BT/set_constants.jl:1-130.
struct Consts {
  double c1, c2, c3, c4, c5, bt, c1c2, c1c5, c3c4, c1345, conz1;
  double dx1, dx2, dx3, dx4, dx5, dy1, dy2, dy3, dy4, dy5, dz1, dz2, dz3, dz4, dz5;
  double dxmax, dymax, dzmax, dssp, c4dssp, c5dssp, c2iv, con43, con16;
  double dnxm1, dnym1, dnzm1, tx1, tx2, tx3, ty1, ty2, ty3, tz1, tz2, tz3;
  double dttx1, dttx2, dtty1, dtty2, dttz1, dttz2, c2dttx1, c2dtty1, c2dttz1;
  double dtdssp, comz1, comz4, comz5, comz6, c3c4tx3, c3c4ty3, c3c4tz3;
  double dx1tx1, dx2tx1, dx3tx1, dx4tx1, dx5tx1, dy1ty1, dy2ty1, dy3ty1, dy4ty1, dy5ty1;
  double dz1tz1, dz2tz1, dz3tz1, dz4tz1, dz5tz1;
  double xxcon1, xxcon2, xxcon3, xxcon4, xxcon5, yycon1, yycon2, yycon3, yycon4, yycon5;
  double zzcon1, zzcon2, zzcon3, zzcon4, zzcon5;
  double dt;
};

// src/BT/set_constants.jl:1-49 (module constants) and :53-130 (set_constants).
// SP's copy (src/SP/set_constants.jl) only differs in dx/dy/dz values, which
// the caller passes through dxyz (15 values, dx1..dx5, dy1..dy5, dz1..dz5).
inline Consts set_constants(double dt, int gx, int gy, int gz, int x_zones, int y_zones,
                            const double dxyz[15]) {
  Consts k{};
  k.c1 = 1.4; k.c2 = 0.4; k.c3 = 0.1; k.c4 = 1.0; k.c5 = 1.4;
  k.bt = std::sqrt(0.5);
  k.c1c2 = k.c1 * k.c2; k.c1c5 = k.c1 * k.c5; k.c3c4 = k.c3 * k.c4;
  k.c1345 = k.c1c5 * k.c3c4; k.conz1 = (1.0 - k.c1c5);
  k.dx1 = dxyz[0]; k.dx2 = dxyz[1]; k.dx3 = dxyz[2]; k.dx4 = dxyz[3]; k.dx5 = dxyz[4];
  k.dy1 = dxyz[5]; k.dy2 = dxyz[6]; k.dy3 = dxyz[7]; k.dy4 = dxyz[8]; k.dy5 = dxyz[9];
  k.dz1 = dxyz[10]; k.dz2 = dxyz[11]; k.dz3 = dxyz[12]; k.dz4 = dxyz[13]; k.dz5 = dxyz[14];
  k.dxmax = std::fmax(k.dx3, k.dx4);
  k.dymax = std::fmax(k.dy2, k.dy4);
  k.dzmax = std::fmax(k.dz2, k.dz3);
  k.dssp = 0.25 * std::fmax(k.dx1, std::fmax(k.dy1, k.dz1));
  k.c4dssp = 4.0 * k.dssp; k.c5dssp = 5.0 * k.dssp;
  k.c2iv = 2.5; k.con43 = 4.0 / 3.0; k.con16 = 1.0 / 6.0;
  k.dnxm1 = 1.0 / (double(gx - 1) / double(x_zones));
  k.dnym1 = 1.0 / (double(gy - 1) / double(y_zones));
  k.dnzm1 = 1.0 / double(gz - 1);
  k.tx1 = 1.0 / (k.dnxm1 * k.dnxm1); k.tx2 = 1.0 / (2.0 * k.dnxm1); k.tx3 = 1.0 / k.dnxm1;
  k.ty1 = 1.0 / (k.dnym1 * k.dnym1); k.ty2 = 1.0 / (2.0 * k.dnym1); k.ty3 = 1.0 / k.dnym1;
  k.tz1 = 1.0 / (k.dnzm1 * k.dnzm1); k.tz2 = 1.0 / (2.0 * k.dnzm1); k.tz3 = 1.0 / k.dnzm1;
  k.dttx1 = dt * k.tx1; k.dttx2 = dt * k.tx2; k.dtty1 = dt * k.ty1; k.dtty2 = dt * k.ty2;
  k.dttz1 = dt * k.tz1; k.dttz2 = dt * k.tz2;
  k.c2dttx1 = 2.0 * k.dttx1; k.c2dtty1 = 2.0 * k.dtty1; k.c2dttz1 = 2.0 * k.dttz1;
  k.dtdssp = dt * k.dssp;
  k.comz1 = k.dtdssp; k.comz4 = 4.0 * k.dtdssp; k.comz5 = 5.0 * k.dtdssp; k.comz6 = 6.0 * k.dtdssp;
  k.c3c4tx3 = k.c3c4 * k.tx3; k.c3c4ty3 = k.c3c4 * k.ty3; k.c3c4tz3 = k.c3c4 * k.tz3;
  k.dx1tx1 = k.dx1 * k.tx1; k.dx2tx1 = k.dx2 * k.tx1; k.dx3tx1 = k.dx3 * k.tx1;
  k.dx4tx1 = k.dx4 * k.tx1; k.dx5tx1 = k.dx5 * k.tx1;
  k.dy1ty1 = k.dy1 * k.ty1; k.dy2ty1 = k.dy2 * k.ty1; k.dy3ty1 = k.dy3 * k.ty1;
  k.dy4ty1 = k.dy4 * k.ty1; k.dy5ty1 = k.dy5 * k.ty1;
  k.dz1tz1 = k.dz1 * k.tz1; k.dz2tz1 = k.dz2 * k.tz1; k.dz3tz1 = k.dz3 * k.tz1;
  k.dz4tz1 = k.dz4 * k.tz1; k.dz5tz1 = k.dz5 * k.tz1;
  k.xxcon1 = k.c3c4tx3 * k.con43 * k.tx3; k.xxcon2 = k.c3c4tx3 * k.tx3;
  k.xxcon3 = k.c3c4tx3 * k.conz1 * k.tx3; k.xxcon4 = k.c3c4tx3 * k.con16 * k.tx3;
  k.xxcon5 = k.c3c4tx3 * k.c1c5 * k.tx3;
  k.yycon1 = k.c3c4ty3 * k.con43 * k.ty3; k.yycon2 = k.c3c4ty3 * k.ty3;
  k.yycon3 = k.c3c4ty3 * k.conz1 * k.ty3; k.yycon4 = k.c3c4ty3 * k.con16 * k.ty3;
  k.yycon5 = k.c3c4ty3 * k.c1c5 * k.ty3;
  k.zzcon1 = k.c3c4tz3 * k.con43 * k.tz3; k.zzcon2 = k.c3c4tz3 * k.tz3;
  k.zzcon3 = k.c3c4tz3 * k.conz1 * k.tz3; k.zzcon4 = k.c3c4tz3 * k.con16 * k.tz3;
  k.zzcon5 = k.c3c4tz3 * k.c1c5 * k.tz3;
  k.dt = dt;
  return k;
}

}  // namespace npb
