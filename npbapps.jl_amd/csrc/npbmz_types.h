// Types shared by the host side of libnpbmz and its gfx950 kernels.
#pragma once
#include <stdint.h>

namespace npbmz {

// One zone as the kernels see it.  Fields live in HBM as one array per field
// for all zones; zone z's grid point (i, j, k), 0 <= i < nx etc., is at point
// index base + (k*ny + j)*nx + i and its 5 components are contiguous
// (the reference's [m,i,j,k] order with m fastest).
struct ZoneDev {
  int32_t nx, ny, nz, pad;
  int64_t base;
};

// A unit of work: `count` consecutive items (points or lines) of zone `zone`
// starting at item `first`, in the zone's own item numbering.
struct Task {
  int32_t zone;
  int32_t first;
};

// Stencil constants for BT and SP (src/BT/set_constants.jl:1-130 and
// src/SP/set_constants.jl), grouped by direction d = 0 (xi), 1 (eta),
// 2 (zeta) so a kernel templated on the direction indexes them statically.
struct StencilK {
  double dt;
  double c1, c2, c3, c4, c5, c1c2, c1c5, c3c4, c1345, con43, bt, c2iv, con16;
  double dssp, c4dssp, c5dssp, dtdssp, comz1, comz4, comz5, comz6;
  double dnm1[3];     // dnxm1, dnym1, dnzm1
  double t1[3];       // tx1, ty1, tz1
  double t2[3];       // tx2, ty2, tz2
  double t3[3];       // tx3, ty3, tz3
  double dd[3][5];    // dx1..dx5, dy1..dy5, dz1..dz5
  double ddt1[3][5];  // dx1tx1..dx5tx1, dy1ty1.., dz1tz1..
  double con1[3];     // xxcon1, yycon1, zzcon1
  double con2[3];     // xxcon2, ...
  double con3[3];
  double con4[3];
  double con5[3];
  double dtt1[3];     // dttx1, dtty1, dttz1
  double dtt2[3];     // dttx2, dtty2, dttz2
  double c2dtt1[3];   // c2dttx1, ...
  double dmax[3];     // dxmax, dymax, dzmax
};

// LU-MZ constants: setcoeff (src/LU/setcoeff.jl:43-70) for the zone size,
// omega (src/LU/lu_data.jl:8) and the jacld/jacu scalars r43, c34, c1345
// (src/LU/jacld.jl:22-24).  LU zones all have the same size, so one set
// serves every zone.
struct LuK {
  double dt, omega, tmp;  // tmp = 1/(omega*(2-omega)), src/LU/ssor.jl:57
  double tx1, tx2, tx3, ty1, ty2, ty3, tz1, tz2, tz3;
  double dxi, deta, dzeta;
  double r43, c34, c1345;
  // jacld/jacu's constant factors collapsed (FMA arithmetic only): per
  // direction d with t1, t2 = tx1, tx2 (ty, tz) and dd = dxc (dyc, dzc),
  // a = dt*t2 (jacu's sign; jacld's is -a), v = dt*t1:
  struct Nb {
    double a, a_c2, a_2mc2, a_c1, a_c2h;       // a, a*c2, a*(2-c2), a*c1, a*c2/2
    double v_dd, v_r43c34, v_c34, v_c1345;     // v*dd, v*r43*c34, v*c34, v*c1345
    double v_ra, v_rb;                         // v*(r43*c34-c1345), v*(c34-c1345)
  } nb[3];
  // the diagonal block: d0 = 1 + 2dt(tx1 dxc + ty1 dyc + tz1 dzc); g[i] =
  // 2dt c34 (tx1 + ty1 + tz1 with r43 on direction i); h[i] = 2dt((r43 c34 -
  // c1345) on direction i's t1 + (c34 - c1345) on the others); h5 = 2dt c1345
  // (tx1 + ty1 + tz1)
  double d0, g[3], h[3], h5;
};

}  // namespace npbmz
