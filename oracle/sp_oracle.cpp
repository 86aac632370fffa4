// Test infrastructure only: CPU oracle for the SP-MZ timestep.
//
// A restatement of the reference's Julia SP path with every zone on one rank
// (one cell per zone).  The zone-level pieces (initialize, exact_rhs,
// compute_rhs, add, exch_qbc, norms) are the BT oracle's: the reference's SP
// copies compute the same values, except for the halo fill, the speed/ainv
// derived fields and the mirrored y exchange, which zone_ops.h takes as
// flags.  Fields are stored [m,i,j,k] like BT's; the reference's SP arrays are
// [i,j,k,m] (src/SP/sp_data.jl:115), which only changes memory order.  Built
// with -ffp-contract=off, every expression keeps the reference's order.
#include <cmath>
#include <vector>

#include "dual.h"
#include "zone_ops.h"

namespace npb {
namespace sp {

using bt::Zone;

// The pentadiagonal coefficients lhs(i,j,k,1:15) of a zone
// (src/SP/sp_data.jl:170-171), persistent across solves like the reference's.
template <class T>
struct Lhs {
  int nx, ny, nz;
  std::vector<T> v;
  void alloc(int a, int b, int c) {
    nx = a; ny = b; nz = c;
    v.assign((size_t)15 * a * b * c, T(0));
  }
  T& operator()(int n, int i, int j, int k) {
    return v[n + 15 * (i + (size_t)nx * (j + (size_t)ny * k))];
  }
};

// lhsinit (src/SP/initialize.jl:209-260): zero, then 1 on the three diagonals.
template <class T>
void lhsinit(Lhs<T>& L) {
  std::fill(L.v.begin(), L.v.end(), T(0.0));
  for (int n = 1; n <= 3; ++n)
    for (int k = 0; k < L.nz; ++k)
      for (int j = 0; j < L.ny; ++j)
        for (int i = 0; i < L.nx; ++i) L(5 * n - 3, i, j, k) = T(1.0);
}

// txinvr (src/SP/txinvr.jl:5-56) on every interior point.
template <class T>
void txinvr(Zone<T>& z, const Consts& K) {
  for (int k = 1; k <= z.nz - 2; ++k)
    for (int j = 1; j <= z.ny - 2; ++j)
      for (int i = 1; i <= z.nx - 2; ++i) {
        size_t d = z.didx(i, j, k);
        T ru1 = z.rho_i[d], uu = z.us[d], vv = z.vs[d], ww = z.ws[d], ac = z.speed[d];
        T ac2inv = 1.0 / (ac * ac);
        T r1 = z.R(0, i, j, k), r2 = z.R(1, i, j, k), r3 = z.R(2, i, j, k),
          r4 = z.R(3, i, j, k), r5 = z.R(4, i, j, k);
        T t1 = K.c2 * ac2inv * (z.qs[d] * r1 - uu * r2 - vv * r3 - ww * r4 + r5);
        T t2 = K.bt * ru1 * (uu * r1 - r2);
        T t3 = (K.bt * ru1 * ac) * t1;
        z.R(0, i, j, k) = r1 - t1;
        z.R(1, i, j, k) = -ru1 * (ww * r1 - r4);
        z.R(2, i, j, k) = ru1 * (vv * r1 - r3);
        z.R(3, i, j, k) = -t2 + t3;
        z.R(4, i, j, k) = t2 + t3;
      }
}

// ninvr (src/SP/ninvr.jl:5-38), pinvr (src/SP/pinvr.jl:5-39).
template <class T>
void ninvr(Zone<T>& z, const Consts& K) {
  for (int k = 1; k <= z.nz - 2; ++k)
    for (int j = 1; j <= z.ny - 2; ++j)
      for (int i = 1; i <= z.nx - 2; ++i) {
        T r1 = z.R(0, i, j, k), r2 = z.R(1, i, j, k), r3 = z.R(2, i, j, k),
          r4 = z.R(3, i, j, k), r5 = z.R(4, i, j, k);
        T t1 = K.bt * r3;
        T t2 = 0.5 * (r4 + r5);
        z.R(0, i, j, k) = -r2;
        z.R(1, i, j, k) = r1;
        z.R(2, i, j, k) = K.bt * (r4 - r5);
        z.R(3, i, j, k) = -t1 + t2;
        z.R(4, i, j, k) = t1 + t2;
      }
}

template <class T>
void pinvr(Zone<T>& z, const Consts& K) {
  for (int k = 1; k <= z.nz - 2; ++k)
    for (int j = 1; j <= z.ny - 2; ++j)
      for (int i = 1; i <= z.nx - 2; ++i) {
        T r1 = z.R(0, i, j, k), r2 = z.R(1, i, j, k), r3 = z.R(2, i, j, k),
          r4 = z.R(3, i, j, k), r5 = z.R(4, i, j, k);
        T t1 = K.bt * r1;
        T t2 = 0.5 * (r4 + r5);
        z.R(0, i, j, k) = K.bt * (r4 - r5);
        z.R(1, i, j, k) = -r3;
        z.R(2, i, j, k) = r2;
        z.R(3, i, j, k) = -t1 + t2;
        z.R(4, i, j, k) = t1 + t2;
      }
}

// tzetar (src/SP/tzetar.jl:5-59).
template <class T>
void tzetar(Zone<T>& z, const Consts& K) {
  for (int k = 1; k <= z.nz - 2; ++k)
    for (int j = 1; j <= z.ny - 2; ++j)
      for (int i = 1; i <= z.nx - 2; ++i) {
        size_t d = z.didx(i, j, k);
        T xvel = z.us[d], yvel = z.vs[d], zvel = z.ws[d], ac = z.speed[d], acinv = z.ainv[d];
        T ac2u = ac * ac;
        T r1 = z.R(0, i, j, k), r2 = z.R(1, i, j, k), r3 = z.R(2, i, j, k),
          r4 = z.R(3, i, j, k), r5 = z.R(4, i, j, k);
        T uzik1 = z.U(0, i, j, k);
        T btuz = K.bt * uzik1;
        T t1 = btuz * acinv * (r4 + r5);
        T t2 = r3 + t1;
        T t3 = btuz * (r4 - r5);
        z.R(0, i, j, k) = t2;
        z.R(1, i, j, k) = -uzik1 * r2 + xvel * t2;
        z.R(2, i, j, k) = uzik1 * r1 + yvel * t2;
        z.R(3, i, j, k) = zvel * t2 + t3;
        z.R(4, i, j, k) = uzik1 * (-xvel * r2 + yvel * r1) + z.qs[d] * t2 + K.c2iv * ac2u * t1 +
                          zvel * t3;
      }
}

inline void transverse(int dir, int& da, int& db) {
  if (dir == 0) { da = 2; db = 1; }
  else if (dir == 1) { da = 2; db = 0; }
  else { da = 1; db = 0; }
}

// Point l of line (a, b) along dir.
inline void line_point(int dir, int a, int b, int l, int& i, int& j, int& k) {
  int da, db, idx[3];
  transverse(dir, da, db);
  idx[da] = a; idx[db] = b; idx[dir] = l;
  i = idx[0]; j = idx[1]; k = idx[2];
}

// lhsx/lhsy/lhsz (src/SP/lhsx.jl:5-114, lhsy.jl:5-118, lhsz.jl:5-116) for
// one line: the 15 pentadiagonal coefficients of rows 1..nl-2.
template <class T>
void assemble_line(Zone<T>& z, Lhs<T>& L, const Consts& K, int dir, int a, int b) {
  const int n[3] = {z.nx, z.ny, z.nz};
  const int nl = n[dir];
  const double dtt1 = dir == 0 ? K.dttx1 : dir == 1 ? K.dtty1 : K.dttz1;
  const double dtt2 = dir == 0 ? K.dttx2 : dir == 1 ? K.dtty2 : K.dttz2;
  const double c2dtt1 = dir == 0 ? K.c2dttx1 : dir == 1 ? K.c2dtty1 : K.c2dttz1;
  // rhon = max(d_v + con43*ru1, d5 + c1c5*ru1, dmax + ru1, d1) with d_v the
  // width of the line's own velocity component (dx2, dy3, dz4)
  const double dv = dir == 0 ? K.dx2 : dir == 1 ? K.dy3 : K.dz4;
  const double d5 = dir == 0 ? K.dx5 : dir == 1 ? K.dy5 : K.dz5;
  const double d1 = dir == 0 ? K.dx1 : dir == 1 ? K.dy1 : K.dz1;
  const double dmax = dir == 0 ? K.dxmax : dir == 1 ? K.dymax : K.dzmax;
  std::vector<T> cv(nl), rhon(nl), spd(nl);
  auto LL = [&](int c, int l) -> T& {
    int i, j, k;
    line_point(dir, a, b, l, i, j, k);
    return L(c, i, j, k);
  };
  for (int l = 0; l <= nl - 1; ++l) {
    int i, j, k;
    line_point(dir, a, b, l, i, j, k);
    size_t d = z.didx(i, j, k);
    T ru1 = K.c3c4 * z.rho_i[d];
    cv[l] = dir == 0 ? z.us[d] : dir == 1 ? z.vs[d] : z.ws[d];
    using std::fmax;
    rhon[l] = fmax(fmax(fmax(dv + K.con43 * ru1, d5 + K.c1c5 * ru1), dmax + ru1), d1);
    spd[l] = z.speed[d];
  }
  for (int l = 1; l <= nl - 2; ++l) {
    LL(0, l) = 0.0;
    LL(1, l) = -dtt2 * cv[l - 1] - dtt1 * rhon[l - 1];
    LL(2, l) = 1.0 + c2dtt1 * rhon[l];
    LL(3, l) = dtt2 * cv[l + 1] - dtt1 * rhon[l + 1];
    LL(4, l) = 0.0;
  }
  // fourth-order dissipation rows, in the reference's loop order
  {
    int l = 1;
    LL(2, l) += K.comz5; LL(3, l) -= K.comz4; LL(4, l) += K.comz1;
    LL(1, l + 1) -= K.comz4; LL(2, l + 1) += K.comz6; LL(3, l + 1) -= K.comz4;
    LL(4, l + 1) += K.comz1;
  }
  for (int l = 3; l <= nl - 4; ++l) {
    LL(0, l) += K.comz1; LL(1, l) -= K.comz4; LL(2, l) += K.comz6; LL(3, l) -= K.comz4;
    LL(4, l) += K.comz1;
  }
  {
    int l = nl - 3;
    LL(0, l) += K.comz1; LL(1, l) -= K.comz4; LL(2, l) += K.comz6; LL(3, l) -= K.comz4;
    LL(0, l + 1) += K.comz1; LL(1, l + 1) -= K.comz4; LL(2, l + 1) += K.comz5;
  }
  for (int l = 1; l <= nl - 2; ++l) {
    LL(5, l) = LL(0, l);
    LL(6, l) = LL(1, l) - dtt2 * spd[l - 1];
    LL(7, l) = LL(2, l);
    LL(8, l) = LL(3, l) + dtt2 * spd[l + 1];
    LL(9, l) = LL(4, l);
    LL(10, l) = LL(0, l);
    LL(11, l) = LL(1, l) + dtt2 * spd[l - 1];
    LL(12, l) = LL(2, l);
    LL(13, l) = LL(3, l) - dtt2 * spd[l + 1];
    LL(14, l) = LL(4, l);
  }
}

// The pentadiagonal sweeps of x/y/z_solve for one cell (src/SP/x_solve.jl:
// 203-313 forward, :450-512 back), then ninvr / pinvr / tzetar.
template <class T>
void solve_dir(Zone<T>& z, Lhs<T>& L, const Consts& K, int dir) {
  const int n[3] = {z.nx, z.ny, z.nz};
  const int nl = n[dir];
  int da, db;
  transverse(dir, da, db);
  for (int a = 1; a <= n[da] - 2; ++a)
    for (int b = 1; b <= n[db] - 2; ++b) {
      assemble_line(z, L, K, dir, a, b);
      auto LL = [&](int c, int l) -> T& {
        int i, j, k;
        line_point(dir, a, b, l, i, j, k);
        return L(c, i, j, k);
      };
      auto RR = [&](int m, int l) -> T& {
        int i, j, k;
        line_point(dir, a, b, l, i, j, k);
        return z.R(m, i, j, k);
      };
      // three systems: components 1..3 share lhs(1:5), 4 uses lhs(6:10),
      // 5 uses lhs(11:15)
      for (int sys = 0; sys < 3; ++sys) {
        const int o = 5 * sys;
        const int m0 = sys == 0 ? 0 : sys + 2, m1 = sys == 0 ? 2 : sys + 2;
        for (int l = 0; l <= nl - 3; ++l) {
          const int l1 = l + 1, l2 = l + 2;
          T fac1 = 1.0 / LL(o + 2, l);
          LL(o + 3, l) *= fac1;
          LL(o + 4, l) *= fac1;
          for (int m = m0; m <= m1; ++m) RR(m, l) *= fac1;
          LL(o + 2, l1) -= LL(o + 1, l1) * LL(o + 3, l);
          LL(o + 3, l1) -= LL(o + 1, l1) * LL(o + 4, l);
          for (int m = m0; m <= m1; ++m) RR(m, l1) -= LL(o + 1, l1) * RR(m, l);
          LL(o + 1, l2) -= LL(o + 0, l2) * LL(o + 3, l);
          LL(o + 2, l2) -= LL(o + 0, l2) * LL(o + 4, l);
          for (int m = m0; m <= m1; ++m) RR(m, l2) -= LL(o + 0, l2) * RR(m, l);
        }
        {
          const int l = nl - 2, l1 = nl - 1;
          T fac1 = 1.0 / LL(o + 2, l);
          LL(o + 3, l) *= fac1;
          LL(o + 4, l) *= fac1;
          for (int m = m0; m <= m1; ++m) RR(m, l) *= fac1;
          LL(o + 2, l1) -= LL(o + 1, l1) * LL(o + 3, l);
          LL(o + 3, l1) -= LL(o + 1, l1) * LL(o + 4, l);
          for (int m = m0; m <= m1; ++m) RR(m, l1) -= LL(o + 1, l1) * RR(m, l);
          T fac2 = 1.0 / LL(o + 2, l1);
          LL(o + 3, l1) *= fac2;
          LL(o + 4, l1) *= fac2;
          for (int m = m0; m <= m1; ++m) RR(m, l1) *= fac2;
        }
      }
      for (int sys = 0; sys < 3; ++sys) {
        const int o = 5 * sys;
        const int m0 = sys == 0 ? 0 : sys + 2, m1 = sys == 0 ? 2 : sys + 2;
        for (int m = m0; m <= m1; ++m) RR(m, nl - 2) -= LL(o + 3, nl - 2) * RR(m, nl - 1);
        for (int m = m0; m <= m1; ++m)
          for (int l = nl - 3; l >= 0; --l)
            RR(m, l) -= LL(o + 3, l) * RR(m, l + 1) + LL(o + 4, l) * RR(m, l + 2);
      }
    }
  if (dir == 0) ninvr(z, K);
  else if (dir == 1) pinvr(z, K);
  else tzetar(z, K);
}

struct Run {
  Zones G;
  Consts K;
  std::vector<Zone<double>> Z;
  std::vector<Lhs<double>> L;
  int nthreads = 1;
};

// SP stencil widths: src/SP/set_constants.jl:16-30 (the same values as BT).
static const double SP_DXYZ[15] = {0.75, 0.75, 0.75, 0.75, 0.75, 0.75, 0.75, 0.75,
                                   0.75, 0.75, 1.0,  1.0,  1.0,  1.0,  1.0};

}  // namespace sp
}  // namespace npb

using namespace npb;
using namespace npb::sp;

extern "C" {

void* oracle_sp_create(int x_zones, int y_zones, int gx, int gy, int gz, double ratio, double dt) {
  Run* r = new Run();
  r->G = zone_setup(x_zones, y_zones, gx, gy, gz, ratio, false);
  r->K = set_constants(dt, gx, gy, gz, x_zones, y_zones, SP_DXYZ);
  int n = r->G.count();
  r->Z.resize(n);
  r->L.resize(n);
  for (int z = 0; z < n; ++z) {
    Zone<double>& zz = r->Z[z];
    zz.nx = r->G.nx[z]; zz.ny = r->G.ny[z]; zz.nz = r->G.nz[z];
    zz.alloc();
    bt::initialize(zz, r->K, true);
    r->L[z].alloc(zz.nx, zz.ny, zz.nz);
    lhsinit(r->L[z]);
    bt::exact_rhs(zz, r->K);
  }
  return r;
}

void oracle_sp_destroy(void* h) { delete (Run*)h; }
void oracle_sp_set_threads(void* h, int n) { ((Run*)h)->nthreads = n < 1 ? 1 : n; }
int oracle_sp_nzones(void* h) { return ((Run*)h)->G.count(); }

// Phases: 0 exch_qbc, 1 compute_rhs + txinvr, 2 x_solve (+ninvr),
// 3 y_solve (+pinvr), 4 z_solve (+tzetar), 5 add.
void oracle_sp_phase(void* h, int phase) {
  Run* r = (Run*)h;
  if (phase == 0) { bt::exch_qbc(r->Z, r->G, true); return; }
  int nzn = r->G.count();
#pragma omp parallel for schedule(dynamic, 1) num_threads(r->nthreads)
  for (int z = 0; z < nzn; ++z) {
    Zone<double>& zz = r->Z[z];
    switch (phase) {
      case 1: bt::compute_rhs(zz, r->K, true); txinvr(zz, r->K); break;
      case 2: solve_dir(zz, r->L[z], r->K, 0); break;
      case 3: solve_dir(zz, r->L[z], r->K, 1); break;
      case 4: solve_dir(zz, r->L[z], r->K, 2); break;
      case 5: bt::add(zz); break;
    }
  }
}

// nsteps iterations of the timestep loop body (src/SP/sp-node.jl:252-391):
// exch_qbc, then per zone adi (src/SP/adi.jl:6-258).
void oracle_sp_step(void* h, int nsteps) {
  Run* r = (Run*)h;
  int nzn = r->G.count();
  for (int s = 0; s < nsteps; ++s) {
    if (nzn > 1) bt::exch_qbc(r->Z, r->G, true);
#pragma omp parallel for schedule(dynamic, 1) num_threads(r->nthreads)
    for (int z = 0; z < nzn; ++z) {
      Zone<double>& zz = r->Z[z];
      bt::compute_rhs(zz, r->K, true);
      txinvr(zz, r->K);
      solve_dir(zz, r->L[z], r->K, 0);
      solve_dir(zz, r->L[z], r->K, 1);
      solve_dir(zz, r->L[z], r->K, 2);
      bt::add(zz);
    }
  }
}

// Node-side verify (src/SP/verify.jl:411-515), the same sequence as BT's.
void oracle_sp_norms(void* h, double* xcr, double* xce) {
  Run* r = (Run*)h;
  int nzn = r->G.count();
  std::vector<double> ce((size_t)nzn * 5), cr((size_t)nzn * 5);
#pragma omp parallel for schedule(dynamic, 1) num_threads(r->nthreads)
  for (int z = 0; z < nzn; ++z) {
    bt::error_norm(r->Z[z], r->K, &ce[(size_t)z * 5]);
    bt::compute_rhs(r->Z[z], r->K, true);
    bt::rhs_norm(r->Z[z], &cr[(size_t)z * 5]);
  }
  for (int m = 0; m < 5; ++m) { xcr[m] = 0.0; xce[m] = 0.0; }
  for (int z = 0; z < nzn; ++z)
    for (int m = 0; m < 5; ++m) {
      xcr[m] = xcr[m] + cr[(size_t)z * 5 + m] / r->K.dt;
      xce[m] = xce[m] + ce[(size_t)z * 5 + m];
    }
}

void oracle_sp_get(void* h, int z, int field, double* out) {
  Run* r = (Run*)h;
  Zone<double>& a = r->Z[z];
  size_t p = 0;
  for (int k = 0; k < a.nz; ++k)
    for (int j = 0; j < a.ny; ++j)
      for (int i = 0; i < a.nx; ++i)
        for (int m = 0; m < 5; ++m, ++p)
          out[p] = field == 0 ? a.U(m, i, j, k) : field == 1 ? a.F(m, i, j, k) : a.R(m, i, j, k);
}

void oracle_sp_set(void* h, int z, int field, const double* in) {
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

// The 15 pentadiagonal coefficients of line (a, b) of zone z in direction dir
// for rows 0..nl-1, as assembled from the derived fields of the last
// compute_rhs (rows 0 and nl-1 keep lhsinit's values).  Test hook; the
// zone's own lhs is not modified.  out: nl x 15 row-major.
// Forward-mode dual numbers through nsteps SP steps from the current state,
// as oracle_bt_dual: out[0] = loss = sum(w*u_N^2), out[1+d] its derivative
// along (vu[d] on u0, vf[d] on the forcing at interior points).  Arrays in
// the whole-run layout; the state is left unchanged.
void oracle_sp_dual(void* h, int nsteps, int ndir, const double* w, const double* vu,
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
    std::vector<Lhs<D>> L(nzn);
    for (int z = 0; z < nzn; ++z) {
      Zone<double>& a = r->Z[z];
      Zone<D>& b = Z[z];
      b.nx = a.nx; b.ny = a.ny; b.nz = a.nz;
      b.alloc();
      L[z].alloc(b.nx, b.ny, b.nz);
      lhsinit(L[z]);
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
      if (nzn > 1) bt::exch_qbc(Z, r->G, true);
#pragma omp parallel for schedule(dynamic, 1) num_threads(r->nthreads)
      for (int z = 0; z < nzn; ++z) {
        bt::compute_rhs(Z[z], r->K, true);
        txinvr(Z[z], r->K);
        solve_dir(Z[z], L[z], r->K, 0);
        solve_dir(Z[z], L[z], r->K, 1);
        solve_dir(Z[z], L[z], r->K, 2);
        bt::add(Z[z]);
      }
    }
    D loss(0.0);
    for (int z = 0; z < nzn; ++z) {
      Zone<D>& b = Z[z];
      size_t p = zoff[z];
      for (int k = 0; k < b.nz; ++k)
        for (int j = 0; j < b.ny; ++j)
          for (int i = 0; i < b.nx; ++i)
            for (int m = 0; m < 5; ++m, ++p) {
              const D& x = b.U(m, i, j, k);
              loss = loss + w[p] * (x * x);
            }
    }
    out[0] = loss.v;
    for (int e = 0; e < ND && d0 + e < ndir; ++e) out[1 + d0 + e] = loss.d[e];
  }
}

void oracle_sp_line_lhs(void* h, int z, int dir, int a, int b, double* out) {
  Run* r = (Run*)h;
  Lhs<double> L = r->L[z];
  assemble_line(r->Z[z], L, r->K, dir, a, b);
  const int n[3] = {r->Z[z].nx, r->Z[z].ny, r->Z[z].nz};
  for (int l = 0; l < n[dir]; ++l) {
    int i, j, k;
    line_point(dir, a, b, l, i, j, k);
    for (int c = 0; c < 15; ++c) out[l * 15 + c] = L(c, i, j, k);
  }
}

}  // extern "C"
