// libnpbmz: MI355X (gfx950) multi-zone BT/SP/LU timestep behind a C ABI.
//
// All zones of a run live in HBM at once.  Fields are stored per point with
// the 5 components contiguous (the reference's [m,i,j,k] order), zone after
// zone.  Every phase of the timestep runs either as one launch per zone or as
// one batched launch over a task table that covers all zones; the batched
// timestep can be captured in a hipGraph.  Kernels exist twice: npbmz::exact
// (no FMA contraction, bit-identical to the reference's CPU arithmetic) and
// npbmz::fast (FMA contraction), chosen per context.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/npbmz.h"
#include "npbmz_types.h"

namespace npbmz {

// ce table of the exact solution, src/BT/set_constants.jl:47-51 (the same
// matrix in SP and LU).  kCE[m][n] is the reference's ce[m+1, n+1].
__constant__ double kCE[5][13] = {
    {2.0, 0.0, 0.0, 4.0, 5.0, 3.0, 0.5, 0.02, 0.01, 0.03, 0.5, 0.4, 0.3},
    {1.0, 0.0, 0.0, 0.0, 1.0, 2.0, 3.0, 0.01, 0.03, 0.02, 0.4, 0.3, 0.5},
    {2.0, 2.0, 0.0, 0.0, 0.0, 2.0, 3.0, 0.04, 0.03, 0.05, 0.3, 0.5, 0.4},
    {2.0, 2.0, 0.0, 0.0, 0.0, 2.0, 3.0, 0.03, 0.05, 0.04, 0.2, 0.1, 0.3},
    {5.0, 4.0, 3.0, 2.0, 0.1, 0.4, 0.3, 0.05, 0.04, 0.03, 0.1, 0.3, 0.2}};

// A zone face copied by exch_qbc: the 5 components of every point of the
// destination zone's boundary plane `type` (0 west i=0, 1 east i=nx-1,
// 2 south j=0, 3 north j=ny-1) come from the source (neighbour) zone's
// plane next to the shared boundary.
struct Face {
  int32_t dst, src, type, npts;
};

// exch_qbc for BT and SP (src/BT/exchange_qbc.jl:145-466, one rank): the
// reference packs every face before unpacking any, and sources are interior
// planes while destinations are boundary planes, so one kernel over all
// faces reproduces it.  Copies are bit exact.
template <int NT>
__global__ __launch_bounds__(NT) void exch_kernel(double* __restrict__ u,
                                                  const ZoneDev* __restrict__ zones,
                                                  const Face* __restrict__ faces,
                                                  const Task* __restrict__ tasks) {
  const Task T = tasks[blockIdx.x];
  const Face F = faces[T.zone];
  const int idx = T.first + (int)threadIdx.x;
  if (idx >= F.npts) return;
  const ZoneDev D = zones[F.dst], S = zones[F.src];
  int64_t dp, sp;
  if (F.type < 2) {
    const int j = 1 + idx % (D.ny - 2), k = 1 + idx / (D.ny - 2);
    const int di = F.type == 0 ? 0 : D.nx - 1;
    const int si = F.type == 0 ? S.nx - 2 : 1;
    dp = D.base + ((int64_t)k * D.ny + j) * D.nx + di;
    sp = S.base + ((int64_t)k * S.ny + j) * S.nx + si;
  } else {
    const int i = 1 + idx % (D.nx - 2), k = 1 + idx / (D.nx - 2);
    const int dj = F.type == 2 ? 0 : D.ny - 1;
    const int sj = F.type == 2 ? S.ny - 2 : 1;
    dp = D.base + ((int64_t)k * D.ny + dj) * D.nx + i;
    sp = S.base + ((int64_t)k * S.ny + sj) * S.nx + i;
  }
  const double* __restrict__ s = u + sp * 5;
  double* __restrict__ d = u + dp * 5;
#pragma unroll
  for (int m = 0; m < 5; ++m) d[m] = s[m];
}

// Streaming copy used to calibrate byte counters (npbmz_probe_copy).
__global__ __launch_bounds__(256) void copy_probe_kernel(const double4* __restrict__ src,
                                                         double4* __restrict__ dst, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

namespace exact {
#pragma clang fp contract(off)
#include "bt_kernels.inc"
#include "sp_kernels.inc"
}  // namespace exact

namespace fast {
#pragma clang fp contract(fast)
#include "bt_kernels.inc"
#include "sp_kernels.inc"
}  // namespace fast

}  // namespace npbmz

using namespace npbmz;

// ---------------------------------------------------------------- host side
namespace {

constexpr int kPtNT = 128;   // threads per point task (rhs, add, init, norms)
constexpr int kExchNT = 128; // threads per face task
constexpr int kLineNT = 64;  // lanes (lines) per solve task: one wave

// A task table on the device plus, per zone, the range of its tasks.
struct Table {
  Task* d = nullptr;
  int n = 0;
  std::vector<int> zfirst, zcount;  // by zone number
};

}  // namespace

struct npbmz_ctx {
  int bench = 0, flags = 0, nz_ = 0;
  int x_zones = 0, y_zones = 0, gx = 0, gy = 0, gz = 0;
  double dt = 0.0, omega = 0.0;
  std::vector<int> nx, ny, nz, west, east, south, north;
  std::vector<int> order;  // zone order of the batched tables
  StencilK K{};
  std::vector<ZoneDev> zh;
  ZoneDev* zd = nullptr;
  int64_t npoints = 0;
  double *u = nullptr, *frc = nullptr, *rhs = nullptr;
  double* cscr = nullptr;
  size_t cscr_slot = 0;  // doubles per scratch slot (max over directions)
  int maxrows[3] = {0, 0, 0};
  int solve_grid[3] = {0, 0, 0};  // 0: one block per task, else persistent grid
  Table t_int, t_all, t_line[3], t_face;
  Face* faces = nullptr;
  double* partial = nullptr;
  std::vector<double> zn_rhs, zn_err;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipGraphExec_t gexec = nullptr;
  hipGraph_t graph = nullptr;
  std::string err;
  std::string sched;
};

namespace {

std::string g_create_err;

#define CK(call)                                                                        \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess) {                                                             \
      ctx->err = std::string(#call) + ": " + hipGetErrorString(e_);                     \
      return NPBMZ_EHIP;                                                                \
    }                                                                                   \
  } while (0)

int fail(npbmz_ctx* ctx, int code, const std::string& msg) {
  if (ctx) ctx->err = msg;
  else g_create_err = msg;
  return code;
}

// set_constants for BT (src/BT/set_constants.jl:1-130) and SP
// (src/SP/set_constants.jl), which differ only in the dx/dy/dz widths.
StencilK make_constants(int bench, double dt, int gx, int gy, int gz, int x_zones, int y_zones) {
  StencilK k{};
  k.dt = dt;
  k.c1 = 1.4; k.c2 = 0.4; k.c3 = 0.1; k.c4 = 1.0; k.c5 = 1.4;
  k.bt = std::sqrt(0.5);
  k.c1c2 = k.c1 * k.c2; k.c1c5 = k.c1 * k.c5; k.c3c4 = k.c3 * k.c4;
  k.c1345 = k.c1c5 * k.c3c4;
  const double conz1 = (1.0 - k.c1c5);
  double d[3][5];
  if (bench == NPBMZ_SP) {
    // src/SP/set_constants.jl: dx = 0.75/0.75/0.75/0.75/0.75, dy = .75,
    // dz = 1.0 (same values as BT)
    for (int m = 0; m < 5; ++m) { d[0][m] = 0.75; d[1][m] = 0.75; d[2][m] = 1.0; }
  } else {
    for (int m = 0; m < 5; ++m) { d[0][m] = 0.75; d[1][m] = 0.75; d[2][m] = 1.0; }
  }
  for (int a = 0; a < 3; ++a)
    for (int m = 0; m < 5; ++m) k.dd[a][m] = d[a][m];
  k.dmax[0] = std::fmax(d[0][2], d[0][3]);
  k.dmax[1] = std::fmax(d[1][1], d[1][3]);
  k.dmax[2] = std::fmax(d[2][1], d[2][2]);
  k.dssp = 0.25 * std::fmax(d[0][0], std::fmax(d[1][0], d[2][0]));
  k.c4dssp = 4.0 * k.dssp; k.c5dssp = 5.0 * k.dssp;
  k.c2iv = 2.5; k.con43 = 4.0 / 3.0; k.con16 = 1.0 / 6.0;
  k.dnm1[0] = 1.0 / (double(gx - 1) / double(x_zones));
  k.dnm1[1] = 1.0 / (double(gy - 1) / double(y_zones));
  k.dnm1[2] = 1.0 / double(gz - 1);
  for (int a = 0; a < 3; ++a) {
    k.t1[a] = 1.0 / (k.dnm1[a] * k.dnm1[a]);
    k.t2[a] = 1.0 / (2.0 * k.dnm1[a]);
    k.t3[a] = 1.0 / k.dnm1[a];
    k.dtt1[a] = dt * k.t1[a];
    k.dtt2[a] = dt * k.t2[a];
    k.c2dtt1[a] = 2.0 * k.dtt1[a];
    for (int m = 0; m < 5; ++m) k.ddt1[a][m] = k.dd[a][m] * k.t1[a];
    const double c3c4t3 = k.c3c4 * k.t3[a];
    k.con1[a] = c3c4t3 * k.con43 * k.t3[a];
    k.con2[a] = c3c4t3 * k.t3[a];
    k.con3[a] = c3c4t3 * conz1 * k.t3[a];
    k.con4[a] = c3c4t3 * k.con16 * k.t3[a];
    k.con5[a] = c3c4t3 * k.c1c5 * k.t3[a];
  }
  k.dtdssp = dt * k.dssp;
  k.comz1 = k.dtdssp; k.comz4 = 4.0 * k.dtdssp; k.comz5 = 5.0 * k.dtdssp;
  k.comz6 = 6.0 * k.dtdssp;
  return k;
}

int lines_of(const npbmz_ctx* c, int z, int dir) {
  return dir == 0 ? (c->ny[z] - 2) * (c->nz[z] - 2)
                  : dir == 1 ? (c->nx[z] - 2) * (c->nz[z] - 2) : (c->nx[z] - 2) * (c->ny[z] - 2);
}

// Build a task table: zones in `order`, each zone's `count(z)` items cut into
// tasks of `per` items.
template <class F>
int build_table(npbmz_ctx* ctx, Table& t, const std::vector<int>& order, int per, F count) {
  std::vector<Task> h;
  int nzn = (int)order.size();
  t.zfirst.assign(nzn, 0);
  t.zcount.assign(nzn, 0);
  for (int z : order) {
    int n = count(z);
    int nt = (n + per - 1) / per;
    t.zfirst[z] = (int)h.size();
    t.zcount[z] = nt;
    for (int q = 0; q < nt; ++q) h.push_back(Task{z, q * per});
  }
  if (t.d) { (void)hipFree(t.d); t.d = nullptr; }
  t.n = (int)h.size();
  if (t.n == 0) return NPBMZ_OK;
  CK(hipMalloc(&t.d, sizeof(Task) * h.size()));
  CK(hipMemcpy(t.d, h.data(), sizeof(Task) * h.size(), hipMemcpyHostToDevice));
  return NPBMZ_OK;
}

int build_tables(npbmz_ctx* ctx) {
  int rc;
  auto& c = *ctx;
  if ((rc = build_table(ctx, c.t_int, c.order, kPtNT, [&](int z) {
         return (c.nx[z] - 2) * (c.ny[z] - 2) * (c.nz[z] - 2);
       })))
    return rc;
  if ((rc = build_table(ctx, c.t_all, c.order, kPtNT,
                        [&](int z) { return c.nx[z] * c.ny[z] * c.nz[z]; })))
    return rc;
  size_t slot = 0;
  for (int d = 0; d < 3; ++d) {
    if ((rc = build_table(ctx, c.t_line[d], c.order, kLineNT,
                          [&](int z) { return lines_of(ctx, z, d); })))
      return rc;
    int mr = 0;
    for (int z = 0; z < c.nz_; ++z) {
      int nl = d == 0 ? c.nx[z] : d == 1 ? c.ny[z] : c.nz[z];
      mr = std::max(mr, nl - 2);
    }
    c.maxrows[d] = mr;
    int slots = c.t_line[d].n;
    for (int z = 0; z < c.nz_; ++z) slots = std::max(slots, c.t_line[d].zcount[z]);
    if (c.solve_grid[d] > 0) slots = std::max(slots, c.solve_grid[d]);
    // reduced coefficients kept per row for the back sweep: BT's 5x5 C'
    // block, SP's lhs(4:5) of three systems
    const int per_row = c.bench == NPBMZ_SP ? 6 : 25;
    slot = std::max(slot, (size_t)slots * mr * per_row * 64);
  }
  if (slot > c.cscr_slot) {
    if (c.cscr) (void)hipFree(c.cscr);
    c.cscr = nullptr;
    CK(hipMalloc(&c.cscr, slot * sizeof(double)));
    c.cscr_slot = slot;
  }
  // faces: west/east take (ny-2)(nz-2) points, south/north (nx-2)(nz-2)
  std::vector<Face> fh;
  for (int z = 0; z < c.nz_; ++z) {
    int nxf = (c.ny[z] - 2) * (c.nz[z] - 2), nyf = (c.nx[z] - 2) * (c.nz[z] - 2);
    // SP's exchange fills j = 0 from the north neighbour and j = ny-1 from
    // the south one (src/SP/exchange_qbc.jl:247-295, 409-455), the mirror of
    // BT's (src/BT/exchange_qbc.jl:247-295, 409-455)
    const bool yswap = c.bench == NPBMZ_SP;
    fh.push_back(Face{z, c.west[z], 0, nxf});
    fh.push_back(Face{z, c.east[z], 1, nxf});
    fh.push_back(Face{z, yswap ? c.north[z] : c.south[z], 2, nyf});
    fh.push_back(Face{z, yswap ? c.south[z] : c.north[z], 3, nyf});
  }
  if (!c.faces) CK(hipMalloc(&c.faces, sizeof(Face) * fh.size()));
  CK(hipMemcpy(c.faces, fh.data(), sizeof(Face) * fh.size(), hipMemcpyHostToDevice));
  std::vector<int> forder(fh.size());
  std::iota(forder.begin(), forder.end(), 0);
  if ((rc = build_table(ctx, c.t_face, forder, kExchNT, [&](int f) { return fh[f].npts; })))
    return rc;
  int pmax = std::max(c.t_int.n, c.t_all.n);
  if (c.partial) (void)hipFree(c.partial);
  c.partial = nullptr;
  CK(hipMalloc(&c.partial, sizeof(double) * 5 * (size_t)std::max(pmax, 1)));
  if (c.gexec) { (void)hipGraphExecDestroy(c.gexec); c.gexec = nullptr; }
  if (c.graph) { (void)hipGraphDestroy(c.graph); c.graph = nullptr; }
  return NPBMZ_OK;
}

bool strict(const npbmz_ctx* c) { return (c->flags & NPBMZ_FLAG_STRICT_FP) != 0; }

// ------------------------------------------------------------ launches
// Every launcher covers either a whole task table (zone < 0, the batched
// form) or one zone's range of it (the per-zone form).

// compute_rhs.  For SP the timestep's rhs phase is compute_rhs + txinvr;
// plain = true gives compute_rhs alone (SP's verify, src/SP/verify.jl:436).
int k_rhs(npbmz_ctx* ctx, int zone, bool plain = false) {
  auto& c = *ctx;
  const Table& t = c.t_int;
  int first = zone < 0 ? 0 : t.zfirst[zone], n = zone < 0 ? t.n : t.zcount[zone];
  if (n <= 0) return NPBMZ_OK;
  const bool tx = c.bench == NPBMZ_SP && !plain;
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3(n), dim3(kPtNT), 0, c.stream, c.u, c.frc, c.rhs, c.zd,
                       t.d + first, c.K);
  };
  if (strict(ctx)) {
    if (tx) go(exact::bt::rhs_kernel<kPtNT, true>);
    else go(exact::bt::rhs_kernel<kPtNT, false>);
  } else {
    if (tx) go(fast::bt::rhs_kernel<kPtNT, true>);
    else go(fast::bt::rhs_kernel<kPtNT, false>);
  }
  CK(hipGetLastError());
  return NPBMZ_OK;
}

int k_add(npbmz_ctx* ctx, int zone) {
  auto& c = *ctx;
  const Table& t = c.t_int;
  int first = zone < 0 ? 0 : t.zfirst[zone], n = zone < 0 ? t.n : t.zcount[zone];
  if (n <= 0) return NPBMZ_OK;
  hipLaunchKernelGGL(exact::bt::add_kernel<kPtNT>, dim3(n), dim3(kPtNT), 0, c.stream, c.u,
                     c.rhs, c.zd, t.d + first);
  CK(hipGetLastError());
  return NPBMZ_OK;
}

template <int DIR, bool FUSE>
int k_solve_t(npbmz_ctx* ctx, int zone) {
  auto& c = *ctx;
  const Table& t = c.t_line[DIR];
  int first = zone < 0 ? 0 : t.zfirst[zone], n = zone < 0 ? t.n : t.zcount[zone];
  if (n <= 0) return NPBMZ_OK;
  int grid = n;
  if (zone < 0 && c.solve_grid[DIR] > 0) grid = std::min(n, c.solve_grid[DIR]);
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3(grid), dim3(kLineNT), 0, c.stream, c.u, c.rhs, c.cscr, c.zd,
                       t.d + first, n, c.maxrows[DIR], c.K);
  };
  if (c.bench == NPBMZ_SP) {
    if (strict(ctx)) go(exact::sp::solve_kernel<DIR, FUSE>);
    else go(fast::sp::solve_kernel<DIR, FUSE>);
  } else {
    if (strict(ctx)) go(exact::bt::solve_kernel<DIR, FUSE>);
    else go(fast::bt::solve_kernel<DIR, FUSE>);
  }
  CK(hipGetLastError());
  return NPBMZ_OK;
}

// x/y/z_solve (BT block tridiagonal, SP pentadiagonal with its inverse
// transforms); fuse folds add into the z sweep.
int k_solve(npbmz_ctx* ctx, int dir, bool fuse, int zone) {
  if (dir == 0) return k_solve_t<0, false>(ctx, zone);
  if (dir == 1) return k_solve_t<1, false>(ctx, zone);
  return fuse ? k_solve_t<2, true>(ctx, zone) : k_solve_t<2, false>(ctx, zone);
}

int exch(npbmz_ctx* ctx) {
  auto& c = *ctx;
  if (c.nz_ <= 1) return NPBMZ_OK;  // bt-node.jl:286: exchange only with >1 zone
  if (c.t_face.n <= 0) return NPBMZ_OK;
  hipLaunchKernelGGL(exch_kernel<kExchNT>, dim3(c.t_face.n), dim3(kExchNT), 0, c.stream, c.u,
                     c.zd, c.faces, c.t_face.d);
  CK(hipGetLastError());
  return NPBMZ_OK;
}

// One BT timestep: exch_qbc, then per zone compute_rhs, x/y/z solve, add.
// Per-zone mode launches each phase zone by zone (the reference's per-zone
// adi call); batched mode launches each phase once over all zones.  The
// per-zone order (all zones' rhs, then all x solves, ...) is equivalent to
// the reference's zone-by-zone adi because zones are independent after the
// exchange.
int adi_step(npbmz_ctx* ctx, int mode) {
  int rc;
  if ((rc = exch(ctx))) return rc;
  if (mode == NPBMZ_MODE_PERZONE) {
    for (int z = 0; z < ctx->nz_; ++z) {
      if ((rc = k_rhs(ctx, z))) return rc;
      if ((rc = k_solve(ctx, 0, false, z))) return rc;
      if ((rc = k_solve(ctx, 1, false, z))) return rc;
      if ((rc = k_solve(ctx, 2, true, z))) return rc;
    }
    return NPBMZ_OK;
  }
  if ((rc = k_rhs(ctx, -1))) return rc;
  if ((rc = k_solve(ctx, 0, false, -1))) return rc;
  if ((rc = k_solve(ctx, 1, false, -1))) return rc;
  if ((rc = k_solve(ctx, 2, true, -1))) return rc;
  return NPBMZ_OK;
}

int run_phase(npbmz_ctx* ctx, int phase, int mode) {
  int zlo = 0, zhi = ctx->nz_;
  if (mode != NPBMZ_MODE_PERZONE) { zlo = -1; zhi = 0; }
  int rc = NPBMZ_OK;
  if (phase == NPBMZ_PHASE_EXCH) return exch(ctx);
  for (int z = zlo; z < zhi; ++z) {
    switch (phase) {
      case NPBMZ_PHASE_RHS: rc = k_rhs(ctx, z); break;
      case NPBMZ_PHASE_XSOLVE: rc = k_solve(ctx, 0, false, z); break;
      case NPBMZ_PHASE_YSOLVE: rc = k_solve(ctx, 1, false, z); break;
      case NPBMZ_PHASE_ZSOLVE: rc = k_solve(ctx, 2, false, z); break;
      case NPBMZ_PHASE_ADD: rc = k_add(ctx, z); break;
      default: return fail(ctx, NPBMZ_EINVAL, "unknown phase");
    }
    if (rc) return rc;
  }
  return rc;
}

int ensure_graph(npbmz_ctx* ctx) {
  auto& c = *ctx;
  if (c.gexec) return NPBMZ_OK;
  CK(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
  int rc = adi_step(ctx, NPBMZ_MODE_BATCHED);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(c.stream, &g);
  if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
  if (e != hipSuccess) return fail(ctx, NPBMZ_EHIP, std::string("capture: ") + hipGetErrorString(e));
  c.graph = g;
  CK(hipGraphInstantiate(&c.gexec, c.graph, nullptr, nullptr, 0));
  return NPBMZ_OK;
}

void describe(npbmz_ctx* c) {
  char b[256];
  snprintf(b, sizeof b, "zones=%d order=%s solve_grid=%d/%d/%d fp=%s", c->nz_,
           c->order.empty() || c->order[0] == 0 ? "natural" : "largest-first", c->solve_grid[0],
           c->solve_grid[1], c->solve_grid[2], strict(c) ? "strict" : "fma");
  c->sched = b;
}

}  // namespace

// ===================================================================== C ABI
extern "C" {

const char* npbmz_last_error(const npbmz_ctx* ctx) {
  return ctx ? ctx->err.c_str() : g_create_err.c_str();
}

int npbmz_create(const npbmz_problem* p, int32_t flags, npbmz_ctx** out) {
  if (!p || !out) return fail(nullptr, NPBMZ_EINVAL, "null argument");
  *out = nullptr;
  if (p->bench != NPBMZ_BT && p->bench != NPBMZ_SP)
    return fail(nullptr, NPBMZ_ENOTSUP, "bench must be NPBMZ_BT or NPBMZ_SP");
  if (p->nzones <= 0 || p->nzones != p->x_zones * p->y_zones)
    return fail(nullptr, NPBMZ_EINVAL, "nzones must equal x_zones*y_zones and be positive");
  if (!p->nx || !p->ny || !p->nz || !p->iz_west || !p->iz_east || !p->iz_south || !p->iz_north)
    return fail(nullptr, NPBMZ_EINVAL, "null zone table");
  if (!(p->dt > 0.0)) return fail(nullptr, NPBMZ_EINVAL, "dt must be positive");
  if (p->gx_size < 2 || p->gy_size < 2 || p->gz_size < 2)
    return fail(nullptr, NPBMZ_EINVAL, "grid sizes must be at least 2");
  int n = p->nzones;
  for (int z = 0; z < n; ++z) {
    // below 5 points the reference's dissipation stencils read its halo
    // values (rhs.jl:158-210, lhsx.jl:48-90), which this library does not keep
    if (p->nx[z] < 5 || p->ny[z] < 5 || p->nz[z] < 5)
      return fail(nullptr, NPBMZ_EINVAL, "zone dimension below 5");
    if ((int64_t)p->nx[z] * p->ny[z] * p->nz[z] > (1ll << 30))
      return fail(nullptr, NPBMZ_EINVAL, "zone too large");
    int nb[4] = {p->iz_west[z], p->iz_east[z], p->iz_south[z], p->iz_north[z]};
    for (int q = 0; q < 4; ++q)
      if (nb[q] < 0 || nb[q] >= n) return fail(nullptr, NPBMZ_EINVAL, "neighbour out of range");
    // faces must match the neighbour's plane
    for (int q = 0; q < 2; ++q)
      if (p->ny[nb[q]] != p->ny[z] || p->nz[nb[q]] != p->nz[z])
        return fail(nullptr, NPBMZ_EINVAL, "east/west neighbours differ in ny or nz");
    for (int q = 2; q < 4; ++q)
      if (p->nx[nb[q]] != p->nx[z] || p->nz[nb[q]] != p->nz[z])
        return fail(nullptr, NPBMZ_EINVAL, "north/south neighbours differ in nx or nz");
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
    return fail(nullptr, NPBMZ_ENODEV, "no HIP device");
  int dev = 0;
  (void)hipGetDevice(&dev);
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) != hipSuccess)
    return fail(nullptr, NPBMZ_ENODEV, "hipGetDeviceProperties failed");
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(nullptr, NPBMZ_ENODEV,
                std::string("device is ") + prop.gcnArchName + ", library is built for gfx950");

  npbmz_ctx* ctx = new npbmz_ctx();
  auto& c = *ctx;
  c.bench = p->bench; c.flags = flags; c.nz_ = n;
  c.x_zones = p->x_zones; c.y_zones = p->y_zones;
  c.gx = p->gx_size; c.gy = p->gy_size; c.gz = p->gz_size;
  c.dt = p->dt; c.omega = p->omega;
  c.nx.assign(p->nx, p->nx + n); c.ny.assign(p->ny, p->ny + n); c.nz.assign(p->nz, p->nz + n);
  c.west.assign(p->iz_west, p->iz_west + n); c.east.assign(p->iz_east, p->iz_east + n);
  c.south.assign(p->iz_south, p->iz_south + n); c.north.assign(p->iz_north, p->iz_north + n);
  c.K = make_constants(c.bench, c.dt, c.gx, c.gy, c.gz, c.x_zones, c.y_zones);
  c.zh.resize(n);
  int64_t base = 0;
  for (int z = 0; z < n; ++z) {
    c.zh[z] = ZoneDev{c.nx[z], c.ny[z], c.nz[z], 0, base};
    base += (int64_t)c.nx[z] * c.ny[z] * c.nz[z];
  }
  c.npoints = base;
  // largest zones first: the batched launches then schedule longest work first
  c.order.resize(n);
  std::iota(c.order.begin(), c.order.end(), 0);
  std::stable_sort(c.order.begin(), c.order.end(), [&](int a, int b) {
    return (int64_t)c.nx[a] * c.ny[a] * c.nz[a] > (int64_t)c.nx[b] * c.ny[b] * c.nz[b];
  });
  int rc = NPBMZ_OK;
  auto bad = [&](int code) {
    g_create_err = c.err;
    npbmz_destroy(ctx);
    return code;
  };
#define CKC(call)                                                                     \
  do {                                                                                \
    hipError_t e_ = (call);                                                           \
    if (e_ != hipSuccess) {                                                           \
      c.err = std::string(#call) + ": " + hipGetErrorString(e_);                      \
      return bad(e_ == hipErrorOutOfMemory ? NPBMZ_ENOMEM : NPBMZ_EHIP);              \
    }                                                                                 \
  } while (0)
  CKC(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
  CKC(hipEventCreate(&c.ev0));
  CKC(hipEventCreate(&c.ev1));
  CKC(hipMalloc(&c.zd, sizeof(ZoneDev) * n));
  CKC(hipMemcpy(c.zd, c.zh.data(), sizeof(ZoneDev) * n, hipMemcpyHostToDevice));
  size_t fb = sizeof(double) * 5 * (size_t)c.npoints;
  CKC(hipMalloc(&c.u, fb));
  CKC(hipMalloc(&c.frc, fb));
  CKC(hipMalloc(&c.rhs, fb));
  CKC(hipMemset(c.u, 0, fb));
  CKC(hipMemset(c.frc, 0, fb));
  CKC(hipMemset(c.rhs, 0, fb));
  if ((rc = build_tables(ctx))) return bad(rc);
  describe(ctx);
#undef CKC
  *out = ctx;
  return NPBMZ_OK;
}

void npbmz_destroy(npbmz_ctx* ctx) {
  if (!ctx) return;
  auto& c = *ctx;
  if (c.stream) (void)hipStreamSynchronize(c.stream);
  if (c.gexec) (void)hipGraphExecDestroy(c.gexec);
  if (c.graph) (void)hipGraphDestroy(c.graph);
  Table* ts[] = {&c.t_int, &c.t_all, &c.t_line[0], &c.t_line[1], &c.t_line[2], &c.t_face};
  for (Table* t : ts)
    if (t->d) (void)hipFree(t->d);
  void* ps[] = {c.zd, c.u, c.frc, c.rhs, c.cscr, c.faces, c.partial};
  for (void* q : ps)
    if (q) (void)hipFree(q);
  if (c.ev0) (void)hipEventDestroy(c.ev0);
  if (c.ev1) (void)hipEventDestroy(c.ev1);
  if (c.stream) (void)hipStreamDestroy(c.stream);
  delete ctx;
}

static int field_ptr(npbmz_ctx* ctx, int field, double** f) {
  if (field == NPBMZ_FIELD_U) *f = ctx->u;
  else if (field == NPBMZ_FIELD_FORCING) *f = ctx->frc;
  else if (field == NPBMZ_FIELD_RHS) *f = ctx->rhs;
  else return fail(ctx, NPBMZ_EINVAL, "unknown field");
  return NPBMZ_OK;
}

// Map (m, i, j, k) of the zone's grid (0-based i, j, k) to an offset into the
// host parent array described by lo/hi/order.  Grid index g maps to the
// reference's array index g + origin (0 for BT and SP, 1 for LU).
static int host_map(npbmz_ctx* ctx, int zone, const int64_t lo[4], const int64_t hi[4],
                    int order, int64_t dims[4]) {
  for (int d = 0; d < 4; ++d) {
    dims[d] = hi[d] - lo[d] + 1;
    if (dims[d] <= 0) return fail(ctx, NPBMZ_EINVAL, "empty axis");
  }
  const int origin = ctx->bench == NPBMZ_LU ? 1 : 0;
  int ax_m = order == NPBMZ_ORDER_IJKM ? 3 : 0;
  int ax_i = order == NPBMZ_ORDER_IJKM ? 0 : 1;
  if (lo[ax_m] > 1 || hi[ax_m] < 5) return fail(ctx, NPBMZ_EINVAL, "component axis lacks 1:5");
  const int n[3] = {ctx->nx[zone], ctx->ny[zone], ctx->nz[zone]};
  for (int d = 0; d < 3; ++d) {
    if (lo[ax_i + d] > origin || hi[ax_i + d] < n[d] - 1 + origin)
      return fail(ctx, NPBMZ_EINVAL, "host array does not cover the zone's grid");
  }
  return NPBMZ_OK;
}

static int transfer(npbmz_ctx* ctx, int32_t zone, int32_t field, double* host,
                    const int64_t lo[4], const int64_t hi[4], int32_t order, bool put) {
  if (!ctx) return NPBMZ_EINVAL;
  if (!host || !lo || !hi) return fail(ctx, NPBMZ_EINVAL, "null argument");
  if (zone < 0 || zone >= ctx->nz_) return fail(ctx, NPBMZ_EINVAL, "zone out of range");
  if (order != NPBMZ_ORDER_MIJK && order != NPBMZ_ORDER_IJKM)
    return fail(ctx, NPBMZ_EINVAL, "unknown axis order");
  double* f = nullptr;
  int rc;
  if ((rc = field_ptr(ctx, field, &f))) return rc;
  int64_t dims[4];
  if ((rc = host_map(ctx, zone, lo, hi, order, dims))) return rc;
  const ZoneDev Z = ctx->zh[zone];
  const int origin = ctx->bench == NPBMZ_LU ? 1 : 0;
  const size_t npt = (size_t)Z.nx * Z.ny * Z.nz;
  std::vector<double> stage(npt * 5);
  double* dev = f + Z.base * 5;
  if (!put) CK(hipMemcpyAsync(stage.data(), dev, stage.size() * sizeof(double),
                              hipMemcpyDeviceToHost, ctx->stream));
  if (!put) CK(hipStreamSynchronize(ctx->stream));
  for (int k = 0; k < Z.nz; ++k)
    for (int j = 0; j < Z.ny; ++j)
      for (int i = 0; i < Z.nx; ++i)
        for (int m = 0; m < 5; ++m) {
          int64_t ii = i + origin, jj = j + origin, kk = k + origin, mm = m + 1;
          int64_t off;
          if (order == NPBMZ_ORDER_MIJK)
            off = (mm - lo[0]) + dims[0] * ((ii - lo[1]) + dims[1] * ((jj - lo[2]) + dims[2] * (kk - lo[3])));
          else
            off = (ii - lo[0]) + dims[0] * ((jj - lo[1]) + dims[1] * ((kk - lo[2]) + dims[2] * (mm - lo[3])));
          size_t s = ((size_t)(k * Z.ny + j) * Z.nx + i) * 5 + m;
          if (put) stage[s] = host[off];
          else host[off] = stage[s];
        }
  if (put) {
    CK(hipMemcpyAsync(dev, stage.data(), stage.size() * sizeof(double), hipMemcpyHostToDevice,
                      ctx->stream));
    CK(hipStreamSynchronize(ctx->stream));
  }
  return NPBMZ_OK;
}

int npbmz_put_field(npbmz_ctx* ctx, int32_t zone, int32_t field, const double* host,
                    const int64_t lo[4], const int64_t hi[4], int32_t order) {
  return transfer(ctx, zone, field, const_cast<double*>(host), lo, hi, order, true);
}

int npbmz_get_field(npbmz_ctx* ctx, int32_t zone, int32_t field, double* host,
                    const int64_t lo[4], const int64_t hi[4], int32_t order) {
  return transfer(ctx, zone, field, host, lo, hi, order, false);
}

int npbmz_init_exact(npbmz_ctx* ctx) {
  if (!ctx) return NPBMZ_EINVAL;
  auto& c = *ctx;
  const Table& t = c.t_all;
  if (t.n <= 0) return NPBMZ_OK;
  // setup arithmetic always runs without contraction: it is the reference's
  // initial state, not part of the timed step
  hipLaunchKernelGGL(exact::bt::init_kernel<kPtNT>, dim3(t.n), dim3(kPtNT), 0, c.stream, c.u,
                     c.zd, t.d, c.K);
  CK(hipGetLastError());
  hipLaunchKernelGGL(exact::bt::exact_rhs_kernel<kPtNT>, dim3(t.n), dim3(kPtNT), 0, c.stream,
                     c.frc, c.zd, t.d, c.K);
  CK(hipGetLastError());
  CK(hipMemsetAsync(c.rhs, 0, sizeof(double) * 5 * (size_t)c.npoints, c.stream));
  CK(hipStreamSynchronize(c.stream));
  return NPBMZ_OK;
}

int npbmz_step(npbmz_ctx* ctx, int32_t nsteps, int32_t mode, double* elapsed_ms) {
  if (!ctx) return NPBMZ_EINVAL;
  if (nsteps < 0) return fail(ctx, NPBMZ_EINVAL, "nsteps < 0");
  if (mode < NPBMZ_MODE_PERZONE || mode > NPBMZ_MODE_GRAPH)
    return fail(ctx, NPBMZ_EINVAL, "unknown mode");
  auto& c = *ctx;
  int rc;
  if (mode == NPBMZ_MODE_GRAPH && (rc = ensure_graph(ctx))) return rc;
  CK(hipEventRecord(c.ev0, c.stream));
  for (int s = 0; s < nsteps; ++s) {
    if (mode == NPBMZ_MODE_GRAPH) {
      CK(hipGraphLaunch(c.gexec, c.stream));
    } else if ((rc = adi_step(ctx, mode))) {
      return rc;
    }
  }
  CK(hipEventRecord(c.ev1, c.stream));
  CK(hipEventSynchronize(c.ev1));
  if (elapsed_ms) {
    float ms = 0.f;
    CK(hipEventElapsedTime(&ms, c.ev0, c.ev1));
    *elapsed_ms = ms;
  }
  return NPBMZ_OK;
}

int npbmz_run_phase(npbmz_ctx* ctx, int32_t phase, int32_t mode) {
  if (!ctx) return NPBMZ_EINVAL;
  int rc = run_phase(ctx, phase, mode);
  if (rc) return rc;
  CK(hipStreamSynchronize(ctx->stream));
  return NPBMZ_OK;
}

int npbmz_phase_times(npbmz_ctx* ctx, int32_t nsteps, double phase_ms[6]) {
  if (!ctx) return NPBMZ_EINVAL;
  if (nsteps <= 0 || !phase_ms) return fail(ctx, NPBMZ_EINVAL, "nsteps <= 0 or null output");
  auto& c = *ctx;
  const int nph = 5;  // exch, rhs, x, y, z(+add)
  std::vector<hipEvent_t> ev((size_t)nsteps * nph * 2, nullptr);
  int rc = NPBMZ_OK;
  for (auto& e : ev)
    if (hipEventCreate(&e) != hipSuccess) { rc = fail(ctx, NPBMZ_EHIP, "hipEventCreate"); break; }
  for (int s = 0; s < nsteps && rc == NPBMZ_OK; ++s) {
    for (int ph = 0; ph < nph && rc == NPBMZ_OK; ++ph) {
      hipEvent_t* e = &ev[((size_t)s * nph + ph) * 2];
      if (hipEventRecord(e[0], c.stream) != hipSuccess) { rc = fail(ctx, NPBMZ_EHIP, "record"); break; }
      if (ph == 0) rc = exch(ctx);
      else if (ph == 1) rc = k_rhs(ctx, -1);
      else rc = k_solve(ctx, ph - 2, true, -1);
      if (rc == NPBMZ_OK && hipEventRecord(e[1], c.stream) != hipSuccess)
        rc = fail(ctx, NPBMZ_EHIP, "record");
    }
  }
  if (rc == NPBMZ_OK && hipStreamSynchronize(c.stream) != hipSuccess)
    rc = fail(ctx, NPBMZ_EHIP, "hipStreamSynchronize");
  for (int p = 0; p < 6; ++p) phase_ms[p] = 0.0;
  if (rc == NPBMZ_OK) {
    for (int s = 0; s < nsteps; ++s)
      for (int ph = 0; ph < nph; ++ph) {
        float ms = 0.f;
        hipEvent_t* e = &ev[((size_t)s * nph + ph) * 2];
        if (hipEventElapsedTime(&ms, e[0], e[1]) != hipSuccess) {
          rc = fail(ctx, NPBMZ_EHIP, "hipEventElapsedTime");
          break;
        }
        phase_ms[ph] += ms;
      }
  }
  for (auto& e : ev)
    if (e) (void)hipEventDestroy(e);
  return rc;
}

int npbmz_norms(npbmz_ctx* ctx, double xcr[5], double xce[5], double* xci) {
  if (!ctx) return NPBMZ_EINVAL;
  if (!xcr || !xce) return fail(ctx, NPBMZ_EINVAL, "null output");
  auto& c = *ctx;
  (void)xci;
  const int nzn = c.nz_;
  std::vector<double> pe((size_t)c.t_all.n * 5), pr((size_t)c.t_int.n * 5);
  // error_norm on every zone's current u
  if (strict(ctx))
    hipLaunchKernelGGL(exact::bt::norm_kernel<kPtNT>, dim3(c.t_all.n), dim3(kPtNT), 0, c.stream,
                       c.u, c.rhs, c.zd, c.t_all.d, c.partial, 0, c.K);
  else
    hipLaunchKernelGGL(fast::bt::norm_kernel<kPtNT>, dim3(c.t_all.n), dim3(kPtNT), 0, c.stream,
                       c.u, c.rhs, c.zd, c.t_all.d, c.partial, 0, c.K);
  CK(hipGetLastError());
  CK(hipMemcpyAsync(pe.data(), c.partial, pe.size() * sizeof(double), hipMemcpyDeviceToHost,
                    c.stream));
  // compute_rhs without a new exchange, then rhs_norm
  int rc = k_rhs(ctx, -1, true);
  if (rc) return rc;
  hipLaunchKernelGGL(exact::bt::norm_kernel<kPtNT>, dim3(c.t_int.n), dim3(kPtNT), 0, c.stream,
                     c.u, c.rhs, c.zd, c.t_int.d, c.partial, 1, c.K);
  CK(hipGetLastError());
  CK(hipMemcpyAsync(pr.data(), c.partial, pr.size() * sizeof(double), hipMemcpyDeviceToHost,
                    c.stream));
  CK(hipStreamSynchronize(c.stream));
  c.zn_rhs.assign((size_t)nzn * 5, 0.0);
  c.zn_err.assign((size_t)nzn * 5, 0.0);
  for (int m = 0; m < 5; ++m) { xcr[m] = 0.0; xce[m] = 0.0; }
  for (int z = 0; z < nzn; ++z) {
    double se[5] = {0, 0, 0, 0, 0}, sr[5] = {0, 0, 0, 0, 0};
    for (int b = 0; b < c.t_all.zcount[z]; ++b)
      for (int m = 0; m < 5; ++m) se[m] += pe[(size_t)(c.t_all.zfirst[z] + b) * 5 + m];
    for (int b = 0; b < c.t_int.zcount[z]; ++b)
      for (int m = 0; m < 5; ++m) sr[m] += pr[(size_t)(c.t_int.zfirst[z] + b) * 5 + m];
    const double vol =
        double(c.nx[z] - 2) * double(c.ny[z] - 2) * double(c.nz[z] - 2);
    for (int m = 0; m < 5; ++m) {
      c.zn_err[(size_t)z * 5 + m] = std::sqrt(se[m] / vol);
      c.zn_rhs[(size_t)z * 5 + m] = std::sqrt(sr[m] / vol);
    }
  }
  for (int z = 0; z < nzn; ++z)
    for (int m = 0; m < 5; ++m) {
      xcr[m] = xcr[m] + c.zn_rhs[(size_t)z * 5 + m] / c.dt;
      xce[m] = xce[m] + c.zn_err[(size_t)z * 5 + m];
    }
  return NPBMZ_OK;
}

int npbmz_zone_norms(npbmz_ctx* ctx, double* rms_rhs, double* rms_err) {
  if (!ctx) return NPBMZ_EINVAL;
  if (ctx->zn_rhs.empty()) return fail(ctx, NPBMZ_EINVAL, "npbmz_norms has not run");
  if (rms_rhs) std::memcpy(rms_rhs, ctx->zn_rhs.data(), ctx->zn_rhs.size() * sizeof(double));
  if (rms_err) std::memcpy(rms_err, ctx->zn_err.data(), ctx->zn_err.size() * sizeof(double));
  return NPBMZ_OK;
}

int npbmz_tune(npbmz_ctx* ctx, int32_t budget_ms) {
  if (!ctx) return NPBMZ_EINVAL;
  (void)budget_ms;
  describe(ctx);
  return NPBMZ_OK;
}

int npbmz_describe(const npbmz_ctx* ctx, char* buf, int32_t cap) {
  if (!ctx || !buf || cap <= 0) return NPBMZ_EINVAL;
  std::snprintf(buf, (size_t)cap, "%s", ctx->sched.c_str());
  return NPBMZ_OK;
}

int npbmz_size(const npbmz_ctx* ctx, int32_t* nzones, int64_t* npoints) {
  if (!ctx) return NPBMZ_EINVAL;
  if (nzones) *nzones = ctx->nz_;
  if (npoints) *npoints = ctx->npoints;
  return NPBMZ_OK;
}

int npbmz_probe_copy(npbmz_ctx* ctx, int64_t nbytes, int32_t reps, double* ms_per_launch) {
  if (!ctx) return NPBMZ_EINVAL;
  if (nbytes < 32 || nbytes > (int64_t(16) << 30) || reps <= 0 || !ms_per_launch)
    return fail(ctx, NPBMZ_EINVAL, "nbytes must be in [32, 16 GiB], reps > 0");
  auto& c = *ctx;
  const int64_t n = nbytes / 32;
  double4 *a = nullptr, *b = nullptr;
  CK(hipMalloc(&a, n * 32));
  if (hipMalloc(&b, n * 32) != hipSuccess) {
    (void)hipFree(a);
    return fail(ctx, NPBMZ_ENOMEM, "probe buffer");
  }
  int rc = NPBMZ_OK;
  const int grid = (int)std::min<int64_t>((n + 255) / 256, 256 * 64);
  hipLaunchKernelGGL(copy_probe_kernel, dim3(grid), dim3(256), 0, c.stream, a, b, n);
  (void)hipEventRecord(c.ev0, c.stream);
  for (int r = 0; r < reps; ++r)
    hipLaunchKernelGGL(copy_probe_kernel, dim3(grid), dim3(256), 0, c.stream, a, b, n);
  (void)hipEventRecord(c.ev1, c.stream);
  float ms = 0.f;
  if (hipEventSynchronize(c.ev1) != hipSuccess || hipEventElapsedTime(&ms, c.ev0, c.ev1) != hipSuccess)
    rc = fail(ctx, NPBMZ_EHIP, "probe timing");
  *ms_per_launch = ms / reps;
  (void)hipFree(a);
  (void)hipFree(b);
  return rc;
}

}  // extern "C"
