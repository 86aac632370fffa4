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
#include <mutex>
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

// A zone face copied by exch_qbc: the 5 components of the destination zone's
// boundary plane `type` (0 x-minus i=0, 1 x-plus i=nx-1, 2 y-minus j=0,
// 3 y-plus j=ny-1) come from the source (neighbour) zone's plane next to the
// shared boundary.  The copied points are the transverse index t0..t0+tn-1
// (j for x faces, i for y faces) on planes k0, k0+1, ..., npts/tn of them.
struct Face {
  int32_t dst, src, type, npts, t0, tn, k0, pad;
};

// exch_qbc (src/BT/exchange_qbc.jl:145-466 for BT and SP, one rank): the
// reference packs every face before unpacking any, and sources are interior
// planes while destinations are boundary planes, so one kernel over all
// faces reproduces it.  LU's faces overlap at the zone corners and its
// sources include the corners of other faces, so LU runs two of these
// launches (see build_tables).  Copies are bit exact.
// Copy t of face task T.
__device__ __forceinline__ void exch_copy(double* u, const ZoneDev* __restrict__ zones,
                                          const Face* __restrict__ faces, const Task T, int t) {
  const Face F = faces[T.zone];
  const int idx = T.first + t;
  if (idx >= F.npts) return;
  const ZoneDev D = zones[F.dst], S = zones[F.src];
  int64_t dp, sp;
  if (F.type < 2) {
    const int j = F.t0 + idx % F.tn, k = F.k0 + idx / F.tn;
    const int di = F.type == 0 ? 0 : D.nx - 1;
    const int si = F.type == 0 ? S.nx - 2 : 1;
    dp = D.base + ((int64_t)k * D.ny + j) * D.nx + di;
    sp = S.base + ((int64_t)k * S.ny + j) * S.nx + si;
  } else {
    const int i = F.t0 + idx % F.tn, k = F.k0 + idx / F.tn;
    const int dj = F.type == 2 ? 0 : D.ny - 1;
    const int sj = F.type == 2 ? S.ny - 2 : 1;
    dp = D.base + ((int64_t)k * D.ny + dj) * D.nx + i;
    sp = S.base + ((int64_t)k * S.ny + sj) * S.nx + i;
  }
  const double* s = u + sp * 5;
  double* d = u + dp * 5;
#pragma unroll
  for (int m = 0; m < 5; ++m) d[m] = s[m];
}

template <int NT>
__global__ __launch_bounds__(NT) void exch_kernel(double* __restrict__ u,
                                                  const ZoneDev* __restrict__ zones,
                                                  const Face* __restrict__ faces,
                                                  const Task* __restrict__ tasks) {
  exch_copy(u, zones, faces, tasks[blockIdx.x], (int)threadIdx.x);
}

// Where a BT/SP zone's halo cells come from in the exchange, per zone and
// side (0 i = 0, 1 i = nx-1, 2 j = 0, 3 j = ny-1): the halo cell at
// transverse index t (j for x sides, i for y sides) and plane k holds the
// source zone's point base + t * st + k * sk (exch_copy's source).  A side
// that receives nothing has st = sk = 0 and base = -1.
struct HaloSrc {
  int64_t base, st, sk;
};

// Each workgroup's XCD: the XCC_ID hardware register (20, bits 0..3) of
// block b into xcc[b].
__global__ __launch_bounds__(64) void xcc_probe_kernel(int* __restrict__ xcc) {
  if (threadIdx.x == 0) xcc[blockIdx.x] = (int)__builtin_amdgcn_s_getreg(20 | (0 << 6) | ((4 - 1) << 11));
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
// Reciprocals of the solves: IEEE division, as the reference computes them.
constexpr bool kIeeeDiv = true;
// LU's jacld/jacu blocks with the reference's own operand grouping
constexpr bool kCollapsedBlocks = false;
// compute_rhs's dissipation loop by loop, as the reference adds it
constexpr bool kBlendedDiss = false;
__device__ __forceinline__ double recip(double x) { return 1.0 / x; }
#include "bt_kernels.inc"
#include "sp_kernels.inc"
#include "lu_kernels.inc"
}  // namespace exact

namespace fast {
#pragma clang fp contract(fast)
// Reciprocals of the solves: the hardware reciprocal refined by two
// Newton-Raphson steps, within about an ulp of 1/x for the normal,
// moderate pivots and densities the solves divide by, and a much shorter
// dependent chain than IEEE division (scaling, reciprocal, refinement and
// fix-up).
constexpr bool kIeeeDiv = false;
// LU's jacld/jacu blocks from per-point velocities and the collapsed
// constant factors of LuK (about half the arithmetic)
constexpr bool kCollapsedBlocks = true;
// compute_rhs's dissipation as one five-point formula whose weights are the
// reference's loops' coefficients summed for the point's line index: no
// branch, where the loop form runs every loop's branch in most waves
constexpr bool kBlendedDiss = true;
__device__ __forceinline__ double recip(double x) {
  double r = __builtin_amdgcn_rcp(x);
  r = __builtin_fma(__builtin_fma(-x, r, 1.0), r, r);
  return __builtin_fma(__builtin_fma(-x, r, 1.0), r, r);
}
#include "bt_kernels.inc"
#include "bt_rows.inc"
#include "bt_adjoint.inc"
#include "sp_kernels.inc"
#include "sp_adjoint.inc"
#include "lu_kernels.inc"
#include "lu_adjoint.inc"
}  // namespace fast

// --------------------------------------------- the allocation-free norms
// Scratch of npbmz_norms_noalloc: the norm kernels' per-block partial sums,
// each zone's block ranges, per-zone norms and the final sums.  One copy per
// GPU, held for the life of the process; host code reaches it only through
// the address hipGetSymbolAddress returns for the current device.
constexpr size_t kScratchDoubles = (size_t(2) << 20) / sizeof(double);  // 2 MiB
static __device__ __align__(256) double scratch[kScratchDoubles];

// Final reductions of npbmz_norms on the device, in the host path's order:
// per zone, the blocks' partial sums in block order from 0.0; the zone's
// rms; then the sums over zones in zone order.  BT/SP (src/BT/verify.jl:
// 434-499): pe, pr are error_norm's (all points) and rhs_norm's (interior)
// partials, 5 per block; xcr = sum of rms_rhs/dt, xce = sum of rms_err.  LU
// (src/LU/lu-node.jl:433-458): pe holds 10 per block (ERROR, then l2norm of
// rsd), zp each zone's surface integral; xcr and xce are plain sums.
// rng: first block and block count per zone of pe's table, then of pr's.
// out: xcr[5], xce[5], xci.  One workgroup; zn holds nz x 10 values.
__global__ __launch_bounds__(256) void norms_reduce_kernel(
    const double* __restrict__ pe, const double* __restrict__ pr, const double* __restrict__ zp,
    const int32_t* __restrict__ rng, const ZoneDev* __restrict__ zones, int nz, int lu,
    double dt, double* __restrict__ zn, double* __restrict__ out) {
#pragma clang fp contract(off)
  for (int z = threadIdx.x; z < nz; z += blockDim.x) {
    const ZoneDev Z = zones[z];
    const double vol = double(Z.nx - 2) * double(Z.ny - 2) * double(Z.nz - 2);
    double se[10];
    for (int q = 0; q < 10; ++q) se[q] = 0.0;
    const int fa = rng[z], ca = rng[nz + z];
    for (int b = 0; b < ca; ++b)
      for (int q = 0; q < (lu ? 10 : 5); ++q) se[q] += pe[(size_t)(fa + b) * (lu ? 10 : 5) + q];
    if (!lu) {
      const int fi = rng[2 * nz + z], ci = rng[3 * nz + z];
      for (int b = 0; b < ci; ++b)
        for (int m = 0; m < 5; ++m) se[5 + m] += pr[(size_t)(fi + b) * 5 + m];
    }
    for (int q = 0; q < 10; ++q) zn[(size_t)z * 10 + q] = sqrt(se[q] / vol);
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  double xcr[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, xce[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, frc = 0.0;
  for (int z = 0; z < nz; ++z) {
    for (int m = 0; m < 5; ++m) {
      xcr[m] = lu ? xcr[m] + zn[(size_t)z * 10 + 5 + m] : xcr[m] + zn[(size_t)z * 10 + 5 + m] / dt;
      xce[m] = xce[m] + zn[(size_t)z * 10 + m];
    }
    if (lu) frc = frc + zp[z];
  }
  for (int m = 0; m < 5; ++m) {
    out[m] = xcr[m];
    out[5 + m] = xce[m];
  }
  out[10] = frc;
}

}  // namespace npbmz

using namespace npbmz;

// ---------------------------------------------------------------- host side
namespace {

constexpr int kPtNT = 128;   // threads per point task (rhs, add, init, norms)
constexpr int kExchNT = 128; // threads per face task
constexpr int kLineNT = 64;  // lanes (lines) per solve task: one wave
constexpr int kRowR = 6;     // reduced rows per half line kept in registers (rows kernel)
constexpr int kRowRLong = 8; // the same where a direction's tasks need more than one round of
                             // waves: the rows kernel with its block rows read as applied
                             // (fewer registers live), less spill traffic
constexpr int kPlaneNT = 64; // threads per block of an LU hyperplane launch
constexpr int kGraphGroup = 8; // timesteps per replay of the grouped graph
constexpr int kMaxSolveGrid = 65536; // largest persistent solve grid accepted

// LU sweep schedules: one launch per hyperplane over all zones, one
// workgroup per zone walking all hyperplanes of the sweep, or a zone's
// hyperplanes split over pipelined workgroups (tiles).
enum { kSweepPlane = 0, kSweepZone = 1, kSweepTiles = 2 };
constexpr int kMaxTiles = 8;  // tiles per zone of the pipelined sweep

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
  LuK lk{};
  std::vector<ZoneDev> zh;
  ZoneDev* zd = nullptr;
  int64_t npoints = 0;
  double *u = nullptr, *frc = nullptr, *rhs = nullptr;
  double* cscr = nullptr;
  size_t cscr_slot = 0;  // doubles per scratch slot (max over directions)
  int maxrows[3] = {0, 0, 0};
  int solve_grid[3] = {0, 0, 0};  // 0: one block per task, else persistent grid
  // BT line-solve kernels per direction.  The defaults are the fastest
  // measured on MI355X for BT-MZ classes A and B: the two-sided solve in all
  // three directions with FMA arithmetic; with strict arithmetic (no
  // two-sided solve) the wave pair in y.
  int solve_pair = 2;             // BT: bit d set: direction d uses solve_pair_kernel
  int solve_twist = 7;            // BT/SP, FMA only: bit d set: solve_twist_kernel (wins over pair)
  bool sp_twist_ok[3] = {false, false, false};  // SP: every line of direction d has >= 4 unknowns
  int solve_rows = 0;             // BT, FMA only: bit d set: solve_rows_kernel (wins over twist);
  int cus = 256;                  // the device's compute units
                                  // set at creation by size
  int halfmax[3] = {0, 0, 0};     // rows of the longer half of the longest line per direction
  bool largest_first = true;      // task order: largest zones / longest lines first
  int lu_sweep = kSweepZone;      // LU sweep schedule
  int lu_nt = 256;                // threads per workgroup of the per-zone sweep
  int lu_fwd = 3;                 // per-zone sweep: forward each hyperplane's rsd (1) and u (2)
                                  // through LDS; 3: the padded-halo sweep where it applies
  size_t max_lds = 65536;         // dynamic LDS a zone-sweep workgroup may take
  int lu_tiles = 4;               // pipelined sweep: workgroups (tiles) per zone
  int lu_kb = 4;                  // pipelined sweep: hyperplanes per published count
  int lu_spin = 1 << 20;          // pipelined sweep: polls before a wait gives up
  int lu_handoff = -1;            // pipelined sweep: 0 agent-scope, 1 XCD-local, 2 XCD-local
                                  // with every placement check failing; -1 XCD-local where
                                  // it applies
  bool tile_local_ok = false;     // this device places blocks of equal b mod 8 on one XCD
  double* txbuf = nullptr;        // pipelined sweep: tile edge rows of every hyperplane
  int* tcnt = nullptr;            // pipelined sweep: finished hyperplanes per tile,
                                  // then per zone the tiles done, then the error flag,
                                  // then each tile's XCD plus one (XCD-local hand-off)
  Table t_int, t_all, t_line[3], t_face, t_face2;
  Table t_row[3];  // BT: tasks of fast::bt::kRowLines lines for solve_rows_kernel
  HaloSrc* hsrc = nullptr;  // BT/SP: where each zone side's halo cells come from
  bool hsrc_ok = false;     // hsrc describes the face table exactly
  int rhs_exch = 1;         // BT/SP batched steps: 1 exchange inside compute_rhs's launch
  Face* faces = nullptr;
  double* partial = nullptr;
  size_t partial_n = 0;     // blocks the partial sums are sized for
  double* zpart = nullptr;  // one value per zone (LU surface integrals)
  std::vector<double> zn_rhs, zn_err;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipGraphExec_t gexec = nullptr;  // one timestep
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec_n = nullptr;  // kGraphGroup timesteps back to back
  hipGraph_t graph_n = nullptr;
  std::string err;
  std::string sched;
  // npbmz_norms_noalloc: this device's copy of the static scratch, and the
  // per-zone block ranges it uploads there (built with the task tables)
  double* scratch = nullptr;
  std::string scratch_err;
  std::vector<int32_t> noalloc_rng;
};

namespace {

std::string g_create_err;

void drop_graphs(npbmz_ctx& c) {
  if (c.gexec) { (void)hipGraphExecDestroy(c.gexec); c.gexec = nullptr; }
  if (c.graph) { (void)hipGraphDestroy(c.graph); c.graph = nullptr; }
  if (c.gexec_n) { (void)hipGraphExecDestroy(c.gexec_n); c.gexec_n = nullptr; }
  if (c.graph_n) { (void)hipGraphDestroy(c.graph_n); c.graph_n = nullptr; }
}

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

// setcoeff (src/LU/setcoeff.jl:43-70) for nx0 x ny0 x nz0 zones, ssor's
// relaxation factor (src/LU/ssor.jl:57) and jacld's scalars
// (src/LU/jacld.jl:22-24).
LuK make_lu_constants(double dt, double omega, int nx0, int ny0, int nz0) {
  LuK k{};
  k.dt = dt;
  k.omega = omega;
  k.tmp = 1.0 / (omega * (2.0 - omega));
  k.dxi = 1.0 / (nx0 - 1);
  k.deta = 1.0 / (ny0 - 1);
  k.dzeta = 1.0 / (nz0 - 1);
  k.tx1 = 1.0 / (k.dxi * k.dxi);
  k.tx2 = 1.0 / (2.0 * k.dxi);
  k.tx3 = 1.0 / k.dxi;
  k.ty1 = 1.0 / (k.deta * k.deta);
  k.ty2 = 1.0 / (2.0 * k.deta);
  k.ty3 = 1.0 / k.deta;
  k.tz1 = 1.0 / (k.dzeta * k.dzeta);
  k.tz2 = 1.0 / (2.0 * k.dzeta);
  k.tz3 = 1.0 / k.dzeta;
  const double c1 = 1.40, c3 = 1.00e-01, c4 = 1.00, c5 = 1.40;
  k.r43 = (4.0 / 3.0);
  k.c1345 = c1 * c3 * c4 * c5;
  k.c34 = c3 * c4;
  const double c2 = 0.40, dd[3] = {0.75, 0.75, 1.00};  // dx1, dy1, dz1
  const double t1[3] = {k.tx1, k.ty1, k.tz1}, t2[3] = {k.tx2, k.ty2, k.tz2};
  for (int d = 0; d < 3; ++d) {
    LuK::Nb& n = k.nb[d];
    const double a = dt * t2[d], v = dt * t1[d];
    n.a = a;
    n.a_c2 = a * c2;
    n.a_2mc2 = a * (2.0 - c2);
    n.a_c1 = a * c1;
    n.a_c2h = a * c2 * 0.5;
    n.v_dd = v * dd[d];
    n.v_r43c34 = v * k.r43 * k.c34;
    n.v_c34 = v * k.c34;
    n.v_c1345 = v * k.c1345;
    n.v_ra = v * (k.r43 * k.c34 - k.c1345);
    n.v_rb = v * (k.c34 - k.c1345);
  }
  k.d0 = 1.0 + 2.0 * dt * (t1[0] * dd[0] + t1[1] * dd[1] + t1[2] * dd[2]);
  for (int i = 0; i < 3; ++i) {
    double gs = 0.0, hs = 0.0;
    for (int d = 0; d < 3; ++d) {
      gs += t1[d] * (d == i ? k.r43 : 1.0);
      hs += t1[d] * (d == i ? k.r43 * k.c34 - k.c1345 : k.c34 - k.c1345);
    }
    k.g[i] = 2.0 * dt * k.c34 * gs;
    k.h[i] = 2.0 * dt * hs;
  }
  k.h5 = 2.0 * dt * k.c1345 * (t1[0] + t1[1] + t1[2]);
  return k;
}

// The zones by decreasing cost nx*ny*nz, exactly as map_zones orders them
// before balancing them over processes (src/BT/setup_mpi.jl:123-157): a
// selection sort that swaps the first strictly largest remaining zone to
// the front.
std::vector<int> cost_order(const npbmz_ctx& c) {
  const int n = c.nz_;
  std::vector<int> z(n);
  std::iota(z.begin(), z.end(), 0);
  auto size = [&](int q) { return double(c.nx[q]) * c.ny[q] * c.nz[q]; };
  for (int i = 0; i + 1 < n; ++i) {
    int m = i;
    double cur = size(z[i]);
    for (int j = i + 1; j < n; ++j)
      if (cur < size(z[j])) {
        cur = size(z[j]);
        m = j;
      }
    if (m != i) std::swap(z[i], z[m]);
  }
  return z;
}

int lines_of(const npbmz_ctx* c, int z, int dir) {
  return dir == 0 ? (c->ny[z] - 2) * (c->nz[z] - 2)
                  : dir == 1 ? (c->nx[z] - 2) * (c->nz[z] - 2) : (c->nx[z] - 2) * (c->ny[z] - 2);
}

// Build a task table: zones in `order`, each zone's `count(z)` items cut into
// tasks of `per` items.  The new device table is allocated before the old
// one is freed, and t changes only on success, so a failure leaves the
// previous table in place and usable.
template <class F>
int build_table(npbmz_ctx* ctx, Table& t, const std::vector<int>& order, int per, F count) {
  std::vector<Task> h;
  int nid = 0;
  for (int z : order) nid = std::max(nid, z + 1);
  std::vector<int> zfirst(nid, 0), zcount(nid, 0);
  for (int z : order) {
    int n = count(z);
    int nt = (n + per - 1) / per;
    zfirst[z] = (int)h.size();
    zcount[z] = nt;
    for (int q = 0; q < nt; ++q) h.push_back(Task{z, q * per});
  }
  Task* d = nullptr;
  if (!h.empty()) {
    CK(hipMalloc(&d, sizeof(Task) * h.size()));
    hipError_t e = hipMemcpy(d, h.data(), sizeof(Task) * h.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      (void)hipFree(d);
      return fail(ctx, NPBMZ_EHIP, std::string("task table upload: ") + hipGetErrorString(e));
    }
  }
  if (t.d) (void)hipFree(t.d);
  t.d = d;
  t.n = (int)h.size();
  t.zfirst.swap(zfirst);
  t.zcount.swap(zcount);
  return NPBMZ_OK;
}

// (Re)build every task table and the buffers sized by them.  Captured graphs
// hold the old tables' addresses, so they are dropped first, whatever the
// outcome.  Each buffer is replaced only once its successor is allocated: a
// failure (EHIP/ENOMEM) leaves a consistent, usable context.
int build_tables(npbmz_ctx* ctx) {
  int rc;
  auto& c = *ctx;
  drop_graphs(c);
  if ((rc = build_table(ctx, c.t_int, c.order, kPtNT, [&](int z) {
         return (c.nx[z] - 2) * (c.ny[z] - 2) * (c.nz[z] - 2);
       })))
    return rc;
  if ((rc = build_table(ctx, c.t_all, c.order, kPtNT,
                        [&](int z) { return c.nx[z] * c.ny[z] * c.nz[z]; })))
    return rc;
  size_t slot = 0;
  for (int d = 0; d < 3 && c.bench != NPBMZ_LU; ++d) {
    // Line tasks: with the largest-first order, zones with the longest lines
    // first (then the most lines), so that when a direction has more tasks
    // than the GPU holds at once the slowest ones start first.
    std::vector<int> ord = c.order;
    if (c.largest_first) {
      auto len = [&](int z) { return d == 0 ? c.nx[z] : d == 1 ? c.ny[z] : c.nz[z]; };
      std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) {
        return len(a) != len(b) ? len(a) > len(b) : lines_of(ctx, a, d) > lines_of(ctx, b, d);
      });
    }
    if ((rc = build_table(ctx, c.t_line[d], ord, kLineNT,
                          [&](int z) { return lines_of(ctx, z, d); })))
      return rc;
    int mr = 0;
    for (int z = 0; z < c.nz_; ++z) {
      int nl = d == 0 ? c.nx[z] : d == 1 ? c.ny[z] : c.nz[z];
      mr = std::max(mr, nl - 2);
    }
    c.maxrows[d] = mr;
    // one scratch slice per workgroup: a launch never has more workgroups
    // than tasks (the persistent grid is min(tasks, solve_grid))
    const int slots = c.t_line[d].n;
    // reduced coefficients kept per row for the back sweep: BT's 5x5 C'
    // block and reduced rhs (fast::bt::kRowPairs double pairs), SP's
    // lhs(4:5) of three systems
    const int per_row = c.bench == NPBMZ_SP ? 6 : 2 * fast::bt::kRowPairs;
    slot = std::max(slot, (size_t)slots * mr * per_row * 64);
    if (c.bench == NPBMZ_BT) {
      // the rows kernel: its own tasks (same zone and line order), and its
      // spill of reduced rows beyond the last kRowR of each half, 6 doubles
      // per lane and row for both waves of a task
      if ((rc = build_table(ctx, c.t_row[d], ord, fast::bt::kRowLines,
                            [&](int z) { return lines_of(ctx, z, d); })))
        return rc;
      c.halfmax[d] = (mr + 1) / 2;
      const size_t spill_rows = (size_t)std::max(0, c.halfmax[d] - kRowR);
      slot = std::max(slot, (size_t)c.t_row[d].n * spill_rows * 2 * 6 * 64);
    }
  }
  if (slot > c.cscr_slot) {
    double* n = nullptr;
    hipError_t e = hipMalloc(&n, slot * sizeof(double));
    if (e != hipSuccess)
      return fail(ctx, e == hipErrorOutOfMemory ? NPBMZ_ENOMEM : NPBMZ_EHIP,
                  std::string("line scratch: ") + hipGetErrorString(e));
    if (c.cscr) (void)hipFree(c.cscr);
    c.cscr = n;
    c.cscr_slot = slot;
  }
  // Faces.  BT and SP copy the interior of each face, j (or i) in 1..n-2 on
  // planes k = 1..nz-2, all in one launch.
  std::vector<Face> fh;
  std::vector<int> fa, fb;  // faces of the first and the second launch
  for (int z = 0; z < c.nz_; ++z) {
    const int nx = c.nx[z], ny = c.ny[z], nz = c.nz[z];
    if (c.bench == NPBMZ_LU) {
      // LU (src/LU/exchange_qbc.jl:69-288) copies i faces over j = 0..ny-1
      // and j faces over i = 0..nx-1, on planes k = 0..nz-3 as written or
      // 1..nz-2 with NPBMZ_FLAG_LU_NPB_EXCH, and unpacks the j faces last, so
      // they own the corners; every source plane is read before any face is
      // written.  Launch 1 copies the j faces (their sources i = 0, nx-1 at
      // j = 1, ny-2 are written by i faces); launch 2 copies the i faces
      // without the corner rows (their sources j = 1..ny-2 are untouched by
      // launch 1, and the corner rows would be overwritten anyway).
      const int k0 = (c.flags & NPBMZ_FLAG_LU_NPB_EXCH) ? 1 : 0;
      fa.push_back((int)fh.size());
      fh.push_back(Face{z, c.south[z], 2, nx * (nz - 2), 0, nx, k0, 0});
      fa.push_back((int)fh.size());
      fh.push_back(Face{z, c.north[z], 3, nx * (nz - 2), 0, nx, k0, 0});
      fb.push_back((int)fh.size());
      fh.push_back(Face{z, c.west[z], 0, (ny - 2) * (nz - 2), 1, ny - 2, k0, 0});
      fb.push_back((int)fh.size());
      fh.push_back(Face{z, c.east[z], 1, (ny - 2) * (nz - 2), 1, ny - 2, k0, 0});
      continue;
    }
    const int nxf = (ny - 2) * (nz - 2), nyf = (nx - 2) * (nz - 2);
    // SP's exchange fills j = 0 from the north neighbour and j = ny-1 from
    // the south one (src/SP/exchange_qbc.jl:247-295, 409-455), the mirror of
    // BT's (src/BT/exchange_qbc.jl:247-295, 409-455)
    const bool yswap = c.bench == NPBMZ_SP;
    for (int q = 0; q < 4; ++q) fa.push_back((int)fh.size() + q);
    fh.push_back(Face{z, c.west[z], 0, nxf, 1, ny - 2, 1, 0});
    fh.push_back(Face{z, c.east[z], 1, nxf, 1, ny - 2, 1, 0});
    fh.push_back(Face{z, yswap ? c.north[z] : c.south[z], 2, nyf, 1, nx - 2, 1, 0});
    fh.push_back(Face{z, yswap ? c.south[z] : c.north[z], 3, nyf, 1, nx - 2, 1, 0});
  }
  if (!c.faces) CK(hipMalloc(&c.faces, sizeof(Face) * fh.size()));
  CK(hipMemcpy(c.faces, fh.data(), sizeof(Face) * fh.size(), hipMemcpyHostToDevice));
  // BT/SP: the face table by destination side, for compute_rhs with the
  // exchange in its launch (exch_copy's source of each halo cell)
  if (c.bench != NPBMZ_LU && !c.hsrc) {
    std::vector<HaloSrc> hs((size_t)4 * c.nz_, HaloSrc{-1, 0, 0});
    bool ok = true;
    for (const Face& f : fh) {
      const int D = f.dst, S = f.src;
      const ZoneDev& zs = c.zh[S];
      const bool xf = f.type < 2;
      const int tn = xf ? c.ny[D] - 2 : c.nx[D] - 2;
      ok = ok && hs[4 * D + f.type].base < 0 && f.t0 == 1 && f.tn == tn && f.k0 == 1 &&
           f.npts == tn * (c.nz[D] - 2) && c.nz[S] == c.nz[D] &&
           (xf ? c.ny[S] == c.ny[D] : c.nx[S] == c.nx[D]);
      if (!ok) break;
      const int64_t sk = (int64_t)zs.nx * zs.ny;
      hs[4 * D + f.type] = f.type == 0   ? HaloSrc{zs.base + zs.nx - 2, zs.nx, sk}
                           : f.type == 1 ? HaloSrc{zs.base + 1, zs.nx, sk}
                           : f.type == 2 ? HaloSrc{zs.base + (int64_t)(zs.ny - 2) * zs.nx, 1, sk}
                                         : HaloSrc{zs.base + zs.nx, 1, sk};
    }
    HaloSrc* d = nullptr;
    CK(hipMalloc(&d, sizeof(HaloSrc) * hs.size()));
    CK(hipMemcpy(d, hs.data(), sizeof(HaloSrc) * hs.size(), hipMemcpyHostToDevice));
    c.hsrc = d;
    c.hsrc_ok = ok;
  }
  if ((rc = build_table(ctx, c.t_face, fa, kExchNT, [&](int f) { return fh[f].npts; })))
    return rc;
  if ((rc = build_table(ctx, c.t_face2, fb, kExchNT, [&](int f) { return fh[f].npts; })))
    return rc;
  // the norms' per-block partial sums; the point tables' sizes do not depend
  // on the order, so this is allocated once
  const size_t pmax = (size_t)std::max(std::max(c.t_int.n, c.t_all.n), 1);
  if (pmax > c.partial_n) {
    double* n = nullptr;
    CK(hipMalloc(&n, sizeof(double) * 10 * pmax));
    if (c.partial) (void)hipFree(c.partial);
    c.partial = n;
    c.partial_n = pmax;
  }
  if (!c.zpart) CK(hipMalloc(&c.zpart, sizeof(double) * (size_t)c.nz_));
  // npbmz_norms_noalloc's block ranges: first block and count per zone of
  // the error norm's table, then of the residual norm's (LU: one table)
  const Table& ta = c.bench == NPBMZ_LU ? c.t_int : c.t_all;
  c.noalloc_rng.assign((size_t)4 * c.nz_, 0);
  for (int z = 0; z < c.nz_; ++z) {
    c.noalloc_rng[z] = ta.zfirst[z];
    c.noalloc_rng[c.nz_ + z] = ta.zcount[z];
    c.noalloc_rng[2 * c.nz_ + z] = c.t_int.zfirst[z];
    c.noalloc_rng[3 * c.nz_ + z] = c.t_int.zcount[z];
  }
  return NPBMZ_OK;
}

bool strict(const npbmz_ctx* c) { return (c->flags & NPBMZ_FLAG_STRICT_FP) != 0; }

// ------------------------------------------------------------ launches
// Every launcher covers either a whole task table (zone < 0, the batched
// form) or one zone's range of it (the per-zone form).

// compute_rhs.  For SP the timestep's rhs phase is compute_rhs + txinvr;
// plain = true gives compute_rhs alone (SP's verify, src/SP/verify.jl:436).
// u, frc, rhs default to the context's fields.
// With_exch (all zones, the context's fields, rhs_exch_fusable): the step's
// zone-face exchange runs in the same launch.
int k_rhs(npbmz_ctx* ctx, int zone, bool plain = false, double* r0out = nullptr,
          const double* u = nullptr, const double* frc = nullptr, double* rhs = nullptr,
          bool with_exch = false) {
  auto& c = *ctx;
  const Table& t = c.t_int;
  int first = zone < 0 ? 0 : t.zfirst[zone], n = zone < 0 ? t.n : t.zcount[zone];
  if (n <= 0) return NPBMZ_OK;
  const bool tx = c.bench == NPBMZ_SP && !plain;
  if (with_exch) {
    auto gx = [&](auto kern) {
      hipLaunchKernelGGL(kern, dim3(t.n + c.t_face.n), dim3(kPtNT), 0, c.stream, c.u, c.frc,
                         c.rhs, c.zd, t.d, t.n, c.K, tx ? r0out : nullptr, c.hsrc, c.faces,
                         c.t_face.d);
    };
    if (strict(ctx)) {
      if (tx) gx(exact::bt::rhs_exch_kernel<kPtNT, true>);
      else gx(exact::bt::rhs_exch_kernel<kPtNT, false>);
    } else {
      if (tx) gx(fast::bt::rhs_exch_kernel<kPtNT, true>);
      else gx(fast::bt::rhs_exch_kernel<kPtNT, false>);
    }
    CK(hipGetLastError());
    return NPBMZ_OK;
  }
  if (!u) u = c.u;
  if (!frc) frc = c.frc;
  if (!rhs) rhs = c.rhs;
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3(n), dim3(kPtNT), 0, c.stream, u, frc, rhs, c.zd,
                       t.d + first, c.K, tx ? r0out : nullptr);
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

// xraw (SP only) also receives each line's solution before its transform.
// Whether the rows kernel runs its long-ring form in direction d: where the
// direction's 12-line tasks (two waves each) exceed one round of two waves
// per SIMD, it is bound by throughput and spill traffic rather than by one
// wave's latency, and where its half lines are longer than the long ring
// (shorter ones spill little either way, and the short form's reads are
// issued together).
bool rows_long(const npbmz_ctx& c, int d) {
  return c.t_row[d].n > 4 * c.cus && c.halfmax[d] > kRowRLong;
}

template <int DIR, bool FUSE>
int k_solve_t(npbmz_ctx* ctx, int zone, double* xraw = nullptr) {
  auto& c = *ctx;
  const Table& t = c.t_line[DIR];
  int first = zone < 0 ? 0 : t.zfirst[zone], n = zone < 0 ? t.n : t.zcount[zone];
  if (n <= 0) return NPBMZ_OK;
  int grid = n;
  if (zone < 0 && c.solve_grid[DIR] > 0) grid = std::min(n, c.solve_grid[DIR]);
  auto go = [&](auto kern, int nt) {
    hipLaunchKernelGGL(kern, dim3(grid), dim3(nt), 0, c.stream, c.u, c.rhs, c.cscr, c.zd,
                       t.d + first, n, c.maxrows[DIR], c.K);
  };
  auto go_sp = [&](auto kern, int nt) {
    hipLaunchKernelGGL(kern, dim3(grid), dim3(nt), 0, c.stream, c.u, c.rhs, c.cscr, c.zd,
                       t.d + first, n, c.maxrows[DIR], c.K, xraw);
  };
  if (c.bench == NPBMZ_SP) {
    if (strict(ctx)) go_sp(exact::sp::solve_kernel<DIR, FUSE>, kLineNT);
    else if (((c.solve_twist >> DIR) & 1) && c.sp_twist_ok[DIR])
      go_sp(fast::sp::solve_twist_kernel<DIR, FUSE>, 2 * kLineNT);
    else go_sp(fast::sp::solve_kernel<DIR, FUSE>, kLineNT);
  } else if (!strict(ctx) && ((c.solve_rows >> DIR) & 1)) {
    const Table& tr = c.t_row[DIR];
    const int f0 = zone < 0 ? 0 : tr.zfirst[zone], nr = zone < 0 ? tr.n : tr.zcount[zone];
    if (nr <= 0) return NPBMZ_OK;
    int gr = nr;
    if (zone < 0 && c.solve_grid[DIR] > 0) gr = std::min(nr, c.solve_grid[DIR]);
    // the same form for every launch mode (the forms round differently)
    if (rows_long(c, DIR))
      hipLaunchKernelGGL((fast::bt::solve_rows_kernel<DIR, FUSE, kRowRLong, true>), dim3(gr),
                         dim3(128), 0, c.stream, c.u, c.rhs, c.cscr, c.zd, tr.d + f0, nr,
                         c.halfmax[DIR], c.K);
    else
      hipLaunchKernelGGL((fast::bt::solve_rows_kernel<DIR, FUSE, kRowR, false>), dim3(gr),
                         dim3(128), 0, c.stream, c.u, c.rhs, c.cscr, c.zd, tr.d + f0, nr,
                         c.halfmax[DIR], c.K);
  } else if (!strict(ctx) && ((c.solve_twist >> DIR) & 1)) {
    go(fast::bt::solve_twist_kernel<DIR, FUSE>, 2 * kLineNT);
  } else if ((c.solve_pair >> DIR) & 1) {
    if (strict(ctx)) go(exact::bt::solve_pair_kernel<DIR, FUSE>, 2 * kLineNT);
    else go(fast::bt::solve_pair_kernel<DIR, FUSE>, 2 * kLineNT);
  } else {
    if (strict(ctx)) go(exact::bt::solve_kernel<DIR, FUSE>, kLineNT);
    else go(fast::bt::solve_kernel<DIR, FUSE>, kLineNT);
  }
  CK(hipGetLastError());
  return NPBMZ_OK;
}

// x/y/z_solve (BT block tridiagonal, SP pentadiagonal with its inverse
// transforms); fuse folds add into the z sweep.
int k_solve(npbmz_ctx* ctx, int dir, bool fuse, int zone, double* xraw = nullptr) {
  if (dir == 0) return k_solve_t<0, false>(ctx, zone, xraw);
  if (dir == 1) return k_solve_t<1, false>(ctx, zone, xraw);
  return fuse ? k_solve_t<2, true>(ctx, zone, xraw) : k_solve_t<2, false>(ctx, zone, xraw);
}

int exch(npbmz_ctx* ctx) {
  auto& c = *ctx;
  // bt-node.jl:286 and sp-node.jl:258 exchange only with more than one zone;
  // lu-node.jl:344 always does
  if (c.nz_ <= 1 && c.bench != NPBMZ_LU) return NPBMZ_OK;
  const Table* ts[2] = {&c.t_face, &c.t_face2};
  for (const Table* t : ts) {
    if (t->n <= 0) continue;
    hipLaunchKernelGGL(exch_kernel<kExchNT>, dim3(t->n), dim3(kExchNT), 0, c.stream, c.u, c.zd,
                       c.faces, t->d);
    CK(hipGetLastError());
  }
  return NPBMZ_OK;
}

// ------------------------------------------------------------------- LU
// rhs (erhs = false) or erhs (erhs = true, src = the staged polynomial in
// rsd, out = frct) over every point of one zone (zone >= 0) or all zones.
int k_lu_rhs(npbmz_ctx* ctx, int zone, bool erhs = false) {
  auto& c = *ctx;
  const Table& t = c.t_all;
  int first = zone < 0 ? 0 : t.zfirst[zone], n = zone < 0 ? t.n : t.zcount[zone];
  if (n <= 0) return NPBMZ_OK;
  auto go = [&](auto kern, const double* src, double* out) {
    hipLaunchKernelGGL(kern, dim3(n), dim3(kPtNT), 0, c.stream, src, c.frc, out, c.zd, t.d + first,
                       c.lk);
  };
  if (erhs) go(exact::lu::rhs_kernel<kPtNT, true>, c.rhs, c.frc);  // setup: always exact
  else if (strict(ctx)) go(exact::lu::rhs_kernel<kPtNT, false>, c.u, c.rhs);
  else go(fast::lu::rhs_kernel<kPtNT, false>, c.u, c.rhs);
  CK(hipGetLastError());
  return NPBMZ_OK;
}

int k_lu_update(npbmz_ctx* ctx, int zone) {
  auto& c = *ctx;
  const Table& t = c.t_int;
  int first = zone < 0 ? 0 : t.zfirst[zone], n = zone < 0 ? t.n : t.zcount[zone];
  if (n <= 0) return NPBMZ_OK;
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3(n), dim3(kPtNT), 0, c.stream, c.u, c.rhs, c.zd, t.d + first, c.lk);
  };
  if (strict(ctx)) go(exact::lu::update_kernel<kPtNT>);
  else go(fast::lu::update_kernel<kPtNT>);
  CK(hipGetLastError());
  return NPBMZ_OK;
}

// The zone sweeps' hyperplane forwarding through LDS: the requested level
// (2 rsd and u, 1 rsd only, 0 none) lowered until the two hyperplanes'
// buffers fit the workgroup's LDS; and the dynamic LDS it takes.
size_t lu_fwd_lds(const npbmz_ctx* c, int lvl) {
  const size_t one = sizeof(double) * 2 * 5 * (size_t)(c->ny[0] - 2) * (c->nz[0] - 2);
  return lvl == 2 ? 2 * one : lvl == 1 ? one : 0;
}

int lu_fwd_level(const npbmz_ctx* c) {
  int lvl = std::min(c->lu_fwd, 2);
  while (lvl > 0 && lu_fwd_lds(c, lvl) > c->max_lds) --lvl;
  return lvl;
}

// Level 3, the padded-halo sweep with loads a hyperplane ahead: FMA
// arithmetic, one padded (j, k) slot per thread, its LDS within the device's.
size_t lu_halo_lds(const npbmz_ctx* c) {
  return sizeof(double) * 2 * 10 * (size_t)(c->ny[0] - 1) * (c->nz[0] - 1);
}

bool lu_halo_ok(const npbmz_ctx* c) {
  return c->lu_fwd == 3 && !(c->flags & NPBMZ_FLAG_STRICT_FP) &&
         (c->ny[0] - 1) * (c->nz[0] - 1) <= c->lu_nt && lu_halo_lds(c) <= c->max_lds;
}

// The pipelined tile sweep's shape: tiles per zone (none empty), slots per
// tile, threads per workgroup (0 when a tile's slots exceed 512), hyperplane
// steps and dynamic LDS.
struct TileGeom {
  int W, NS, nt, steps, SJ;
  size_t lds;
};
TileGeom tile_geom(const npbmz_ctx* c) {
  TileGeom g{};
  const int NI = c->nx[0] - 2, NJ = c->ny[0] - 2, NK = c->nz[0] - 2;
  const int NR = NK + 1;
  g.SJ = NJ + 1;
  int W = std::max(1, std::min(c->lu_tiles, NR));
  const int rpt = (NR + W - 1) / W;
  g.W = (NR + rpt - 1) / rpt;
  g.NS = rpt * g.SJ;
  g.nt = g.NS <= 128 ? 128 : g.NS <= 256 ? 256 : g.NS <= 512 ? 512 : 0;
  g.steps = (NI + NJ + NK) & ~1;  // the halo hyperplane and ns interior ones, made even
  g.lds = sizeof(double) * (2 * 10 * (size_t)g.NS + 1);  // the planes and one int
  return g;
}

// The tile sweep applies with FMA arithmetic when a tile's slots fit a
// workgroup and its LDS the device's.
bool lu_tiles_ok(const npbmz_ctx* c) {
  const TileGeom g = tile_geom(c);
  return !(c->flags & NPBMZ_FLAG_STRICT_FP) && g.nt > 0 && g.lds <= c->max_lds;
}

// Whether the device places workgroups of equal b mod 8 on one XCD, as the
// pipelined sweep's XCD-local hand-off needs: 1024 one-wave workgroups
// report their XCC_ID register.  Measured once per device and process; the
// sweep itself checks every pair of tiles it hands rows between.
bool xcd_round_robin() {
  static std::mutex mu;
  static int state[64];        // 0 unknown, 1 round robin, 2 not
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return false;
  std::lock_guard<std::mutex> lock(mu);
  if (state[dev] == 0) {
    constexpr int nb = 1024;
    int* d = nullptr;
    std::vector<int> h(nb, -1);
    bool ok = hipMalloc(&d, nb * sizeof(int)) == hipSuccess;
    if (ok) {
      hipLaunchKernelGGL(xcc_probe_kernel, dim3(nb), dim3(64), 0, 0, d);
      ok = hipGetLastError() == hipSuccess &&
           hipMemcpy(h.data(), d, nb * sizeof(int), hipMemcpyDeviceToHost) == hipSuccess;
    }
    if (d) (void)hipFree(d);
    for (int b = 0; ok && b < nb; ++b) ok = h[b] >= 0 && h[b] < 16 && h[b] == h[b & 7];
    state[dev] = ok ? 1 : 2;
  }
  return state[dev] == 1;
}

// Exchange rows and counters of the tile sweep, allocated once per context
// for up to kMaxTiles tiles per zone, zeroed (each launch leaves them zero),
// and the device's workgroup placement for the XCD-local hand-off.
int ensure_tiles(npbmz_ctx* ctx) {
  auto& c = *ctx;
  if (c.txbuf) return NPBMZ_OK;
  c.tile_local_ok = xcd_round_robin();
  const TileGeom g = tile_geom(&c);
  const size_t nx = (size_t)c.nz_ * kMaxTiles * g.steps * g.SJ * 10;
  const size_t nc = (size_t)c.nz_ * kMaxTiles * 2 + c.nz_ + 1;
  double* xb = nullptr;
  int* cn = nullptr;
  hipError_t e = hipMalloc(&xb, nx * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&cn, nc * sizeof(int));
  if (e == hipSuccess) e = hipMemset(cn, 0, nc * sizeof(int));
  if (e != hipSuccess) {
    if (xb) (void)hipFree(xb);
    if (cn) (void)hipFree(cn);
    return fail(ctx, e == hipErrorOutOfMemory ? NPBMZ_ENOMEM : NPBMZ_EHIP,
                std::string("pipelined sweep buffers: ") + hipGetErrorString(e));
  }
  c.txbuf = xb;
  c.tcnt = cn;
  return NPBMZ_OK;
}

// After work that may have run tile sweeps: a wait that hit its bound shows
// as the error flag, which is reported and cleared.
int check_tiles(npbmz_ctx* ctx) {
  auto& c = *ctx;
  if (!c.tcnt) return NPBMZ_OK;
  int* ef = c.tcnt + (size_t)c.nz_ * kMaxTiles + c.nz_;
  int v = 0;
  CK(hipMemcpyAsync(&v, ef, sizeof v, hipMemcpyDeviceToHost, c.stream));
  CK(hipStreamSynchronize(c.stream));
  if (v == 0) return NPBMZ_OK;
  const int zero = 0;
  CK(hipMemcpyAsync(ef, &zero, sizeof zero, hipMemcpyHostToDevice, c.stream));
  CK(hipStreamSynchronize(c.stream));
  return fail(ctx, NPBMZ_EHIP,
              (v & 2) ? "LU pipelined sweep: a tile ran off its zone's XCD (XCD-local hand-off)"
                      : "LU pipelined sweep: a tile's wait for its upstream tile exceeded its bound");
}

// The hand-off the pipelined sweep runs: XCD-local (1; 2 with a map that
// no tile matches) or agent-scope (0).
int tile_handoff(const npbmz_ctx* c) {
  if (c->lu_handoff >= 0) return c->lu_handoff;
  return c->tile_local_ok ? 1 : 0;
}

// One SSOR sweep (lower: rsd *= dt, jacld, blts; upper: jacu, buts) of one
// zone (zone >= 0) or all zones, in the context's schedule.
// upd (upper sweep, pipelined tiles or padded-halo zone sweep): the sweep
// also takes ssor's update and stores u instead of rsd; *upd tells whether
// it did (the other schedules leave the update to k_lu_update).
template <bool UPPER>
int k_lu_sweep_t(npbmz_ctx* ctx, int zone, bool* upd = nullptr) {
  auto& c = *ctx;
  const int z0 = zone < 0 ? 0 : zone, nzl = zone < 0 ? c.nz_ : 1;
  const int NI = c.nx[0] - 2, NJ = c.ny[0] - 2, NK = c.nz[0] - 2;
  const int ns = NI + NJ + NK - 2;
  if (c.lu_sweep == kSweepTiles && lu_tiles_ok(&c)) {
    // the exchange buffers come with the schedule (npbmz_set_schedule,
    // npbmz_create): nothing is allocated here, where a graph may be capturing
    if (!c.txbuf) return fail(ctx, NPBMZ_EHIP, "pipelined sweep buffers missing");
    const TileGeom g = tile_geom(&c);
    int* const cn = c.tcnt;
    int* const dn = cn + (size_t)c.nz_ * kMaxTiles;
    int* const ef = dn + c.nz_;
    // tile w of zone zl is block w * nzp + zl: the same residue mod 8 for all
    // tiles of a zone
    const int nzp = (nzl + 7) & ~7;
    const int ho = tile_handoff(&c);
    int* const xc = ef + 1;
    auto go = [&](auto kern) {
      hipLaunchKernelGGL(kern, dim3(nzp * g.W), dim3(g.nt), g.lds, c.stream, c.u, c.rhs, c.zd,
                         z0, nzl, nzp, g.W, c.lu_kb, g.steps, c.txbuf, cn, dn, ef, c.lu_spin,
                         xc, ho == 2 ? 1 : 0, c.lk);
    };
    if (UPPER && upd) {
      *upd = true;
      if (ho != 0) {
        switch (g.nt) {
          case 128: go(fast::lu::sweep_tile_kernel<UPPER, 128, true, true>); break;
          case 256: go(fast::lu::sweep_tile_kernel<UPPER, 256, true, true>); break;
          default: go(fast::lu::sweep_tile_kernel<UPPER, 512, true, true>); break;
        }
      } else {
        switch (g.nt) {
          case 128: go(fast::lu::sweep_tile_kernel<UPPER, 128, false, true>); break;
          case 256: go(fast::lu::sweep_tile_kernel<UPPER, 256, false, true>); break;
          default: go(fast::lu::sweep_tile_kernel<UPPER, 512, false, true>); break;
        }
      }
    } else if (ho != 0) {
      switch (g.nt) {
        case 128: go(fast::lu::sweep_tile_kernel<UPPER, 128, true>); break;
        case 256: go(fast::lu::sweep_tile_kernel<UPPER, 256, true>); break;
        default: go(fast::lu::sweep_tile_kernel<UPPER, 512, true>); break;
      }
    } else {
      switch (g.nt) {
        case 128: go(fast::lu::sweep_tile_kernel<UPPER, 128, false>); break;
        case 256: go(fast::lu::sweep_tile_kernel<UPPER, 256, false>); break;
        default: go(fast::lu::sweep_tile_kernel<UPPER, 512, false>); break;
      }
    }
    CK(hipGetLastError());
    return NPBMZ_OK;
  }
  if (c.lu_sweep == kSweepZone) {
    const int lvl = lu_fwd_level(&c);
    const size_t lds = lu_fwd_lds(&c, lvl);
    auto go = [&](auto kern, int nt) {
      hipLaunchKernelGGL(kern, dim3(nzl), dim3(nt), lds, c.stream, c.u, c.rhs, c.zd, z0, c.lk);
    };
    auto pick = [&](auto k0, auto k1, auto k2, int nt) {
      if (lvl == 2) go(k2, nt);
      else if (lvl == 1) go(k1, nt);
      else go(k0, nt);
    };
    const bool s = strict(ctx);
    if (lu_halo_ok(&c)) {
      auto goh = [&](auto kern, int nt) {
        hipLaunchKernelGGL(kern, dim3(nzl), dim3(nt), lu_halo_lds(&c), c.stream, c.u, c.rhs, c.zd,
                           z0, c.lk);
      };
      if (UPPER && upd) {
        *upd = true;
        switch (c.lu_nt) {
          case 128: goh(fast::lu::sweep_zone_halo_kernel<UPPER, 128, true>, 128); break;
          case 512: goh(fast::lu::sweep_zone_halo_kernel<UPPER, 512, true>, 512); break;
          default: goh(fast::lu::sweep_zone_halo_kernel<UPPER, 256, true>, 256); break;
        }
      } else {
        switch (c.lu_nt) {
          case 128: goh(fast::lu::sweep_zone_halo_kernel<UPPER, 128>, 128); break;
          case 512: goh(fast::lu::sweep_zone_halo_kernel<UPPER, 512>, 512); break;
          default: goh(fast::lu::sweep_zone_halo_kernel<UPPER, 256>, 256); break;
        }
      }
      CK(hipGetLastError());
      return NPBMZ_OK;
    }
    switch (c.lu_nt) {
      case 128:
        if (s) pick(exact::lu::sweep_zone_kernel<UPPER, 128, 0>, exact::lu::sweep_zone_kernel<UPPER, 128, 1>, exact::lu::sweep_zone_kernel<UPPER, 128, 2>, 128);
        else pick(fast::lu::sweep_zone_kernel<UPPER, 128, 0>, fast::lu::sweep_zone_kernel<UPPER, 128, 1>, fast::lu::sweep_zone_kernel<UPPER, 128, 2>, 128);
        break;
      case 512:
        if (s) pick(exact::lu::sweep_zone_kernel<UPPER, 512, 0>, exact::lu::sweep_zone_kernel<UPPER, 512, 1>, exact::lu::sweep_zone_kernel<UPPER, 512, 2>, 512);
        else pick(fast::lu::sweep_zone_kernel<UPPER, 512, 0>, fast::lu::sweep_zone_kernel<UPPER, 512, 1>, fast::lu::sweep_zone_kernel<UPPER, 512, 2>, 512);
        break;
      default:
        if (s) pick(exact::lu::sweep_zone_kernel<UPPER, 256, 0>, exact::lu::sweep_zone_kernel<UPPER, 256, 1>, exact::lu::sweep_zone_kernel<UPPER, 256, 2>, 256);
        else pick(fast::lu::sweep_zone_kernel<UPPER, 256, 0>, fast::lu::sweep_zone_kernel<UPPER, 256, 1>, fast::lu::sweep_zone_kernel<UPPER, 256, 2>, 256);
        break;
    }
    CK(hipGetLastError());
    return NPBMZ_OK;
  }
  for (int q = 0; q < ns; ++q) {
    const int s = UPPER ? ns - 1 - q : q;
    const int c_lo = std::max(0, s - (NI - 1) - (NJ - 1)), c_hi = std::min(NK - 1, s);
    const int npair = (c_hi - c_lo + 1) * NJ;
    const dim3 grid((npair + kPlaneNT - 1) / kPlaneNT, nzl);
    if (strict(ctx))
      hipLaunchKernelGGL((exact::lu::sweep_plane_kernel<UPPER, kPlaneNT>), grid, dim3(kPlaneNT), 0,
                         c.stream, c.u, c.rhs, c.zd, z0, s, c.lk);
    else
      hipLaunchKernelGGL((fast::lu::sweep_plane_kernel<UPPER, kPlaneNT>), grid, dim3(kPlaneNT), 0,
                         c.stream, c.u, c.rhs, c.zd, z0, s, c.lk);
    CK(hipGetLastError());
  }
  return NPBMZ_OK;
}

int k_lu_sweep(npbmz_ctx* ctx, bool upper, int zone, bool* upd = nullptr) {
  return upper ? k_lu_sweep_t<true>(ctx, zone, upd) : k_lu_sweep_t<false>(ctx, zone);
}

// One LU iteration (src/LU/lu-node.jl:338-428): exch_qbc, then per zone
// ssor (src/LU/ssor.jl:9-254): lower sweep with the dt scaling folded in,
// upper sweep, update, rhs.
int ssor_step(npbmz_ctx* ctx, int mode) {
  int rc;
  if ((rc = exch(ctx))) return rc;
  // the upper sweep takes the update where its schedule can (upd)
  if (mode == NPBMZ_MODE_PERZONE) {
    for (int z = 0; z < ctx->nz_; ++z) {
      bool upd = false;
      if ((rc = k_lu_sweep(ctx, false, z))) return rc;
      if ((rc = k_lu_sweep(ctx, true, z, &upd))) return rc;
      if (!upd && (rc = k_lu_update(ctx, z))) return rc;
      if ((rc = k_lu_rhs(ctx, z))) return rc;
    }
    return NPBMZ_OK;
  }
  bool upd = false;
  if ((rc = k_lu_sweep(ctx, false, -1))) return rc;
  if ((rc = k_lu_sweep(ctx, true, -1, &upd))) return rc;
  if (!upd && (rc = k_lu_update(ctx, -1))) return rc;
  return k_lu_rhs(ctx, -1);
}

// One BT timestep: exch_qbc, then per zone compute_rhs, x/y/z solve, add.
// Per-zone mode launches each phase zone by zone (the reference's per-zone
// adi call); batched mode launches each phase once over all zones.  The
// per-zone order (all zones' rhs, then all x solves, ...) is equivalent to
// the reference's zone-by-zone adi because zones are independent after the
// exchange.
// Whether a batched BT/SP step runs its exchange inside compute_rhs's
// launch: more than one zone (so there is an exchange), one exchange launch
// (kExchNT-point face tasks, as exch_copy takes) and a face table hsrc
// describes.
bool rhs_exch_fusable(const npbmz_ctx* c) {
  return c->bench != NPBMZ_LU && c->rhs_exch == 1 && c->nz_ > 1 && c->hsrc_ok &&
         c->t_face2.n == 0 && c->t_face.n > 0 && kExchNT == kPtNT;
}

int adi_step(npbmz_ctx* ctx, int mode) {
  int rc;
  if (mode != NPBMZ_MODE_PERZONE && rhs_exch_fusable(ctx)) {
    // per-zone mode keeps them apart: a zone's add would otherwise reach
    // another zone's rhs in the same step
    if ((rc = k_rhs(ctx, -1, false, nullptr, nullptr, nullptr, nullptr, true))) return rc;
    if ((rc = k_solve(ctx, 0, false, -1))) return rc;
    if ((rc = k_solve(ctx, 1, false, -1))) return rc;
    return k_solve(ctx, 2, true, -1);
  }
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

// ------------------------------------------------------------ BT adjoint
// One batched BT step with the z solve and add as separate launches, copying
// the post-exchange u and the outputs of the three solves when the targets
// are given (the forward pass and the segment recomputation of the adjoint
// run the same launches, so they follow the same trajectory).
int bt_step_store(npbmz_ctx* ctx, double* us, double* x1, double* x2, double* x3) {
  auto& c = *ctx;
  const size_t fb = sizeof(double) * 5 * (size_t)c.npoints;
  int rc;
  if ((rc = exch(ctx))) return rc;
  if (us) CK(hipMemcpyAsync(us, c.u, fb, hipMemcpyDeviceToDevice, c.stream));
  if ((rc = k_rhs(ctx, -1))) return rc;
  if ((rc = k_solve(ctx, 0, false, -1))) return rc;
  if (x1) CK(hipMemcpyAsync(x1, c.rhs, fb, hipMemcpyDeviceToDevice, c.stream));
  if ((rc = k_solve(ctx, 1, false, -1))) return rc;
  if (x2) CK(hipMemcpyAsync(x2, c.rhs, fb, hipMemcpyDeviceToDevice, c.stream));
  if ((rc = k_solve(ctx, 2, false, -1))) return rc;
  if (x3) CK(hipMemcpyAsync(x3, c.rhs, fb, hipMemcpyDeviceToDevice, c.stream));
  return k_add(ctx, -1);
}

// The transposed solve of direction DIR (lam = A^-T gin), then the
// derivative of its blocks into ubar.
template <int DIR>
int k_adj_solve(npbmz_ctx* ctx, const double* us, const double* x, const double* gin, double* lam,
                double* ubar) {
  auto& c = *ctx;
  const Table& t = c.t_line[DIR];
  if (t.n <= 0) return NPBMZ_OK;
  int grid = t.n;
  if (c.solve_grid[DIR] > 0) grid = std::min(t.n, c.solve_grid[DIR]);
  if ((c.solve_twist >> DIR) & 1)
    hipLaunchKernelGGL(fast::bt::adj_solve_twist_kernel<DIR>, dim3(grid), dim3(2 * kLineNT), 0,
                       c.stream, us, gin, lam, c.cscr, c.zd, t.d, t.n, c.maxrows[DIR], c.K);
  else
    hipLaunchKernelGGL(fast::bt::adj_solve_kernel<DIR>, dim3(grid), dim3(kLineNT), 0, c.stream,
                       us, gin, lam, c.cscr, c.zd, t.d, t.n, c.maxrows[DIR], c.K);
  CK(hipGetLastError());
  const Table& ti = c.t_int;
  if (ti.n > 0)
    hipLaunchKernelGGL((fast::bt::adj_jac_kernel<DIR, kPtNT>), dim3(ti.n), dim3(kPtNT), 0,
                       c.stream, us, x, lam, ubar, c.zd, ti.d, c.K);
  CK(hipGetLastError());
  return NPBMZ_OK;
}

// Reverse one BT step from its stored forward data (us: u after the
// exchange, x1..x3: the solve outputs): ubar holds dL/du after the step on
// entry and dL/du before it on return; fbar accumulates dL/dforcing.
int bt_adj_step(npbmz_ctx* ctx, const double* us, const double* x1, const double* x2,
                const double* x3, double* ubar, double* lam, double* fbar) {
  auto& c = *ctx;
  int rc;
  if ((rc = k_adj_solve<2>(ctx, us, x3, ubar, lam, ubar))) return rc;
  if ((rc = k_adj_solve<1>(ctx, us, x2, lam, lam, ubar))) return rc;
  if ((rc = k_adj_solve<0>(ctx, us, x1, lam, lam, ubar))) return rc;
  const Table& t = c.t_all;
  hipLaunchKernelGGL(fast::bt::adj_rhs_kernel<kPtNT>, dim3(t.n), dim3(kPtNT), 0, c.stream, us, lam,
                     ubar, fbar, c.zd, t.d, c.K);
  CK(hipGetLastError());
  if (c.nz_ > 1 && c.t_face.n > 0) {
    for (int xf = 1; xf >= 0; --xf) {
      hipLaunchKernelGGL(fast::bt::adj_exch_kernel<kExchNT>, dim3(c.t_face.n), dim3(kExchNT), 0,
                         c.stream, ubar, c.zd, c.faces, c.t_face.d, xf);
      CK(hipGetLastError());
    }
  }
  return NPBMZ_OK;
}

// ------------------------------------------------------------ SP adjoint
// One batched SP step with the production kernels, keeping what the reverse
// pass reads when the targets are given: the post-exchange u, r0 =
// dt*compute_rhs before txinvr, and each direction's line solutions before
// their transforms.
int sp_step_store(npbmz_ctx* ctx, double* us, double* r0, double* x1, double* x2, double* x3) {
  auto& c = *ctx;
  const size_t fb = sizeof(double) * 5 * (size_t)c.npoints;
  int rc;
  if ((rc = exch(ctx))) return rc;
  if (us) CK(hipMemcpyAsync(us, c.u, fb, hipMemcpyDeviceToDevice, c.stream));
  if ((rc = k_rhs(ctx, -1, false, r0))) return rc;
  if ((rc = k_solve(ctx, 0, false, -1, x1))) return rc;
  if ((rc = k_solve(ctx, 1, false, -1, x2))) return rc;
  return k_solve(ctx, 2, true, -1, x3);
}

// The transposed solve of direction DIR: lam = A^-T (T^T gin), tzetar's
// u-term (DIR 2) added to ubar.
template <int DIR>
int k_sp_adj_solve(npbmz_ctx* ctx, const double* us, const double* x, const double* gin,
                   double* lam, double* ubar) {
  auto& c = *ctx;
  const Table& t = c.t_line[DIR];
  if (t.n <= 0) return NPBMZ_OK;
  int grid = t.n;
  if (c.solve_grid[DIR] > 0) grid = std::min(t.n, c.solve_grid[DIR]);
  if (((c.solve_twist >> DIR) & 1) && c.sp_twist_ok[DIR])
    hipLaunchKernelGGL(fast::sp::adj_solve_twist_kernel<DIR>, dim3(grid), dim3(2 * kLineNT), 0,
                       c.stream, us, x, gin, lam, ubar, c.cscr, c.zd, t.d, t.n, c.maxrows[DIR],
                       c.K);
  else
    hipLaunchKernelGGL(fast::sp::adj_solve_kernel<DIR>, dim3(grid), dim3(kLineNT), 0, c.stream,
                       us, x, gin, lam, ubar, c.cscr, c.zd, t.d, t.n, c.maxrows[DIR], c.K);
  CK(hipGetLastError());
  return NPBMZ_OK;
}

// Reverse one SP step from its stored forward data; ubar holds dL/du after
// the step on entry and dL/du before it (before the exchange) on return;
// fbar accumulates dL/dforcing; lz, ly, lx (each direction's transposed
// solution) are workspace.  The solves leave ubar alone apart from tzetar's
// u-term, so the three matrices' derivatives and txinvr's adjoint follow in
// one pass after the last solve.
int sp_adj_step(npbmz_ctx* ctx, const double* us, const double* r0, const double* x1,
                const double* x2, const double* x3, double* ubar, double* lz, double* ly,
                double* lx, double* fbar) {
  auto& c = *ctx;
  int rc;
  if ((rc = k_sp_adj_solve<2>(ctx, us, x3, ubar, lz, ubar))) return rc;
  if ((rc = k_sp_adj_solve<1>(ctx, us, x2, lz, ly, ubar))) return rc;
  if ((rc = k_sp_adj_solve<0>(ctx, us, x1, ly, lx, ubar))) return rc;
  // txinvr's adjoint goes to the step workspace rhs, free during the reverse
  // pass (off the interior it is never read: adj_rhs_kernel reads the
  // interior only)
  const Table& ti = c.t_int;
  if (ti.n > 0)
    hipLaunchKernelGGL(fast::sp::adj_deriv_kernel<kPtNT>, dim3(ti.n), dim3(kPtNT), 0, c.stream,
                       us, x1, x2, x3, lx, ly, lz, r0, c.rhs, ubar, c.zd, ti.d, c.K);
  CK(hipGetLastError());
  const Table& t = c.t_all;
  hipLaunchKernelGGL(fast::bt::adj_rhs_kernel<kPtNT>, dim3(t.n), dim3(kPtNT), 0, c.stream, us,
                     c.rhs, ubar, fbar, c.zd, t.d, c.K);
  CK(hipGetLastError());
  if (c.nz_ > 1 && c.t_face.n > 0) {
    for (int xf = 1; xf >= 0; --xf) {
      hipLaunchKernelGGL(fast::bt::adj_exch_kernel<kExchNT>, dim3(c.t_face.n), dim3(kExchNT), 0,
                         c.stream, ubar, c.zd, c.faces, c.t_face.d, xf);
      CK(hipGetLastError());
    }
  }
  return NPBMZ_OK;
}

// ------------------------------------------------------------ LU adjoint
// One batched LU iteration (exch_qbc, lower and upper sweeps, update, rhs),
// copying u before and after the exchange and the outputs of the two sweeps
// when the targets are given.  Without r2 the upper sweep may take the
// update itself, as in ssor_step.
int lu_step_store(npbmz_ctx* ctx, double* un, double* ue, double* r1, double* r2) {
  auto& c = *ctx;
  const size_t fb = sizeof(double) * 5 * (size_t)c.npoints;
  int rc;
  bool upd = false;
  if (un) CK(hipMemcpyAsync(un, c.u, fb, hipMemcpyDeviceToDevice, c.stream));
  if ((rc = exch(ctx))) return rc;
  if (ue) CK(hipMemcpyAsync(ue, c.u, fb, hipMemcpyDeviceToDevice, c.stream));
  if ((rc = k_lu_sweep(ctx, false, -1))) return rc;
  if (r1) CK(hipMemcpyAsync(r1, c.rhs, fb, hipMemcpyDeviceToDevice, c.stream));
  if ((rc = k_lu_sweep(ctx, true, -1, r2 ? nullptr : &upd))) return rc;
  if (r2) CK(hipMemcpyAsync(r2, c.rhs, fb, hipMemcpyDeviceToDevice, c.stream));
  if (!upd && (rc = k_lu_update(ctx, -1))) return rc;
  return k_lu_rhs(ctx, -1);
}

// Both transposed sweeps in the context's LU sweep schedule: one workgroup
// per zone (of lu_nt threads) walking the hyperplanes, or one launch per
// hyperplane; ascending for the transposed upper sweep, descending for the
// transposed lower one.
template <bool ASC>
int k_lu_adj_sweep(npbmz_ctx* ctx, const double* ue, const double* ubar, double* nu, double* rho) {
  auto& c = *ctx;
  // the pipelined tiles have no transposed form: their contexts run the
  // transposed sweeps one workgroup per zone where a zone's pairs fit one
  const bool tiles_as_zone = c.lu_sweep == kSweepTiles &&
                             (c.ny[0] - 2) * (c.nz[0] - 2) <= c.lu_nt;
  if (c.lu_sweep == kSweepZone || tiles_as_zone) {
    const int np = (c.ny[0] - 2) * (c.nz[0] - 2);
    if (c.lu_fwd == 3 && np <= c.lu_nt) {
      // the forward's level 3: hyperplanes handed on through LDS, own data
      // loaded a hyperplane ahead (2*5*np doubles, within 64 KiB for np <= 512)
      const size_t lds = sizeof(double) * 2 * 5 * (size_t)np;
      auto gof = [&](auto kern, int nt) {
        hipLaunchKernelGGL(kern, dim3(c.nz_), dim3(nt), lds, c.stream, ue, ubar, nu, rho, c.zd,
                           c.lk);
      };
      switch (c.lu_nt) {
        case 128: gof(fast::lu::adj_sweep_zone_fwd_kernel<ASC, 128>, 128); break;
        case 512: gof(fast::lu::adj_sweep_zone_fwd_kernel<ASC, 512>, 512); break;
        default: gof(fast::lu::adj_sweep_zone_fwd_kernel<ASC, 256>, 256); break;
      }
      CK(hipGetLastError());
      return NPBMZ_OK;
    }
    auto go = [&](auto kern, int nt) {
      hipLaunchKernelGGL(kern, dim3(c.nz_), dim3(nt), 0, c.stream, ue, ubar, nu, rho, c.zd, c.lk);
    };
    switch (c.lu_nt) {
      case 128: go(fast::lu::adj_sweep_zone_kernel<ASC, 128>, 128); break;
      case 512: go(fast::lu::adj_sweep_zone_kernel<ASC, 512>, 512); break;
      default: go(fast::lu::adj_sweep_zone_kernel<ASC, 256>, 256); break;
    }
    CK(hipGetLastError());
    return NPBMZ_OK;
  }
  const int NI = c.nx[0] - 2, NJ = c.ny[0] - 2, NK = c.nz[0] - 2;
  const int ns = NI + NJ + NK - 2;
  for (int q = 0; q < ns; ++q) {
    const int s = ASC ? q : ns - 1 - q;
    const int c_lo = std::max(0, s - (NI - 1) - (NJ - 1)), c_hi = std::min(NK - 1, s);
    const int npair = (c_hi - c_lo + 1) * NJ;
    const dim3 grid((npair + kPlaneNT - 1) / kPlaneNT, c.nz_);
    hipLaunchKernelGGL((fast::lu::adj_sweep_plane_kernel<ASC, kPlaneNT>), grid, dim3(kPlaneNT), 0,
                       c.stream, ue, ubar, nu, rho, c.zd, s, c.lk);
    CK(hipGetLastError());
  }
  return NPBMZ_OK;
}

// Reverse one LU iteration from its stored forward data (un, ue: u before
// and after the exchange; r1, r2: the sweeps' outputs): ubar holds dL/du
// after the iteration on entry and dL/du before it on return; fbar
// accumulates dL/dfrct; nu and rho are workspace.
int lu_adj_step(npbmz_ctx* ctx, const double* un, const double* ue, const double* r1,
                const double* r2, double* ubar, double* nu, double* rho, double* fbar) {
  auto& c = *ctx;
  const size_t fb = sizeof(double) * 5 * (size_t)c.npoints;
  CK(hipMemsetAsync(nu, 0, fb, c.stream));
  CK(hipMemsetAsync(rho, 0, fb, c.stream));
  int rc;
  if ((rc = k_lu_adj_sweep<true>(ctx, ue, ubar, nu, rho))) return rc;
  if ((rc = k_lu_adj_sweep<false>(ctx, ue, ubar, nu, rho))) return rc;
  const Table& t = c.t_all;
  hipLaunchKernelGGL(fast::lu::adj_blocks_kernel<kPtNT>, dim3(t.n), dim3(kPtNT), 0, c.stream, ue,
                     r1, r2, nu, rho, ubar, fbar, c.zd, t.d, c.lk);
  CK(hipGetLastError());
  // the exchange's two launches in reverse order
  const Table* ts[2] = {&c.t_face2, &c.t_face};
  for (const Table* tf : ts) {
    if (tf->n <= 0) continue;
    hipLaunchKernelGGL(fast::bt::adj_exch_kernel<kExchNT>, dim3(tf->n), dim3(kExchNT), 0, c.stream,
                       ubar, c.zd, c.faces, tf->d, -1);
    CK(hipGetLastError());
  }
  hipLaunchKernelGGL(fast::lu::adj_rhs_kernel<kPtNT>, dim3(t.n), dim3(kPtNT), 0, c.stream, un, rho,
                     ubar, c.zd, t.d, c.lk);
  CK(hipGetLastError());
  return NPBMZ_OK;
}

// Sum of w*u^2 over every grid point of the current u (wd on the device);
// with seed, also ubar = 2 w u and fbar = 0.
int bt_loss(npbmz_ctx* ctx, const double* wd, double* ubar, double* fbar, bool seed, double* loss) {
  auto& c = *ctx;
  const Table& t = c.t_all;
  std::vector<double> part((size_t)t.n);
  hipLaunchKernelGGL(fast::bt::loss_kernel<kPtNT>, dim3(t.n), dim3(kPtNT), 0, c.stream, c.u, wd,
                     ubar, fbar, c.zd, t.d, c.partial, seed ? 1 : 0);
  CK(hipGetLastError());
  CK(hipMemcpyAsync(part.data(), c.partial, part.size() * sizeof(double), hipMemcpyDeviceToHost,
                    c.stream));
  CK(hipStreamSynchronize(c.stream));
  double s = 0.0;
  for (double v : part) s += v;
  *loss = s;
  return NPBMZ_OK;
}

int step_once(npbmz_ctx* ctx, int mode) {
  return ctx->bench == NPBMZ_LU ? ssor_step(ctx, mode) : adi_step(ctx, mode);
}

int run_phase(npbmz_ctx* ctx, int phase, int mode) {
  int zlo = 0, zhi = ctx->nz_;
  if (mode != NPBMZ_MODE_PERZONE) { zlo = -1; zhi = 0; }
  int rc = NPBMZ_OK;
  if (phase == NPBMZ_PHASE_EXCH) return exch(ctx);
  for (int z = zlo; z < zhi; ++z) {
    if (ctx->bench == NPBMZ_LU) {
      switch (phase) {
        case NPBMZ_PHASE_RHS: rc = k_lu_rhs(ctx, z); break;
        case NPBMZ_PHASE_SSOR_LOWER: rc = k_lu_sweep(ctx, false, z); break;
        case NPBMZ_PHASE_SSOR_UPPER: rc = k_lu_sweep(ctx, true, z); break;
        case NPBMZ_PHASE_SSOR_UPDATE: rc = k_lu_update(ctx, z); break;
        default: return fail(ctx, NPBMZ_EINVAL, "unknown LU phase");
      }
      if (rc) return rc;
      continue;
    }
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

// Captures nsteps batched timesteps into one graph.  Replaying a group of
// steps as one graph saves the per-launch gap between graphs, which on a
// 0.4 ms step is about a percent.
int capture_steps(npbmz_ctx* ctx, int nsteps, hipGraph_t* graph, hipGraphExec_t* exec) {
  auto& c = *ctx;
  CK(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
  int rc = NPBMZ_OK;
  for (int s = 0; s < nsteps && !rc; ++s) rc = step_once(ctx, NPBMZ_MODE_BATCHED);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(c.stream, &g);
  if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
  if (e != hipSuccess) return fail(ctx, NPBMZ_EHIP, std::string("capture: ") + hipGetErrorString(e));
  *graph = g;
  CK(hipGraphInstantiate(exec, g, nullptr, nullptr, 0));
  return NPBMZ_OK;
}

int ensure_graph(npbmz_ctx* ctx, int nsteps) {
  auto& c = *ctx;
  int rc;
  if (!c.gexec && (rc = capture_steps(ctx, 1, &c.graph, &c.gexec))) return rc;
  if (nsteps >= kGraphGroup && !c.gexec_n &&
      (rc = capture_steps(ctx, kGraphGroup, &c.graph_n, &c.gexec_n)))
    return rc;
  return NPBMZ_OK;
}

// This device's address of the static scratch, looked up once per device
// (with that device current, as in npbmz_create) and kept for every later
// context on it.  A failed lookup is kept as a message: only
// npbmz_norms_noalloc needs the scratch, and it reports the failure.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_dev[kMaxDevices] = {};

void bind_scratch(npbmz_ctx& c, int dev) {
  if (dev < 0 || dev >= kMaxDevices) {
    c.scratch_err = "device ordinal beyond the scratch table";
    return;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      c.scratch_err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
      return;
    }
    g_scratch_dev[dev] = static_cast<double*>(p);
  }
  c.scratch = g_scratch_dev[dev];
}

void describe(npbmz_ctx* c) {
  char b[256];
  if (c->bench == NPBMZ_LU && c->lu_sweep == kSweepTiles && lu_tiles_ok(c)) {
    const TileGeom g = tile_geom(c);
    snprintf(b, sizeof b, "zones=%d sweep=pipelined-tiles tiles=%d batch=%d nt=%d handoff=%s exch=%s fp=%s",
             c->nz_, g.W, c->lu_kb, g.nt, tile_handoff(c) != 0 ? "xcd-local" : "agent",
             (c->flags & NPBMZ_FLAG_LU_NPB_EXCH) ? "npb" : "reference",
             strict(c) ? "strict" : "fma");
  } else if (c->bench == NPBMZ_LU)
    snprintf(b, sizeof b, "zones=%d sweep=%s nt=%d%s exch=%s fp=%s", c->nz_,
             c->lu_sweep == kSweepZone ? "zone-workgroup" : "hyperplane-launch", c->lu_nt,
             c->lu_sweep != kSweepZone ? ""
             : lu_halo_ok(c)           ? " lds-halo+prefetch"
             : lu_fwd_level(c) == 2    ? " lds-forward=rsd+u"
             : lu_fwd_level(c) == 1    ? " lds-forward=rsd"
                                       : "",
             (c->flags & NPBMZ_FLAG_LU_NPB_EXCH) ? "npb" : "reference", strict(c) ? "strict" : "fma");
  else if (c->bench == NPBMZ_SP)
    snprintf(b, sizeof b, "zones=%d order=%s solve_grid=%d/%d/%d twist=%d fp=%s", c->nz_,
             c->largest_first ? "largest-first" : "natural", c->solve_grid[0],
             c->solve_grid[1], c->solve_grid[2],
             strict(c) ? 0
                       : (c->solve_twist & ((c->sp_twist_ok[0] ? 1 : 0) | (c->sp_twist_ok[1] ? 2 : 0) |
                                            (c->sp_twist_ok[2] ? 4 : 0))),
             strict(c) ? "strict" : "fma");
  else
  {
    // the rows kernel's ring of reduced rows per direction (0: not used)
    int ring[3];
    for (int d = 0; d < 3; ++d)
      ring[d] = strict(c) || !((c->solve_rows >> d) & 1) ? 0 : rows_long(*c, d) ? kRowRLong : kRowR;
    snprintf(b, sizeof b,
             "zones=%d order=%s solve_grid=%d/%d/%d pair=%d twist=%d rows=%d ring=%d/%d/%d fp=%s",
             c->nz_, c->largest_first ? "largest-first" : "natural", c->solve_grid[0],
             c->solve_grid[1], c->solve_grid[2], c->solve_pair, strict(c) ? 0 : c->solve_twist,
             strict(c) ? 0 : c->solve_rows, ring[0], ring[1], ring[2], strict(c) ? "strict" : "fma");
  }
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
  if (p->bench != NPBMZ_BT && p->bench != NPBMZ_SP && p->bench != NPBMZ_LU)
    return fail(nullptr, NPBMZ_ENOTSUP, "bench must be NPBMZ_BT, NPBMZ_SP or NPBMZ_LU");
  if (p->nzones <= 0 || p->nzones != p->x_zones * p->y_zones)
    return fail(nullptr, NPBMZ_EINVAL, "nzones must equal x_zones*y_zones and be positive");
  if (!p->nx || !p->ny || !p->nz || !p->iz_west || !p->iz_east || !p->iz_south || !p->iz_north)
    return fail(nullptr, NPBMZ_EINVAL, "null zone table");
  if (!(p->dt > 0.0)) return fail(nullptr, NPBMZ_EINVAL, "dt must be positive");
  if (p->gx_size < 2 || p->gy_size < 2 || p->gz_size < 2)
    return fail(nullptr, NPBMZ_EINVAL, "grid sizes must be at least 2");
  int n = p->nzones;
  if (p->bench == NPBMZ_LU) {
    if (!(p->omega > 0.0 && p->omega < 2.0))
      return fail(nullptr, NPBMZ_EINVAL, "LU needs 0 < omega < 2");
    // src/LU/zone_setup.jl:22-33 splits the grid into equal zones, and the
    // node's setcoeff globals (src/LU/setcoeff.jl:43-70) assume it
    for (int z = 1; z < n; ++z)
      if (p->nx[z] != p->nx[0] || p->ny[z] != p->ny[0] || p->nz[z] != p->nz[0])
        return fail(nullptr, NPBMZ_EINVAL, "LU zones must all have the same size");
  }
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
  if (prop.multiProcessorCount > 0) c.cus = prop.multiProcessorCount;
  bind_scratch(c, dev);
  c.bench = p->bench; c.flags = flags; c.nz_ = n;
  c.x_zones = p->x_zones; c.y_zones = p->y_zones;
  c.gx = p->gx_size; c.gy = p->gy_size; c.gz = p->gz_size;
  c.dt = p->dt; c.omega = p->omega;
  c.nx.assign(p->nx, p->nx + n); c.ny.assign(p->ny, p->ny + n); c.nz.assign(p->nz, p->nz + n);
  c.west.assign(p->iz_west, p->iz_west + n); c.east.assign(p->iz_east, p->iz_east + n);
  c.south.assign(p->iz_south, p->iz_south + n); c.north.assign(p->iz_north, p->iz_north + n);
  c.K = make_constants(c.bench, c.dt, c.gx, c.gy, c.gz, c.x_zones, c.y_zones);
  if (c.bench == NPBMZ_SP) {
    // the two-sided solve needs 4 unknown rows on every line of a direction
    for (int d = 0; d < 3; ++d) {
      int shortest = 1 << 30;
      for (int z = 0; z < n; ++z)
        shortest = std::min(shortest, d == 0 ? c.nx[z] : d == 1 ? c.ny[z] : c.nz[z]);
      c.sp_twist_ok[d] = shortest - 2 >= 4;
    }
  }
  if (c.bench == NPBMZ_LU) {
    c.lk = make_lu_constants(c.dt, c.omega, c.nx[0], c.ny[0], c.nz[0]);
    // default sweep schedule by zone size, as measured on MI355X: one
    // 512-thread workgroup per zone while a zone's interior is small (class
    // A, 30x30x14: 0.54 ms per iteration against 0.58 with 256 threads),
    // hyperplane launches once it is large (class B, 74x50x15: 1.2 ms per
    // iteration against 2.1 with a workgroup per zone)
    const int64_t interior = int64_t(c.nx[0] - 2) * (c.ny[0] - 2) * (c.nz[0] - 2);
    if (interior > 32768) c.lu_sweep = kSweepPlane;
    else c.lu_nt = 512;
    // with FMA arithmetic, zones whose padded hyperplane exceeds 256 slots
    // run as pipelined tiles of at most 256 slots each (below, npbmz_create:
    // class A, 31x15 slots, 2 tiles: 0.300 ms per iteration against 0.327
    // for the one-workgroup padded-halo sweep; class B, 51x16 slots, 4
    // tiles: 0.686 against 1.095 with hyperplane launches; class W, 15x7
    // slots, stays one workgroup per zone: 0.104 against 0.133 with 2 tiles)
    // hyperplane forwarding: with FMA arithmetic the padded-halo sweep in
    // the smallest workgroup that holds a zone's padded hyperplane (class W,
    // 15x7 slots, 128 threads: 0.088 ms per iteration against 0.105 with 512
    // threads from global memory; class S: 0.040 against 0.044; class A, 465
    // slots, 512 threads: 0.351 against 0.533 from global memory); strict
    // arithmetic has no padded-halo form: rsd and u through LDS from 128
    // (j, k) pairs up, else global memory
    const int hslots = (c.ny[0] - 1) * (c.nz[0] - 1);
    if (!strict(ctx) && hslots <= 512) {
      c.lu_fwd = 3;
      c.lu_nt = hslots <= 128 ? 128 : hslots <= 256 ? 256 : 512;
    } else if ((c.ny[0] - 2) * (c.nz[0] - 2) < 128) {
      c.lu_fwd = 0;
    }
  }
  c.zh.resize(n);
  int64_t base = 0;
  for (int z = 0; z < n; ++z) {
    c.zh[z] = ZoneDev{c.nx[z], c.ny[z], c.nz[z], 0, base};
    base += (int64_t)c.nx[z] * c.ny[z] * c.nz[z];
  }
  c.npoints = base;
  // largest zones first, in the reference's cost order: the batched
  // launches then schedule the longest work first
  c.order = cost_order(c);
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
  if (c.bench == NPBMZ_LU) {
    // the zone and tile sweeps' LDS: the device's per-workgroup maximum, granted to
    // every instantiation that forwards hyperplanes (64 KiB if refused)
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess &&
        v > 65536) {
      bool ok = true;
      auto grant = [&](auto kern) {
        ok = ok && hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, v) == hipSuccess;
      };
      grant(exact::lu::sweep_zone_kernel<false, 128, 1>); grant(exact::lu::sweep_zone_kernel<true, 128, 1>);
      grant(exact::lu::sweep_zone_kernel<false, 256, 1>); grant(exact::lu::sweep_zone_kernel<true, 256, 1>);
      grant(exact::lu::sweep_zone_kernel<false, 512, 1>); grant(exact::lu::sweep_zone_kernel<true, 512, 1>);
      grant(exact::lu::sweep_zone_kernel<false, 128, 2>); grant(exact::lu::sweep_zone_kernel<true, 128, 2>);
      grant(exact::lu::sweep_zone_kernel<false, 256, 2>); grant(exact::lu::sweep_zone_kernel<true, 256, 2>);
      grant(exact::lu::sweep_zone_kernel<false, 512, 2>); grant(exact::lu::sweep_zone_kernel<true, 512, 2>);
      grant(fast::lu::sweep_zone_kernel<false, 128, 1>); grant(fast::lu::sweep_zone_kernel<true, 128, 1>);
      grant(fast::lu::sweep_zone_kernel<false, 256, 1>); grant(fast::lu::sweep_zone_kernel<true, 256, 1>);
      grant(fast::lu::sweep_zone_kernel<false, 512, 1>); grant(fast::lu::sweep_zone_kernel<true, 512, 1>);
      grant(fast::lu::sweep_zone_kernel<false, 128, 2>); grant(fast::lu::sweep_zone_kernel<true, 128, 2>);
      grant(fast::lu::sweep_zone_kernel<false, 256, 2>); grant(fast::lu::sweep_zone_kernel<true, 256, 2>);
      grant(fast::lu::sweep_zone_kernel<false, 512, 2>); grant(fast::lu::sweep_zone_kernel<true, 512, 2>);
      grant(fast::lu::sweep_zone_halo_kernel<false, 128>); grant(fast::lu::sweep_zone_halo_kernel<true, 128>);
      grant(fast::lu::sweep_zone_halo_kernel<false, 256>); grant(fast::lu::sweep_zone_halo_kernel<true, 256>);
      grant(fast::lu::sweep_zone_halo_kernel<false, 512>); grant(fast::lu::sweep_zone_halo_kernel<true, 512>);
      grant(fast::lu::sweep_tile_kernel<false, 128, false>); grant(fast::lu::sweep_tile_kernel<true, 128, false>);
      grant(fast::lu::sweep_tile_kernel<false, 256, false>); grant(fast::lu::sweep_tile_kernel<true, 256, false>);
      grant(fast::lu::sweep_tile_kernel<false, 512, false>); grant(fast::lu::sweep_tile_kernel<true, 512, false>);
      grant(fast::lu::sweep_tile_kernel<false, 128, true>); grant(fast::lu::sweep_tile_kernel<true, 128, true>);
      grant(fast::lu::sweep_tile_kernel<false, 256, true>); grant(fast::lu::sweep_tile_kernel<true, 256, true>);
      grant(fast::lu::sweep_tile_kernel<false, 512, true>); grant(fast::lu::sweep_tile_kernel<true, 512, true>);
      grant(fast::lu::sweep_zone_halo_kernel<true, 128, true>);
      grant(fast::lu::sweep_zone_halo_kernel<true, 256, true>);
      grant(fast::lu::sweep_zone_halo_kernel<true, 512, true>);
      grant(fast::lu::sweep_tile_kernel<true, 128, false, true>); grant(fast::lu::sweep_tile_kernel<true, 128, true, true>);
      grant(fast::lu::sweep_tile_kernel<true, 256, false, true>); grant(fast::lu::sweep_tile_kernel<true, 256, true, true>);
      grant(fast::lu::sweep_tile_kernel<true, 512, false, true>); grant(fast::lu::sweep_tile_kernel<true, 512, true, true>);
      (void)hipGetLastError();
      if (ok) c.max_lds = (size_t)v;
    }
    const int slots = (c.ny[0] - 1) * (c.nz[0] - 1);
    if (slots > 256 && !strict(ctx)) {
      const int prev = c.lu_sweep;
      c.lu_sweep = kSweepTiles;
      c.lu_tiles = std::min(kMaxTiles, (slots + 255) / 256);
      if (!lu_tiles_ok(ctx) || ensure_tiles(ctx) != NPBMZ_OK) {
        c.lu_sweep = prev;
        c.lu_tiles = 4;
      }
    }
  }
  if (c.bench == NPBMZ_BT) {
    // default: the rows kernel in every direction, as measured on MI355X
    // (ms/step, two-sided solves / rows kernel in x and y / in x, y and z):
    // class S 0.057 / 0.039 / 0.037, W 0.083 / 0.056 / 0.054, A 0.172 /
    // 0.120 / 0.120, B 0.397 / 0.304 / 0.296, C 1.421 / 1.243 / 1.236
    c.solve_rows = 7;
  }
  if (c.bench == NPBMZ_SP) {
    // default: two-sided solves in the directions with fewer line tasks
    // than the chip has CUs, as measured on MI355X (class A, at most 225
    // tasks per direction: 0.112 -> 0.096 ms/step; class B, 360 to 864
    // tasks: 0.203 -> 0.209, so one wave per 64 lines there)
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess)
      (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    c.solve_twist = 0;
    for (int d = 0; d < 3; ++d)
      if (c.sp_twist_ok[d] && c.t_line[d].n < cus) c.solve_twist |= 1 << d;
  }
  describe(ctx);
#undef CKC
  *out = ctx;
  return NPBMZ_OK;
}

void npbmz_destroy(npbmz_ctx* ctx) {
  if (!ctx) return;
  auto& c = *ctx;
  if (c.stream) (void)hipStreamSynchronize(c.stream);
  drop_graphs(c);
  Table* ts[] = {&c.t_int,    &c.t_all,    &c.t_line[0], &c.t_line[1], &c.t_line[2],
                 &c.t_face,   &c.t_face2,  &c.t_row[0],  &c.t_row[1],  &c.t_row[2]};
  for (Table* t : ts)
    if (t->d) (void)hipFree(t->d);
  void* ps[] = {c.zd, c.u, c.frc, c.rhs, c.cscr, c.faces, c.partial, c.zpart, c.txbuf, c.tcnt,
                c.hsrc};
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
  if (c.bench == NPBMZ_LU) {
    // setbv, setiv, erhs and the first rhs (src/LU/lu-node.jl:145-196); the
    // warm-up iteration and re-initialisation that follow (:203-311) end in
    // exactly this state again
    hipLaunchKernelGGL(exact::lu::init_kernel<kPtNT>, dim3(t.n), dim3(kPtNT), 0, c.stream, c.u,
                       c.rhs, c.frc, c.zd, t.d);
    CK(hipGetLastError());
    int rc;
    if ((rc = k_lu_rhs(ctx, -1, true))) return rc;
    hipLaunchKernelGGL((exact::lu::rhs_kernel<kPtNT, false>), dim3(t.n), dim3(kPtNT), 0, c.stream,
                       c.u, c.frc, c.rhs, c.zd, t.d, c.lk);
    CK(hipGetLastError());
    CK(hipStreamSynchronize(c.stream));
    return NPBMZ_OK;
  }
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
  if (mode == NPBMZ_MODE_GRAPH && (rc = ensure_graph(ctx, nsteps))) return rc;
  CK(hipEventRecord(c.ev0, c.stream));
  for (int s = 0; s < nsteps;) {
    if (mode != NPBMZ_MODE_GRAPH) {
      if ((rc = step_once(ctx, mode))) return rc;
      s += 1;
    } else if (nsteps - s >= kGraphGroup) {
      CK(hipGraphLaunch(c.gexec_n, c.stream));
      s += kGraphGroup;
    } else {
      CK(hipGraphLaunch(c.gexec, c.stream));
      s += 1;
    }
  }
  CK(hipEventRecord(c.ev1, c.stream));
  CK(hipEventSynchronize(c.ev1));
  if (elapsed_ms) {
    float ms = 0.f;
    CK(hipEventElapsedTime(&ms, c.ev0, c.ev1));
    *elapsed_ms = ms;
  }
  return check_tiles(ctx);
}

int npbmz_run_phase(npbmz_ctx* ctx, int32_t phase, int32_t mode) {
  if (!ctx) return NPBMZ_EINVAL;
  int rc = run_phase(ctx, phase, mode);
  if (rc) return rc;
  CK(hipStreamSynchronize(ctx->stream));
  return check_tiles(ctx);
}

int npbmz_phase_times(npbmz_ctx* ctx, int32_t nsteps, double phase_ms[6]) {
  if (!ctx) return NPBMZ_EINVAL;
  if (nsteps <= 0 || !phase_ms) return fail(ctx, NPBMZ_EINVAL, "nsteps <= 0 or null output");
  auto& c = *ctx;
  const bool lu = c.bench == NPBMZ_LU;
  // BT/SP: exch, rhs, x, y, z(+add); LU: exch, lower, upper, update, rhs
  const int nph = 5;
  std::vector<hipEvent_t> ev((size_t)nsteps * nph * 2, nullptr);
  int rc = NPBMZ_OK;
  for (auto& e : ev)
    if (hipEventCreate(&e) != hipSuccess) { rc = fail(ctx, NPBMZ_EHIP, "hipEventCreate"); break; }
  for (int s = 0; s < nsteps && rc == NPBMZ_OK; ++s) {
    for (int ph = 0; ph < nph && rc == NPBMZ_OK; ++ph) {
      hipEvent_t* e = &ev[((size_t)s * nph + ph) * 2];
      if (hipEventRecord(e[0], c.stream) != hipSuccess) { rc = fail(ctx, NPBMZ_EHIP, "record"); break; }
      if (ph == 0) rc = exch(ctx);
      else if (lu) rc = ph == 1 ? k_lu_sweep(ctx, false, -1)
                       : ph == 2 ? k_lu_sweep(ctx, true, -1)
                       : ph == 3 ? k_lu_update(ctx, -1) : k_lu_rhs(ctx, -1);
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
        static const int lu_slot[5] = {NPBMZ_PHASE_EXCH, NPBMZ_PHASE_SSOR_LOWER,
                                       NPBMZ_PHASE_SSOR_UPPER, NPBMZ_PHASE_SSOR_UPDATE,
                                       NPBMZ_PHASE_RHS};
        phase_ms[lu ? lu_slot[ph] : ph] += ms;
      }
  }
  for (auto& e : ev)
    if (e) (void)hipEventDestroy(e);
  if (rc == NPBMZ_OK) rc = check_tiles(ctx);
  return rc;
}

int npbmz_norms(npbmz_ctx* ctx, double xcr[5], double xce[5], double* xci) {
  if (!ctx) return NPBMZ_EINVAL;
  if (!xcr || !xce) return fail(ctx, NPBMZ_EINVAL, "null output");
  auto& c = *ctx;
  const int nzn = c.nz_;
  if (c.bench == NPBMZ_LU) {
    // src/LU/lu-node.jl:433-458: per zone ERROR, l2norm of the current rsd
    // (the last ssor ended with rhs) and pintgr, summed in zone order
    const Table& t = c.t_int;
    std::vector<double> pp((size_t)t.n * 10), fz((size_t)nzn);
    hipLaunchKernelGGL(exact::lu::norm_kernel<kPtNT>, dim3(t.n), dim3(kPtNT), 0, c.stream, c.u,
                       c.rhs, c.zd, t.d, c.partial);
    CK(hipGetLastError());
    hipLaunchKernelGGL(exact::lu::pintgr_kernel<256>, dim3(nzn), dim3(256), 0, c.stream, c.u, c.zd,
                       c.lk, c.zpart);
    CK(hipGetLastError());
    CK(hipMemcpyAsync(pp.data(), c.partial, pp.size() * sizeof(double), hipMemcpyDeviceToHost,
                      c.stream));
    CK(hipMemcpyAsync(fz.data(), c.zpart, fz.size() * sizeof(double), hipMemcpyDeviceToHost,
                      c.stream));
    CK(hipStreamSynchronize(c.stream));
    c.zn_rhs.assign((size_t)nzn * 5, 0.0);
    c.zn_err.assign((size_t)nzn * 5, 0.0);
    for (int m = 0; m < 5; ++m) { xcr[m] = 0.0; xce[m] = 0.0; }
    double frc = 0.0;
    for (int z = 0; z < nzn; ++z) {
      double s[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
      for (int b = 0; b < t.zcount[z]; ++b)
        for (int q = 0; q < 10; ++q) s[q] += pp[(size_t)(t.zfirst[z] + b) * 10 + q];
      const double vol = double(c.nx[z] - 2) * (c.ny[z] - 2) * (c.nz[z] - 2);
      for (int m = 0; m < 5; ++m) {
        c.zn_err[(size_t)z * 5 + m] = std::sqrt(s[m] / vol);
        c.zn_rhs[(size_t)z * 5 + m] = std::sqrt(s[5 + m] / vol);
        xce[m] = xce[m] + c.zn_err[(size_t)z * 5 + m];
        xcr[m] = xcr[m] + c.zn_rhs[(size_t)z * 5 + m];
      }
      frc = frc + fz[(size_t)z];
    }
    if (xci) *xci = frc;
    return NPBMZ_OK;
  }
  if (xci) *xci = 0.0;
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

int npbmz_norms_noalloc(npbmz_ctx* ctx, const double* u, const double* forcing, double* rhs,
                        double xcr[5], double xce[5], double* xci) {
  if (!ctx) return NPBMZ_EINVAL;
  auto& c = *ctx;
  const bool lu = c.bench == NPBMZ_LU;
  if (!u || !rhs || (!lu && !forcing) || !xcr || !xce)
    return fail(ctx, NPBMZ_EINVAL, "null argument");
  if (!c.scratch) return fail(ctx, NPBMZ_EHIP, "no scratch on this device: " + c.scratch_err);
  const int nzn = c.nz_;
  const Table& ta = lu ? c.t_int : c.t_all;
  // scratch layout (doubles): partials A (LU 10 per block, else 5), partials B
  // (BT/SP, 5 per interior block), per-zone pintgr (LU), per-zone norms
  // (10 each), results (11), block ranges (4 int32 per zone)
  const size_t na = (size_t)ta.n * (lu ? 10 : 5), nb = lu ? 0 : (size_t)c.t_int.n * 5;
  const size_t nzp = lu ? (size_t)nzn : 0, nzn10 = (size_t)nzn * 10, nrng = ((size_t)4 * nzn + 1) / 2;
  const size_t need = na + nb + nzp + nzn10 + 11 + nrng;
  if (need > kScratchDoubles)
    return fail(ctx, NPBMZ_EINVAL,
                "problem too large for the no-allocation scratch (" + std::to_string(need) +
                    " doubles needed, " + std::to_string(kScratchDoubles) + " available)");
  double* pa = c.scratch;
  double* pb = pa + na;
  double* zp = pb + nb;
  double* zn = zp + nzp;
  double* res = zn + nzn10;
  int32_t* rng = reinterpret_cast<int32_t*>(res + 11);
  CK(hipMemcpyAsync(rng, c.noalloc_rng.data(), c.noalloc_rng.size() * sizeof(int32_t),
                    hipMemcpyHostToDevice, c.stream));
  if (lu) {
    hipLaunchKernelGGL(exact::lu::norm_kernel<kPtNT>, dim3(ta.n), dim3(kPtNT), 0, c.stream, u,
                       rhs, c.zd, ta.d, pa);
    CK(hipGetLastError());
    hipLaunchKernelGGL(exact::lu::pintgr_kernel<256>, dim3(nzn), dim3(256), 0, c.stream, u, c.zd,
                       c.lk, zp);
    CK(hipGetLastError());
  } else {
    // error_norm on u, compute_rhs of u without an exchange into rhs, then
    // rhs_norm: the launches of npbmz_norms, on the caller's fields
    if (strict(ctx))
      hipLaunchKernelGGL(exact::bt::norm_kernel<kPtNT>, dim3(ta.n), dim3(kPtNT), 0, c.stream, u,
                         rhs, c.zd, ta.d, pa, 0, c.K);
    else
      hipLaunchKernelGGL(fast::bt::norm_kernel<kPtNT>, dim3(ta.n), dim3(kPtNT), 0, c.stream, u,
                         rhs, c.zd, ta.d, pa, 0, c.K);
    CK(hipGetLastError());
    int rc = k_rhs(ctx, -1, true, nullptr, u, forcing, rhs);
    if (rc) return rc;
    hipLaunchKernelGGL(exact::bt::norm_kernel<kPtNT>, dim3(c.t_int.n), dim3(kPtNT), 0, c.stream,
                       u, rhs, c.zd, c.t_int.d, pb, 1, c.K);
    CK(hipGetLastError());
  }
  hipLaunchKernelGGL(norms_reduce_kernel, dim3(1), dim3(256), 0, c.stream, pa, pb, zp, rng, c.zd,
                     nzn, lu ? 1 : 0, c.dt, zn, res);
  CK(hipGetLastError());
  double out[11];
  CK(hipMemcpyAsync(out, res, sizeof out, hipMemcpyDeviceToHost, c.stream));
  CK(hipStreamSynchronize(c.stream));
  for (int m = 0; m < 5; ++m) {
    xcr[m] = out[m];
    xce[m] = out[5 + m];
  }
  if (xci) *xci = out[10];
  return NPBMZ_OK;
}

int npbmz_tune(npbmz_ctx* ctx, int32_t budget_ms) {
  if (!ctx) return NPBMZ_EINVAL;
  if (budget_ms < 0) return fail(ctx, NPBMZ_EINVAL, "budget_ms < 0");
  auto& c = *ctx;
  // Schedule knobs, tuned one after another: each keeps its fastest value
  // while the next is tried.  LU: the sweep form (hyperplane launches, or one
  // workgroup of 128/256/512 threads per zone).  BT/SP: the task order
  // (largest zones and longest lines first, or natural), the line-solve
  // kernel per direction (BT) and the solve grid.  Trial steps run as graph
  // replays; the state is saved and restored around them.
  struct Knob { int which; std::vector<int> values; };
  std::vector<Knob> knobs;
  if (c.bench == NPBMZ_LU) {
    knobs = {{NPBMZ_SCHED_LU_SWEEP, {1, 0, 2}}, {NPBMZ_SCHED_LU_THREADS, {128, 256, 512}},
             {NPBMZ_SCHED_LU_FORWARD, {3, 2, 1, 0}}, {NPBMZ_SCHED_LU_TILES, {2, 4, 8}},
             {NPBMZ_SCHED_LU_TILE_BATCH, {2, 4, 8}}};
  } else {
    knobs.push_back({-1, {1, 0}});
    if (c.bench == NPBMZ_BT) knobs.push_back({NPBMZ_SCHED_SOLVE_PAIR, {0, 2, 3, 7}});
    if (!strict(ctx)) knobs.push_back({NPBMZ_SCHED_SOLVE_TWIST, {0, 1, 3, 7}});
    if (c.bench == NPBMZ_BT && !strict(ctx))
      knobs.push_back({NPBMZ_SCHED_SOLVE_ROWS, {0, 1, 3, 7}});
    knobs.push_back({NPBMZ_SCHED_SOLVE_GRID, {0, 1024}});
    knobs.push_back({NPBMZ_SCHED_RHS_EXCH, {1, 0}});
  }
  auto get = [&](int which) -> int {
    switch (which) {
      case NPBMZ_SCHED_LU_SWEEP: return c.lu_sweep;
      case NPBMZ_SCHED_LU_THREADS: return c.lu_nt;
      case NPBMZ_SCHED_LU_FORWARD: return c.lu_fwd;
      case NPBMZ_SCHED_LU_TILES: return c.lu_tiles;
      case NPBMZ_SCHED_LU_TILE_BATCH: return c.lu_kb;
      case NPBMZ_SCHED_SOLVE_PAIR: return c.solve_pair;
      case NPBMZ_SCHED_SOLVE_TWIST: return c.solve_twist;
      case NPBMZ_SCHED_SOLVE_ROWS: return c.solve_rows;
      case NPBMZ_SCHED_SOLVE_GRID: return c.solve_grid[0];
      case NPBMZ_SCHED_RHS_EXCH: return c.rhs_exch;
      default: return c.largest_first ? 1 : 0;
    }
  };
  auto set = [&](int which, int v) -> int {
    if (which >= 0) {
      if ((which == NPBMZ_SCHED_LU_THREADS || which == NPBMZ_SCHED_LU_FORWARD) &&
          c.lu_sweep != kSweepZone)
        return NPBMZ_OK;
      if ((which == NPBMZ_SCHED_LU_TILES || which == NPBMZ_SCHED_LU_TILE_BATCH) &&
          c.lu_sweep != kSweepTiles)
        return NPBMZ_OK;
      return npbmz_set_schedule(ctx, which, v);
    }
    c.largest_first = v == 1;
    if (c.largest_first) c.order = cost_order(c);
    else std::iota(c.order.begin(), c.order.end(), 0);
    drop_graphs(c);
    return build_tables(ctx);
  };
  const size_t fb = sizeof(double) * 5 * (size_t)c.npoints;
  double *su = nullptr, *sr = nullptr;
  CK(hipMalloc(&su, fb));
  if (hipMalloc(&sr, fb) != hipSuccess) { (void)hipFree(su); return fail(ctx, NPBMZ_ENOMEM, "tune buffers"); }
  int rc = NPBMZ_OK;
  if (hipMemcpyAsync(su, c.u, fb, hipMemcpyDeviceToDevice, c.stream) != hipSuccess ||
      hipMemcpyAsync(sr, c.rhs, fb, hipMemcpyDeviceToDevice, c.stream) != hipSuccess)
    rc = fail(ctx, NPBMZ_EHIP, "tune save");
  size_t trials = 0;
  for (auto& k : knobs) trials += k.values.size();
  const double per = std::max(1.0, double(budget_ms) / trials);
  auto trial = [&](double& t) -> int {
    double ms = 0.0, tot = 0.0;
    int steps = 0, r;
    if ((r = npbmz_step(ctx, 1, NPBMZ_MODE_GRAPH, &ms))) return r;  // capture + warm-up
    do {
      if ((r = npbmz_step(ctx, 1, NPBMZ_MODE_GRAPH, &ms))) return r;
      tot += ms;
      ++steps;
    } while (tot < per && steps < 50);
    t = tot / steps;
    if (hipMemcpyAsync(c.u, su, fb, hipMemcpyDeviceToDevice, c.stream) != hipSuccess ||
        hipMemcpyAsync(c.rhs, sr, fb, hipMemcpyDeviceToDevice, c.stream) != hipSuccess)
      return fail(ctx, NPBMZ_EHIP, "tune restore");
    return NPBMZ_OK;
  };
  for (auto& k : knobs) {
    if (rc) break;
    int best_v = get(k.which);
    double best = 1e300;
    for (int v : k.values) {
      double t = 0.0;
      if ((rc = set(k.which, v)) || (rc = trial(t))) break;
      if (t < best) { best = t; best_v = v; }
    }
    if (rc == NPBMZ_OK) rc = set(k.which, best_v);
  }
  if (hipStreamSynchronize(c.stream) != hipSuccess && rc == NPBMZ_OK)
    rc = fail(ctx, NPBMZ_EHIP, "tune sync");
  (void)hipFree(su);
  (void)hipFree(sr);
  describe(ctx);
  return rc;
}

int npbmz_set_schedule(npbmz_ctx* ctx, int32_t which, int32_t value) {
  if (!ctx) return NPBMZ_EINVAL;
  auto& c = *ctx;
  switch (which) {
    case NPBMZ_SCHED_LU_SWEEP:
      if (c.bench != NPBMZ_LU || value < 0 || value > 2)
        return fail(ctx, NPBMZ_EINVAL,
                    "LU sweep schedule is 0 (hyperplane), 1 (zone) or 2 (pipelined tiles)");
      if (value == kSweepTiles) {
        const int rc = ensure_tiles(ctx);
        if (rc) return rc;
      }
      c.lu_sweep = value;
      break;
    case NPBMZ_SCHED_LU_TILES:
      if (c.bench != NPBMZ_LU || value < 1 || value > kMaxTiles)
        return fail(ctx, NPBMZ_EINVAL, "LU tiles per zone is 1..8");
      c.lu_tiles = value;
      break;
    case NPBMZ_SCHED_LU_TILE_BATCH:
      if (c.bench != NPBMZ_LU || value < 1 || value > 64)
        return fail(ctx, NPBMZ_EINVAL, "LU tile batch is 1..64 hyperplanes");
      c.lu_kb = value;
      break;
    case NPBMZ_SCHED_LU_SPIN:
      if (c.bench != NPBMZ_LU || value < -1)
        return fail(ctx, NPBMZ_EINVAL, "LU tile wait bound is -1 or more polls");
      c.lu_spin = value;
      break;
    case NPBMZ_SCHED_LU_HANDOFF:
      if (c.bench != NPBMZ_LU || value < 0 || value > 2)
        return fail(ctx, NPBMZ_EINVAL,
                    "LU tile hand-off is 0 (agent scope), 1 (XCD-local) or 2 (XCD-local, "
                    "placement check forced to fail)");
      {
        const int rc = ensure_tiles(ctx);
        if (rc) return rc;
      }
      if (value == 1 && !c.tile_local_ok)
        return fail(ctx, NPBMZ_EINVAL,
                    "LU tile hand-off 1 needs workgroup b on the XCD of b mod 8, which this "
                    "device does not do");
      c.lu_handoff = value;
      break;
    case NPBMZ_SCHED_LU_THREADS:
      if (c.bench != NPBMZ_LU || (value != 128 && value != 256 && value != 512))
        return fail(ctx, NPBMZ_EINVAL, "LU workgroup size is 128, 256 or 512");
      c.lu_nt = value;
      break;
    case NPBMZ_SCHED_LU_FORWARD:
      if (c.bench != NPBMZ_LU || value < 0 || value > 3)
        return fail(ctx, NPBMZ_EINVAL, "LU hyperplane forwarding is 0, 1, 2 or 3");
      c.lu_fwd = value;
      break;
    case NPBMZ_SCHED_SOLVE_PAIR:
      if (c.bench != NPBMZ_BT || value < 0 || value > 7)
        return fail(ctx, NPBMZ_EINVAL, "BT solve pair mask is 0..7");
      c.solve_pair = value;
      break;
    case NPBMZ_SCHED_SOLVE_TWIST:
      if (c.bench == NPBMZ_LU || value < 0 || value > 7)
        return fail(ctx, NPBMZ_EINVAL, "BT/SP two-sided solve mask is 0..7");
      c.solve_twist = value;
      break;
    case NPBMZ_SCHED_SOLVE_ROWS:
      if (c.bench != NPBMZ_BT || value < 0 || value > 7)
        return fail(ctx, NPBMZ_EINVAL, "BT rows-over-lanes solve mask is 0..7");
      c.solve_rows = value;
      break;
    case NPBMZ_SCHED_RHS_EXCH:
      if (c.bench == NPBMZ_LU || value < 0 || value > 1)
        return fail(ctx, NPBMZ_EINVAL, "BT/SP exchange inside compute_rhs is 0 or 1");
      c.rhs_exch = value;
      break;
    case NPBMZ_SCHED_SOLVE_GRID:
      // a launch takes min(tasks, grid) workgroups, so the line scratch
      // (one slice per task) never depends on it
      if (c.bench == NPBMZ_LU || value < 0 || value > kMaxSolveGrid)
        return fail(ctx, NPBMZ_EINVAL, "solve grid must be 0..65536 (BT/SP)");
      for (int d = 0; d < 3; ++d) c.solve_grid[d] = value;
      break;
    default: return fail(ctx, NPBMZ_EINVAL, "unknown schedule knob");
  }
  drop_graphs(c);
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

int npbmz_loss(npbmz_ctx* ctx, const double* w, double* loss) {
  if (!ctx) return NPBMZ_EINVAL;
  if (!w || !loss) return fail(ctx, NPBMZ_EINVAL, "null argument");
  auto& c = *ctx;
  const size_t fb = sizeof(double) * 5 * (size_t)c.npoints;
  double* wd = nullptr;
  CK(hipMalloc(&wd, fb));
  int rc = NPBMZ_OK;
  if (hipMemcpyAsync(wd, w, fb, hipMemcpyHostToDevice, c.stream) != hipSuccess)
    rc = fail(ctx, NPBMZ_EHIP, "loss weights upload");
  if (rc == NPBMZ_OK) rc = bt_loss(ctx, wd, nullptr, nullptr, false, loss);
  (void)hipFree(wd);
  return rc;
}

int npbmz_adjoint(npbmz_ctx* ctx, int32_t nsteps, int32_t ckpt_every, const double* w,
                  double* grad_u0, double* grad_forcing, double* loss, double* fwd_ms,
                  double* adj_ms) {
  if (!ctx) return NPBMZ_EINVAL;
  auto& c = *ctx;
  const bool sp = c.bench == NPBMZ_SP, lu = c.bench == NPBMZ_LU;
  if (nsteps < 1 || ckpt_every < 1) return fail(ctx, NPBMZ_EINVAL, "nsteps and ckpt_every must be >= 1");
  if (!w || !grad_u0 || !grad_forcing || !loss) return fail(ctx, NPBMZ_EINVAL, "null argument");
  const size_t nd = 5 * (size_t)c.npoints, fb = sizeof(double) * nd;
  const int k = std::min(ckpt_every, nsteps);
  const int ncp = (nsteps + k - 1) / k;
  // device buffers: w, ubar, lam, fbar, the checkpoints, and per step of a
  // segment the post-exchange u and the three solve outputs (SP: also r0;
  // LU: the pre-exchange u in x3 and the sweep outputs in x1, x2, with one
  // more workspace field)
  std::vector<double*> bufs;
  auto alloc = [&](double** p) -> bool {
    if (hipMalloc(p, fb) != hipSuccess) return false;
    bufs.push_back(*p);
    return true;
  };
  auto release = [&]() {
    for (double* p : bufs) (void)hipFree(p);
  };
  double *wd, *ubar, *lam, *fbar, *aux = nullptr, *rs0 = nullptr;
  std::vector<double*> ck(ncp), us(k), x1(k), x2(k), x3(k), r0(k, nullptr);
  // LU: a workspace field and the saved rsd; SP: the y and x transposed
  // solutions (lam holds z's)
  bool ok = alloc(&wd) && alloc(&ubar) && alloc(&lam) && alloc(&fbar) &&
            (!(lu || sp) || (alloc(&aux) && alloc(&rs0)));
  for (int q = 0; q < ncp && ok; ++q) ok = alloc(&ck[q]);
  for (int q = 0; q < k && ok; ++q) ok = alloc(&us[q]) && alloc(&x1[q]) && alloc(&x2[q]) && alloc(&x3[q]);
  for (int q = 0; q < k && ok && sp; ++q) ok = alloc(&r0[q]);
  if (!ok) {
    release();
    return fail(ctx, NPBMZ_ENOMEM, "adjoint buffers");
  }
  int rc = NPBMZ_OK;
  auto hip = [&](hipError_t e, const char* what) {
    if (e != hipSuccess && rc == NPBMZ_OK) rc = fail(ctx, NPBMZ_EHIP, std::string(what) + ": " + hipGetErrorString(e));
  };
  hip(hipMemcpyAsync(wd, w, fb, hipMemcpyHostToDevice, c.stream), "weights upload");
  // LU's rsd at an iteration's start is rhs(u): recomputed, so that the loss
  // is a function of u0 and frct alone (the context's own rsd is put back at
  // the end)
  if (lu) hip(hipMemcpyAsync(rs0, c.rhs, fb, hipMemcpyDeviceToDevice, c.stream), "save rsd");
  if (lu && rc == NPBMZ_OK) rc = k_lu_rhs(ctx, -1);
  hipEvent_t e[4] = {nullptr, nullptr, nullptr, nullptr};
  for (auto& x : e) hip(hipEventCreate(&x), "hipEventCreate");
  // forward pass, keeping u every k steps
  if (rc == NPBMZ_OK) hip(hipEventRecord(e[0], c.stream), "record");
  for (int n = 0; n < nsteps && rc == NPBMZ_OK; ++n) {
    if (n % k == 0) hip(hipMemcpyAsync(ck[n / k], c.u, fb, hipMemcpyDeviceToDevice, c.stream), "checkpoint");
    if (rc == NPBMZ_OK)
      rc = sp   ? sp_step_store(ctx, nullptr, nullptr, nullptr, nullptr, nullptr)
           : lu ? lu_step_store(ctx, nullptr, nullptr, nullptr, nullptr)
                : bt_step_store(ctx, nullptr, nullptr, nullptr, nullptr);
  }
  if (rc == NPBMZ_OK) hip(hipEventRecord(e[1], c.stream), "record");
  if (rc == NPBMZ_OK) rc = bt_loss(ctx, wd, ubar, fbar, true, loss);
  // reverse pass: recompute each segment from its checkpoint, then reverse
  // its steps from the last to the first
  if (rc == NPBMZ_OK) hip(hipEventRecord(e[2], c.stream), "record");
  for (int sgm = ncp - 1; sgm >= 0 && rc == NPBMZ_OK; --sgm) {
    const int n0 = sgm * k, n1 = std::min(nsteps, n0 + k);
    hip(hipMemcpyAsync(c.u, ck[sgm], fb, hipMemcpyDeviceToDevice, c.stream), "restore");
    if (lu && rc == NPBMZ_OK) rc = k_lu_rhs(ctx, -1);
    for (int n = n0; n < n1 && rc == NPBMZ_OK; ++n) {
      const int q = n - n0;
      rc = sp   ? sp_step_store(ctx, us[q], r0[q], x1[q], x2[q], x3[q])
           : lu ? lu_step_store(ctx, x3[q], us[q], x1[q], x2[q])
                : bt_step_store(ctx, us[q], x1[q], x2[q], x3[q]);
    }
    for (int n = n1 - 1; n >= n0 && rc == NPBMZ_OK; --n) {
      const int q = n - n0;
      rc = sp   ? sp_adj_step(ctx, us[q], r0[q], x1[q], x2[q], x3[q], ubar, lam, aux, rs0, fbar)
           : lu ? lu_adj_step(ctx, x3[q], us[q], x1[q], x2[q], ubar, lam, aux, fbar)
                : bt_adj_step(ctx, us[q], x1[q], x2[q], x3[q], ubar, lam, fbar);
    }
  }
  if (rc == NPBMZ_OK) hip(hipEventRecord(e[3], c.stream), "record");
  // results, and u back to the initial state
  if (rc == NPBMZ_OK) {
    hip(hipMemcpyAsync(grad_u0, ubar, fb, hipMemcpyDeviceToHost, c.stream), "gradient download");
    hip(hipMemcpyAsync(grad_forcing, fbar, fb, hipMemcpyDeviceToHost, c.stream), "gradient download");
    hip(hipMemcpyAsync(c.u, ck[0], fb, hipMemcpyDeviceToDevice, c.stream), "restore u0");
    if (lu) hip(hipMemcpyAsync(c.rhs, rs0, fb, hipMemcpyDeviceToDevice, c.stream), "restore rsd");
    hip(hipStreamSynchronize(c.stream), "synchronize");
  }
  if (rc == NPBMZ_OK) {
    float a = 0.f, b = 0.f;
    hip(hipEventElapsedTime(&a, e[0], e[1]), "elapsed");
    hip(hipEventElapsedTime(&b, e[2], e[3]), "elapsed");
    if (fwd_ms) *fwd_ms = a;
    if (adj_ms) *adj_ms = b;
  }
  (void)hipStreamSynchronize(c.stream);
  for (auto& x : e)
    if (x) (void)hipEventDestroy(x);
  release();
  if (rc == NPBMZ_OK) rc = check_tiles(ctx);
  return rc;
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
