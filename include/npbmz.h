/* npbmz.h - C ABI of libnpbmz, the MI355X (gfx950) multi-zone BT/SP/LU timestep.
 *
 * The reference (NPBApps.jl) has no FFI: its node loop calls exch_qbc and a
 * threaded per-zone adi/ssor directly in Julia.  These entry points replace
 * exactly that loop body and the node-side verify numerics, so that the Julia
 * driver (go, zone_setup, initialize/exact_rhs, class verification) can stay
 * as it is and reach the GPU through ccall.  Citations below are to the
 * reference's files.
 *
 * Conventions
 *   - Every function returns 0 on success and a negative NPBMZ_E* code on
 *     failure; nothing aborts or throws across the ABI.  npbmz_last_error()
 *     gives a message for the last failure on a context.
 *   - Host pointers are borrowed for the duration of the call only.
 *   - A context owns device memory, one HIP stream and captured graphs; it is
 *     not thread safe (one caller thread per context).
 *   - Zones are numbered from 0 (the reference numbers them from 1).
 *   - All arithmetic is IEEE binary64.
 */
#ifndef NPBMZ_H
#define NPBMZ_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* benchmark selector */
#define NPBMZ_BT 0
#define NPBMZ_SP 1
#define NPBMZ_LU 2

/* error codes */
#define NPBMZ_OK 0
#define NPBMZ_EINVAL -1   /* bad argument */
#define NPBMZ_EHIP -2     /* HIP runtime failure (message holds the HIP error) */
#define NPBMZ_ENOMEM -3   /* device allocation failed */
#define NPBMZ_ENOTSUP -4  /* feature not built for this benchmark */
#define NPBMZ_ENODEV -5   /* no usable gfx950 device */

/* creation flags */
#define NPBMZ_FLAG_STRICT_FP 1 /* kernels built without fused multiply-add:
                                  bit-identical to the reference's CPU path */
#define NPBMZ_FLAG_LU_NPB_EXCH 2 /* LU: exchange face planes k = 2..nz-1 (1-based)
                                    as NPB-MZ does, instead of the reference's
                                    k = 1..nz-2 (src/LU/exchange_qbc.jl:85,109,
                                    135,160,205,226,247,268); only this choice
                                    reproduces src/LU/verify.jl's tables */

/* fields for put/get */
#define NPBMZ_FIELD_U 0       /* conserved variables u (BT/SP/LU) */
#define NPBMZ_FIELD_FORCING 1 /* forcing (BT/SP), frct (LU) */
#define NPBMZ_FIELD_RHS 2     /* rhs (BT/SP), rsd (LU) */

/* host array axis order for put/get */
#define NPBMZ_ORDER_MIJK 0 /* Julia [m,i,j,k] (BT u/forcing/rhs, LU u/rsd/frct) */
#define NPBMZ_ORDER_IJKM 1 /* Julia [i,j,k,m] (SP u/forcing/rhs) */

/* step modes */
#define NPBMZ_MODE_PERZONE 0 /* one launch per zone per phase */
#define NPBMZ_MODE_BATCHED 1 /* all zones in one launch per phase */
#define NPBMZ_MODE_GRAPH 2   /* batched, replayed from hipGraphs of 8 timesteps and of 1 */

/* phases for npbmz_run_phase (testing and profiling) */
#define NPBMZ_PHASE_EXCH 0  /* exch_qbc face copy */
#define NPBMZ_PHASE_RHS 1   /* compute_rhs (BT/SP), rhs (LU) */
#define NPBMZ_PHASE_XSOLVE 2
#define NPBMZ_PHASE_YSOLVE 3
#define NPBMZ_PHASE_ZSOLVE 4
#define NPBMZ_PHASE_ADD 5
/* LU's phases (exch and rhs as above) */
#define NPBMZ_PHASE_SSOR_LOWER 2  /* rsd *= dt, jacld + blts (src/LU/ssor.jl:62-108) */
#define NPBMZ_PHASE_SSOR_UPPER 3  /* jacu + buts (src/LU/ssor.jl:114-182) */
#define NPBMZ_PHASE_SSOR_UPDATE 4 /* u += rsd/(omega(2-omega)) (src/LU/ssor.jl:193-202) */

typedef struct npbmz_ctx npbmz_ctx;

/* One multi-zone problem.  nx/ny/nz/iz_* are per zone, in the order that
 * zone_setup produces them (src/BT/zone_setup.jl:53-74).  gx/gy/gz and
 * x_zones/y_zones feed set_constants (src/BT/set_constants.jl:53-130). */
typedef struct {
  int32_t bench;   /* NPBMZ_BT / NPBMZ_SP / NPBMZ_LU */
  int32_t nzones;  /* x_zones * y_zones */
  int32_t x_zones, y_zones;
  int32_t gx_size, gy_size, gz_size;
  const int32_t *nx, *ny, *nz;
  const int32_t *iz_west, *iz_east, *iz_south, *iz_north; /* 0-based */
  double dt;
  double omega; /* LU only (src/LU/lu_data.jl:8); ignored otherwise */
} npbmz_problem;

/* Create a context on the current HIP device: validates the zone table,
 * computes the stencil constants, allocates every zone's fields in HBM.
 * Replaces the allocation part of perform (src/BT/bt-node.jl:63-93). */
int npbmz_create(const npbmz_problem* prob, int32_t flags, npbmz_ctx** out);

/* Release everything the context owns. */
void npbmz_destroy(npbmz_ctx* ctx);

/* Message for the last failure on ctx (or a creation failure if ctx is NULL). */
const char* npbmz_last_error(const npbmz_ctx* ctx);

/* Copy one zone's field between a host array in the reference's layout and
 * HBM.  host points at the parent array of the Julia OffsetArray; lo[d]/hi[d]
 * are its axis bounds (d = 0..3 in `order`), e.g. BT u is
 * [1:5, -2:nx+2, -2:ny+2, -2:nz+2] (src/BT/bt_data.jl:143-144).  Only the
 * zone's grid points are transferred (get leaves the rest untouched).
 * Replaces the Julia arrays being owned by the node (bt_data.jl:270-326). */
int npbmz_put_field(npbmz_ctx* ctx, int32_t zone, int32_t field, const double* host,
                    const int64_t lo[4], const int64_t hi[4], int32_t order);
int npbmz_get_field(npbmz_ctx* ctx, int32_t zone, int32_t field, double* host,
                    const int64_t lo[4], const int64_t hi[4], int32_t order);

/* Device-side initial condition: initialize + exact_rhs (BT/SP,
 * src/BT/initialize.jl:26-225, src/BT/exact_rhs.jl:5-370) or
 * setbv + setiv + erhs (LU).  Same values as the host routines. */
int npbmz_init_exact(npbmz_ctx* ctx);

/* Run nsteps timesteps: for each, exch_qbc then every zone's adi (BT/SP) or
 * ssor (LU).  Replaces the STEP loop body of perform
 * (src/BT/bt-node.jl:280-417, src/SP/sp-node.jl:252-391,
 * src/LU/lu-node.jl:338-428).  If elapsed_ms is not NULL it receives the
 * HIP-event time of the whole call (inputs already resident in HBM).  In
 * NPBMZ_MODE_GRAPH the first call captures a one-step graph, and the first
 * call with nsteps >= 8 an eight-step graph, before the timed region; a
 * schedule change discards both. */
int npbmz_step(npbmz_ctx* ctx, int32_t nsteps, int32_t mode, double* elapsed_ms);

/* Run one phase over all zones (testing and profiling). */
int npbmz_run_phase(npbmz_ctx* ctx, int32_t phase, int32_t mode);

/* Run nsteps batched timesteps with a HIP event pair around every phase's
 * launch on the context's stream.  phase_ms[p] (p = NPBMZ_PHASE_*) receives
 * the summed time of phase p over the steps; the z sweep folds add, so
 * phase_ms[NPBMZ_PHASE_ADD] is 0 for BT.  Per-phase counterpart of the
 * reference's itimer timers (src/BT/bt_data.jl:50-76). */
int npbmz_phase_times(npbmz_ctx* ctx, int32_t nsteps, double phase_ms[6]);

/* Node-side verification numerics on the current state, in the reference's
 * order: per zone error_norm, compute_rhs without a new exchange, rhs_norm;
 * xcr = sum over zones of rms_rhs/dt, xce = sum of rms_err
 * (src/BT/verify.jl:411-501, src/SP/verify.jl:411-515).  For LU, xcr and xce
 * are the summed l2norm of rsd and ERROR, and xci (may be NULL) the summed
 * surface integral (src/LU/lu-node.jl:430-464). */
int npbmz_norms(npbmz_ctx* ctx, double xcr[5], double xce[5], double* xci);

/* npbmz_norms on caller-owned device fields, allocating nothing.  No
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  u, forcing and rhs are device arrays in the
 * library's whole-run layout (every zone's nx*ny*nz*5 values, zone after
 * zone, each zone [k][j][i][m] with m fastest).  BT/SP: rhs receives
 * compute_rhs of u without an exchange, as npbmz_norms leaves it in the
 * context's rhs; LU: rhs holds rsd (read only) and forcing is not used.  The
 * results equal npbmz_norms' on the same state bit for bit.  The kernels'
 * partial sums live in a static 2 MiB device array, one per GPU, shared by
 * every call on that GPU: calls that overlap on one GPU race, and problems
 * whose partial sums exceed it (about 23,000 blocks of 128 points for BT/SP,
 * e.g. BT-MZ class C) fail with NPBMZ_EINVAL.  npbmz_zone_norms does not see
 * these calls.  No hipMalloc, hipFree or other allocation. */
int npbmz_norms_noalloc(npbmz_ctx* ctx, const double* u, const double* forcing, double* rhs,
                        double xcr[5], double xce[5], double* xci);

/* Per-zone raw residual/error norms of the last npbmz_norms call
 * (nzones x 5 each, zone-major); either pointer may be NULL. */
int npbmz_zone_norms(npbmz_ctx* ctx, double* rms_rhs, double* rms_err);

/* Autotune the launch schedule by timing graph replays of candidate
 * schedules for about budget_ms in total, one knob after another (each knob
 * keeps its fastest value while the next is tried): for BT/SP the task order
 * (largest zones first, and in each solve direction the zones with the
 * longest lines first; or natural order), the line-solve kernel per
 * direction (BT: one wave per 64 lines, a builder/eliminator wave pair, or,
 * with FMA arithmetic, the two-sided elimination)
 * and the solve grid (one block per 64 lines, or a persistent grid of 1024
 * blocks striding over them); for LU the SSOR sweep schedule (one workgroup
 * of 128/256/512 threads per zone walking all hyperplanes, or one launch per
 * i+j+k hyperplane over all zones).  The state is restored afterwards; the
 * chosen schedule is kept in the context for all modes.  No counterpart in
 * the reference (its zone->thread schedule is Threads.@threads over zones,
 * src/BT/bt-node.jl:318). */
int npbmz_tune(npbmz_ctx* ctx, int32_t budget_ms);

/* Set one schedule knob directly (the tuner's candidates, for tests and
 * profiling).  No counterpart in the reference. */
#define NPBMZ_SCHED_LU_SWEEP 0   /* 0 hyperplane launches, 1 workgroup per zone,
                                    2 each zone's hyperplanes split by k rows over
                                    pipelined workgroups (tiles; FMA arithmetic,
                                    where a tile's padded slots fit 512 threads;
                                    otherwise hyperplane launches).  Default
                                    with FMA arithmetic: 2 where a zone's padded
                                    hyperplane has more than 256 slots, else 1;
                                    strict: 1 for zones of at most 32768
                                    interior points, else 0 */
#define NPBMZ_SCHED_LU_THREADS 1 /* 128, 256 or 512 threads per zone workgroup
                                    (default with FMA arithmetic: the smallest that
                                    holds a zone's hyperplane padded by a halo row
                                    and column, up to 512; strict: 512) */
#define NPBMZ_SCHED_SOLVE_GRID 2 /* BT/SP: 0 one block per task, else the size
                                    (1..65536) of a persistent grid striding
                                    over the tasks; NPBMZ_EINVAL otherwise */
#define NPBMZ_SCHED_SOLVE_PAIR 3 /* BT: bit d set: direction d solves each 64 lines
                                    with a builder and an eliminator wave (default 2) */
#define NPBMZ_SCHED_SOLVE_TWIST 4 /* BT and SP, FMA arithmetic only (ignored with
                                     NPBMZ_FLAG_STRICT_FP): bit d set: direction d
                                     eliminates each line from both ends with two
                                     waves meeting in the middle (BT: overrides
                                     PAIR, default 7; SP: for directions whose
                                     lines all have >= 6 points, default on where
                                     the direction has fewer line tasks of 64
                                     lines than the device has CUs); the
                                     adjoint's transposed solves follow the
                                     same bits */
#define NPBMZ_SCHED_LU_FORWARD 5 /* LU workgroup-per-zone sweeps: each hyperplane
                                    hands the next its new rsd (1), or its rsd and
                                    u (2), through LDS, as far as two hyperplanes
                                    fit the device's LDS per workgroup; 0 reads
                                    them from global memory (same values).  3:
                                    with FMA arithmetic and at most one (j, k)
                                    slot of a hyperplane padded by a halo row and
                                    column per thread, the zone's boundary values
                                    pass through LDS too and each thread's global
                                    loads are issued a hyperplane ahead;
                                    otherwise as 2.  Default: with FMA arithmetic
                                    3 where the padded hyperplane fits 512
                                    threads; otherwise (and strict) 3 from 128
                                    (j, k) pairs per hyperplane up, else 0 */
#define NPBMZ_SCHED_SOLVE_ROWS 6 /* BT, FMA arithmetic only (ignored with
                                    NPBMZ_FLAG_STRICT_FP): bit d set: direction d
                                    solves 12 lines per wave with each line's five
                                    block rows on five lanes, two-sided, keeping
                                    the last reduced rows of each half in
                                    registers (overrides TWIST and PAIR).
                                    Default: all three directions (7) */
#define NPBMZ_SCHED_LU_TILES 7 /* LU pipelined sweep: tiles per zone, 1..8 (default:
                                  the padded slots over 256, rounded up, at most 8;
                                  4 where the tiles are not the default; fewer if a
                                  zone has fewer padded k rows) */
#define NPBMZ_SCHED_LU_TILE_BATCH 8 /* LU pipelined sweep: hyperplanes a tile finishes
                                       between publishing its progress, 1..64
                                       (default 4) */
#define NPBMZ_SCHED_LU_SPIN 9 /* LU pipelined sweep: polls a tile's wait for its
                                 upstream tile may take before the step fails with
                                 NPBMZ_EHIP (default 2^20; -1 fails every wait at
                                 once, for tests of the error path) */
#define NPBMZ_SCHED_LU_HANDOFF 10 /* LU pipelined sweep, how a tile hands its edge rows
                                     on: 0 agent-scope release and acquire (an L2
                                     write-back and invalidation per published count,
                                     correct wherever the tiles run); 1 XCD-local: the
                                     tiles of a zone share an XCD and its L2, and the
                                     rows need no cache maintenance (the default where
                                     the device places workgroup b on the XCD of
                                     workgroup b mod 8, which is checked once per
                                     device; EINVAL elsewhere); 2 as 1 with every tile's
                                     placement check failing, for tests of the error
                                     path (the step fails with NPBMZ_EHIP) */
#define NPBMZ_SCHED_RHS_EXCH 11 /* BT and SP with more than one zone: 1 batched and
                                   graph steps run the zone-face exchange in
                                   compute_rhs's launch (the rhs reads each face
                                   neighbour from the cell the exchange copies, the
                                   exchange's blocks fill the halos for the line
                                   solves; same bits); 0 a launch of its own;
                                   default 1 */
int npbmz_set_schedule(npbmz_ctx* ctx, int32_t which, int32_t value);

/* Describe the schedule currently in use (for logs); writes a NUL-terminated
 * string of at most cap bytes. */
int npbmz_describe(const npbmz_ctx* ctx, char* buf, int32_t cap);

/* Number of zones and total grid points held by ctx. */
int npbmz_size(const npbmz_ctx* ctx, int32_t* nzones, int64_t* npoints);

/* Adjoint of the BT, SP and LU timesteps (no counterpart in the reference,
 * which has no derivatives; a ChainRulesCore rrule's pullback would call it).
 * From the context's current u (u0) and forcing (LU: frct), runs nsteps
 * timesteps keeping u every ckpt_every steps, evaluates loss = sum over every
 * grid point and component of w*u^2 on the final u, then runs the steps
 * backwards (each segment between checkpoints recomputed forward first):
 * BT and SP: the transposed z, y and x solves (BT block-tridiagonal; SP
 * pentadiagonal, after the adjoints of tzetar, pinvr and ninvr) with the
 * derivatives of their matrices, the adjoint of txinvr (SP), of compute_rhs
 * and of the face exchange; LU: the transposed upper and lower SSOR sweeps
 * (hyperplanes in the opposite order) with the derivatives of jacld's and
 * jacu's blocks, the adjoint of the exchange and of rhs.  For LU, rsd is
 * first recomputed as rhs(u0), the value an iteration starts from, so the
 * loss is a function of u0 and frct alone, and frct's gradient covers every
 * point (rsd = -frct on the boundary enters the sweeps).  grad_u0 and
 * grad_forcing receive dloss/du0 and dloss/dforcing; fwd_ms and adj_ms (may
 * be NULL) the HIP-event times of the forward pass and of the reverse pass
 * including its recomputation.  w and both gradients hold every zone's
 * nx*ny*nz*5 values, zone after zone, each zone in [k][j][i][m] order (m
 * fastest).  On return u (LU: and rsd) hold their values from before the
 * call. */
int npbmz_adjoint(npbmz_ctx* ctx, int32_t nsteps, int32_t ckpt_every, const double* w,
                  double* grad_u0, double* grad_forcing, double* loss, double* fwd_ms,
                  double* adj_ms);

/* The same loss, sum of w*u^2, of the current u (layout as above). */
int npbmz_loss(npbmz_ctx* ctx, const double* w, double* loss);

/* Diagnostic with no counterpart in the reference: copy nbytes between two
 * scratch buffers with a plain streaming kernel, reps times, and return the
 * average kernel time.  Its known traffic (nbytes read + nbytes written per
 * launch) calibrates profiler byte counters and gives the achievable HBM
 * bandwidth on the device. */
int npbmz_probe_copy(npbmz_ctx* ctx, int64_t nbytes, int32_t reps, double* ms_per_launch);

#ifdef __cplusplus
}
#endif

#endif /* NPBMZ_H */
