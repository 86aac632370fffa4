/* npbmz_run: one whole NPB-MZ class run through the C ABI of libnpbmz alone
 * (no Python, no torch): the caller side that a C or C++ driver, or a
 * foreign-function binding, writes around the library.
 *
 * It does what the reference's go() does around its node loop
 * (src/BT/bt-driver.jl, src/SP/sp-driver.jl, src/LU/lu-driver.jl): the zone
 * split (src/BT/zone_setup.jl:1-77; LU rounds instead of truncating,
 * src/LU/zone_setup.jl:15,19), the class parameters (src/BT/classes.jl:95-104,
 * src/SP/classes.jl:14-24, src/LU/classes.jl:14-24), initialisation, one
 * untimed step followed by re-initialisation (src/BT/bt-node.jl:138-253),
 * niter timed steps, the node-side verification numerics and the class
 * verification (src/BT/verify.jl:355-400, src/LU/verify.jl:355-420: dt and
 * niter as tabulated, every norm within relative 1e-8 of the table).  It
 * prints one JSON line: the norms, the verdict, the time and Mop/s with the
 * reference's flop counts (src/BT/bt-driver.jl:131-140,
 * src/LU/lu-driver.jl:142-151).
 *
 *   usage: npbmz_run bt|sp|lu S|W|A|B [perzone|batched|graph] [strict] [npbexch]
 *          npbmz_run bt|sp|lu S|W|A|B zones   (the zone table only, no GPU)
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "npbmz.h"
#include "verify_tables.h"

typedef struct {
  const char* name;
  int niter;
  double dt, ratio;
  int x_zones, y_zones, gx, gy, gz;
} klass;

static const klass bt_classes[] = {
    {"S", 60, 0.010, 3.0, 2, 2, 24, 24, 6},
    {"W", 200, 0.0008, 4.5, 4, 4, 64, 64, 8},
    {"A", 200, 0.0008, 4.5, 4, 4, 128, 128, 16},
    {"B", 200, 0.0003, 4.5, 8, 8, 304, 208, 17},
};
static const klass sp_classes[] = {
    {"S", 100, 0.015, 1.0, 2, 2, 24, 24, 6},
    {"W", 400, 0.0015, 1.0, 4, 4, 64, 64, 8},
    {"A", 400, 0.0015, 1.0, 4, 4, 128, 128, 16},
    {"B", 400, 0.001, 1.0, 8, 8, 304, 208, 17},
};
static const klass lu_classes[] = {
    {"S", 50, 0.5, 1.0, 4, 4, 24, 24, 6},
    {"W", 300, 1.5e-3, 1.0, 4, 4, 64, 64, 8},
    {"A", 250, 2.0, 1.0, 4, 4, 128, 128, 16},
    {"B", 250, 2.0, 1.0, 4, 4, 304, 208, 17},
};

/* End indices of the zones along one axis (1-based, inclusive); geo selects
 * the geometric widths (ratio != 1 and more than one zone on both axes). */
static void zone_ends(int nzones, int g, double ratio, int geo, int use_round, int* end) {
  end[0] = 0;
  if (geo) {
    const double r = exp(log(ratio) / (nzones - 1));
    const double smallest = (double)g * (r - 1.0) / (pow(r, nzones) - 1.0);
    for (int i = 1; i <= nzones; ++i) {
      const double v = smallest * (pow(r, i) - 1.0) / (r - 1.0) + 0.45;
      end[i] = use_round ? (int)nearbyint(v) : (int)trunc(v);
    }
  } else {
    for (int i = 1; i <= nzones; ++i) end[i] = (i * g) / nzones;
  }
}

static double mflops_per_step(int bench, int n, const int32_t* nx, const int32_t* ny,
                              const int32_t* nz) {
  const double a = bench == NPBMZ_LU ? 1984.77 : 881.174, b = bench == NPBMZ_LU ? 10923.3 : 4683.91,
               c = bench == NPBMZ_LU ? 27770.9 : 11484.5, d = bench == NPBMZ_LU ? 144010.0 : 19272.4;
  double tot = 0.0;
  for (int z = 0; z < n; ++z) {
    const double n3 = (double)nx[z] * ny[z] * nz[z];
    const double navg = (nx[z] + ny[z] + nz[z]) / 3.0;
    const double nsur = ((double)nx[z] * ny[z] + (double)nx[z] * nz[z] + (double)ny[z] * nz[z]) / 3.0;
    tot += 1.0e-6 * (a * n3 - b * nsur + c * navg - d);
  }
  return tot;
}

/* The class table's verdict on a run (1 verified, 0 not, -1 no table). */
static int verified(const char* bench, const char* cls, double dt, int niter, const double xcr[5],
                    const double xce[5], double xci) {
  for (size_t t = 0; t < sizeof npbmz_vtables / sizeof npbmz_vtables[0]; ++t) {
    const npbmz_vtable* v = &npbmz_vtables[t];
    if (strcmp(v->bench, bench) || strcmp(v->cls, cls)) continue;
    int ok = fabs(dt - v->dt) <= npbmz_veps && niter == v->niter;
    for (int m = 0; m < 5; ++m) {
      ok = ok && fabs((xcr[m] - v->xcr[m]) / v->xcr[m]) <= npbmz_veps;
      ok = ok && fabs((xce[m] - v->xce[m]) / v->xce[m]) <= npbmz_veps;
    }
    if (v->has_xci) ok = ok && fabs((xci - v->xci) / v->xci) <= npbmz_veps;
    return ok;
  }
  return -1;
}

static int check(int rc, npbmz_ctx* ctx, const char* what) {
  if (rc != NPBMZ_OK) {
    fprintf(stderr, "npbmz_run: %s failed (%d): %s\n", what, rc, npbmz_last_error(ctx));
    exit(1);
  }
  return rc;
}

int main(int argc, char** argv) {
  if (argc < 3) {
    fprintf(stderr, "usage: %s bt|sp|lu S|W|A|B [perzone|batched|graph] [strict] [npbexch]\n", argv[0]);
    return 2;
  }
  int bench;
  const klass* table;
  if (!strcmp(argv[1], "bt")) { bench = NPBMZ_BT; table = bt_classes; }
  else if (!strcmp(argv[1], "sp")) { bench = NPBMZ_SP; table = sp_classes; }
  else if (!strcmp(argv[1], "lu")) { bench = NPBMZ_LU; table = lu_classes; }
  else { fprintf(stderr, "unknown benchmark %s\n", argv[1]); return 2; }
  const klass* k = NULL;
  for (int i = 0; i < 4; ++i)
    if (!strcmp(argv[2], table[i].name)) k = &table[i];
  if (!k) { fprintf(stderr, "unknown class %s\n", argv[2]); return 2; }
  int mode = NPBMZ_MODE_GRAPH, flags = 0, zones_only = 0;
  for (int a = 3; a < argc; ++a) {
    if (!strcmp(argv[a], "zones")) zones_only = 1;
    else if (!strcmp(argv[a], "perzone")) mode = NPBMZ_MODE_PERZONE;
    else if (!strcmp(argv[a], "batched")) mode = NPBMZ_MODE_BATCHED;
    else if (!strcmp(argv[a], "graph")) mode = NPBMZ_MODE_GRAPH;
    else if (!strcmp(argv[a], "strict")) flags |= NPBMZ_FLAG_STRICT_FP;
    else if (!strcmp(argv[a], "npbexch")) flags |= NPBMZ_FLAG_LU_NPB_EXCH;
    else { fprintf(stderr, "unknown option %s\n", argv[a]); return 2; }
  }

  /* zone split, numbered from 0 with x fastest, periodic neighbours */
  const int xz = k->x_zones, yz = k->y_zones, n = xz * yz;
  int xe[xz + 1], ye[yz + 1];
  const int geo = fabs(k->ratio - 1.0) > 1.0e-10 && xz > 1 && yz > 1;
  zone_ends(xz, k->gx, k->ratio, geo, bench == NPBMZ_LU, xe);
  zone_ends(yz, k->gy, k->ratio, geo, bench == NPBMZ_LU, ye);
  int32_t nx[n], ny[n], nz[n], w[n], e[n], s[n], no[n];
  for (int j = 1; j <= yz; ++j)
    for (int i = 1; i <= xz; ++i) {
      const int z = (i - 1) + (j - 1) * xz;
      nx[z] = xe[i] - xe[i - 1];
      ny[z] = ye[j] - ye[j - 1];
      nz[z] = k->gz;
      w[z] = (i - 2 + xz) % xz + (j - 1) * xz;
      e[z] = i % xz + (j - 1) * xz;
      s[z] = (i - 1) + ((j - 2 + yz) % yz) * xz;
      no[z] = (i - 1) + (j % yz) * xz;
    }
  if (zones_only) {
    const int32_t* t[7] = {nx, ny, nz, w, e, s, no};
    const char* names[7] = {"nx", "ny", "nz", "west", "east", "south", "north"};
    printf("{");
    for (int q = 0; q < 7; ++q) {
      printf("%s\"%s\": [", q ? ", " : "", names[q]);
      for (int z = 0; z < n; ++z) printf("%s%d", z ? ", " : "", t[q][z]);
      printf("]");
    }
    printf("}\n");
    return 0;
  }
  npbmz_problem p = {bench, n, xz, yz, k->gx, k->gy, k->gz, nx, ny, nz, w, e, s, no, k->dt,
                     bench == NPBMZ_LU ? 1.2 : 0.0 /* src/LU/lu_data.jl:8 */};

  npbmz_ctx* ctx = NULL;
  check(npbmz_create(&p, flags, &ctx), NULL, "npbmz_create");
  check(npbmz_init_exact(ctx), ctx, "npbmz_init_exact");
  check(npbmz_step(ctx, 1, mode, NULL), ctx, "warm-up step");
  check(npbmz_init_exact(ctx), ctx, "npbmz_init_exact");
  double ms = 0.0;
  check(npbmz_step(ctx, k->niter, mode, &ms), ctx, "npbmz_step");
  double xcr[5], xce[5], xci = 0.0;
  check(npbmz_norms(ctx, xcr, xce, &xci), ctx, "npbmz_norms");
  npbmz_destroy(ctx);

  const double mops = mflops_per_step(bench, n, nx, ny, nz) * k->niter / (ms * 1e-3);
  const char* bname = bench == NPBMZ_BT ? "BT" : bench == NPBMZ_SP ? "SP" : "LU";
  const int ok = verified(bname, k->name, k->dt, k->niter, xcr, xce, xci);
  printf("{\"bench\": \"%s\", \"class\": \"%s\", \"niter\": %d, \"dt\": %.17g, \"ms\": %.6f, "
         "\"mops\": %.3f, \"xcr\": [%.17g, %.17g, %.17g, %.17g, %.17g], "
         "\"xce\": [%.17g, %.17g, %.17g, %.17g, %.17g], \"xci\": %.17g, \"verified\": %s}\n",
         bname, k->name, k->niter, k->dt, ms, mops, xcr[0], xcr[1], xcr[2], xcr[3], xcr[4], xce[0],
         xce[1], xce[2], xce[3], xce[4], xci, ok < 0 ? "null" : ok ? "true" : "false");
  return 0;
}
