#!/usr/bin/env python3
"""BT-MZ timestep throughput on one MI355X.

Workload: BT-MZ class B, 304x208x17 grid in 8x8 uneven zones, FP64, dt=0.0003
(the configuration the headline metric is quoted on).  A "step" is one
timestep over all zones (exch_qbc + compute_rhs + x/y/z block-tridiagonal
solves + add).  The timed region is --steps timesteps of the batched,
hipGraph-replayed form, with every field already resident in HBM; the same
count of per-zone-launch steps is timed for the batched/per-zone ratio, and a
batched pass with HIP events around each phase gives the dominant kernel's
duration for the roofline.  One JSON line is printed.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
# Algorithmic bytes per grid point and launch: 5-vector fields (40 B per
# point) each read or written once: rhs reads u and forcing and writes rhs;
# each solve reads u and rhs and writes rhs; the z solve with add folded in
# reads u and rhs and writes u.
FIELDS_PER_PHASE = {"exch": None, "rhs": 3, "x_solve": 3, "y_solve": 3, "z_solve+add": 3}
PHASES = ["exch", "rhs", "x_solve", "y_solve", "z_solve+add"]


def face_points(zones):
    n = 0
    for z in range(zones.count):
        n += 2 * (zones.ny[z] - 2) * (zones.nz[z] - 2) + 2 * (zones.nx[z] - 2) * (zones.nz[z] - 2)
    return n


def host_cpu():
    """This host's CPU model, the CPUs this process may run on, and the thread
    budget the environment grants (OMP_NUM_THREADS, if set)."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    avail = len(os.sched_getaffinity(0))
    cap = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return model, avail, (min(avail, cap) if cap > 0 else avail)


def oracle_full_run(params, threads):
    """The CPU oracle (C++ restatement of the reference's Julia BT path) over
    the whole class run: one warm-up step, re-initialisation (as the
    reference's node does, src/BT/bt-node.jl:138-253), then niter timed
    steps and the node-side norms.  Returns (seconds, xcr, xce)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import OracleBT
    o = OracleBT.for_class(params, threads=threads)
    o.step(1)
    o.close()
    o = OracleBT.for_class(params, threads=threads)
    t0 = time.perf_counter()
    o.step(params.niter)
    t = time.perf_counter() - t0
    xcr, xce = o.norms()
    o.close()
    return t, xcr, xce


def oracle_sample(params, budget_s, threads):
    """The oracle timed over a bounded number of steps (about budget_s)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import OracleBT
    o = OracleBT.for_class(params, threads=threads)
    t0 = time.perf_counter()
    o.step(1)
    t1 = time.perf_counter() - t0
    n = max(1, int(budget_s / max(t1, 1e-3)) - 1)
    t0 = time.perf_counter()
    o.step(n)
    t = time.perf_counter() - t0
    o.close()
    return n, t


def cpu_baseline(P, params, budget_s):
    """The reported CPU baseline: the oracle's whole class B run (200 steps)
    on this host's cores, OpenMP over zones as the reference threads over
    zones, plus a bounded single-core sample.  Also returns the oracle's
    end-of-run norms, which the GPU run is checked against."""
    zones, _ = P.zones_for("BT", "B")
    mflop = P.mflops_per_step("BT", zones)
    model, avail, threads = host_cpu()
    t, xcr, xce = oracle_full_run(params, threads)
    n1, t1 = oracle_sample(params, budget_s, 1)
    base = {"value": round(mflop * params.niter / t, 2), "unit": "Mop/s", "cores": threads,
            "kind": "port",
            "sample": f"BT-MZ class B, the whole {params.niter}-step run after 1 warm-up step, "
                      f"{threads} OpenMP threads over zones, {t:.2f} s",
            "cpu_model": model, "cpus_available": avail,
            "single_core": {"value": round(mflop * n1 / t1, 2), "unit": "Mop/s", "cores": 1,
                            "sample": f"BT-MZ class B, {n1} timesteps after 1 warm-up step, "
                                      f"1 thread, {t1:.2f} s"}}
    return base, xcr, xce


def adjoint_case(P):
    """The adjoint case: BT-MZ class A (128x128x16, 4x4 uneven zones), 200
    steps at dt=0.0008, loss sum(w u^2) with w uniform [0,1) (Philox seed 20),
    checkpoints every 20 steps.  The reverse time includes recomputing every
    segment from its checkpoint."""
    import numpy as np
    zones, p = P.zones_for("BT", "A")
    ctx = P.Context("BT", zones, p.dt)
    ctx.init_exact()
    w = np.random.Generator(np.random.Philox(20)).random(ctx.size)
    ctx.adjoint(p.niter, 20, w)  # warm-up
    r = ctx.adjoint(p.niter, 20, w)
    ctx.close()
    return {"workload": "BT-MZ class A, 4x4 uneven zones, 200 steps, checkpoint every 20",
            "forward_ms_per_step": round(r["fwd_s"] / p.niter * 1e3, 5),
            "adjoint_ms_per_step": round(r["adj_s"] / p.niter * 1e3, 5),
            "ratio": round(r["adj_s"] / r["fwd_s"], 3)}


def sp_adjoint_case(P):
    """The SP adjoint on class A (128x128x16, 4x4 equal zones), 100 steps,
    checkpoints every 20, loss sum(w u^2) with w uniform [0,1) (Philox seed
    20)."""
    import numpy as np
    zones, p = P.zones_for("SP", "A")
    ctx = P.Context("SP", zones, p.dt)
    ctx.init_exact()
    w = np.random.Generator(np.random.Philox(20)).random(ctx.size)
    ctx.adjoint(100, 20, w)  # warm-up
    r = ctx.adjoint(100, 20, w)
    ctx.close()
    return {"workload": "SP-MZ class A, 4x4 zones, 100 steps, checkpoint every 20",
            "forward_ms_per_step": round(r["fwd_s"] / 100 * 1e3, 5),
            "adjoint_ms_per_step": round(r["adj_s"] / 100 * 1e3, 5),
            "ratio": round(r["adj_s"] / r["fwd_s"], 3)}


def lu_adjoint_case(P):
    """The LU adjoint on class A (128x128x16, 4x4 equal zones, NPB's exchange
    planes), 100 iterations, checkpoints every 20, loss sum(w u^2) with w
    uniform [0,1) (Philox seed 20)."""
    import numpy as np
    zones, p = P.zones_for("LU", "A")
    ctx = P.Context("LU", zones, p.dt, omega=1.2, npb_exchange=True)
    ctx.init_exact()
    w = np.random.Generator(np.random.Philox(20)).random(ctx.size)
    ctx.adjoint(100, 20, w)  # warm-up
    r = ctx.adjoint(100, 20, w)
    ctx.close()
    return {"workload": "LU-MZ class A, 4x4 zones, 100 iterations, checkpoint every 20",
            "forward_ms_per_step": round(r["fwd_s"] / 100 * 1e3, 5),
            "adjoint_ms_per_step": round(r["adj_s"] / 100 * 1e3, 5),
            "ratio": round(r["adj_s"] / r["fwd_s"], 3)}


def class_run(P, bench, cls, npb_exchange=False):
    """A whole class run (graph mode) and the same steps launched per zone."""
    zones, p = P.zones_for(bench, cls)
    r = P.go(bench, cls, mode="graph", verbose=False, npb_exchange=npb_exchange, tune_ms=1000)
    ctx = P.Context(bench, zones, p.dt, omega=1.2, npb_exchange=npb_exchange)
    ctx.init_exact()
    n = min(p.niter, 50)
    ctx.step(2, "perzone")
    t_pz = ctx.step(n, "perzone") / n
    ctx.close()
    t = r["time_s"] / p.niter
    return {"mops": round(r["mops"], 2), "ms_per_step": round(t * 1e3, 5),
            "grid_pt_steps_per_s": round(zones.npoints / t, 1), "verified": r["verified"],
            "perzone_ms_per_step": round(t_pz * 1e3, 5), "graph_vs_perzone": round(t_pz / t, 3),
            "schedule": r["schedule"]}


SIDE_WORKLOADS = {
    "sp-a": ("SP", "A"), "sp-b": ("SP", "B"), "lu-a": ("LU", "A"), "lu-b": ("LU", "B"),
    "bt-a": ("BT", "A"),
}


def side_workload(P, name, tune_ms):
    """One secondary workload on its own (for its own timing and profile): a
    whole class run in graph mode with the library's default schedule (LU
    with NPB's exchange planes), or one of the three adjoint cases."""
    if name == "bt-a-adjoint":
        return {"workload": name, **adjoint_case(P)}
    if name == "sp-a-adjoint":
        return {"workload": name, **sp_adjoint_case(P)}
    if name == "lu-a-adjoint":
        return {"workload": name, **lu_adjoint_case(P)}
    bench, cls = SIDE_WORKLOADS[name]
    zones, p = P.zones_for(bench, cls)
    r = P.go(bench, cls, mode="graph", verbose=False, npb_exchange=bench == "LU",
             tune_ms=tune_ms)
    t = r["time_s"] / p.niter
    return {"workload": f"{bench}-MZ class {cls}, {zones.count} zones, {p.niter} steps",
            "mops": round(r["mops"], 2), "ms_per_step": round(t * 1e3, 5),
            "grid_pt_steps_per_s": round(zones.npoints / t, 1), "verified": r["verified"],
            "schedule": r["schedule"]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-perzone", action="store_true")
    ap.add_argument("--tune-ms", type=int, default=0,
                    help="autotune budget before timing (0: the library's default schedule, "
                         "the fastest the tuner finds for this class on MI355X)")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the adjoint ratio and the SP/LU class A runs")
    ap.add_argument("--copy-probe-mib", type=int, default=1024,
                    help="size of the plain copy kernel run for the achievable-bandwidth "
                         "reference and counter calibration (0 = skip)")
    ap.add_argument("--sustain-seconds", type=float, default=8.0,
                    help="after the timed region, repeat the whole class run for this long "
                         "and report the sustained rate and whether every repeat's norms "
                         "are bit-identical (0 = skip)")
    ap.add_argument("--workload", default="bt-b",
                    choices=["bt-b", *SIDE_WORKLOADS, "bt-a-adjoint", "sp-a-adjoint",
                             "lu-a-adjoint"],
                    help="bt-b (default): the headline line below; any other runs that "
                         "workload alone and prints its own JSON line")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for {args.gpus}); the "
              "multi-zone path would shard by zone across GPUs, which is not built here.",
              file=sys.stderr)
        sys.exit(2)

    import npbapps_amd as P
    if args.workload != "bt-b":
        print(json.dumps(side_workload(P, args.workload, args.tune_ms)))
        return
    zones, p = P.zones_for("BT", "B")
    K, W = args.steps, args.warmup
    mflop_step = P.mflops_per_step("BT", zones)
    npts = zones.npoints

    ctx = P.Context("BT", zones, p.dt)
    ctx.init_exact()
    if args.tune_ms > 0:
        ctx.tune(args.tune_ms)
    # the whole class run on the same context and schedule, outside the timed
    # region and before it (its replays also instantiate both graphs on the
    # device): its norms are checked against the class table and the oracle
    ctx.init_exact()
    t_full = ctx.step(p.niter, "graph")
    gxcr, gxce, _ = ctx.norms()
    verified, rel = P.verify.verify("BT", "B", p.dt, p.niter, gxcr, gxce)

    # the class run repeated back to back: the rate over seconds of
    # continuous work, and every repeat's norms against the first run's
    sustained = None
    if args.sustain_seconds > 0:
        import numpy as np
        runs, t_sum, same = 0, 0.0, True
        t0 = time.perf_counter()
        while time.perf_counter() - t0 < args.sustain_seconds:
            ctx.init_exact()
            t_sum += ctx.step(p.niter, "graph")
            r1, r2, _ = ctx.norms()
            same = same and np.array_equal(r1, gxcr) and np.array_equal(r2, gxce)
            runs += 1
        sustained = {"seconds": round(time.perf_counter() - t0, 2), "class_runs": runs,
                     "steps": runs * p.niter,
                     "mops": round(mflop_step * runs * p.niter / t_sum, 2),
                     "norms_identical": bool(same)}

    # the measurement: W untimed warm-up steps, then exactly K timed steps
    ctx.init_exact()
    ctx.step(W, "graph")
    ctx.init_exact()
    t_graph = ctx.step(K, "graph")

    t_perzone = None
    if not args.no_perzone:
        ctx.init_exact()
        ctx.step(W, "perzone")
        ctx.init_exact()
        t_perzone = ctx.step(K, "perzone")

    ctx.init_exact()
    ctx.step(W, "batched")
    ctx.init_exact()
    t_batched = ctx.step(K, "batched")
    # per-phase kernel times (HIP events around each phase's launch on the
    # library's stream) over at least 100 batched steps, whatever K is
    KP = max(K, 100)
    ctx.init_exact()
    ph = ctx.phase_times(KP)[:5]
    sched = ctx.describe()
    copy = None
    if args.copy_probe_mib > 0:
        nb = args.copy_probe_mib << 20
        tc = ctx.probe_copy(nb, 20)
        copy = {"bytes_per_launch": 2 * nb, "GBps": round(2 * nb / tc / 1e9, 1)}
    ctx.close()

    per_launch_s = {name: ph[i] / KP for i, name in enumerate(PHASES)}
    dom = max(PHASES[1:], key=lambda n: per_launch_s[n])
    alg_bytes = FIELDS_PER_PHASE[dom] * 40 * npts
    achieved = alg_bytes / per_launch_s[dom] / 1e9
    # measured HBM bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, scaled
    # on the copy kernel), used only while the kernel sources hash as they did
    # when it was measured
    traffic, traffic_src = None, None
    tfile = os.path.join(ROOT, "profiles", "bt_class_b_traffic.json")
    if os.path.exists(tfile):
        from npbapps_amd.rocprof_summary import kernel_sources_sha256
        rec = json.load(open(tfile))
        if rec.get("kernels_sha256") == kernel_sources_sha256():
            traffic = rec.get(dom)
            traffic_src = {"head": rec.get("head"), "kernel": rec.get("kernels", {}).get(dom)}
        else:
            traffic_src = {"head": rec.get("head"), "stale": True}

    extra = {}
    if not args.no_extra:
        extra["adjoint"] = adjoint_case(P)
        extra["sp_adjoint"] = sp_adjoint_case(P)
        extra["lu_adjoint"] = lu_adjoint_case(P)
        extra["sp_class_A"] = class_run(P, "SP", "A")
        extra["sp_class_B"] = class_run(P, "SP", "B")
        extra["lu_class_A"] = class_run(P, "LU", "A", npb_exchange=True)
        extra["lu_class_B"] = class_run(P, "LU", "B", npb_exchange=True)

    value = mflop_step * K / t_graph
    out = {
        "metric": "Mop/s & grid-pt-steps/s FP64 1 GPU: batched vs per-zone; adjoint/forward ratio",
        "value": round(value, 2),
        "unit": "Mop/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": round(t_graph / K * 1e3, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: the reference's exact-solution initial field and forcing "
                "(deterministic, no dataset)",
        "config": {"workload": "BT-MZ class B: 304x208x17 grid, 8x8 uneven zones "
                               "(x 16..72, y 11..49), 200 timesteps at dt=0.0003",
                   "mode": "batched + hipGraph", "fp": "fma", "schedule": sched},
        "grid_pt_steps_per_s": round(npts * K / t_graph, 1),
        "verified": verified,
        "oracle_rel_max": None,
        "batched_mops": round(mflop_step * K / t_batched, 2),
        "perzone": None if t_perzone is None else {
            "value": round(mflop_step * K / t_perzone, 2), "unit": "Mop/s",
            "ms_per_step": round(t_perzone / K * 1e3, 5)},
        "graph_vs_perzone": None if t_perzone is None else round(t_perzone / t_graph, 3),
        "adjoint_forward_ratio": extra.get("adjoint", {}).get("ratio"),
        "adjoint": extra.get("adjoint"),
        "sp_adjoint": extra.get("sp_adjoint"),
        "lu_adjoint": extra.get("lu_adjoint"),
        "sp_class_A": extra.get("sp_class_A"),
        "sp_class_B": extra.get("sp_class_B"),
        "lu_class_A": extra.get("lu_class_A"),
        "lu_class_B": extra.get("lu_class_B"),
        "phase_ms_per_step": {n: round(per_launch_s[n] * 1e3, 5) for n in PHASES},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "traffic_measured_on": traffic_src,
                     "algorithmic_bytes": alg_bytes, "copy_kernel": copy,
                     # the same launch priced with its measured HBM bytes (workspace included)
                     "measured_GBps": None if traffic is None
                     else round(traffic / per_launch_s[dom] / 1e9, 1),
                     "measured_frac": None if traffic is None
                     else round(traffic / per_launch_s[dom] / 1e9 / HBM_PEAK_GBS, 4)},
        "cpu_baseline": None,
    }
    out["sustained"] = sustained
    out["full_run"] = {"steps": p.niter, "ms_per_step": round(t_full / p.niter * 1e3, 5),
                       "verified": verified,
                       "class_table_rel_max": max(max(rel["xcr"]), max(rel["xce"])),
                       "oracle_rel_max": None}
    if not args.no_cpu:
        import numpy as np
        base, oxcr, oxce = cpu_baseline(P, p, args.cpu_seconds)
        out["cpu_baseline"] = base
        orel = max(np.abs(np.asarray(gxcr) / oxcr - 1).max(),
                   np.abs(np.asarray(gxce) / oxce - 1).max())
        out["full_run"]["oracle_rel_max"] = float(orel)
        out["oracle_rel_max"] = float(orel)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
