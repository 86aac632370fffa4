"""Host driver mirroring the reference's run/verify API (``go(class)``).

The reference's ``go`` (src/BT/bt-driver.jl:55-176) sets up zones, hands the
node loop every zone, times the step loop (timer 1, src/BT/bt-node.jl:266-420),
verifies the norms against the class table and prints Mop/s.  Here the node
loop is libnpbmz on one GPU; everything else keeps the reference's meaning.
"""
import time

from . import native, verify as _verify
from .zones import CLASSES, mflops_per_step, zones_for


def go(bench="BT", cls="S", mode="graph", niter=None, strict=False, warmup=True,
       verbose=True, tune_ms=0, npb_exchange=False):
    """Run one benchmark class end to end on the current GPU.

    bench: "BT", "SP" or "LU".  npb_exchange (LU): exchange planes
    k = 2..nz-1 as NPB-MZ does; the reference's own planes k = 1..nz-2
    (the default) do not reproduce its verification tables.
    mode: "perzone" (one launch per zone per phase), "batched" (one launch per
    phase over all zones) or "graph" (batched timesteps replayed from
    hipGraphs of eight steps and of one).  strict: use the kernels built without FMA contraction
    (bit-identical arithmetic to the reference's CPU path).
    Returns a dict with the norms, verification verdict, time and Mop/s.
    """
    zones, p = zones_for(bench, cls)
    niter = p.niter if niter is None else int(niter)
    t0 = time.time()
    # omega: src/LU/lu_data.jl:8 (BT and SP ignore it)
    ctx = native.Context(bench, zones, p.dt, omega=1.2, strict=strict, npb_exchange=npb_exchange)
    try:
        ctx.init_exact()
        if tune_ms:
            ctx.tune(tune_ms)
        if warmup:
            # bt-node.jl:138-253 (lu-node.jl:203-311): one untimed step, then
            # the state is re-initialised
            ctx.step(1, mode)
            ctx.init_exact()
        setup_s = time.time() - t0
        t = ctx.step(niter, mode)
        xcr, xce, xci = ctx.norms()
        sched = ctx.describe()
    finally:
        ctx.close()
    ok, rel = _verify.verify(bench, cls, p.dt, niter, xcr, xce, xci)
    mflops = mflops_per_step(bench, zones) * niter / t if t > 0 else 0.0
    res = {
        "bench": bench, "class": cls, "mode": mode, "niter": niter, "dt": p.dt,
        "grid": (p.gx, p.gy, p.gz), "zones": (p.x_zones, p.y_zones), "npoints": zones.npoints,
        "time_s": t, "mops": mflops, "gpts_per_s": zones.npoints * niter / t if t > 0 else 0.0,
        "xcr": list(xcr), "xce": list(xce), "xci": xci, "verified": ok, "rel": rel,
        "setup_s": setup_s, "schedule": sched, "strict": strict, "npb_exchange": npb_exchange,
    }
    if verbose:
        print_results(res)
    return res


def print_results(r):
    """Result banner in the spirit of src/common/print_results.jl:3-50."""
    print(f"\n {r['bench']}-MZ Benchmark Completed.")
    print(f" Class           =                        {r['class']}")
    print(f" Size            =            {r['grid'][0]:4d}x{r['grid'][1]:4d}x{r['grid'][2]:4d}")
    print(f" Zones           =                    {r['zones'][0]:3d}x{r['zones'][1]:3d}")
    print(f" Iterations      =                    {r['niter']:6d}")
    print(f" Time in seconds =             {r['time_s']:12.6f}")
    print(f" Mop/s total     =             {r['mops']:12.2f}")
    print(f" Operation type  =           floating point")
    print(f" Verification    =             {'SUCCESSFUL' if r['verified'] else 'UNSUCCESSFUL'}")
    print(f" Mode            =             {r['mode']} ({r['schedule']})")
    if r["rel"]:
        for m in range(5):
            print(f"   {m + 1} xcr {r['xcr'][m]:20.13E} rel {r['rel']['xcr'][m]:.2E}   "
                  f"xce {r['xce'][m]:20.13E} rel {r['rel']['xce'][m]:.2E}")
        if r["rel"].get("xci") is not None:
            print(f"     xci {r['xci']:20.13E} rel {r['rel']['xci']:.2E}")
