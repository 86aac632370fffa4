"""Summaries of rocprofv3 output for libnpbmz's kernels.

rocprofv3 writes CSV files under its -d directory (per host and process):
``*_kernel_stats.csv`` for ``--kernel-trace --stats`` and
``*_counter_collection.csv`` for ``--pmc``.  This module turns them into the
per-kernel tables kept under profiles/ and into the HBM-traffic record that
bench.py reads for ``roofline.traffic``.

The traffic record carries the commit and a hash of the kernel sources it was
measured on; bench.py uses it only while the sources still hash the same, so
a kernel change can never be priced with another kernel's traffic.

Usage (after two --pmc runs, FETCH_SIZE then WRITE_SIZE, of the same command):
    python -m npbapps_amd.rocprof_summary traffic FETCH_DIR WRITE_DIR OUT.json
    python -m npbapps_amd.rocprof_summary stats TRACE_DIR
    python -m npbapps_amd.rocprof_summary f64 F64_DIR TRACE_DIR
    python -m npbapps_amd.rocprof_summary sq PATTERN PMC_DIR [PMC_DIR ...]
"""
import csv
import glob
import hashlib
import json
import os
import re
import subprocess
import sys
from collections import defaultdict

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
# the sources whose text determines the BT kernels' memory traffic
KERNEL_SOURCES = ("csrc/npbmz.hip", "csrc/bt_kernels.inc", "csrc/bt_rows.inc",
                  "csrc/npbmz_types.h")
COPY_KERNEL = "copy_probe_kernel"
# phase name (bench.py) -> pattern of the kernel that runs the phase (any of
# the line-solve variants, by direction and whether add is folded in)
PHASE_KERNELS = {
    "exch": r"npbmz::exch_kernel<",
    "rhs": r"npbmz::fast::bt::rhs_kernel<",
    "x_solve": r"npbmz::fast::bt::solve\w*_kernel<0, false[,>]",
    "y_solve": r"npbmz::fast::bt::solve\w*_kernel<1, false[,>]",
    "z_solve+add": r"npbmz::fast::bt::solve\w*_kernel<2, true[,>]",
}


def kernel_sources_sha256():
    h = hashlib.sha256()
    for rel in KERNEL_SOURCES:
        with open(os.path.join(HERE, rel), "rb") as f:
            h.update(rel.encode() + b"\0" + f.read())
    return h.hexdigest()


def git_head():
    try:
        return subprocess.run(["git", "-C", ROOT, "rev-parse", "--short=12", "HEAD"],
                              capture_output=True, text=True, check=True).stdout.strip()
    except (OSError, subprocess.CalledProcessError):
        return None


def _rows(d, suffix):
    files = sorted(glob.glob(os.path.join(d, "**", f"*{suffix}"), recursive=True))
    if not files:
        raise FileNotFoundError(f"no *{suffix} under {d}")
    for fn in files:
        with open(fn, newline="") as f:
            yield from csv.DictReader(f)


def counter_per_dispatch(d, counter):
    """Average value of `counter` per dispatch, by kernel name."""
    tot, cnt = defaultdict(float), defaultdict(set)
    for r in _rows(d, "counter_collection.csv"):
        if r.get("Counter_Name") != counter:
            continue
        k = r["Kernel_Name"]
        tot[k] += float(r["Counter_Value"])
        cnt[k].add(r.get("Dispatch_Id") or r.get("Correlation_Id"))
    return {k: tot[k] / max(1, len(cnt[k])) for k in tot}


def kernel_stats(d):
    """(name, calls, avg us, min us, max us, percent) rows, slowest total first."""
    out = []
    for r in _rows(d, "kernel_stats.csv"):
        out.append((r["Name"], int(r["Calls"]), float(r["AverageNs"]) / 1e3,
                    float(r["MinNs"]) / 1e3, float(r["MaxNs"]) / 1e3, float(r["Percentage"])))
    return sorted(out, key=lambda x: -x[1] * x[2])


def traffic(fetch_dir, write_dir, copy_bytes):
    """Scaled HBM bytes per dispatch of each BT phase kernel.  FETCH_SIZE and
    WRITE_SIZE (KiB per dispatch) are each scaled by what they read for the
    copy kernel, whose reads and writes are copy_bytes each."""
    fe = counter_per_dispatch(fetch_dir, "FETCH_SIZE")
    wr = counter_per_dispatch(write_dir, "WRITE_SIZE")
    ck = [k for k in fe if COPY_KERNEL in k]
    if not ck:
        raise ValueError("the copy kernel is not in the FETCH_SIZE run")
    fscale = copy_bytes / (fe[ck[0]] * 1024.0)
    wscale = copy_bytes / (wr[ck[0]] * 1024.0)
    rec = {"head": git_head(), "kernels_sha256": kernel_sources_sha256(),
           "fetch_scale": round(fscale, 4), "write_scale": round(wscale, 4), "kernels": {}}
    for phase, pat in PHASE_KERNELS.items():
        kf = [k for k in fe if re.search(pat, k)]
        if not kf:
            continue
        k = max(kf, key=lambda name: fe[name])
        b = fe[k] * 1024.0 * fscale + wr.get(k, 0.0) * 1024.0 * wscale
        rec[phase] = int(round(b))
        rec["kernels"][phase] = k
    return rec


def f64_rates(f64_dir, trace_dir):
    """(kernel, FMA, ADD, MUL, TRANS, GFLOP per launch, avg us, TFLOP/s)
    with flops = 64 x (2 FMA + ADD + MUL + TRANS) per dispatch: every lane of
    every issued FP64 wave instruction."""
    names = ("SQ_INSTS_VALU_FMA_F64", "SQ_INSTS_VALU_ADD_F64", "SQ_INSTS_VALU_MUL_F64",
             "SQ_INSTS_VALU_TRANS_F64")
    c = [counter_per_dispatch(f64_dir, n) for n in names]
    dur = {n: a for n, _, a, _, _, _ in kernel_stats(trace_dir)}
    out = []
    for k in c[0]:
        if k not in dur:
            continue
        fma, add, mul, tr = (c[i].get(k, 0.0) for i in range(4))
        gf = 64.0 * (2 * fma + add + mul + tr) / 1e9
        out.append((k, fma, add, mul, tr, gf, dur[k], gf / dur[k] * 1e6 / 1e3))
    return sorted(out, key=lambda x: -x[5])


# SQ counters shown as shares of the waves' lifetime (SQ_WAVE_CYCLES)
SQ_SHARES = (("SQ_ACTIVE_INST_VALU", "issuing VALU"), ("SQ_ACTIVE_INST_LDS", "issuing LDS"),
             ("SQ_ACTIVE_INST_SCA", "issuing scalar"), ("SQ_WAIT_INST_ANY", "ready, not issued"),
             ("SQ_WAIT_ANY", "waiting on counters or barrier"))


def sq_table(pattern, dirs):
    """Per kernel matching `pattern`: each SQ_SHARES counter as a share of
    SQ_WAVE_CYCLES, and VALU and LDS instructions per wave, from --pmc runs
    of the same command (one or more passes)."""
    names = ["SQ_WAVES", "SQ_WAVE_CYCLES", "SQ_INSTS_VALU", "SQ_INSTS_LDS"] + [n for n, _ in SQ_SHARES]
    val = defaultdict(dict)
    for d in dirs:
        for n in names:
            try:
                for k, v in counter_per_dispatch(d, n).items():
                    if re.search(pattern, k):
                        val[k][n] = v
            except FileNotFoundError:
                pass
    rows = []
    for k, v in val.items():
        wc, wv = v.get("SQ_WAVE_CYCLES"), v.get("SQ_WAVES")
        if not wc:
            continue
        shares = [100.0 * v[n] / wc if n in v else None for n, _ in SQ_SHARES]
        per = [v[n] / wv if wv and n in v else None for n in ("SQ_INSTS_VALU", "SQ_INSTS_LDS")]
        rows.append((k.split("(")[0].replace("void ", ""), wv, shares, per))
    return rows


def main(argv):
    if len(argv) >= 4 and argv[0] == "traffic":
        copy = int(argv[4]) if len(argv) > 4 else 512 << 20
        rec = traffic(argv[1], argv[2], copy)
        with open(argv[3], "w") as f:
            json.dump(rec, f, indent=1)
            f.write("\n")
        print(json.dumps(rec, indent=1))
        return 0
    if len(argv) == 2 and argv[0] == "stats":
        print("| kernel | calls | avg us | min us | max us | % |\n|---|---|---|---|---|---|")
        for n, c, a, lo, hi, pc in kernel_stats(argv[1]):
            print(f"| `{n}` | {c} | {a:.1f} | {lo:.1f} | {hi:.1f} | {pc:.1f} |")
        return 0
    if len(argv) == 3 and argv[0] == "f64":
        print("| kernel | FMA | ADD | MUL | TRANS | GFLOP per launch | avg us | TFLOP/s | of peak |")
        print("|---|---|---|---|---|---|---|---|---|")
        for k, fma, add, mul, tr, gf, us, tf in f64_rates(argv[1], argv[2]):
            if gf < 0.01:
                continue
            print(f"| `{k.split('(')[0].replace('void ', '')}` | {fma:.3g} | {add:.3g} | {mul:.3g} | "
                  f"{tr:.3g} | {gf:.3f} | {us:.1f} | {tf:.2f} | {tf / 78.6 * 100:.1f}% |")
        return 0
    if len(argv) >= 3 and argv[0] == "sq":
        head = " | ".join(f"{label} %" for _, label in SQ_SHARES)
        print(f"| kernel | waves | {head} | VALU insts per wave | LDS insts per wave |")
        print("|---" * (4 + len(SQ_SHARES)) + "|")
        f = lambda x, p: "-" if x is None else f"{x:.{p}f}"
        for k, wv, shares, per in sq_table(argv[1], argv[2:]):
            print(f"| `{k}` | {wv:.0f} | " + " | ".join(f(x, 1) for x in shares) +
                  f" | {f(per[0], 0)} | {f(per[1], 0)} |")
        return 0
    print(__doc__)
    return 2


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
