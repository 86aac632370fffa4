"""The C caller of the ABI (npbapps.jl_amd/harness/npbmz_run.c) on the GPU:
whole class runs through the C entry points alone verify against the
reference's class tables (by the harness's own check and by verify.py) and
reproduce the Python binding's norms bit for bit (same library, same
launches)."""
import json
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "npbapps.jl_amd", "harness", "npbmz_run")


@pytest.mark.parametrize("bench,cls,opts", [("bt", "S", []), ("sp", "S", []),
                                            ("lu", "S", ["npbexch"]), ("bt", "W", ["perzone"]),
                                            ("bt", "A", ["strict"])])
def test_c_harness_runs_a_class_and_verifies(gpu, bench, cls, opts):
    if not os.path.exists(EXE):
        pytest.fail("npbmz_run is not built (run __graft_entry__.build())")
    print("harness", bench, cls, opts, flush=True)
    r = subprocess.run([EXE, bench, cls, *opts], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    got = json.loads(r.stdout)
    B = bench.upper()
    zones, p = gpu.zones_for(B, cls)
    ok, rel = gpu.verify.verify(B, cls, p.dt, p.niter, got["xcr"], got["xce"],
                                got["xci"] if B == "LU" else None)
    assert ok, rel
    assert got["verified"] is True  # the harness's own check of the class table
    mode = "perzone" if "perzone" in opts else "graph"
    ref = gpu.go(B, cls, mode=mode, verbose=False, strict="strict" in opts,
                 npb_exchange="npbexch" in opts)
    assert got["xcr"] == list(ref["xcr"]) and got["xce"] == list(ref["xce"])
