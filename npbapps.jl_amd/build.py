"""Build libnpbmz for gfx950 with hipcc, in place (the .so sits next to this file)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc", "npbmz.hip")
OUT = os.path.join(HERE, "libnpbmz.so")

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # FMA contraction is controlled per kernel set by #pragma clang fp contract
    "-ffp-contract=fast-honor-pragmas",
    # keep hipcc from emitting packed FP32 arithmetic
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def build_library(verbose=False):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    tmp = OUT + ".tmp"
    cmd = [hipcc, *HIPCC_FLAGS, SRC, "-o", tmp]
    if verbose:
        print(" ".join(cmd))
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed ({r.returncode}):\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, OUT)
    return OUT


HARNESS_SRC = os.path.join(HERE, "harness", "npbmz_run.c")
HARNESS_OUT = os.path.join(HERE, "harness", "npbmz_run")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")


def build_harness(verbose=False):
    """The C caller of the ABI (harness/npbmz_run.c), linked against the
    in-tree libnpbmz with an rpath to it."""
    cc = os.environ.get("CC", "gcc")
    cmd = [cc, "-O2", "-std=c11", "-Wall", "-Wextra", "-I", INCLUDE, HARNESS_SRC, "-L", HERE,
           "-lnpbmz", "-Wl,-rpath,$ORIGIN/..", "-lm", "-o", HARNESS_OUT]
    if verbose:
        print(" ".join(cmd))
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"{cc} failed ({r.returncode}):\n{r.stdout}\n{r.stderr}")
    return HARNESS_OUT
