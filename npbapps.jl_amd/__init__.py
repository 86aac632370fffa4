"""npbapps_amd: MI355X (gfx950) multi-zone BT/SP/LU timestep.

The hot path (exch_qbc + per-zone adi/ssor of NPBApps.jl's node loop) runs as
hand-written HIP kernels in libnpbmz.so behind the C ABI of include/npbmz.h.
This package holds the ctypes binding and a host driver that mirrors the
reference's go(class) run/verify API.
"""
from . import native, verify, zones
from .driver import go, print_results
from .native import Context, NpbmzError
from .zones import CLASSES, mflops_per_step, zone_setup, zones_for

__all__ = ["go", "print_results", "Context", "NpbmzError", "CLASSES", "mflops_per_step",
           "zone_setup", "zones_for", "native", "verify", "zones"]
