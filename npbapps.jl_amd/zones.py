"""Zone geometry and class tables of the multi-zone benchmarks.

Host-side setup, as in the reference: the Julia driver computes the zone
split and the class parameters and hands them to the node loop, which is the
part libnpbmz replaces.
"""
import math
from dataclasses import dataclass, field
from typing import List


@dataclass
class Params:
    """One benchmark class (src/BT/classes.jl:83-104, src/SP/classes.jl,
    src/LU/classes.jl)."""
    niter: int
    dt: float
    ratio: float
    x_zones: int
    y_zones: int
    gx: int
    gy: int
    gz: int


# src/BT/classes.jl:95-104 (sp_class, as the reference names it)
BT_CLASSES = {
    "S": Params(60, 0.010, 3.0, 2, 2, 24, 24, 6),
    "W": Params(200, 0.0008, 4.5, 4, 4, 64, 64, 8),
    "A": Params(200, 0.0008, 4.5, 4, 4, 128, 128, 16),
    "B": Params(200, 0.0003, 4.5, 8, 8, 304, 208, 17),
    "C": Params(200, 0.0001, 4.5, 16, 16, 480, 320, 28),
    "D": Params(250, 0.00002, 4.5, 32, 32, 1632, 1216, 34),
    "E": Params(250, 0.000004, 4.5, 64, 64, 4224, 3456, 92),
    "F": Params(250, 0.000001, 4.5, 128, 128, 12032, 8960, 250),
}

# src/SP/classes.jl:14-24
SP_CLASSES = {
    "S": Params(100, 0.015, 1.0, 2, 2, 24, 24, 6),
    "W": Params(400, 0.0015, 1.0, 4, 4, 64, 64, 8),
    "A": Params(400, 0.0015, 1.0, 4, 4, 128, 128, 16),
    "B": Params(400, 0.001, 1.0, 8, 8, 304, 208, 17),
    "C": Params(400, 0.00067, 1.0, 16, 16, 480, 320, 28),
    "D": Params(500, 0.0003, 1.0, 32, 32, 1632, 1216, 34),
    "E": Params(500, 0.0002, 1.0, 64, 64, 4224, 3456, 92),
    "F": Params(500, 0.0001, 1.0, 128, 128, 12032, 8960, 250),
}

# src/LU/classes.jl:14-24 (itmax as niter; inorm is not used on this path)
LU_CLASSES = {
    "S": Params(50, 0.5, 1.0, 4, 4, 24, 24, 6),
    "W": Params(300, 1.5e-3, 1.0, 4, 4, 64, 64, 8),
    "A": Params(250, 2.0, 1.0, 4, 4, 128, 128, 16),
    "B": Params(250, 2.0, 1.0, 4, 4, 304, 208, 17),
    "C": Params(250, 2.0, 1.0, 4, 4, 480, 320, 28),
    "D": Params(300, 1.0, 1.0, 4, 4, 1632, 1216, 34),
    "E": Params(300, 0.5, 1.0, 4, 4, 4224, 3456, 92),
    "F": Params(300, 0.2, 1.0, 4, 4, 12032, 8960, 250),
}

CLASSES = {"BT": BT_CLASSES, "SP": SP_CLASSES, "LU": LU_CLASSES}


@dataclass
class Zones:
    x_zones: int
    y_zones: int
    gx: int
    gy: int
    gz: int
    nx: List[int] = field(default_factory=list)
    ny: List[int] = field(default_factory=list)
    nz: List[int] = field(default_factory=list)
    west: List[int] = field(default_factory=list)
    east: List[int] = field(default_factory=list)
    south: List[int] = field(default_factory=list)
    north: List[int] = field(default_factory=list)

    @property
    def count(self):
        return self.x_zones * self.y_zones

    @property
    def npoints(self):
        return sum(a * b * c for a, b, c in zip(self.nx, self.ny, self.nz))


def zone_setup(x_zones, y_zones, gx, gy, gz, ratio, use_round=False):
    """src/BT/zone_setup.jl:1-77 (SP identical; LU rounds instead of
    truncating, src/LU/zone_setup.jl:15,19).  Zones are numbered from 0 with
    i (x) fastest; neighbours are periodic."""
    x_end = [0] * (x_zones + 1)
    y_end = [0] * (y_zones + 1)
    if abs(ratio - 1.0) > 1.0e-10 and x_zones > 1 and y_zones > 1:
        x_r = math.exp(math.log(ratio) / (x_zones - 1))
        y_r = math.exp(math.log(ratio) / (y_zones - 1))
        x_smallest = float(gx) * (x_r - 1.0) / (x_r ** x_zones - 1.0)
        y_smallest = float(gy) * (y_r - 1.0) / (y_r ** y_zones - 1.0)
        rnd = (lambda v: int(round(v))) if use_round else (lambda v: int(math.trunc(v)))
        for i in range(1, x_zones + 1):
            x_end[i] = rnd(x_smallest * (x_r ** i - 1.0) / (x_r - 1.0) + 0.45)
        for j in range(1, y_zones + 1):
            y_end[j] = rnd(y_smallest * (y_r ** j - 1.0) / (y_r - 1.0) + 0.45)
    else:
        for i in range(1, x_zones + 1):
            x_end[i] = (i * gx) // x_zones
        for j in range(1, y_zones + 1):
            y_end[j] = (j * gy) // y_zones
    x_start = [0] * (x_zones + 2)
    y_start = [0] * (y_zones + 2)
    x_start[1] = 1
    x_size = [0] * (x_zones + 1)
    y_size = [0] * (y_zones + 1)
    for i in range(1, x_zones + 1):
        if i != x_zones:
            x_start[i + 1] = x_end[i] + 1
        x_size[i] = x_end[i] - x_start[i] + 1
    y_start[1] = 1
    for j in range(1, y_zones + 1):
        if j != y_zones:
            y_start[j + 1] = y_end[j] + 1
        y_size[j] = y_end[j] - y_start[j] + 1
    z = Zones(x_zones, y_zones, gx, gy, gz)
    n = x_zones * y_zones
    for lst in (z.nx, z.ny, z.nz, z.west, z.east, z.south, z.north):
        lst.extend([0] * n)
    for j in range(1, y_zones + 1):
        for i in range(1, x_zones + 1):
            zone = (i - 1) + (j - 1) * x_zones
            z.nx[zone] = x_size[i]
            z.ny[zone] = y_size[j]
            z.nz[zone] = gz
            z.west[zone] = (i - 2 + x_zones) % x_zones + (j - 1) * x_zones
            z.east[zone] = i % x_zones + (j - 1) * x_zones
            z.south[zone] = (i - 1) + ((j - 2 + y_zones) % y_zones) * x_zones
            z.north[zone] = (i - 1) + (j % y_zones) * x_zones
    return z


def zones_for(bench, cls):
    p = CLASSES[bench][cls]
    return zone_setup(p.x_zones, p.y_zones, p.gx, p.gy, p.gz, p.ratio,
                      use_round=(bench == "LU")), p


def mflops_per_step(bench, zones):
    """Repo flop count per step, summed over zones: src/BT/bt-driver.jl:131-140,
    src/SP/sp-driver.jl:143-151, src/LU/lu-driver.jl:142-151 (LU evaluated with
    the zone's nx; the reference reads a never-filled nx1 there)."""
    if bench == "LU":
        a, b, c, d = 1984.77, 10923.3, 27770.9, 144010.0
    else:
        a, b, c, d = 881.174, 4683.91, 11484.5, 19272.4
    tot = 0.0
    for nx, ny, nz in zip(zones.nx, zones.ny, zones.nz):
        n3 = float(nx) * ny * nz
        navg = (nx + ny + nz) / 3.0
        nsur = (nx * ny + nx * nz + ny * nz) / 3.0
        tot += 1.0e-6 * (a * n3 - b * nsur + c * navg - d)
    return tot
