"""The mixed-case workload: many grounding cases of uneven size from one seed.

Each case is drawn as follows (the batching configuration of the solve):

* type uniform over {vertical rod, horizontal counterpoise, rectangular grid};
* rod: length U(1, 30) m, from (0, 0, -0.5) straight down; counterpoise:
  length U(10, 100) m along x at depth 0.5 m; grid: sides U(10, 40) m, mesh
  U(2, 8) m (round(side / mesh) + 1 conductors each way), at depth 0.5 m;
* segments of about 0.5 m, their count kept in [20, 1500] (a grid's edge
  segments are reduced while it has more than 1500, raised while it has fewer
  than 20);
* radius U(5, 12) mm; resistivity log-uniform in [50, 5000] Ohm m; relative
  permittivity U(5, 40);
* nf uniform in {64, ..., 512} frequencies, logspace(2, 7, nf) Hz;
* unit current injected at node 0 (the rod's top, the counterpoise's start,
  the grid's corner).

Geometry comes from the library's host code (segment_electrode,
electrode_grid: src/electrode.c:104-127, src/grid.c:30-91), so the segment
and node order is the reference's."""
from __future__ import annotations

import numpy as np

SEG_LEN = 0.5
M_MIN, M_MAX = 20, 1500
DEPTH = 0.5


def _logspace(a, b, n):
    # src/auxiliary.c:39-50: 10^(a + i (b - a) / (n - 1)), both ends included
    return 10.0 ** np.linspace(a, b, n)


def _grid_shape(lx, ly, mesh):
    vx = max(2, int(round(lx / mesh)) + 1)
    vy = max(2, int(round(ly / mesh)) + 1)
    ex = max(1, int(round(lx / (vx - 1) / SEG_LEN)))
    ey = max(1, int(round(ly / (vy - 1) / SEG_LEN)))

    def segs(ex, ey):
        return ex * (vx - 1) * vy + ey * (vy - 1) * vx

    while segs(ex, ey) > M_MAX and max(ex, ey) > 1:
        if ex >= ey:
            ex -= 1
        else:
            ey -= 1
    while segs(ex, ey) < M_MIN:
        ex += 1
        ey += 1
    return vx, vy, ex, ey


def case_params(seed=2024, count=512):
    """The parameters of `count` cases drawn from `seed` (no geometry)."""
    rng = np.random.default_rng(seed)
    out = []
    for c in range(count):
        kind = ("rod", "counterpoise", "grid")[int(rng.integers(0, 3))]
        radius = float(rng.uniform(5e-3, 12e-3))
        rho = float(np.exp(rng.uniform(np.log(50.0), np.log(5000.0))))
        epsr = float(rng.uniform(5.0, 40.0))
        nf = int(rng.integers(64, 513))
        p = dict(index=c, kind=kind, radius=radius, rho=rho, sigma=1.0 / rho, epsr=epsr, nf=nf)
        if kind == "grid":
            lx, ly = (float(v) for v in rng.uniform(10.0, 40.0, 2))
            mesh = float(rng.uniform(2.0, 8.0))
            vx, vy, ex, ey = _grid_shape(lx, ly, mesh)
            p.update(lx=lx, ly=ly, mesh=mesh, vx=vx, vy=vy, ex=ex, ey=ey)
            p["m"] = ex * (vx - 1) * vy + ey * (vy - 1) * vx
            p["n"] = vx * vy + (ex - 1) * (vx - 1) * vy + (ey - 1) * (vy - 1) * vx
        else:
            length = float(rng.uniform(1.0, 30.0) if kind == "rod" else rng.uniform(10.0, 100.0))
            p["length"] = length
            p["m"] = int(np.clip(round(length / SEG_LEN), M_MIN, M_MAX))
            p["n"] = p["m"] + 1
        p["N"] = p["n"] + 2 * p["m"]
        p["freqs"] = _logspace(2.0, 7.0, nf)
        out.append(p)
    return out


def build_case(T, p):
    """Electrodes and nodes of one case (T: the package module)."""
    if p["kind"] == "grid":
        rc, e, nodes = T.electrode_grid(p["vx"], p["vy"], p["lx"], p["ly"], p["ex"], p["ey"],
                                        p["radius"], -DEPTH)
    elif p["kind"] == "rod":
        rc, e, nodes = T.segment_electrode([0.0, 0.0, -DEPTH], [0.0, 0.0, -DEPTH - p["length"]],
                                           p["m"], p["radius"])
    else:
        rc, e, nodes = T.segment_electrode([0.0, 0.0, -DEPTH], [p["length"], 0.0, -DEPTH],
                                           p["m"], p["radius"])
    if rc:
        raise RuntimeError("geometry of case %d failed: %d" % (p["index"], rc))
    return e, nodes


def case_family(T, seed=2024, count=512, freq_step=1):
    """Cases ready for T.Cases: dicts with electrodes, nodes, inj_node, s and
    kappa (every freq_step-th frequency of each case)."""
    out = []
    for p in case_params(seed, count):
        e, nodes = build_case(T, p)
        f = p["freqs"][::freq_step]
        s = 2j * np.pi * f
        kap = p["sigma"] + s * p["epsr"] * T.EPS0
        q = dict(p)
        q.update(electrodes=e, nodes=nodes, inj_node=0, f=f, s=s, kappa=kap)
        out.append(q)
    return out
