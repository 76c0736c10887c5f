"""GPU parity on a seeded sample of the mixed-case family: vertical rods,
horizontal counterpoises and rectangular grids with random sizes, radii, depths
and soils (the case mix of a many-case batch), each swept at a few frequencies
through the device-resident sweep in both precisions, against the oracle's
pipeline (Genz-Malik cubature at rel tol 1e-9, LAPACK zsysv) within 1e-7, and
mixed against FP64 within 1e-8. Sizes are kept to m <= 120 so that the oracle
finishes in seconds.

The vertical rods exercise segments along z (their images flip direction
only in z) and the injection at the top node."""
import numpy as np
import pytest

import tags_pkg

T = tags_pkg.load()
O = tags_pkg.oracle()

pytestmark = pytest.mark.gpu


def _case(seed):
    rng = np.random.default_rng(1000 + seed)
    kind = ("rod", "counterpoise", "grid")[seed % 3]
    radius = rng.uniform(5e-3, 12e-3)
    depth = 0.5
    if kind == "rod":
        length = rng.uniform(1.0, 30.0)
        nseg = int(np.clip(round(length / 0.5), 4, 60))
        e, nodes = O.segment_electrode([0, 0, -depth], [0, 0, -depth - length], nseg, radius)
    elif kind == "counterpoise":
        length = rng.uniform(10.0, 100.0)
        nseg = int(np.clip(round(length / 1.0), 10, 100))
        e, nodes = O.segment_electrode([0, 0, -depth], [length, 0, -depth], nseg, radius)
    else:
        lx, ly = rng.uniform(10.0, 40.0, 2)
        mesh = rng.uniform(4.0, 8.0)
        vx, vy = max(2, int(lx / mesh) + 1), max(2, int(ly / mesh) + 1)
        e, nodes = O.grid(vx, vy, lx, ly, 2, 2, radius, -depth)
    rho = float(np.exp(rng.uniform(np.log(50.0), np.log(5000.0))))
    epsr = rng.uniform(5.0, 40.0)
    return kind, e, nodes, 1.0 / rho, epsr


@pytest.mark.parametrize("seed", range(6))
def test_case_family_sweep_vs_oracle(seed):
    kind, e, nodes, sigma, epsr = _case(seed)
    m = len(e)
    assert m <= 120, (kind, m)
    f = np.array([100.0, 3e4, 2e6])
    print("case: seed %d %s m=%d sigma=%.3g epsr=%.1f" % (seed, kind, m, sigma, epsr),
          flush=True)
    s = 2j * np.pi * f
    kap = sigma + s * epsr * T.EPS0
    sw = T.Sweep(e, nodes)
    u64, _ = sw.solve(s, kap, 0, T.PREC_FP64)
    umx, st = sw.solve(s, kap, 0, T.PREC_MIXED)
    sw.close()
    zo = O.harmonic_impedance(e, nodes, f, sigma, epsr, 0, rel_tol=1e-9, threads=8)
    assert np.max(np.abs(u64 - zo) / np.abs(zo)) <= 1e-7
    assert np.max(np.abs(umx - u64) / np.abs(u64)) <= 1e-8
    assert np.all(u64.real > 0)
