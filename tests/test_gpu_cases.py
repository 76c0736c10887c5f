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


# ---------------------------------------------------------------------------
# many cases at once (tags_cases_*): the seeded 512-case workload
# ---------------------------------------------------------------------------
C = tags_pkg.cases()


def test_cases_subset_vs_oracle():
    """The first 14 cases of the 512-case family with m <= 120 (rods,
    counterpoises, small grids; random radii and soils), four frequencies each
    across 100 Hz - 10 MHz, solved together through tags_cases in both
    precisions: FP64 against the oracle's pipeline (Genz-Malik cubature at rel
    tol 1e-9, LAPACK zsysv) within 1e-7, mixed against FP64 within 1e-8."""
    fam = [c for c in C.case_family(T) if c["m"] <= 120][:14]
    cases = []
    for c in fam:
        sel = np.unique([0, c["nf"] // 3, 2 * c["nf"] // 3, c["nf"] - 1])
        cases.append(dict(c, s=c["s"][sel], kappa=c["kappa"][sel], f=c["f"][sel]))
    print("cases: %s" % [(c["index"], c["kind"], c["m"]) for c in cases], flush=True)
    cs = T.Cases(cases)
    u64, st64 = cs.solve(T.PREC_FP64)
    umx, stmx = cs.solve(T.PREC_MIXED)
    cs.close()
    for c, a, b, st in zip(cases, u64, umx, stmx):
        zo = O.harmonic_impedance(c["electrodes"], c["nodes"], c["f"], c["sigma"], c["epsr"], 0,
                                  rel_tol=1e-9, threads=8)
        assert np.max(np.abs(a - zo) / np.abs(zo)) <= 1e-7, c["index"]
        assert np.max(np.abs(b - a) / np.abs(a)) <= 1e-8, c["index"]
        assert all(q["precision_used"] in (0, 1) for q in st)


def test_cases_full_family_mixed_vs_fp64():
    """All 512 cases (N = 61 .. 4246), every 16th frequency of each (9,300
    samples), through one multi-case solve per precision: every mixed sample
    within 1e-8 of its FP64 value, and the mixed samples meet the acceptance
    rule."""
    fam = C.case_family(T, freq_step=16)
    cs = T.Cases(fam)
    u64, _ = cs.solve(T.PREC_FP64)
    umx, stmx = cs.solve(T.PREC_MIXED)
    cs.close()
    worst = max(float(np.max(np.abs(b - a) / np.abs(a))) for a, b in zip(u64, umx))
    nfb = sum(q["precision_used"] == 0 for st in stmx for q in st)
    print("full family: %d samples, worst mixed vs fp64 %.3g, fp64 fallbacks %d" %
          (sum(len(u) for u in u64), worst, nfb), flush=True)
    assert worst <= 1e-8
    for st in stmx:
        for q in st:
            if q["precision_used"] == 1:
                assert q["backward_err"] <= 1e-15 and q["forward_est"] <= 1e-10, q
    assert all(np.all(np.isfinite(u)) for u in u64)


def test_cases_match_one_sweep_each():
    """A multi-case solve gives each case what its own sweep gives (the same
    groups and kernels: FP64 bit for bit; the mixed form within 1e-9, since the
    hybrid plan's switch may rise at other points of the lanes' order)."""
    fam = C.case_family(T, freq_step=8)
    pick = [fam[i] for i in (6, 11, 23, 40)]
    cs = T.Cases(pick)
    u64, _ = cs.solve(T.PREC_FP64)
    umx, _ = cs.solve(T.PREC_MIXED)
    cs.close()
    for c, a, b in zip(pick, u64, umx):
        sw = T.Sweep(c["electrodes"], c["nodes"])
        r64, _ = sw.solve(c["s"], c["kappa"], 0, T.PREC_FP64)
        rmx, _ = sw.solve(c["s"], c["kappa"], 0, T.PREC_MIXED)
        sw.close()
        assert np.array_equal(a, r64), c["index"]
        assert np.max(np.abs(b - rmx) / np.abs(rmx)) <= 1e-9, c["index"]
