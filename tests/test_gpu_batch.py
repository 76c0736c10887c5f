"""GPU parity of batched frequency samples: a sweep assembles, factors, solves
and refines a group of samples with one set of kernel launches (each launch
covers every sample of the group). Each sample's arithmetic is its own, so
groups of any size give the one-sample results to rounding (the GEMM tile
shape follows the size of the whole launch): FP64 within 1e-12, the mixed form
within 1e-9 (both far inside the 1e-8 bar; the mixed form accepts an iterate
at an estimated error <= 1e-10). Ragged last groups, groups larger than
the sweep, and the hybrid plan's change of complex64 factor inside a group's
frequency range are covered, on the 20 x 20 m grid (N = 1185) for every
plan, and the result of a batched sweep against the oracle at three
frequencies."""
import numpy as np
import pytest

import tags_pkg

T = tags_pkg.load()
O = tags_pkg.oracle()

pytestmark = pytest.mark.gpu


def _c2():
    rc, e, nodes = T.electrode_grid(5, 5, 20.0, 20.0, 10, 10, 6e-3, -0.5)
    f = O.logspace(2, 7, 400)[::13]  # 31 samples, 100 Hz .. 10 MHz
    s = 2j * np.pi * f
    kap = 1e-3 + s * 15.0 * T.EPS0
    return e, nodes, f, s, kap


@pytest.mark.parametrize("plan", [0, 1, 2, 3, 6, 7])
def test_batch_sizes_agree(plan):
    e, nodes, f, s, kap = _c2()
    sw = T.Sweep(e, nodes)
    sw.set_plan(plan)
    res = {}
    for b in (1, 7, 0, 64):
        sw.set_batch(b)
        u64, st64 = sw.solve(s, kap, 0, T.PREC_FP64)
        umx, stmx = sw.solve(s, kap, 0, T.PREC_MIXED)
        res[b] = (u64, umx, stmx)
    sw.close()
    ref64, refmx, _ = res[1]
    for b, (u64, umx, stmx) in res.items():
        d64 = np.max(np.abs(u64 - ref64) / np.abs(ref64))
        dmx = np.max(np.abs(umx - refmx) / np.abs(refmx))
        print("plan %d batch %d: fp64 %.2e mixed %.2e steps %s" %
              (plan, b, d64, dmx, [q["refine_iters"] for q in stmx]), flush=True)
        assert d64 <= 1e-12, (b, d64)
        assert dmx <= 1e-9, (b, dmx)
        assert np.max(np.abs(umx - u64) / np.abs(u64)) <= 1e-8
        for q in stmx:
            if q["precision_used"] == 1:
                assert q["backward_err"] <= 1e-15 and q["forward_est"] <= 1e-10, q


def test_batched_sweep_vs_oracle():
    e, nodes, f, s, kap = _c2()
    sw = T.Sweep(e, nodes)
    sw.set_batch(16)
    u64, _, x64 = sw.solve(s, kap, 0, T.PREC_FP64, want_x=True)
    umx, st = sw.solve(s, kap, 0, T.PREC_MIXED)
    sw.close()
    sel = [0, 15, 30]
    zo = O.harmonic_impedance(e, nodes, f[sel], 1e-3, 15.0, 0, rel_tol=1e-8, threads=8)
    assert np.max(np.abs(u64[sel] - zo) / np.abs(zo)) <= 1e-6
    assert np.max(np.abs(umx - u64) / np.abs(u64)) <= 1e-8
    m, n = len(e), len(nodes)
    assert np.array_equal(x64[:, 0], u64)
    assert np.max(np.abs(x64[:, n + m:].sum(1) - 1.0)) < 1e-9  # KCL: sum I_T = I_E
