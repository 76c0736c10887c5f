"""GPU tests of the refinement residual in double-double: the mixed form's
residual b - W x is summed with every product and addition error carried,
so refinement converges to working accuracy whatever the conditioning (as
long as the complex64 factor still contracts), where FP64 residuals stop at
a forward error of about cond(W) eps.

* A dense complex system of order 60 with singular values from 1 to 1e-6
  (cond = 1e6) against its solution to 40 digits (mpmath): the mixed answer
  (accepted at an estimated error <= 1e-10) is within 1e-11 and at least 5x
  closer than the FP64 LU's own answer (about cond eps; measured 5.5e-13
  against 1.0e-11).
* The seeded case family's smallest cases (N < 400, where vertical rods at
  about 1 kHz have cond ~ 1e7): every mixed Z_h within 1e-9 of the FP64 one
  (measured worst 2.6e-10, against 9.0e-9 with FP64 residuals and FGMRES
  steps accepted on the ratio estimate), and the two worst samples of that
  earlier form (a 10 m rod, N = 61, at 1 and 5 kHz) within 1e-12 of the
  oracle (measured 2e-15)."""
import numpy as np
import pytest

import tags_pkg

T = tags_pkg.load()
O = tags_pkg.oracle()

pytestmark = pytest.mark.gpu


def test_mixed_refinement_reaches_working_accuracy():
    import mpmath
    rng = np.random.default_rng(5)
    n = 60

    def unitary(k):
        q, _ = np.linalg.qr(rng.standard_normal((k, k)) + 1j * rng.standard_normal((k, k)))
        return q
    sv = np.logspace(0, -6, n)
    a = (unitary(n) * sv) @ unitary(n)
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    mpmath.mp.dps = 40
    am = mpmath.matrix([[mpmath.mpc(complex(v)) for v in row] for row in a])
    bm = mpmath.matrix([mpmath.mpc(complex(v)) for v in b])
    xm = mpmath.lu_solve(am, bm)
    xref = np.array([complex(xm[i]) for i in range(n)])
    rc64, x64, _ = T.dense_solve(a, b, T.PREC_FP64)
    rcmx, xmx, st = T.dense_solve(a, b, T.PREC_MIXED)
    assert rc64 == 0 and rcmx == 0
    e64 = np.max(np.abs(x64 - xref)) / np.max(np.abs(xref))
    emx = np.max(np.abs(xmx - xref)) / np.max(np.abs(xref))
    print("cond 1e6: FP64 LU %.2e, mixed %.2e (precision used %d, %d steps)"
          % (e64, emx, st.precision_used, st.refine_iters), flush=True)
    assert st.precision_used == T.PREC_MIXED
    assert emx <= 1e-11
    assert emx <= 0.2 * e64


def test_small_cases_mixed_matches_fp64_and_oracle():
    C = tags_pkg.cases()
    fam = C.case_family(T, freq_step=8)
    sub = [c for c in fam if c["N"] < 400]
    res = {}
    for prec in (T.PREC_FP64, T.PREC_MIXED):
        cs = T.Cases(sub)
        res[prec] = cs.solve(prec)
        cs.close()
    worst = (0.0, 0, 0)
    for ci in range(len(sub)):
        u64, umx = res[T.PREC_FP64][0][ci], res[T.PREC_MIXED][0][ci]
        d = np.abs(umx - u64) / np.abs(u64)
        k = int(np.argmax(d))
        if d[k] > worst[0]:
            worst = (float(d[k]), ci, k)
        assert all(q["precision_used"] == 1 for q in res[T.PREC_MIXED][1][ci])
    print("N < 400, %d cases: worst mixed vs FP64 %.2e (case %d, sample %d)"
          % (len(sub), worst[0], worst[1], worst[2]), flush=True)
    assert worst[0] <= 1e-9
    # the rod that carried the largest difference with FP64-only residuals
    rods = [i for i, c in enumerate(sub) if c["N"] == 61 and abs(c["sigma"] - 5.0356e-4) < 1e-7]
    assert rods, "case family changed"
    ci = rods[0]
    c = sub[ci]
    f = c["f"]
    ks = [int(np.argmin(np.abs(f - 1045.0))), int(np.argmin(np.abs(f - 4995.0)))]
    zo = O.harmonic_impedance(c["electrodes"], c["nodes"], f[ks], c["sigma"], c["epsr"], 0,
                              rel_tol=1e-10)
    for j, k in enumerate(ks):
        emx = abs(res[T.PREC_MIXED][0][ci][k] - zo[j]) / abs(zo[j])
        e64 = abs(res[T.PREC_FP64][0][ci][k] - zo[j]) / abs(zo[j])
        print("rod N=61 at %.0f Hz: mixed %.2e, FP64 %.2e vs oracle" % (f[k], emx, e64), flush=True)
        assert emx <= 1e-12 and e64 <= 1e-12
