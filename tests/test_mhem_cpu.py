"""The oracle's INTG_MHEM integral (src/electrode.c:219-235, :288-293): the
receiver-axis integral of log((r1 + r2 + L_s) / (r1 + r2 - L_s)) by adaptive
Gauss-Kronrod (7, 15), times the receiver length. Pinned by:
- the rule itself: its Gauss part equals numpy's 7-point Gauss-Legendre, the
  Kronrod part integrates x^k exactly for k <= 22 (and not k = 24);
- log N_f being the sender-axis integral of 1/r in closed form, the result is
  the double integral of 1/r: the mpmath golden vectors at gamma = 0, closed
  forms (collinear and perpendicular touching, aligned parallel filaments), and
  the oracle's 2-D cubature of e^{-gamma r}/r at gamma = 0;
- the mHEM impedance matrices (calculate_impedances / impedances_images with
  INTG_MHEM) against the INTG_DOUBLE ones at gamma = 0, unit scalars.
"""
import json
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _seg(a, b, r=5e-3):
    return O.populate_electrode(a, b, r)[1]


def test_gk15_rule_is_gauss_kronrod():
    x, wk, wg = O.gk15_rule()
    gx, gw = np.polynomial.legendre.leggauss(7)
    # Gauss nodes: x[1], x[3], x[5] (and their negatives) and 0
    assert np.allclose(np.sort(np.r_[-x[1:7:2], x[1:7:2], 0.0]), np.sort(gx), atol=1e-15)
    assert np.allclose(np.sort(wg[:3]), np.sort(gw[gx > 1e-9]), atol=1e-15)
    assert abs(wg[3] - gw[np.argmin(np.abs(gx))]) <= 1e-15
    xs = np.r_[-x[:7], x[7], x[6::-1]]
    ws = np.r_[wk[:7], wk[7], wk[6::-1]]
    for k in range(0, 23):
        exact = 0.0 if k % 2 else 2.0 / (k + 1)
        assert abs(ws @ xs ** k - exact) <= 2e-15, k
    assert abs(ws @ xs ** 24 - 2.0 / 25) > 1e-10


def test_mhem_vs_golden_zero_gamma():
    with open(os.path.join(GOLD, "pair_integrals.json")) as fh:
        pairs = [p for p in json.load(fh)["pairs"] if p["gamma_name"] == "zero"]
    assert len(pairs) >= 10
    for p in pairs:
        s = _seg(*p["sender"])
        r = _seg(*p["receiver"])
        rc, v, err = O.integral(s, r, 0.0, rel_tol=1e-12, integration_type=O.INTG_MHEM)
        assert rc == 0 and v.imag == 0.0
        ref = p["integral"][0]
        assert abs(v.real - ref) <= 1e-10 * abs(ref), (p["geometry"], v, ref)


@pytest.mark.parametrize("a,b", [(0.5, 0.5), (0.3, 2.0), (4.0, 0.25)])
def test_mhem_closed_forms(a, b):
    # collinear, touching: (a+b) ln(a+b) - a ln a - b ln b
    v = O.integral(_seg([0, 0, -1], [a, 0, -1]), _seg([a, 0, -1], [a + b, 0, -1]), 0.0,
                   rel_tol=1e-13, integration_type=O.INTG_MHEM)[1].real
    ref = (a + b) * np.log(a + b) - a * np.log(a) - b * np.log(b)
    assert abs(v - ref) <= 1e-11 * ref
    # perpendicular, touching at a corner: a asinh(b/a) + b asinh(a/b)
    v = O.integral(_seg([a, 0, -1], [0, 0, -1]), _seg([0, 0, -1], [0, b, -1]), 0.0,
                   rel_tol=1e-13, integration_type=O.INTG_MHEM)[1].real
    ref = a * np.arcsinh(b / a) + b * np.arcsinh(a / b)
    assert abs(v - ref) <= 1e-11 * ref
    # aligned parallel filaments of length a, distance b: 2 (a asinh(a/b) - sqrt(a^2+b^2) + b)
    v = O.integral(_seg([0, 0, -1], [a, 0, -1]), _seg([0, b, -1], [a, b, -1]), 0.0,
                   rel_tol=1e-13, integration_type=O.INTG_MHEM)[1].real
    ref = 2 * (a * np.arcsinh(a / b) - np.hypot(a, b) + b)
    assert abs(v - ref) <= 1e-11 * ref


def test_mhem_matrices_equal_double_at_zero_gamma():
    e, nodes = O.grid(3, 3, 8.0, 8.0, 2, 2, 5e-3, -0.6)
    img = O.images_of(e)
    m = len(e)
    rc, zl, zt = O.calculate_impedances(e, 0.0, 0.0, 1.0, 0.0, rel_tol=1e-11,
                                        integration_type=O.INTG_MHEM, threads=8)
    assert rc == 0
    rc, zl2, zt2 = O.calculate_impedances(e, 0.0, 0.0, 1.0, 0.0, rel_tol=1e-11,
                                          integration_type=O.INTG_NONE, threads=8)
    # INTG_NONE has unit scalars too; its off-diagonal "integral" is the
    # middle-point distance, so only the diagonals (self_integral) agree
    idx = np.arange(m) * (m + 1)
    assert np.array_equal(zl[idx], zl2[idx]) and np.array_equal(zt[idx], zt2[idx])
    # the same pairs as INTG_DOUBLE at gamma = 0 with the scalars of s mu / 4pi = 1
    # and 1 / (4 pi kappa) = 1
    s = 4 * np.pi / O.MU0
    kap = 1 / (4 * np.pi)
    rc, zd, td = O.calculate_impedances(e, 0.0, s, 1.0, kap, rel_tol=1e-11, threads=8)
    assert rc == 0
    lo = np.triu_indices(m)  # flat i*m + k, k >= i
    flat = lo[0] * m + lo[1]
    assert np.abs(zl[flat] - zd[flat]).max() <= 1e-9 * np.abs(zd[flat]).max()
    assert np.abs(zt[flat] - td[flat]).max() <= 1e-9 * np.abs(td[flat]).max()
    zi = np.zeros(m * m, complex)
    ti = np.zeros(m * m, complex)
    assert O.impedances_images(zi, ti, e, img, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, rel_tol=1e-11,
                               integration_type=O.INTG_MHEM, threads=8) == 0
    zdi = np.zeros(m * m, complex)
    tdi = np.zeros(m * m, complex)
    assert O.impedances_images(zdi, tdi, e, img, 0.0, s, 1.0, kap, 1.0, 1.0, rel_tol=1e-11,
                               threads=8) == 0
    assert np.abs(zi[flat] - zdi[flat]).max() <= 1e-9 * np.abs(zdi[flat]).max()
    assert np.abs(ti[flat] - tdi[flat]).max() <= 1e-9 * np.abs(tdi[flat]).max()


def test_mhem_scaling_at_zero_gamma_is_hem():
    """With gamma = 0 the mHEM scaling must give exactly the HEM matrices at
    gamma = 0 (the same 1/r pair integrals, scalars and reflection
    coefficients): pins which potential goes with which index and scalar."""
    e, nodes = O.grid(3, 3, 8.0, 8.0, 2, 2, 5e-3, -0.6)
    img = O.images_of(e)
    m = len(e)
    pots = O.mhem_potentials(e, img, rel_tol=1e-11, threads=8)
    sc = O.soil_scalars(1e4, 0.01, 10.0)
    sc["gamma"] = 0.0
    zl, zt = O.mhem_impedances(pots, e, img, sc)
    rc, zh, th = O.calculate_impedances(e, 0.0, sc["s"], 1.0, sc["kappa"], rel_tol=1e-11,
                                        threads=8)
    assert O.impedances_images(zh, th, e, img, 0.0, sc["s"], 1.0, sc["kappa"], sc["ref_l"],
                               sc["ref_t"], rel_tol=1e-11, threads=8) == 0
    rows, cols = np.triu_indices(m)
    flat = rows * m + cols
    assert np.abs(zl[flat] - zh[flat]).max() <= 1e-9 * np.abs(zh[flat]).max()
    assert np.abs(zt[flat] - th[flat]).max() <= 1e-9 * np.abs(th[flat]).max()


def test_mhem_close_to_hem_at_low_frequency():
    """mHEM replaces e^{-gamma r} by e^{-gamma rbar} inside each pair; where
    |gamma| L << 1 the two harmonic impedances agree closely (a swapped
    distance or scalar would not)."""
    e, nodes = O.segment_electrode([0, 0, -0.5], [15, 0, -0.5], 30, 7e-3)
    f = np.array([100.0, 1e3])
    zm = O.mhem_harmonic_impedance(e, nodes, f, 0.01, 10.0, rel_tol=1e-10, threads=8)
    zh = O.harmonic_impedance(e, nodes, f, 0.01, 10.0, rel_tol=1e-10, threads=8)
    assert np.all(np.abs(zm - zh) / np.abs(zh) <= 1e-3)
    assert np.all(zm.real > 0)
