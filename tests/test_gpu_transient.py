"""GPU parity of the transient wrap: the numerical Laplace transform pair on
the GPU (Stockham radix-2 for power-of-two lengths, direct sum otherwise)
against the oracle's numpy restatement, and a GPR waveform end to end (the
sunjerga173powsys example's route on a small conductor): injected current ->
NLT -> Z_h(s_k) from the GPU sweep with the Alipio-Visacro soil -> INLT.

Tolerances: transforms 1e-12 of the largest value (power of two) and 1e-11
(direct sum, longer rounding chains); inverse transforms and the round trip
10x that, compared without the e^{c t} factor (it grows to nt^2 at tmax and
scales the rounding of both sides alike); the waveform 1e-7 of its peak (Z_h
to 1e-8 from each side)."""
import numpy as np
import pytest

import tags_pkg

T = tags_pkg.load()
O = tags_pkg.oracle()

pytestmark = pytest.mark.gpu


def _double_exp(t):
    return 1e3 * 1.0373 * (np.exp(-1.4659e4 * t) - np.exp(-2.4689e6 * t))


@pytest.mark.parametrize("nt,tol", [(2 ** 14, 1e-12), (1024, 1e-12), (1000, 1e-11), (999, 1e-11)])
def test_nlt_pair_vs_oracle(nt, tol):
    tmax = 1e-4
    rng = np.random.default_rng(nt)
    f = np.vstack([_double_exp(np.linspace(0, tmax, nt)), rng.standard_normal((2, nt))])
    g, s = T.laplace_trans(f, tmax)
    go, so = O.laplace_trans(f, tmax)
    assert np.array_equal(s, so)
    assert np.max(np.abs(g - go)) <= tol * np.max(np.abs(go))
    # inverse results compared after taking out the e^{c t} factor (which
    # grows to nt^2 at tmax and multiplies the transform's own rounding)
    w = np.exp(-s[0].real * np.arange(nt) * tmax / (nt - 1))
    for filt in (T.FILTER_NONE, T.FILTER_BLACKMAN, T.FILTER_HANNING, T.FILTER_RIESZ):
        back = T.inv_laplace_trans(go, so, tmax, nt, filt) * w
        ref = O.inv_laplace_trans(go, so, tmax, nt, filt) * w
        assert np.max(np.abs(back - ref)) <= 10 * tol * np.max(np.abs(ref)), filt
    back = T.inv_laplace_trans(g, s, tmax, nt)
    assert np.max(np.abs((back - f) * w)) <= 10 * tol * np.max(np.abs(f * w))
    assert np.all(np.isnan(T.inv_laplace_trans(g[:1], s, tmax, nt, T.FILTER_LANCZOS)))


def test_single_signal_entry_points():
    nt, tmax = 64, 1e-5
    f = _double_exp(np.linspace(0, tmax, nt))
    g = np.zeros(nt // 2 + 1, complex)
    s = np.zeros(nt // 2 + 1, complex)
    L = T.lib()
    assert L.laplace_trans(T._p(f), T._p(g), T._p(s), tmax, nt) == 0
    go, so = O.laplace_trans(f, tmax)
    assert np.max(np.abs(g - go)) <= 1e-12 * np.max(np.abs(go))
    out = np.zeros(nt)
    assert L.inv_laplace_trans(T._p(out), T._p(g), T._p(s), tmax, nt, T.FILTER_HANNING) == 0
    ref = O.inv_laplace_trans(go, so, tmax, nt, 2)
    assert np.max(np.abs(out - ref)) <= 1e-10 * np.max(np.abs(ref))


def test_gpr_waveform_end_to_end():
    e, nodes = O.segment_electrode([0, 0, -0.5], [15, 0, -0.5], 30, 7e-3)
    nt, tmax = 256, 2e-5
    t = np.linspace(0, tmax, nt)
    inj = _double_exp(t)
    ig, s = T.laplace_trans(inj, tmax)
    sigma0 = 5e-4
    h = 1.26 * (1000 * sigma0) ** -0.73
    kap = np.zeros(len(s), complex)
    for k, sk in enumerate(s):
        sg, er = T.alipio_soil(sigma0, sk, h, 0.54, 12.0)
        kap[k] = sg + sk * er * T.EPS0
    sw = T.Sweep(e, nodes)
    zh, _ = sw.solve(s, kap, 0, T.PREC_FP64)
    zm, _ = sw.solve(s, kap, 0, T.PREC_MIXED)
    sw.close()
    v = T.inv_laplace_trans(zh * ig, s, tmax, nt, T.FILTER_BLACKMAN)
    vm = T.inv_laplace_trans(zm * ig, s, tmax, nt, T.FILTER_BLACKMAN)
    # the oracle's route: its own NLT, soil, immittance solve (LAPACK) and INLT
    igo, so = O.laplace_trans(inj, tmax)
    zo = np.zeros(len(so), complex)
    images = O.images_of(e)
    m, n = len(e), len(nodes)
    for k, sk in enumerate(so):
        sg, er = O.alipio_soil(sigma0, sk, h, 0.54, 12.0)
        kk = sg + sk * er * O.EPS0
        sc = dict(s=sk, kappa=kk, gamma=complex(np.sqrt(sk * O.MU0 * kk + 0j)), ref_l=1.0 + 0j,
                  ref_t=(kk - sk * O.EPS0) / (kk + sk * O.EPS0), mur=1.0)
        wg, _, _ = O.assemble(e, nodes, images, sc, rel_tol=1e-10, threads=8)
        ie = np.zeros(n + 2 * m, complex)
        ie[0] = 1.0
        info, x = O.solve_immittance(wg, ie, m, n)
        assert info == 0
        zo[k] = x[0]
    assert np.max(np.abs(zh - zo) / np.abs(zo)) <= 1e-8
    vo = O.inv_laplace_trans(zo * igo, so, tmax, nt, 1)
    # the last samples of an NLT inversion are not physical (the e^{c t}
    # factor reaches nt^2 at tmax); the waveform is compared before 0.9 tmax
    body = slice(0, int(0.9 * nt))
    peak = np.max(np.abs(vo[body]))
    assert np.max(np.abs(v - vo)[body]) <= 1e-7 * peak
    assert np.max(np.abs(vm - v)[body]) <= 1e-7 * peak
    assert np.max(np.abs(v - vo)) <= 1e-7 * np.max(np.abs(vo))
    # a positive GPR that rises with the current
    assert np.all(vo[1:body.stop] > 0) and vo[body.stop // 4] > vo[1]
