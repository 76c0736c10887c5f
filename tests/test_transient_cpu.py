"""Transient wrap on the CPU: the oracle's soil models and numerical Laplace
transform pair (src/auxiliary.c:61-93, :240-329), pinned by closed forms, a
brute-force DFT and the transform pair's exact round trip; the library's
soil models (host code) against the oracle."""
import numpy as np
import pytest

import tags_pkg

O = tags_pkg.oracle()
T = tags_pkg.load()


def test_alipio_known_values():
    sigma0, h, g, er = 5e-4, 1.26 * (1000 * 5e-4) ** -0.73, 0.54, 12.0
    # at f = 1 MHz (s = j 2 pi 1e6) (f / 1 MHz)^g = 1: sigma = sigma0 (1 + h)
    s1 = 2j * np.pi * 1e6
    sg, ep = O.alipio_soil(sigma0, s1, h, g, er)
    assert abs(sg - sigma0 * (1 + h)) <= 1e-15 * abs(sg)
    t = np.tan(np.pi * g / 2) / (2 * np.pi * O.EPS0 * 1e6 ** g)
    assert abs(ep - (er + t * sigma0 * h * 1e6 ** (g - 1))) <= 1e-13 * abs(ep)
    # real frequency f: sigma grows as f^g
    for f in (1e3, 1e5, 3e6):
        sg, _ = O.alipio_soil(sigma0, 2j * np.pi * f, h, g, er)
        assert abs(sg - sigma0 * (1 + h * (f / 1e6) ** g)) <= 1e-14 * abs(sg)


def test_smith_longmire_limits():
    sigma0, erinf = 1e-3, 5.0
    a = [3.4e6, 2.74e5, 2.58e4, 3.38e3, 5.26e2, 1.33e2, 2.72e1, 1.25e1, 4.8e0, 2.17e0, 9.8e-1,
         3.92e-1, 1.73e-1]
    sg, ep = O.smith_longmire_soil(sigma0, 1e-9j, erinf)  # s -> 0
    assert abs(ep - (erinf + sum(a))) <= 1e-9 * abs(ep)
    assert abs(sg - sigma0) <= 1e-9 * sigma0
    sg, ep = O.smith_longmire_soil(sigma0, 2j * np.pi * 1e15, erinf)  # s -> infinity
    assert abs(ep - erinf) <= 1e-6 * erinf


@pytest.mark.parametrize("s", [2j * np.pi * 60, 2j * np.pi * 1e5, 3e4 + 2j * np.pi * 5e6])
def test_library_soils_match_oracle(s):
    sigma0, h, g, er = 5e-4, 1.26 * (1000 * 5e-4) ** -0.73, 0.54, 12.0
    a = T.alipio_soil(sigma0, s, h, g, er)
    b = O.alipio_soil(sigma0, s, h, g, er)
    assert all(abs(x - y) <= 1e-14 * abs(y) for x, y in zip(a, b))
    a = T.smith_longmire_soil(1e-3, s, 5.0)
    b = O.smith_longmire_soil(1e-3, s, 5.0)
    assert all(abs(x - y) <= 1e-14 * abs(y) for x, y in zip(a, b))


@pytest.mark.parametrize("nt", [7, 8, 12])
def test_nlt_brute_force(nt):
    rng = np.random.default_rng(nt)
    f = rng.standard_normal(nt)
    tmax = 3e-5
    g, s = O.laplace_trans(f, tmax)
    c = np.log(nt ** 2.0) / tmax
    dt = tmax / (nt - 1)
    ref = [sum(f[j] * dt * np.exp(-c * j * dt) * np.exp(-2j * np.pi * j * k / nt)
               for j in range(nt)) for k in range(nt // 2 + 1)]
    assert np.allclose(g, ref, rtol=1e-13, atol=1e-13 * np.abs(ref).max())
    assert np.allclose(s, c + 2j * np.pi / tmax * np.arange(nt // 2 + 1), rtol=1e-15)
    # inverse: the Hermitian extension's real inverse DFT, by brute force
    gi = g * (1 + 0.3j)  # also non-Hermitian DC / Nyquist imaginary parts
    x = np.r_[gi, np.conj(gi[1:(nt + 1) // 2][::-1])]
    x[0] = x[0].real
    if nt % 2 == 0:
        x[nt // 2] = x[nt // 2].real
    out = np.array([sum(x[k] * np.exp(2j * np.pi * j * k / nt) for k in range(nt))
                    for j in range(nt)]).real
    ref_f = out * np.exp(c * np.arange(nt) * dt) / (nt * dt)
    got = O.inv_laplace_trans(gi, s, tmax, nt)
    assert np.allclose(got, ref_f, rtol=1e-12, atol=1e-12 * np.abs(ref_f).max())


def _double_exp(t):
    I0, k, a, b = 1e3, 1.0373, 1.4659e4, 2.4689e6
    return I0 * k * (np.exp(-a * t) - np.exp(-b * t)), (I0, k, a, b)


def test_nlt_of_double_exponential_matches_analytic():
    nt, tmax = 2 ** 14, 1e-4
    t = np.linspace(0, tmax, nt)
    f, (I0, k, a, b) = _double_exp(t)
    g, s = O.laplace_trans(f, tmax)
    G = I0 * k * (1 / (s + a) - 1 / (s + b))
    # the left Riemann sum of a smooth f with f(0) = 0: error O(dt^2) at low
    # frequencies; a wrong weight, sign or scale would be O(1)
    sel = slice(0, 400)
    assert np.max(np.abs(g[sel] - G[sel]) / np.abs(G[sel])) <= 2e-2


def test_nlt_inlt_round_trip():
    nt, tmax = 1024, 1e-4
    t = np.linspace(0, tmax, nt)
    f, _ = _double_exp(t)
    g, s = O.laplace_trans(f, tmax)
    back = O.inv_laplace_trans(g, s, tmax, nt)
    assert np.max(np.abs(back - f)) <= 1e-9 * np.max(np.abs(f))
    # the Lanczos window is 0/0 at w = 0 in the reference
    assert np.all(np.isnan(O.inv_laplace_trans(g, s, tmax, nt, 3)))
