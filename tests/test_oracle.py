"""Pins the CPU oracle (oracle/tags_oracle.c) before it is trusted as the checker.

The reference ships no tests, fixtures or golden data (its Makefile test target
points at a missing file), so each oracle function is pinned here by something
independent of it: mpmath high-precision quadrature, closed forms of the
static (gamma = 0) integrals, brute-force numpy constructions, LAPACK through
scipy, and physics (Sunde's DC resistance of a buried horizontal wire).
"""
import json
import math
import os

import numpy as np
import pytest

import tags_pkg

O = tags_pkg.oracle()
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _seg(p0, p1, radius=0.005):
    rc, e = O.populate_electrode(p0, p1, radius)
    assert rc == 0
    return e


def _pairs():
    with open(os.path.join(GOLD, "pair_integrals.json")) as fh:
        return json.load(fh)["pairs"]


@pytest.mark.parametrize("case", _pairs(), ids=lambda c: c["geometry"] + "-" + c["gamma_name"])
def test_pair_integral_vs_mpmath(case):
    s = _seg(*case["sender"])
    r = _seg(*case["receiver"])
    ref = complex(*case["integral"])
    rc, v, _ = O.integral(s, r, complex(*case["gamma"]), rel_tol=1e-10)
    assert rc == 0
    assert abs(v - ref) <= 1e-9 * abs(ref)
    # at the reference tolerance the adaptive rule stays inside its band
    rc, v6, _ = O.integral(s, r, complex(*case["gamma"]), rel_tol=1e-6)
    assert abs(v6 - ref) <= 5e-6 * abs(ref)


def test_self_integral_vs_mpmath():
    with open(os.path.join(GOLD, "self_integrals.json")) as fh:
        cases = json.load(fh)["cases"]
    for c in cases:
        e = _seg([0, 0, 0], [c["length"], 0, 0], c["radius"])
        assert abs(O.self_integral(e) - c["value"]) <= 1e-13 * c["value"]


def test_static_closed_forms():
    """gamma = 0: touching collinear (a+b)ln(a+b) - a ln a - b ln b, touching
    perpendicular a asinh(b/a) + b asinh(a/b), and aligned parallel filaments
    2 [L asinh(L/d) - sqrt(L^2 + d^2) + d]."""
    a, b = 0.5, 0.3
    e1 = _seg([0, 0, -1], [a, 0, -1])
    e2 = _seg([a, 0, -1], [a + b, 0, -1])
    _, v, _ = O.integral(e1, e2, 0.0, rel_tol=1e-11)
    ref = (a + b) * math.log(a + b) - a * math.log(a) - b * math.log(b)
    assert abs(v - ref) <= 1e-9 * ref
    e3 = _seg([a, 0, -1], [a, b, -1])
    _, v, _ = O.integral(e1, e3, 0.0, rel_tol=1e-11)
    ref = a * math.asinh(b / a) + b * math.asinh(a / b)
    assert abs(v - ref) <= 1e-9 * ref
    L, d = 0.5, 0.2
    e4 = _seg([0, d, -1], [L, d, -1])
    _, v, _ = O.integral(e1, e4, 0.0, rel_tol=1e-11)
    ref = 2 * (L * math.asinh(L / d) - math.sqrt(L * L + d * d) + d)
    assert abs(v - ref) <= 1e-9 * ref


def test_integral_symmetry_and_types():
    e1 = _seg([0, 0, -0.5], [0.5, 0, -0.5])
    e2 = _seg([0.7, 0.2, -0.6], [0.9, 0.8, -0.4])
    g = 0.3 + 0.4j
    _, v12, _ = O.integral(e1, e2, g, rel_tol=1e-10)
    _, v21, _ = O.integral(e2, e1, g, rel_tol=1e-10)
    assert abs(v12 - v21) <= 1e-9 * abs(v12)
    rc, v, _ = O.integral(e1, e2, g, integration_type=O.INTG_NONE)
    assert rc == 0
    assert abs(v - np.linalg.norm(e1["middle_point"][0] - e2["middle_point"][0])) < 1e-15
    rc, _, _ = O.integral(e1, e2, g, integration_type=99)
    assert rc == -10


def test_populate_errors():
    rc, _ = O.populate_electrode([0, 0, 0], [0, 0, 0], 0.01)
    assert rc == 1
    rc, _ = O.populate_electrode([0, 0, 0], [1, 0, 0], 0.0)
    assert rc == 2


@pytest.mark.parametrize("g,m,n", [((5, 5, 20., 20., 10, 10, 6e-3, -0.5), 400, 385),
                                   ((31, 31, 60., 60., 4, 4, 5e-3, -0.6), 7440, 6541),
                                   ((11, 11, 30., 30., 6, 6, 5e-3, -0.5), 1320, 1221),
                                   ((2, 3, 10., 12., 3, 2, 1e-2, -1.0), 17, 16)])
def test_grid_counts_and_order(g, m, n):
    e, nodes = O.grid(*g)
    assert len(e) == m and len(nodes) == n
    # node 0 is the corner (0, 0, depth); horizontal segments come first
    assert np.allclose(nodes[0], [0, 0, g[7]])
    nh = g[4] * g[1] * (g[0] - 1)
    d = e["end_point"] - e["start_point"]
    assert np.all(np.abs(d[:nh, 1]) == 0) and np.all(d[:nh, 0] > 0)
    assert np.all(np.abs(d[nh:, 0]) == 0) and np.all(d[nh:, 1] > 0)
    # every segment end is a node; nodes are unique
    pts = np.concatenate([e["start_point"], e["end_point"]])
    dist = np.abs(pts[:, None, :] - nodes[None, :, :]).max(-1) if m < 2000 else None
    if dist is not None:
        assert np.all(dist.min(1) < 1e-9)
        dn = np.abs(nodes[:, None, :] - nodes[None, :, :]).max(-1) + np.eye(n)
        assert dn.min() > 1e-6


def _grid_small():
    return O.grid(3, 3, 4.0, 4.0, 2, 2, 6e-3, -0.5)


def test_incidence_invariants_and_bruteforce():
    e, nodes = _grid_small()
    m, n = len(e), len(nodes)
    N = n + 2 * m
    rc, wg = O.fill_incidence_imm(e, nodes)
    assert rc == 0
    W = wg.reshape(N, N).T  # W[row, col]
    A = W[n:n + m, :n].real
    B = W[n + m:, :n].real
    assert np.array_equal(A, W[:n, n:n + m].T.real)
    assert np.array_equal(B, W[:n, n + m:].T.real)
    assert A.sum() == 0 and B.sum() == m
    assert np.all((A == 1).sum(1) == 1) and np.all((A == -1).sum(1) == 1)
    assert np.array_equal(B, np.abs(A) / 2)
    # brute force from the geometry
    for s in range(m):
        i0 = np.argmin(np.abs(nodes - e["start_point"][s]).max(1))
        i1 = np.argmin(np.abs(nodes - e["end_point"][s]).max(1))
        assert A[s, i0] == 1 and A[s, i1] == -1
    assert np.all(W[:n, :n] == 0)


def test_incidence_errors():
    e, nodes = _grid_small()
    bad = np.vstack([nodes, [[100.0, 100.0, -0.5]]])
    rc, _ = O.fill_incidence_imm(e, bad)
    assert rc == 1
    dup = np.vstack([nodes, nodes[:1] + 1e-7])
    rc, _ = O.fill_incidence_imm(e, dup)
    assert rc != 0


def test_fill_impedance_blocks():
    m, n = 4, 3
    N = n + 2 * m
    rng = np.random.default_rng(1)
    zl = rng.standard_normal(m * m) + 1j * rng.standard_normal(m * m)
    zt = rng.standard_normal(m * m) + 1j * rng.standard_normal(m * m)
    wg = np.full(N * N, 7.0 + 7.0j)
    O.fill_impedance_imm(wg, zl, zt, m, n)
    W = wg.reshape(N, N).T
    ZL = zl.reshape(m, m).T  # zl[i + m k] at (i, k)
    ZT = zt.reshape(m, m).T
    assert np.array_equal(W[n:n + m, n:n + m], -ZL)
    assert np.array_equal(W[n + m:, n + m:], -ZT)
    assert np.all(W[n:n + m, n + m:] == 0) and np.all(W[n + m:, n:n + m] == 0)
    assert np.all(W[:n, :n] == 0)
    assert np.all(W[n:, :n] == 7 + 7j)  # incidence blocks untouched


def test_impedance_conventions():
    """Lower-half convention, closed-form diagonal, Z_L = 0 for perpendicular
    pairs, symmetry of the image part, and Z = scalar x integral."""
    e, nodes = _grid_small()
    m = len(e)
    sc = O.soil_scalars(1e5, 1e-3, 15.0)
    rc, zl, zt = O.calculate_impedances(e, sc["gamma"], sc["s"], 1.0, sc["kappa"],
                                        rel_tol=1e-10)
    assert rc == 0
    ZL = zl.reshape(m, m)
    ZT = zt.reshape(m, m)
    assert np.all(ZL[np.tril_indices(m, -1)] == 0)  # untouched (zero-initialised)
    iwu = sc["s"] * O.MU0 / O.FOUR_PI
    k4 = 1.0 / (O.FOUR_PI * sc["kappa"])
    for i in range(m):
        si = O.self_integral(e[i:i + 1])
        assert abs(ZL[i, i] - iwu * si) < 1e-15 * abs(iwu * si)
        assert abs(ZT[i, i] - k4 / e["length"][i] ** 2 * si) < 1e-15 * abs(k4 * si)
    d = e["end_point"] - e["start_point"]
    for i in range(m):
        for k in range(i + 1, m):
            _, I, _ = O.integral(e[i:i + 1], e[k:k + 1], sc["gamma"], rel_tol=1e-10)
            assert abs(ZT[i, k] - k4 * I / (e["length"][i] * e["length"][k])) <= 1e-12 * abs(ZT[i, k])
            c = abs(d[i] @ d[k])
            if c < 2.2e-16:
                assert ZL[i, k] == 0
            else:
                ref = iwu * I * c / (e["length"][i] * e["length"][k])
                assert abs(ZL[i, k] - ref) <= 1e-12 * abs(ref)
    img = O.images_of(e)
    zl2 = np.zeros(m * m, complex)
    zt2 = np.zeros(m * m, complex)
    O.impedances_images(zl2, zt2, e, img, sc["gamma"], sc["s"], 1.0, sc["kappa"], 1.0,
                        sc["ref_t"], rel_tol=1e-10)
    # image integral is symmetric: (e_i, img_k) == (e_k, img_i)
    for i, k in [(0, 5), (1, 7), (2, 11)]:
        _, a, _ = O.integral(e[i:i + 1], img[k:k + 1], sc["gamma"], rel_tol=1e-11)
        _, b, _ = O.integral(e[k:k + 1], img[i:i + 1], sc["gamma"], rel_tol=1e-11)
        assert abs(a - b) <= 1e-9 * abs(a)


def test_solve_matches_numpy():
    e, nodes = _grid_small()
    m, n = len(e), len(nodes)
    N = n + 2 * m
    sc = O.soil_scalars(3e4, 1e-3, 15.0)
    wg, _, _ = O.assemble(e, nodes, O.images_of(e), sc, rel_tol=1e-8)
    W = wg.reshape(N, N).T.copy()
    Wsym = np.tril(W) + np.tril(W, -1).T
    b = np.zeros(N, complex)
    b[0] = 1.0
    info, x = O.solve_immittance(wg, b, m, n)
    assert info == 0
    x0 = np.linalg.solve(Wsym, b)
    assert np.abs(x - x0).max() <= 1e-10 * np.abs(x0).max()
    # KCL: the transversal (leakage) currents add up to the injected current
    assert abs(x[n + m:].sum() - 1.0) < 1e-9


def test_sunde_dc_resistance():
    """Low-frequency harmonic impedance of the 15 m horizontal conductor vs
    Sunde's formula R = rho/(pi L) (ln(2L / sqrt(2 a d)) - 1)."""
    e, nodes = O.segment_electrode([0, 0, -0.5], [15, 0, -0.5], 30, 7e-3)
    zh = O.harmonic_impedance(e, nodes, [100.0], 0.01, 10.0, 0, rel_tol=1e-6)[0]
    L, a, d, rho = 15.0, 7e-3, 0.5, 100.0
    R = rho / (math.pi * L) * (math.log(2 * L / math.sqrt(2 * a * d)) - 1)
    assert abs(zh.real - R) < 0.02 * R
    assert abs(zh.imag) < 0.01 * R


def test_z_entries_match_full_matrices():
    """The pair sampler gives exactly the entries of calculate_impedances +
    impedances_images (lower half, k >= i, pairs given either way round),
    including diagonals, touching and perpendicular pairs, on a small grid and
    a vertical rod (images flip the segment direction)."""
    for e in (_grid_small()[0], O.segment_electrode([0, 0, -0.5], [0, 0, -6.5], 12, 8e-3)[0]):
        m = len(e)
        img = O.images_of(e)
        sc = O.soil_scalars(2e5, 1e-3, 15.0)
        rc, zl, zt = O.calculate_impedances(e, sc["gamma"], sc["s"], 1.0, sc["kappa"],
                                            rel_tol=1e-8, threads=4)
        assert rc == 0
        O.impedances_images(zl, zt, e, img, sc["gamma"], sc["s"], 1.0, sc["kappa"], sc["ref_l"],
                            sc["ref_t"], rel_tol=1e-8, threads=4)
        rng = np.random.default_rng(5)
        pairs = np.vstack([np.column_stack([np.arange(m), np.arange(m)]),
                           rng.integers(0, m, (200, 2))])
        ezl, ezt = O.z_entries(e, img, pairs, sc, rel_tol=1e-8, threads=4)
        lo, hi = pairs.min(1), pairs.max(1)
        assert np.array_equal(ezl, zl[lo * m + hi])
        assert np.array_equal(ezt, zt[lo * m + hi])
