"""GPU parity: the HIP path through the C ABI against the CPU oracle, the
mpmath fixtures, LAPACK (scipy) and size-independent properties.

Tolerances:
* pair integrals vs mpmath: 1e-10 relative (fixed-order Gauss-Legendre with
  analytic treatment of touching segments);
* Z_L / Z_T entries vs the oracle's adaptive cubature at rel tol 1e-10:
  1e-8 relative to the largest entry of the matrix;
* harmonic impedance vs the oracle pipeline (Cubature-class integrals + LAPACK
  zsysv): 1e-8 with the oracle at rel tol 1e-10, and 1e-6 with the oracle at
  the reference's own tolerance (1e-6);
* mixed precision (complex64 LU + FP64 refinement) vs the FP64 path: 1e-8.
"""
import json
import os

import numpy as np
import pytest

import tags_pkg

T = tags_pkg.load()
O = tags_pkg.oracle()
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

pytestmark = pytest.mark.gpu


def _c1():
    return O.segment_electrode([0, 0, -0.5], [15, 0, -0.5], 30, 7e-3)


def _small_grid():
    return O.grid(4, 3, 6.0, 4.0, 3, 2, 6e-3, -0.5)


def test_pair_integrals_vs_mpmath():
    with open(os.path.join(GOLD, "pair_integrals.json")) as fh:
        pairs = json.load(fh)["pairs"]
    worst = 0.0
    for p in pairs:
        _, s = T.populate_electrode(p["sender"][0], p["sender"][1], 0.005)
        _, r = T.populate_electrode(p["receiver"][0], p["receiver"][1], 0.005)
        rc, v, err = T.integral(s, r, complex(*p["gamma"]))
        assert rc == 0
        ref = complex(*p["integral"])
        e = abs(v - ref) / abs(ref)
        worst = max(worst, e)
        assert e <= 1e-10, (p["geometry"], p["gamma_name"], e)
        assert abs(err) <= 1e-8 * abs(ref)
    print("worst pair error vs mpmath: %.2e" % worst)


@pytest.mark.parametrize("case", ["c1", "grid"])
@pytest.mark.parametrize("freq", [100.0, 2e5, 1e7])
def test_impedances_vs_oracle(case, freq):
    e, nodes = _c1() if case == "c1" else _small_grid()
    m = len(e)
    sc = O.soil_scalars(freq, 0.01, 10.0)
    img = O.images_of(e)
    rc, zlo, zto = O.calculate_impedances(e, sc["gamma"], sc["s"], 1.0, sc["kappa"],
                                          rel_tol=1e-10, threads=8)
    assert rc == 0
    O.impedances_images(zlo, zto, e, img, sc["gamma"], sc["s"], 1.0, sc["kappa"],
                        sc["ref_l"], sc["ref_t"], rel_tol=1e-10, threads=8)
    # sentinel in the untouched half must survive
    zl = np.full(m * m, 3.0 + 4.0j)
    zt = np.full(m * m, 5.0 - 6.0j)
    rc = T.calculate_impedances(zl, zt, e, sc["gamma"], sc["s"], 1.0, sc["kappa"])
    assert rc == 0
    rc = T.impedances_images(zl, zt, e, img, sc["gamma"], sc["s"], 1.0, sc["kappa"],
                             sc["ref_l"], sc["ref_t"])
    assert rc == 0
    up = np.triu_indices(m)
    lo = np.tril_indices(m, -1)
    ZL, ZT = zl.reshape(m, m), zt.reshape(m, m)
    ZLo, ZTo = zlo.reshape(m, m), zto.reshape(m, m)
    assert np.all(ZL[lo] == 3.0 + 4.0j) and np.all(ZT[lo] == 5.0 - 6.0j)
    for Z, Zo in ((ZL, ZLo), (ZT, ZTo)):
        scale = np.abs(Zo[up]).max()
        err = np.abs(Z[up] - Zo[up]).max() / scale
        assert err <= 1e-8, err


def test_impedances_aliasing_and_none():
    """zl == zt: zl wins off the diagonal, zt on it (src/electrode.c:346-372)."""
    e, _ = _small_grid()
    m = len(e)
    sc = O.soil_scalars(1e5, 1e-3, 15.0)
    z = np.zeros(m * m, complex)
    assert T.calculate_impedances(z, z, e, sc["gamma"], sc["s"], 1.0, sc["kappa"]) == 0
    zl = np.zeros(m * m, complex)
    zt = np.zeros(m * m, complex)
    T.calculate_impedances(zl, zt, e, sc["gamma"], sc["s"], 1.0, sc["kappa"])
    Z, ZL, ZT = z.reshape(m, m), zl.reshape(m, m), zt.reshape(m, m)
    assert np.array_equal(np.diag(Z), np.diag(ZT))
    iu = np.triu_indices(m, 1)
    assert np.array_equal(Z[iu], ZL[iu])
    # INTG_NONE: scalars are 1 and the "integral" is the middle-point distance
    zl = np.zeros(m * m, complex)
    zt = np.zeros(m * m, complex)
    assert T.calculate_impedances(zl, zt, e, 0, 0, 0, 0, integration_type=T.INTG_NONE) == 0
    rc, zlo, zto = O.calculate_impedances(e, 0, 0, 0, 0, integration_type=O.INTG_NONE)
    assert np.allclose(zl, zlo, rtol=1e-14, atol=0) and np.allclose(zt, zto, rtol=1e-14, atol=0)


def test_assembly_vs_oracle():
    e, nodes = _small_grid()
    m, n = len(e), len(nodes)
    N = n + 2 * m
    wg = np.zeros(N * N, complex)
    assert T.fill_incidence_imm(wg, e, nodes) == 0
    rc, wgo = O.fill_incidence_imm(e, nodes)
    assert rc == 0 and np.array_equal(wg, wgo)
    rng = np.random.default_rng(3)
    zl = rng.standard_normal(m * m) + 1j * rng.standard_normal(m * m)
    zt = rng.standard_normal(m * m) + 1j * rng.standard_normal(m * m)
    w1 = np.full(N * N, 2.0 - 1.0j)
    w2 = w1.copy()
    assert T.fill_impedance_imm(w1, zl, zt, m, n) == 0
    O.fill_impedance_imm(w2, zl, zt, m, n)
    assert np.array_equal(w1, w2)
    # error codes: a node without a segment (1), a duplicated node (sum(B) != m)
    bad = np.vstack([nodes, [[50.0, 50.0, -0.5]]])
    Nb = len(bad) + 2 * m
    assert T.fill_incidence_imm(np.zeros(Nb * Nb, complex), e, bad) == \
        O.fill_incidence_imm(e, bad)[0] == 1
    dup = np.vstack([nodes, nodes[2:3] + 1e-7])
    Nd = len(dup) + 2 * m
    wd = np.zeros(Nd * Nd, complex)
    rc = T.fill_incidence_imm(wd, e, dup)
    rco, wdo = O.fill_incidence_imm(e, dup)
    assert rc == rco != 0 and np.array_equal(wd, wdo)


def _assembled(e, nodes, freq, sigma, epsr, rel_tol=1e-10):
    sc = O.soil_scalars(freq, sigma, epsr)
    wg, _, _ = O.assemble(e, nodes, O.images_of(e), sc, rel_tol=rel_tol, threads=8)
    return wg


@pytest.mark.parametrize("case", ["c1", "grid"])
def test_solve_immittance_vs_lapack(case):
    e, nodes = _c1() if case == "c1" else _small_grid()
    m, n = len(e), len(nodes)
    N = n + 2 * m
    wg = _assembled(e, nodes, 3e4, 0.01, 10.0)
    keep = wg.copy()
    ie = np.zeros(N, complex)
    ie[0] = 1.0
    info_o, xo = O.solve_immittance(wg, ie.copy(), m, n)
    x = ie.copy()
    assert T.solve_immittance(wg, x, m, n) == 0
    assert info_o == 0
    assert np.array_equal(wg, keep)
    assert np.abs(x - xo).max() <= 1e-10 * np.abs(xo).max()
    assert abs(x[n + m:].sum() - 1.0) < 1e-9  # KCL


def _dropin_harmonic(e, img, nodes, freq):
    """The reference caller's per-frequency sequence through the drop-in API
    (examples/harmonic_impedance.c:114-164): impedances, images, assembly and
    solve; returns U."""
    m, n = len(e), len(nodes)
    N = n + 2 * m
    sc = O.soil_scalars(freq, 0.01, 10.0)
    zl = np.zeros(m * m, complex)
    zt = np.zeros(m * m, complex)
    assert T.calculate_impedances(zl, zt, e, sc["gamma"], sc["s"], 1.0, sc["kappa"]) == 0
    assert T.impedances_images(zl, zt, e, img, sc["gamma"], sc["s"], 1.0, sc["kappa"],
                               sc["ref_l"], sc["ref_t"]) == 0
    wg = np.zeros(N * N, complex)
    assert T.fill_incidence_imm(wg, e, nodes) == 0
    T.fill_impedance_imm(wg, zl, zt, m, n)
    ie = np.zeros(N, complex)
    ie[0] = 1.0
    assert T.solve_immittance(wg, ie, m, n) == 0
    return ie


def test_dropin_concurrent_threads():
    """The reference's examples call the path from concurrent OpenMP threads
    with per-thread buffers (examples/harmonic_impedance.c:87-176): the same
    calls from 6 threads at once give the serial results."""
    from concurrent.futures import ThreadPoolExecutor
    e, nodes = _small_grid()
    img = O.images_of(e)
    freqs = [100.0, 1e3, 1e4, 1e5, 1e6, 1e7]
    serial = [_dropin_harmonic(e, img, nodes, f) for f in freqs]
    with ThreadPoolExecutor(max_workers=6) as ex:
        par = list(ex.map(lambda f: _dropin_harmonic(e, img, nodes, f), freqs * 2))
    for i, x in enumerate(par):
        ref = serial[i % len(freqs)]
        assert np.abs(x - ref).max() <= 1e-13 * np.abs(ref).max()


def test_c_driver_vs_oracle():
    """The plain-C driver (tests/c_driver/harmonic_c1.c, built by build()),
    linked against libtags_hip.so and calling the reference API exactly as
    the reference caller does, gives the oracle's harmonic impedance of the
    15 m conductor within 1e-8 (oracle cubature at rel tol 1e-10)."""
    import subprocess
    assert os.path.exists(tags_pkg.C_DRIVER_BIN), "run __graft_entry__.build() first"
    print("case: c_driver harmonic_c1 6 freqs", flush=True)
    r = subprocess.run([tags_pkg.C_DRIVER_BIN, "6"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    rows = np.array([[float(v) for v in line.split()] for line in r.stdout.split("\n")
                     if line.strip()])
    assert rows.shape == (6, 3)
    f = rows[:, 0]
    z = rows[:, 1] + 1j * rows[:, 2]
    assert np.allclose(f, np.logspace(2, 7, 6), rtol=1e-10)
    e = _c1()
    zo = O.harmonic_impedance(e[0], e[1], f, 0.01, 10.0, 0, rel_tol=1e-10, threads=8)
    assert np.max(np.abs(z - zo) / np.abs(zo)) <= 1e-8


def test_solve_singular_reports_info():
    a = np.zeros((70, 70), complex)
    a[np.arange(69), np.arange(69)] = 1.0
    rc, _, st = T.dense_solve(a, np.ones(70), T.PREC_FP64)
    assert rc == 70 and st.info == 70


@pytest.mark.parametrize("n", [1, 7, 64, 65, 127, 128, 129, 300, 1000, 2049])
@pytest.mark.parametrize("prec", [0, 1])
def test_dense_solve_random(n, prec):
    rng = np.random.default_rng(n)
    a = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    rc, x, st = T.dense_solve(a, b, prec)
    assert rc == 0
    x0 = np.linalg.solve(a, b)
    assert np.abs(x - x0).max() <= 1e-9 * np.abs(x0).max()
    assert st.precision_used == prec


def test_c1_sweep_vs_oracle():
    """Config 1 (15 m conductor, 200 frequencies) end to end."""
    e, nodes = _c1()
    f = O.logspace(2, 7, 200)
    sel = f[::10]
    zo_tight = O.harmonic_impedance(e, nodes, sel, 0.01, 10.0, 0, rel_tol=1e-10, threads=8)
    zo_ref = O.harmonic_impedance(e, nodes, sel, 0.01, 10.0, 0, rel_tol=1e-6, threads=8)
    sw = T.Sweep(e, nodes)
    s = 2j * np.pi * f
    kap = 0.01 + s * 10.0 * T.EPS0
    for prec in (T.PREC_FP64, T.PREC_MIXED):
        u, st = sw.solve(s, kap, 0, prec)
        us = u[::10]
        assert np.max(np.abs(us - zo_tight) / np.abs(zo_tight)) <= 1e-8
        assert np.max(np.abs(us - zo_ref) / np.abs(zo_ref)) <= 1e-6
        assert np.all(u.real > 0)
    sw.close()


@pytest.mark.parametrize("plan", [0, 1, 2, 3, 4, 5, 6, 7])
def test_c2_sweep_vs_oracle_and_mixed(plan):
    """Config 2 (20 x 20 m grid, N = 1,185), every factorisation plan: FP64 vs
    the oracle at a few frequencies, mixed vs FP64 on 40 frequencies, KCL and
    the reference unknown order of x on every sample."""
    rc, e, nodes = T.electrode_grid(5, 5, 20.0, 20.0, 10, 10, 6e-3, -0.5)
    f = O.logspace(2, 7, 400)[::10]
    s = 2j * np.pi * f
    kap = 1e-3 + s * 15.0 * T.EPS0
    sw = T.Sweep(e, nodes)
    sw.set_plan(plan)
    u64, st64, x64 = sw.solve(s, kap, 0, T.PREC_FP64, want_x=True)
    umx, stmx = sw.solve(s, kap, 0, T.PREC_MIXED)
    # plan 3 pivots only inside the Z_L block in its complex64 factor (refined);
    # its FP64 factor pivots fully (an unrefined factor with that weak pivoting
    # was off by up to ~1e-6 here)
    tol = 1e-8
    assert np.max(np.abs(umx - u64) / np.abs(u64)) <= tol
    if plan in (4, 5):  # the grounded complex64 factors converge on this grid
        assert all(q["precision_used"] == 1 for q in stmx), stmx
    m, n = len(e), len(nodes)
    assert np.array_equal(x64[:, 0], u64)
    assert np.max(np.abs(x64[:, n + m:].sum(1) - 1.0)) < tol  # KCL: sum I_T = I_E
    sel = [0, 20, 39]
    zo = O.harmonic_impedance(e, nodes, f[sel], 1e-3, 15.0, 0, rel_tol=1e-8, threads=8)
    assert np.max(np.abs(u64[sel] - zo) / np.abs(zo)) <= 1e-6
    sw.close()


@pytest.mark.parametrize("plan", [0, 1, 2, 3, 4, 5, 6, 7])
def test_lu_modes_agree(plan):
    """Config 2 sweep under the factorisation schedules: look-ahead (default;
    the next panel on a second stream with its own row map), strictly in
    order, partial pivoting over the whole panel (cooperative launch, in
    order), and look-ahead with 128-column steps instead of 256. Every plan has
    a phase boundary where the next panel permutes rows beyond the current
    step's row range; all must give the same Z_h, and the mixed form must
    converge without the FP64 fallback."""
    rc, e, nodes = T.electrode_grid(5, 5, 20.0, 20.0, 10, 10, 6e-3, -0.5)
    f = O.logspace(2, 7, 400)[::40]
    s = 2j * np.pi * f
    kap = 1e-3 + s * 15.0 * T.EPS0
    res = {}
    try:
        for mode in ("lookahead", "inorder", "partial", "width128"):
            T.set_lookahead(mode in ("lookahead", "width128"))
            T.set_panel(T.PANEL_PARTIAL if mode == "partial" else T.PANEL_TOURNAMENT)
            T.set_step_width(128 if mode == "width128" else 0)
            sw = T.Sweep(e, nodes)
            sw.set_plan(plan)
            u64, _ = sw.solve(s, kap, 0, T.PREC_FP64)
            umx, stmx = sw.solve(s, kap, 0, T.PREC_MIXED)
            sw.close()
            if plan not in (3, 6, 7):  # these pivot within the Z_L block only
                assert all(q["precision_used"] == 1 for q in stmx), (mode, stmx)
            res[mode] = (u64, umx)
    finally:
        T.set_lookahead(True)
        T.set_panel(T.PANEL_TOURNAMENT)
        T.set_step_width(0)
    ref = res["lookahead"][0]
    tol = 1e-9  # plan 3's FP64 factor pivots fully; its weak pivoting is complex64's
    for mode, (u64, umx) in res.items():
        assert np.max(np.abs(u64 - ref) / np.abs(ref)) <= tol, mode
        assert np.max(np.abs(umx - ref) / np.abs(ref)) <= max(tol, 1e-8), mode


def test_c3_mixed_vs_fp64_property():
    """Full 60 x 60 m grid (N = 21,421): mixed vs FP64 at the band edges, and
    the structured factorisation plan vs the dense one, and the grounded
    complex64 factor without row exchanges (no FP64 fallback)."""
    rc, e, nodes = T.electrode_grid(31, 31, 60.0, 60.0, 4, 4, 5e-3, -0.6)
    f = np.array([100.0, 5e6])
    s = 2j * np.pi * f
    kap = 2e-3 + s * 20.0 * T.EPS0
    sw = T.Sweep(e, nodes)
    sw.set_plan(T.PLAN_STRUCTURED)
    u64, _ = sw.solve(s, kap, 0, T.PREC_FP64)
    umx, st = sw.solve(s, kap, 0, T.PREC_MIXED)
    assert np.max(np.abs(umx - u64) / np.abs(u64)) <= 1e-8
    assert all(q["backward_err"] <= 1e-13 for q in st)
    # grounded matrix, tournament pivoting: no FP64 fallback at either edge
    sw.set_plan(T.PLAN_GROUNDED_PIVOTED)
    ug, stg = sw.solve(s, kap, 0, T.PREC_MIXED)
    assert all(q["precision_used"] == 1 and q["backward_err"] <= 1e-13 for q in stg), stg
    assert np.max(np.abs(ug - u64) / np.abs(u64)) <= 1e-8
    # grounded matrix without row exchanges: converges at 5 MHz; at 100 Hz the
    # refinement stalls on this grid and the FP64 fallback gives the answer
    sw.set_plan(T.PLAN_GROUNDED)
    ug, stg = sw.solve(s, kap, 0, T.PREC_MIXED)
    assert stg[1]["precision_used"] == 1 and stg[1]["backward_err"] <= 1e-13, stg
    assert np.max(np.abs(ug - u64) / np.abs(u64)) <= 1e-8
    sw.set_plan(T.PLAN_DENSE)
    ud, _ = sw.solve(s, kap, 0, T.PREC_FP64)
    assert np.max(np.abs(ud - u64) / np.abs(u64)) <= 1e-8
    sw.close()


def test_c3_hybrid_vs_zl_first():
    """Full 60 x 60 m grid (N = 21,421) with the plans the bench runs: the
    hybrid plan's mixed form (the nodal complex64 factor, refined by FGMRES
    below 20 kHz and classically from it up) against the Z_L-first FP64
    factor, at the band edges and on both sides of 20 kHz, each within 1e-8;
    the mixed
    samples meet the acceptance rule (backward error <= 1e-15, forward-error
    estimate <= 1e-10) and the FP64 ones report a backward error below 1e-13
    (measured 1e-19 .. 2.7e-14; the normwise bound of a stable LU grows like
    N eps, 2.4e-12 at this order)."""
    rc, e, nodes = T.electrode_grid(31, 31, 60.0, 60.0, 4, 4, 5e-3, -0.6)
    f = np.array([100.0, 1.9e4, 2.1e4, 5e6])
    s = 2j * np.pi * f
    kap = 2e-3 + s * 20.0 * T.EPS0
    sw = T.Sweep(e, nodes)
    sw.set_plan(T.PLAN_HYBRID)
    sw.set_verify(True)
    u64, st64 = sw.solve(s, kap, 0, T.PREC_FP64)
    umx, st = sw.solve(s, kap, 0, T.PREC_MIXED)
    sw.close()
    d = np.abs(umx - u64) / np.abs(u64)
    print("c3 hybrid vs zl_first: %s steps %s" % (d, [q["refine_iters"] for q in st]),
          flush=True)
    assert d.max() <= 1e-8
    assert all(q["precision_used"] == 1 and q["backward_err"] <= 1e-15 and
               q["forward_est"] <= 1e-10 for q in st), st
    assert all(0 <= q["backward_err"] <= 1e-13 for q in st64), st64


@pytest.mark.parametrize("freq", [100.0, 5e6])
def test_c3_assembly_vs_oracle_pairs(freq):
    """2000 random segment pairs of the 60 x 60 m grid, its diagonal and the
    pairs touching a grid crossing, as assembled in HBM by the sweep, against
    the oracle's entries (calculate_impedances + impedances_images, rel tol
    1e-11): 1e-9 of each entry."""
    e, nodes = O.grid(31, 31, 60.0, 60.0, 4, 4, 5e-3, -0.6)
    m, n = len(e), len(nodes)
    sc = O.soil_scalars(freq, 2e-3, 20.0)
    rng = np.random.default_rng(int(freq))
    ends = np.concatenate([e["start_point"], e["end_point"]])
    s0 = 3000
    d = np.abs(ends[:, None, :] - ends[None, [s0, s0 + m], :]).sum(-1).min(1)
    touch = sorted(set(np.nonzero(d < 1e-9)[0] % m) - {s0})
    assert len(touch) >= 2
    pairs = np.vstack([rng.integers(0, m, (2000, 2)),
                       np.column_stack([np.arange(0, m, 31), np.arange(0, m, 31)]),
                       np.column_stack([np.full(len(touch), s0), touch])])
    zl, zt = O.z_entries(e, O.images_of(e), pairs, sc, rel_tol=1e-11, threads=16)
    sw = T.Sweep(e, nodes)
    wl = sw.assemble_entries(sc["s"], sc["kappa"], n + pairs[:, 0], n + pairs[:, 1])
    wt = sw.assemble_entries(sc["s"], sc["kappa"], n + m + pairs[:, 0], n + m + pairs[:, 1])
    sw.close()
    for got, ref in ((-wl, zl), (-wt, zt)):
        err = np.abs(got - ref)
        assert np.all(err <= 1e-9 * np.abs(ref) + 1e-14 * np.abs(ref).max()), err.max()


@pytest.mark.parametrize("plan", [0, 1, 2, 3, 4, 5, 6, 7])
def test_sweep_assembly_vs_oracle(plan):
    """W assembled in HBM by the sweep (pair integrals, images, incidence) vs
    the oracle pipeline, in the reference unknown order, for every storage
    layout."""
    e, nodes = _small_grid()
    m, n = len(e), len(nodes)
    N = n + 2 * m
    sc = O.soil_scalars(3e5, 1e-3, 15.0)
    wo = _assembled(e, nodes, 3e5, 1e-3, 15.0)
    Wo = wo.reshape(N, N).T
    Wo = np.tril(Wo) + np.tril(Wo, -1).T
    sw = T.Sweep(e, nodes)
    sw.set_plan(plan)
    W = sw.assemble(sc["s"], sc["kappa"]).reshape(N, N).T
    assert np.abs(W - Wo).max() <= 1e-9 * np.abs(Wo).max()
    assert np.array_equal(W, W.T)
    sw.close()


@pytest.mark.parametrize("plan", [0, 1, 2, 3, 4, 5, 6])
@pytest.mark.parametrize("prec", [0, 1])
def test_immittance_solve_plans(plan, prec):
    """Structured factorisation plans on matrices with the immittance zero
    pattern of their storage order, including signed zeros (-0.0) in the
    first columns: a -0.0 must never win the pivot search."""
    rng = np.random.default_rng(11)
    m, n = 150, 97
    N = n + 2 * m
    A = rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))
    if plan in (0, 1):  # [U; I_L; I_T]: Y_E = 0, no Z_L / Z_T coupling
        A[:n, :n] = -0.0
        A[n:n + m, n + m:] = 0
        A[n + m:, n:n + m] = 0
    else:               # [I_L; I_T; U]
        A[:m, m:2 * m] = 0
        A[m:2 * m, :m] = 0
        A[m + 5:m + 40, :3] = -0.0
        A[:3, 10:20] = -0.0
    b = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    rc, x, st = T.immittance_solve(A, b, m, n, prec, plan)
    assert rc == 0 and st.info == 0
    x0 = np.linalg.solve(A, b)
    tol = 1e-9  # plan 3 pivots within the Z_L block in complex64 only (then refines)
    assert np.abs(x - x0).max() <= tol * np.abs(x0).max()


@pytest.mark.parametrize("freq", [100.0, 1e3, 3e5, 1e7])
def test_nodal_vs_lapack(freq):
    """The oracle's W of a small grid in the [I_L; I_T; U] order, mixed form of
    the nodal plan (complex64 factor: Z_L and Z_T factored within their own
    rows, the node block's Schur complement from the incidence rows; FP64
    refinement on W): no FP64 fallback and few refinement steps at high
    frequency (a wrong factor would need many), and x equals LAPACK's
    solution of W. At 100 Hz the complex64 nodal factor's refinement stalls on
    this grid and the FP64 fallback answers (the node block's Schur complement
    is dominated by A^T Z_L^-1 A, about 1/|Z_L|, whose complex64 rounding
    outweighs the grounding terms that fix its smallest modes)."""
    prec = T.PREC_MIXED
    e, nodes = _small_grid()
    m, n = len(e), len(nodes)
    N = n + 2 * m
    wo = _assembled(e, nodes, freq, 1e-3, 15.0).reshape(N, N).T
    Wr = np.tril(wo) + np.tril(wo, -1).T  # reference order [U; I_L; I_T]
    order = np.concatenate([n + np.arange(2 * m), np.arange(n)])
    A = np.ascontiguousarray(Wr[np.ix_(order, order)])
    b = np.zeros(N, complex)
    b[2 * m + 3] = 1.0  # unit current at node 3
    rc, x, st = T.immittance_solve(A, b, m, n, prec, T.PLAN_NODAL)
    assert rc == 0 and st.info == 0
    if freq >= 3e5:
        assert st.precision_used == 1 and st.backward_err <= 1e-13, (
            st.precision_used, st.refine_iters, st.backward_err)
        assert st.refine_iters <= 4, st.refine_iters
    x0 = np.linalg.solve(A, b)
    assert np.abs(x - x0).max() <= 1e-9 * np.abs(x0).max(), (st.precision_used, st.refine_iters)


def test_hybrid_low_switch_retries():
    """Hybrid plan with its switch at 1 kHz and classic refinement at every
    frequency (no FGMRES) on the 20 x 20 m grid: the nodal complex64 samples
    whose refinement stalls start again with the Z_L-first complex64 factor (W
    is intact) and raise the switch; every sample stays within 1e-8 of the FP64
    answer and none needs the FP64 factor."""
    rc, e, nodes = T.electrode_grid(5, 5, 20.0, 20.0, 10, 10, 6e-3, -0.5)
    f = O.logspace(2, 7, 400)[::10]
    s = 2j * np.pi * f
    kap = 1e-3 + s * 15.0 * T.EPS0
    sw = T.Sweep(e, nodes)
    sw.set_plan(T.PLAN_ZL_FIRST)
    u64, _ = sw.solve(s, kap, 0, T.PREC_FP64)
    sw.set_plan(T.PLAN_HYBRID)
    sw.set_nodal_above(1e3)
    sw.set_gmres(0.0, 6)
    umx, st = sw.solve(s, kap, 0, T.PREC_MIXED)
    sw.close()
    assert np.max(np.abs(umx - u64) / np.abs(u64)) <= 1e-8
    assert all(q["precision_used"] == 1 and q["backward_err"] <= 1e-15 for q in st), st


@pytest.mark.parametrize("freq", [100.0, 3e5, 1e7])
def test_grounded_unpivoted_vs_lapack(freq):
    """The oracle's W of a small grid in the [I_L; I_T; U] order, mixed form of
    the grounded plan (complex64 factor of T^T W T without row exchanges, FP64
    refinement on W): no fallback, and x equals LAPACK's solution of W."""
    e, nodes = _small_grid()
    m, n = len(e), len(nodes)
    N = n + 2 * m
    wo = _assembled(e, nodes, freq, 1e-3, 15.0).reshape(N, N).T
    Wr = np.tril(wo) + np.tril(wo, -1).T  # reference order [U; I_L; I_T]
    order = np.concatenate([n + np.arange(2 * m), np.arange(n)])
    A = np.ascontiguousarray(Wr[np.ix_(order, order)])
    b = np.zeros(N, complex)
    b[2 * m + 3] = 1.0  # unit current at node 3
    rc, x, st = T.immittance_solve(A, b, m, n, T.PREC_MIXED, T.PLAN_GROUNDED)
    assert rc == 0 and st.info == 0
    assert st.precision_used == 1 and st.backward_err <= 1e-13, (
        st.precision_used, st.refine_iters, st.backward_err)
    x0 = np.linalg.solve(A, b)
    assert np.abs(x - x0).max() <= 1e-9 * np.abs(x0).max()



@pytest.mark.parametrize("plan", [1, 7])
def test_sweep_inflight_agree(plan):
    """A sweep with 1, 2 and 3 samples in flight (own streams and buffers,
    queued from the calling thread) gives bit-identical results per sample in
    both precisions, including the full solution vectors, and fills every
    output slot; the result of a lone sample equals its result in a sweep. With
    the hybrid plan a raise of its switch during a call may give a sample in
    flight the other complex64 factor: then its result agrees within 1e-8 (the
    refinement's acceptance holds for both), and the next call starts from the
    configured switch again."""
    rc, e, nodes = T.electrode_grid(5, 5, 20.0, 20.0, 10, 10, 6e-3, -0.5)
    f = O.logspace(2, 7, 400)[::37]
    s = 2j * np.pi * f
    kap = 1e-3 + s * 15.0 * T.EPS0
    sw = T.Sweep(e, nodes)
    sw.set_plan(plan)
    sw.set_batch(3)  # four groups of samples, so that the count in flight matters
    res = {}
    for k in (1, 2, 3):
        sw.set_inflight(k)
        res[k] = (sw.solve(s, kap, 0, T.PREC_FP64, want_x=True),
                  sw.solve(s, kap, 0, T.PREC_MIXED, want_x=True))
    lone = sw.solve(s[4:5], kap[4:5], 0, T.PREC_MIXED)[0]
    again = sw.solve(s, kap, 0, T.PREC_MIXED)[0]
    sw.set_inflight(2)
    sw.close()
    for k in (2, 3):
        for a, b in zip(res[1], res[k]):
            if plan == 1:
                assert np.array_equal(a[0], b[0]) and np.array_equal(a[2], b[2])
                assert [q["refine_iters"] for q in a[1]] == [q["refine_iters"] for q in b[1]]
            else:
                assert np.max(np.abs(a[0] - b[0]) / np.abs(a[0])) <= 1e-8
    assert np.all(np.isfinite(res[1][0][0])) and np.all(np.isfinite(res[1][1][0]))
    # a lone sample is a one-sample group: the same arithmetic to rounding
    assert abs(lone[0] - res[2][1][0][4]) <= 1e-9 * abs(lone[0])
    assert np.array_equal(again, res[3][1][0])


def test_gemm_split_bf16_as_accurate_as_f32():
    """The complex64 products on the bf16 MFMA cores (variants 22-32: every
    fp32 value as three bf16 parts, six part products per real product) carry
    the accuracy of the fp32 three-multiply tiles (variant 19): at K = 256 (the
    LU's step) and entries spread over six decades, the largest error against
    complex128 numpy on the same fp32 operands, relative to |A| |B| + |C|, is
    within twice variant 19's (about 1e-6 here, K fp32 roundings); ragged M,
    N, K too."""
    rng = np.random.default_rng(11)

    def spread(shape):
        mag = 10.0 ** rng.uniform(-3, 3, shape)
        return ((rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) * mag
                ).astype(np.complex64)

    for (M, N, K) in [(300, 200, 256), (129, 65, 29), (640, 384, 256)]:
        a, b, c = spread((M, K)), spread((K, N)), spread((M, N))
        ar = rng.permutation(M)
        br = rng.permutation(K)
        cr = rng.permutation(M)
        ref = c.astype(np.complex128).copy()
        ref[cr] -= a[ar].astype(np.complex128) @ b[br].astype(np.complex128)
        bound = np.zeros((M, N))
        bound[cr] = np.abs(a[ar]).astype(np.float64) @ np.abs(b[br]).astype(np.float64)
        bound += np.abs(c)
        errs = {}
        for v in [19] + list(range(22, 33)):
            out = T.gemm_run(a, b, c, 1, v, ar, br, cr, 1)
            errs[v] = float((np.abs(out - ref) / bound).max())
        for v in range(22, 33):
            assert errs[v] <= max(2.0 * errs[19], 1e-7) and errs[v] < 1e-5, (M, N, K, errs)


@pytest.mark.parametrize("prec", [0, 1])
def test_gemm_variants_vs_numpy(prec):
    """Every MFMA GEMM variant of a precision, through tags_gemm_run, against
    numpy on the same (rounded) operands: C -= A B with row maps on A, B and C
    (as the LU's trailing update addresses them), and C = A B with ragged
    M, N, K; elementwise error <= tol (|A| |B| + |C|)."""
    rng = np.random.default_rng(5)
    rt = np.complex64 if prec == 1 else np.complex128
    tol = 1e-5 if prec == 1 else 1e-14
    cases = [(300, 200, 128, True, 1), (129, 65, 29, False, 0), (64, 64, 64, True, 1)]
    nv = T.lib().tags_gemm_variants(prec)
    assert nv >= 1
    for (M, N, K, maps, mode) in cases:
        a = (rng.standard_normal((M, K)) + 1j * rng.standard_normal((M, K))).astype(rt)
        b = (rng.standard_normal((K, N)) + 1j * rng.standard_normal((K, N))).astype(rt)
        c = (rng.standard_normal((M, N)) + 1j * rng.standard_normal((M, N))).astype(rt)
        ar = rng.permutation(M) if maps else None
        br = rng.permutation(K) if maps else None
        cr = rng.permutation(M) if maps else None
        A = a[ar] if maps else a
        B = b[br] if maps else b
        prod = A.astype(np.complex128) @ B.astype(np.complex128)
        ref = c.astype(np.complex128).copy()
        rows = cr if maps else np.arange(M)
        ref[rows] = (ref[rows] - prod) if mode == 1 else prod
        bound = np.abs(A) @ np.abs(B)
        bound_full = np.zeros((M, N))
        bound_full[rows] = bound
        bound_full += np.abs(c)
        for v in range(nv):
            out = T.gemm_run(a, b, c, prec, v, ar, br, cr, mode)
            err = np.abs(out - ref)
            assert np.all(err <= tol * bound_full), (v, M, N, K, float((err / bound_full).max()))
