"""GPU parity of the nodal admittance drop-ins (src/linalg.c:142-350) through
the C ABI against the oracle (numpy/LAPACK restatement) and the immittance
path.

Tolerances:
* incidence: bit-exact (integer-valued entries, same return codes);
* Y_N, Z^-1 A, Z^-1 (complex LU with row exchanges on the GPU vs LAPACK
  getrf/getri-class inverses in numpy): max-abs difference <= 1e-10 of the
  largest entry; both carry errors of about cond(Z) eps (cond <= 1e4 here);
* node potentials from solve_admittance vs zsysv: 1e-10 of the largest;
* the admittance potentials vs the immittance solve of the same physics:
  1e-9 (the two formulations' rounding differs by about cond(W) eps).
"""
import os

import numpy as np
import pytest

import tags_pkg

T = tags_pkg.load()
O = tags_pkg.oracle()

pytestmark = pytest.mark.gpu


def _cm(x, r, c):
    return x.reshape(c, r).T


def _close(x, ref, tol):
    err = np.abs(x - ref).max() / np.abs(ref).max()
    assert err <= tol, err
    return err


def _physical(freq, grid=(4, 3, 6.0, 4.0, 3, 2, 6e-3, -0.5)):
    e, nodes = O.grid(*grid)
    sc = O.soil_scalars(freq, 0.01, 10.0)
    wg, zl, zt = O.assemble(e, nodes, O.images_of(e), sc, rel_tol=1e-10, threads=8)
    return e, nodes, wg, zl, zt


def _synthetic(m, seed):
    """A conductor of m segments and complex symmetric Z_L, Z_T with a
    dominant diagonal (cond ~ 10), lower triangle in the column-major arrays,
    junk in the strict upper triangle (never read)."""
    e, nodes = O.segment_electrode([0, 0, -0.5], [0.5 * m, 0, -0.5], m, 7e-3)
    rng = np.random.default_rng(seed)
    zs = []
    for _ in range(2):
        r = rng.standard_normal((m, m)) + 1j * rng.standard_normal((m, m))
        z = (r + r.T) / 2 + (2.0 + 1.0j) * np.sqrt(m) * np.eye(m)
        flat = z.T.ravel().copy()  # column-major
        v = _cm(flat, m, m)
        v[np.triu_indices(m, 1)] = 99.0 - 7.0j
        zs.append(flat)
    return e, nodes, zs[0], zs[1]


def _incidence(e, nodes):
    m, n = len(e), len(nodes)
    a = np.zeros(m * n, complex)
    b = np.zeros(m * n, complex)
    assert T.fill_incidence_adm(a, b, e, nodes) == 0
    return a, b


def test_incidence_adm_vs_oracle():
    for e, nodes in [O.grid(4, 3, 6.0, 4.0, 3, 2, 6e-3, -0.5),
                     O.segment_electrode([0, 0, -0.5], [15, 0, -0.5], 30, 7e-3)]:
        m, n = len(e), len(nodes)
        a, b = _incidence(e, nodes)
        rc, ao, bo = O.fill_incidence_adm(e, nodes)
        assert rc == 0 and np.array_equal(a, ao) and np.array_equal(b, bo)
    e, nodes = O.segment_electrode([0, 0, -0.5], [15, 0, -0.5], 6, 7e-3)
    m = len(e)
    cases = [np.vstack([nodes[:3], [[5.0, 5.0, 5.0]], nodes[3:]]),  # unconnected -> 1
             np.vstack([nodes, nodes[:1]]),                         # sum(A) = 1 -> -1
             np.vstack([nodes, nodes[2:3] + 1e-7])]                 # sum(B) = m + 1 -> 1
    for nd in cases:
        k = len(nd)
        a = np.full(m * k, 3.0 + 0j)
        b = np.full(m * k, 3.0 + 0j)
        rc = T.fill_incidence_adm(a, b, e, nd)
        rco, ao, bo = O.fill_incidence_adm(e, nd, np.full(m * k, 3.0 + 0j),
                                           np.full(m * k, 3.0 + 0j))
        assert rc == rco != 0 and np.array_equal(a, ao) and np.array_equal(b, bo)
    # aliasing a == b: the array ends as A and B is reported wrong
    a = np.zeros(m * len(nodes), complex)
    ao = np.zeros(m * len(nodes), complex)
    assert T.fill_incidence_adm(a, a, e, nodes) == O.fill_incidence_adm(e, nodes, ao, ao)[0]
    assert np.array_equal(a, ao)
    assert T.fill_incidence_adm(a, a, e, nodes[:0]) == -m


@pytest.mark.parametrize("freq", [100.0, 1e5, 5e6])
def test_fill_impedance_adm_physical(freq):
    e, nodes, wg, zl, zt = _physical(freq)
    m, n = len(e), len(nodes)
    a, b = _incidence(e, nodes)
    zl_g, zt_g = zl.copy(), zt.copy()
    yn = np.zeros(n * n, complex)
    assert T.fill_impedance_adm(yn, zl_g, zt_g, a, b, m, n) == 0
    zl_o, zt_o = zl.copy(), zt.copy()
    yno = O.fill_impedance_adm(zl_o, zt_o, a, b, m, n)
    _close(yn, yno, 1e-10)
    lo = np.tril_indices(m)
    _close(_cm(zl_g, m, m)[lo], _cm(zl_o, m, m)[lo], 1e-10)
    _close(_cm(zt_g, m, m)[lo], _cm(zt_o, m, m)[lo], 1e-10)
    up = np.triu_indices(m, 1)
    assert np.array_equal(_cm(zl_g, m, m)[up], _cm(zl, m, m)[up])
    # admittance solve on the GPU vs the immittance solve of the same W
    ie = np.zeros(n, complex)
    ie[0] = 1.0
    yn_keep = yn.copy()
    assert T.solve_admittance(yn, ie, n) == 0
    assert np.array_equal(yn, yn_keep)
    N = n + 2 * m
    x = np.zeros(N, complex)
    x[0] = 1.0
    assert T.solve_immittance(wg.copy(), x, m, n) == 0
    _close(ie, x[:n], 1e-9)


@pytest.mark.parametrize("m", [129, 700, 1500])
def test_adm_synthetic_sizes(m):
    e, nodes, zl, zt = _synthetic(m, m)
    n = len(nodes)
    a, b = _incidence(e, nodes)
    # calculate_yla_ytb
    yla = np.zeros(m * n, complex)
    ytb = np.zeros(m * n, complex)
    zl_g, zt_g = zl.copy(), zt.copy()
    assert T.calculate_yla_ytb(yla, ytb, zl_g, zt_g, a, b, m, n) == 0
    zl_o, zt_o = zl.copy(), zt.copy()
    ylo, yto = O.calculate_yla_ytb(zl_o, zt_o, a, b, m, n)
    _close(yla, ylo, 1e-10)
    _close(ytb, yto, 1e-10)
    lo, up = np.tril_indices(m), np.triu_indices(m, 1)
    _close(_cm(zl_g, m, m)[lo], _cm(zl_o, m, m)[lo], 1e-10)
    _close(_cm(zt_g, m, m)[lo], _cm(zt_o, m, m)[lo], 1e-10)
    assert np.all(_cm(zt_g, m, m)[up] == 99.0 - 7.0j)
    # fill_impedance_adm2 on the GPU's own yla / ytb, inputs left alone
    keep = [x.copy() for x in (yla, ytb, a, b)]
    yn2 = np.zeros(n * n, complex)
    assert T.fill_impedance_adm2(yn2, yla, ytb, a, b, m, n) == 0
    assert all(np.array_equal(x, k) for x, k in zip((yla, ytb, a, b), keep))
    yno = O.fill_impedance_adm2(ylo, yto, a, b, m, n)
    _close(yn2, yno, 1e-10)
    # fill_impedance_adm in one call
    yn = np.zeros(n * n, complex)
    assert T.fill_impedance_adm(yn, zl.copy(), zt.copy(), a, b, m, n) == 0
    _close(yn, yno, 1e-10)
    # solve_admittance vs zsysv
    ie = np.zeros(n, complex)
    ie[n // 3] = 1.0
    ieo = ie.copy()
    assert T.solve_admittance(yn, ie, n) == 0
    info, xo = O.solve_admittance(yno, ieo, n)
    assert info == 0
    _close(ie, xo, 1e-10)


def test_adm_singular_reports_info():
    e, nodes = O.segment_electrode([0, 0, -0.5], [35, 0, -0.5], 70, 7e-3)
    m, n = len(e), len(nodes)
    a, b = _incidence(e, nodes)
    z = np.zeros((m, m), complex)
    z[np.arange(m - 1), np.arange(m - 1)] = 1.0
    zl = z.T.ravel().copy()
    zl_keep = zl.copy()
    zt = np.eye(m, dtype=complex).ravel()
    yn = np.zeros(n * n, complex)
    assert T.fill_impedance_adm(yn, zl, zt, a, b, m, n) == m
    assert np.array_equal(zl, zl_keep)
    # singular Z_T: zl already holds its (here identity) inverse, zt unchanged
    zl2 = np.eye(m, dtype=complex).ravel() * 2.0
    zt2 = z.T.ravel().copy()
    zt2_keep = zt2.copy()
    assert T.fill_impedance_adm(yn, zl2, zt2, a, b, m, n) == m
    assert np.allclose(zl2, np.eye(m).ravel() * 0.5, rtol=0, atol=1e-15)
    assert np.array_equal(zt2, zt2_keep)


@pytest.mark.parametrize("mode", ["adm", "adm2"])
def test_c_admittance_driver_vs_oracle(mode):
    """The plain-C driver tests/c_driver/admittance_c1.c (built by build()):
    the examples' mHEM admittance route through the reference API gives the
    oracle's mHEM harmonic impedance of the 15 m conductor within 1e-8, and
    (adm2) the leakage currents ytb U add up to the unit injection."""
    import subprocess
    assert os.path.exists(tags_pkg.C_ADM_DRIVER_BIN), "run __graft_entry__.build() first"
    print("case: c_driver admittance_c1 6 freqs %s" % mode, flush=True)
    r = subprocess.run([tags_pkg.C_ADM_DRIVER_BIN, "6", mode], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    rows = np.array([[float(v) for v in line.split()] for line in r.stdout.split("\n")
                     if line.strip()])
    f = rows[:, 0]
    z = rows[:, 1] + 1j * rows[:, 2]
    assert np.allclose(f, np.logspace(2, 7, 6), rtol=1e-10)
    e, nodes = O.segment_electrode([0, 0, -0.5], [15, 0, -0.5], 30, 7e-3)
    zo = O.mhem_harmonic_impedance(e, nodes, f, 0.01, 10.0, rel_tol=1e-11, threads=8)
    assert np.max(np.abs(z - zo) / np.abs(zo)) <= 1e-8
    if mode == "adm2":
        kcl = rows[:, 3] + 1j * rows[:, 4]
        assert np.max(np.abs(kcl - 1.0)) <= 1e-9
