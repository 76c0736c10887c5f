"""GPU parity of the modified HEM (mHEM) path: the INTG_MHEM integral and
impedance drop-ins, and the device-resident sweep with TAGS_FORM_MHEM, against
the oracle (adaptive Gauss-Kronrod of log N_f; the examples' scaling; the
admittance solve with LAPACK) and the mpmath golden vectors.

Tolerances:
* INTG_MHEM pair integrals vs mpmath (gamma = 0): 1e-10 relative;
* mHEM potentials (Z_L, Z_T with INTG_MHEM) vs the oracle at rel tol 1e-11:
  1e-9 of the largest entry;
* the sweep's assembled impedance blocks vs the oracle's scaling of its own
  potentials: 1e-9 of the largest entry;
* harmonic impedance of the mHEM sweep vs the oracle's admittance route:
  1e-8 (FP64 and mixed precision).
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


def test_mhem_integral_vs_mpmath():
    with open(os.path.join(GOLD, "pair_integrals.json")) as fh:
        pairs = [p for p in json.load(fh)["pairs"] if p["gamma_name"] == "zero"]
    for p in pairs:
        _, s = T.populate_electrode(p["sender"][0], p["sender"][1], 0.005)
        _, r = T.populate_electrode(p["receiver"][0], p["receiver"][1], 0.005)
        rc, v, err = T.integral(s, r, 0.3 + 0.7j, integration_type=T.INTG_MHEM)
        assert rc == 0 and v.imag == 0.0 and err.imag == 0.0
        ref = p["integral"][0]
        assert abs(v.real - ref) <= 1e-10 * abs(ref), (p["geometry"], v, ref)


@pytest.mark.parametrize("case", ["c1", "grid"])
def test_mhem_potentials_vs_oracle(case):
    if case == "c1":
        e, nodes = O.segment_electrode([0, 0, -0.5], [15, 0, -0.5], 30, 7e-3)
    else:
        e, nodes = O.grid(4, 3, 6.0, 4.0, 3, 2, 6e-3, -0.5)
    m = len(e)
    img = O.images_of(e)
    zl = np.full(m * m, 5.0 + 5.0j)
    zt = np.full(m * m, 5.0 + 5.0j)
    assert T.calculate_impedances(zl, zt, e, 0, 0, 0, 0, integration_type=T.INTG_MHEM) == 0
    pl, pt, pli, pti = O.mhem_potentials(e, img, rel_tol=1e-11, threads=8)
    rows, cols = np.triu_indices(m)
    flat = rows * m + cols
    for got, ref in ((zl, pl), (zt, pt)):
        assert np.abs(got[flat] - ref[flat]).max() <= 1e-9 * np.abs(ref[flat]).max()
        assert np.all(got[cols * m + rows][rows != cols] == 5.0 + 5.0j)  # upper untouched
    zi = np.zeros(m * m, complex)
    ti = np.zeros(m * m, complex)
    assert T.impedances_images(zi, ti, e, img, 0, 0, 0, 0, 0, 0,
                               integration_type=T.INTG_MHEM) == 0
    assert np.abs(zi[flat] - pli[flat]).max() <= 1e-9 * np.abs(pli[flat]).max()
    assert np.abs(ti[flat] - pti[flat]).max() <= 1e-9 * np.abs(pti[flat]).max()


def test_mhem_sweep_assembly_vs_oracle():
    e, nodes = O.grid(4, 3, 6.0, 4.0, 3, 2, 6e-3, -0.5)
    m, n = len(e), len(nodes)
    N = n + 2 * m
    img = O.images_of(e)
    pots = O.mhem_potentials(e, img, rel_tol=1e-11, threads=8)
    sw = T.Sweep(e, nodes)
    sw.set_formulation(T.FORM_MHEM)
    for f in (100.0, 3e5, 5e6):
        sc = O.soil_scalars(f, 0.01, 10.0)
        w = sw.assemble(sc["s"], sc["kappa"]).reshape(N, N).T  # (row, col)
        zl, zt = O.mhem_impedances(pots, e, img, sc)
        ZL = O._sym_lower(zl, m)
        ZT = O._sym_lower(zt, m)
        assert np.abs(-w[n:n + m, n:n + m] - ZL).max() <= 1e-9 * np.abs(ZL).max()
        assert np.abs(-w[n + m:, n + m:] - ZT).max() <= 1e-9 * np.abs(ZT).max()
    sw.close()


@pytest.mark.parametrize("case", ["c1", "grid"])
def test_mhem_sweep_vs_oracle(case):
    if case == "c1":
        e, nodes = O.segment_electrode([0, 0, -0.5], [15, 0, -0.5], 30, 7e-3)
    else:
        e, nodes = O.grid(5, 5, 20.0, 20.0, 2, 2, 6e-3, -0.5)
    f = np.array([100.0, 1e4, 3e5, 2e6, 1e7])
    s = 2j * np.pi * f
    kap = 0.01 + s * 10.0 * T.EPS0
    zo = O.mhem_harmonic_impedance(e, nodes, f, 0.01, 10.0, rel_tol=1e-11, threads=8)
    sw = T.Sweep(e, nodes)
    sw.set_formulation(T.FORM_MHEM)
    u64, _ = sw.solve(s, kap, 0, T.PREC_FP64)
    umx, _ = sw.solve(s, kap, 0, T.PREC_MIXED)
    sw.set_formulation(T.FORM_HEM)
    uh, _ = sw.solve(s, kap, 0, T.PREC_FP64)
    sw.close()
    assert np.max(np.abs(u64 - zo) / np.abs(zo)) <= 1e-8
    assert np.max(np.abs(umx - u64) / np.abs(u64)) <= 1e-8
    # the HEM sweep on the same object is a different (not identical) model
    assert np.max(np.abs(uh - u64) / np.abs(u64)) > 1e-8
