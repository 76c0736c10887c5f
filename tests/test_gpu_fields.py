"""GPU parity of the potentials, fields and voltages (src/electrode.c:433-668)
and the examples' mHEM ground potential against the oracle (adaptive
Genz-Malik / Gauss-Kronrod at rel tol 1e-12, the reference's own routes).

The oracle's field and voltage integrals get an absolute tolerance of 1e-13 as
well: a field component that integrates to about zero never meets a pure
relative one. Tolerances: potentials and fields 1e-10 of the largest component over the
point set (graded Gauss-Legendre vs adaptive quadrature, both ~1e-12 or
better off the segment axes); voltages 1e-9; mHEM ground potential 1e-12
(closed form on both sides)."""
import numpy as np
import pytest

import tags_pkg

T = tags_pkg.load()
O = tags_pkg.oracle()

pytestmark = pytest.mark.gpu


def _case():
    e, nodes = O.grid(3, 3, 8.0, 8.0, 2, 2, 5e-3, -0.6)
    rng = np.random.default_rng(11)
    m = len(e)
    it = rng.standard_normal((2, m)) + 1j * rng.standard_normal((2, m))
    il = rng.standard_normal((2, m)) + 1j * rng.standard_normal((2, m))
    pts = np.array([[1.0, 1.0, 0.0], [4.0, 3.7, 0.0], [9.5, -1.0, 0.0], [2.0, 0.05, -0.6],
                    [4.0, 4.0, -0.3], [25.0, 30.0, 0.0], [1.3, 2.0, -0.61]])
    return e, it, il, pts


@pytest.mark.parametrize("freq", [100.0, 1e5, 5e6])
def test_fields_batch_vs_oracle(freq):
    e, it, il, pts = _case()
    sc = O.soil_scalars(freq, 0.002, 20.0)
    g, s, kap = sc["gamma"], sc["s"], sc["kappa"]
    pot, va, ve = T.fields(pts, e, il, it, g, s, kap, mur=1.0)
    for k in range(2):
        po = np.array([O.electric_potential(P, e, it[k], g, kap, rel_tol=1e-12) for P in pts])
        ao = np.array([O.magnetic_potential(P, e, il[k], g, rel_tol=1e-12) for P in pts])
        eo = np.array([O.electric_field(P, e, il[k], it[k], g, s, kap, rel_tol=1e-12, abs_tol=1e-13)
                       for P in pts])
        assert np.max(np.abs(pot[:, k] - po)) <= 1e-10 * np.max(np.abs(po))
        assert np.max(np.abs(va[:, k] - ao)) <= 1e-10 * np.max(np.abs(ao))
        assert np.max(np.abs(ve[:, k] - eo)) <= 1e-10 * np.max(np.abs(eo))


def test_single_point_drop_ins():
    e, it, il, pts = _case()
    sc = O.soil_scalars(2e4, 0.002, 20.0)
    g, s, kap = sc["gamma"], sc["s"], sc["kappa"]
    L = T.lib()
    P = np.ascontiguousarray(pts[1])
    it0, il0 = np.ascontiguousarray(it[0]), np.ascontiguousarray(il[0])
    u = L.electric_potential(T._p(P), T._p(e), len(e), T._p(it0), T.CDouble.of(g),
                             T.CDouble.of(kap), 0, 0.0, 1e-6).value()
    uo = O.electric_potential(P, e, it0, g, kap, rel_tol=1e-12)
    assert abs(u - uo) <= 1e-10 * abs(uo)
    va = np.zeros(3, complex)
    assert L.magnetic_potential(T._p(P), T._p(e), len(e), T._p(il0), T.CDouble.of(g), 1.0, 0,
                                0.0, 1e-6, T._p(va)) == 0
    ao = O.magnetic_potential(P, e, il0, g, rel_tol=1e-12)
    assert np.max(np.abs(va - ao)) <= 1e-10 * np.max(np.abs(ao))
    # s = 0: the conservative part is added to what ve holds
    ve = np.array([1.0, 2.0, 3.0], complex)
    assert L.electric_field(T._p(P), T._p(e), len(e), T._p(il0), T._p(it0), T.CDouble.of(g),
                            T.CDouble.of(0.0), 1.0, T.CDouble.of(kap), 0, 0.0, 1e-6,
                            T._p(ve)) == 0
    eo = O.electric_field(P, e, il0, it0, g, 0.0, kap, rel_tol=1e-12, abs_tol=1e-13)
    assert np.max(np.abs(ve - (eo + [1.0, 2.0, 3.0]))) <= 1e-10 * np.max(np.abs(eo))
    # s != 0: ve is overwritten
    ve = np.array([1.0, 2.0, 3.0], complex)
    assert L.electric_field(T._p(P), T._p(e), len(e), T._p(il0), T._p(it0), T.CDouble.of(g),
                            T.CDouble.of(s), 1.0, T.CDouble.of(kap), 0, 0.0, 1e-6,
                            T._p(ve)) == 0
    eo = O.electric_field(P, e, il0, it0, g, s, kap, rel_tol=1e-12, abs_tol=1e-13)
    assert np.max(np.abs(ve - eo)) <= 1e-10 * np.max(np.abs(eo))


@pytest.mark.parametrize("freq", [1e3, 1e6])
def test_voltage_vs_oracle(freq):
    e, it, il, pts = _case()
    sc = O.soil_scalars(freq, 0.002, 20.0)
    g, s, kap = sc["gamma"], sc["s"], sc["kappa"]
    for p1, p2 in (([2.0, 2.0, 0.0], [3.0, 2.0, 0.0]), ([-1.0, 4.0, 0.0], [6.0, 5.0, 0.0])):
        v = T.voltage(p1, p2, e, il[0], it[0], g, s, kap)
        vo = O.voltage(np.array(p1), np.array(p2), e, il[0], it[0], g, s, kap, rel_tol=1e-11,
                       abs_tol=1e-13)
        assert abs(v - vo) <= 1e-9 * abs(vo), (p1, p2, v, vo)


def test_mhem_gpd_vs_oracle():
    e, it, il, pts = _case()
    sc = O.soil_scalars(3e5, 0.002, 20.0)
    out = T.mhem_gpd(pts, e, it, sc["gamma"], sc["kappa"])
    for k in range(2):
        ref = O.mhem_gpd(pts, e, it[k], sc["gamma"], sc["kappa"])
        assert np.max(np.abs(out[:, k] - ref)) <= 1e-12 * np.max(np.abs(ref))


@pytest.mark.parametrize("gamma", [0.0, 0.05 + 0.04j, 1.5 + 1.2j])
def test_intg_single_vs_oracle(gamma):
    e, it, il, pts = _case()
    img = O.images_of(e)
    for k in (1, 5, len(e) - 1):
        rc, v, _ = T.integral(e[0], img[k], gamma, integration_type=T.INTG_SINGLE)
        rco, vo, _ = O.integral(e[0], img[k], gamma, rel_tol=1e-13,
                                integration_type=O.INTG_SINGLE)
        assert rc == rco == 0 and abs(v - vo) <= 1e-11 * abs(vo)
    m = len(e)
    sc = O.soil_scalars(2e5, 0.002, 20.0)
    zl = np.zeros(m * m, complex)
    zt = np.zeros(m * m, complex)
    assert T.calculate_impedances(zl, zt, e, sc["gamma"], sc["s"], 1.0, sc["kappa"],
                                  integration_type=T.INTG_SINGLE) == 0
    rc, zlo, zto = O.calculate_impedances(e, sc["gamma"], sc["s"], 1.0, sc["kappa"],
                                          rel_tol=1e-12, integration_type=O.INTG_SINGLE,
                                          threads=8)
    rows, cols = np.triu_indices(m)
    flat = rows * m + cols
    assert rc == 0
    assert np.abs(zl[flat] - zlo[flat]).max() <= 1e-10 * np.abs(zlo[flat]).max()
    assert np.abs(zt[flat] - zto[flat]).max() <= 1e-10 * np.abs(zto[flat]).max()


def test_fields_many_right_hand_sides():
    """Five right-hand sides (two work-group rows of at most four) against
    the same sides one at a time, and linearity in the currents."""
    e, it, il, pts = _case()
    rng = np.random.default_rng(3)
    m = len(e)
    it5 = rng.standard_normal((5, m)) + 1j * rng.standard_normal((5, m))
    il5 = rng.standard_normal((5, m)) + 1j * rng.standard_normal((5, m))
    sc = O.soil_scalars(7e4, 0.002, 20.0)
    g, s, kap = sc["gamma"], sc["s"], sc["kappa"]
    pot, va, ve = T.fields(pts, e, il5, it5, g, s, kap)
    for k in range(5):
        p1, a1, e1 = T.fields(pts, e, il5[k], it5[k], g, s, kap)
        assert np.max(np.abs(pot[:, k] - p1[:, 0])) <= 1e-13 * np.max(np.abs(p1))
        assert np.max(np.abs(va[:, k] - a1[:, 0])) <= 1e-13 * np.max(np.abs(a1))
        assert np.max(np.abs(ve[:, k] - e1[:, 0])) <= 1e-13 * np.max(np.abs(e1))
    p2, a2, e2 = T.fields(pts, e, il5[0] + 2 * il5[1], it5[0] + 2 * it5[1], g, s, kap)
    assert np.max(np.abs(p2[:, 0] - (pot[:, 0] + 2 * pot[:, 1]))) <= 1e-12 * np.max(np.abs(p2))
    assert np.max(np.abs(e2[:, 0] - (ve[:, 0] + 2 * ve[:, 1]))) <= 1e-12 * np.max(np.abs(e2))
