"""The oracle's potentials, fields and voltages (src/electrode.c:433-668),
pinned by what the mathematics fixes:
- gamma = 0: the potential of a segment is log((r1 + r2 + L)/(r1 + r2 - L))/L
  per unit current, the field that of a finite line charge in closed form;
- any gamma: the conservative field is minus the gradient of the potential
  (central differences); the magnetic potential is the potential's line
  integral along each segment's direction (il in place of it, mu/4pi);
- the voltage of a purely conservative field (s = 0) is minus the potential
  difference divided by the path length (the reference's unit-interval form);
- the mHEM ground potential equals the potential at gamma = 0."""
import numpy as np
import pytest

from oracle import oracle as O


def _grid():
    e, nodes = O.grid(3, 3, 8.0, 8.0, 2, 2, 5e-3, -0.6)
    rng = np.random.default_rng(5)
    it = rng.standard_normal(len(e)) + 1j * rng.standard_normal(len(e))
    il = rng.standard_normal(len(e)) + 1j * rng.standard_normal(len(e))
    return e, it, il


def test_static_potential_and_field_closed_forms():
    e = O.segment_electrode([1.0, -0.5, -0.7], [3.0, 0.5, -0.2], 1, 5e-3)[0]
    P = np.array([1.7, 1.2, 0.0])
    kappa = 1.0 / (4 * np.pi)  # 1 / (4 pi kappa) = 1
    a, b = e[0]["start_point"], e[0]["end_point"]
    L = e[0]["length"]
    r1, r2 = np.linalg.norm(P - a), np.linalg.norm(P - b)
    u = O.electric_potential(P, e, [1.0], 0.0, kappa, rel_tol=1e-13)
    assert abs(u - np.log((r1 + r2 + L) / (r1 + r2 - L)) / L) <= 1e-12 * abs(u)
    E = O.electric_field(P, e, None, [1.0], 0.0, 0.0, kappa, rel_tol=1e-13)
    uh = (b - a) / L
    foot = a + np.dot(P - a, uh) * uh
    rho = np.linalg.norm(P - foot)
    nh = (P - foot) / rho
    s1, s2 = np.dot(a - foot, uh), np.dot(b - foot, uh)
    ref = (nh * (s2 / r2 - s1 / r1) / rho + uh * (1 / r2 - 1 / r1)) / L
    assert np.max(np.abs(E - ref)) <= 1e-11 * np.max(np.abs(ref))


@pytest.mark.parametrize("gamma", [0.05 + 0.04j, 0.6 + 0.5j])
def test_field_is_minus_gradient_of_potential(gamma):
    e, it, il = _grid()
    kappa = 0.002 + 1e-6j
    P = np.array([3.3, 2.1, 0.0])
    E = O.electric_field(P, e, il, it, gamma, 0.0, kappa, rel_tol=1e-12)
    h = 1e-4
    grad = np.zeros(3, complex)
    for i in range(3):
        d = np.zeros(3)
        d[i] = h
        grad[i] = (O.electric_potential(P + d, e, it, gamma, kappa, rel_tol=1e-13) -
                   O.electric_potential(P - d, e, it, gamma, kappa, rel_tol=1e-13)) / (2 * h)
    assert np.max(np.abs(E + grad)) <= 1e-6 * np.max(np.abs(E))
    # with s: -s A added (and the A of il along each direction)
    s = 2j * np.pi * 1e5
    Es = O.electric_field(P, e, il, it, gamma, s, kappa, rel_tol=1e-12)
    A = O.magnetic_potential(P, e, il, gamma, rel_tol=1e-12)
    assert np.max(np.abs(Es - (E - s * A))) <= 1e-12 * np.max(np.abs(Es))


def test_magnetic_potential_is_directed_potential():
    e, it, il = _grid()
    gamma = 0.2 + 0.1j
    P = np.array([1.1, 5.2, -0.1])
    A = O.magnetic_potential(P, e, il, gamma, mur=2.0, rel_tol=1e-12)
    ref = np.zeros(3, complex)
    for k in range(len(e)):
        cur = np.zeros(len(e), complex)
        cur[k] = il[k]
        # electric_potential with kappa = 1/(4 pi) is sum it int e^{-gamma r}/r dt
        pk = O.electric_potential(P, e, cur, gamma, 1 / (4 * np.pi), rel_tol=1e-13)
        ref += pk * (e[k]["end_point"] - e[k]["start_point"])
    ref *= 2.0 * O.MU0 / (4 * np.pi)
    assert np.max(np.abs(A - ref)) <= 1e-11 * np.max(np.abs(ref))


def test_conservative_voltage_is_potential_difference():
    e, it, il = _grid()
    gamma, kappa = 0.1 + 0.08j, 0.002 + 0.0j
    p1, p2 = np.array([2.0, 2.0, 0.0]), np.array([5.0, 2.5, 0.0])
    v = O.voltage(p1, p2, e, il, it, gamma, 0.0, kappa, rel_tol=1e-10)
    du = (O.electric_potential(p2, e, it, gamma, kappa, rel_tol=1e-13) -
          O.electric_potential(p1, e, it, gamma, kappa, rel_tol=1e-13))
    assert abs(v + du / np.linalg.norm(p2 - p1)) <= 1e-8 * abs(du)


def test_mhem_gpd_is_static_potential_at_zero_gamma():
    e, it, il = _grid()
    kappa = 0.004 + 0.0j
    pts = np.array([[1.0, 1.0, 0.0], [4.0, 7.0, 0.0], [9.0, 9.0, 0.0]])
    g = O.mhem_gpd(pts, e, it, 0.0, kappa)
    ref = [O.electric_potential(P, e, it, 0.0, kappa, rel_tol=1e-13) for P in pts]
    assert np.max(np.abs(g - ref)) <= 1e-11 * np.max(np.abs(ref))


def test_intg_single_closed_form_at_zero_gamma():
    """INTG_SINGLE (src/electrode.c:279-284): L_s times the sender-axis integral
    of e^{-gamma r}/r to the receiver's middle point; at gamma = 0 that is
    log((r1 + r2 + L)/(r1 + r2 - L)) with r1, r2 from the middle point."""
    s = O.segment_electrode([0.0, 0.0, -0.5], [2.0, 0.5, -0.5], 1, 5e-3)[0]
    r = O.segment_electrode([1.0, 2.0, -0.5], [3.0, 2.0, -0.5], 1, 5e-3)[0]
    rc, v, err = O.integral(s, r, 0.0, rel_tol=1e-13, integration_type=O.INTG_SINGLE)
    M = r[0]["middle_point"]
    r1 = np.linalg.norm(M - s[0]["start_point"])
    r2 = np.linalg.norm(M - s[0]["end_point"])
    L = s[0]["length"]
    assert rc == 0 and abs(v - np.log((r1 + r2 + L) / (r1 + r2 - L))) <= 1e-12 * abs(v)
