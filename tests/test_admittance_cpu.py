"""The oracle's nodal admittance formulation (src/linalg.c:142-350), pinned on
the CPU by what the mathematics fixes rather than by retyping its formulas:
- fill_incidence_adm against fill_incidence_imm (two restatements of two
  reference functions with different layouts) and the incidence invariants;
- the admittance solve Y_N U = I_E against the immittance solve of W built by
  the reference pipeline: the same node potentials U, and I_L = yla U,
  I_T = ytb U (the block elimination of W); a transposed operand, a dropped
  B term or a wrong sign changes U;
- the inverses written into zl / zt: Z Z^-1 = I, strict upper triangle kept;
- fill_impedance_adm2(yla, ytb) = fill_impedance_adm; solve_admittance's
  backward error.
"""
import numpy as np
import pytest

from oracle import oracle as O


def _case(freq, grid=(3, 3, 8.0, 8.0, 2, 2, 5e-3, -0.6)):
    e, nodes = O.grid(*grid)
    m, n = len(e), len(nodes)
    sc = O.soil_scalars(freq, 1 / 500.0, 20.0)
    wg, zl, zt = O.assemble(e, nodes, O.images_of(e), sc, rel_tol=1e-6, threads=8)
    return e, nodes, m, n, wg, zl, zt


def test_incidence_adm_matches_imm_layout():
    e, nodes = O.grid(4, 3, 12.0, 8.0, 2, 3, 5e-3, -0.5)
    m, n = len(e), len(nodes)
    N = n + 2 * m
    rc, a, b = O.fill_incidence_adm(e, nodes)
    rci, wg = O.fill_incidence_imm(e, nodes)
    assert rc == rci == 0
    W = wg.reshape(N, N).T
    A, B = O._cm(a, m, n), O._cm(b, m, n)
    assert np.array_equal(A, W[n:n + m, :n]) and np.array_equal(B, W[n + m:, :n])
    # each segment: one +1 (start node) and one -1 (end node); B = |A| / 2
    assert np.array_equal((A == 1).sum(1), np.ones(m)) and np.array_equal((A == -1).sum(1),
                                                                          np.ones(m))
    assert np.array_equal(B, np.abs(A) / 2)
    for s in range(m):
        assert np.allclose(nodes[np.argmax(A[s].real)], e[s]["start_point"])
        assert np.allclose(nodes[np.argmin(A[s].real)], e[s]["end_point"])


def test_incidence_adm_error_codes():
    e, nodes = O.segment_electrode([0, 0, -0.5], [15, 0, -0.5], 6, 7e-3)
    m = len(e)
    bad = np.vstack([nodes[:3], [[5.0, 5.0, 5.0]], nodes[3:]])
    rc, a, b = O.fill_incidence_adm(e, bad)
    assert rc == 1
    # written up to and including the unconnected node (index 3), not beyond
    assert np.count_nonzero(O._cm(a, m, len(bad))[:, 3]) == 0
    # a repeated end node: sum(A) = +1 -> -1; a repeated inner node: sum(A) = 0,
    # sum(B) = m + 1 -> 1
    assert O.fill_incidence_adm(e, np.vstack([nodes, nodes[:1]]))[0] == -1
    assert O.fill_incidence_adm(e, np.vstack([nodes, nodes[2:3]]))[0] == 1
    # aliasing a == b: the array ends as A, and sum(B) = sum(A) = 0 != m
    a2 = np.zeros(m * len(nodes), complex)
    rc2, _, _ = O.fill_incidence_adm(e, nodes, a2, a2)
    assert rc2 == -m
    assert np.array_equal(a2, O.fill_incidence_adm(e, nodes)[1])
    assert O.fill_incidence_adm(e, nodes[:0])[0] == -m


@pytest.mark.parametrize("freq", [100.0, 1e5, 3e6])
def test_admittance_solve_matches_immittance(freq):
    e, nodes, m, n, wg, zl, zt = _case(freq)
    N = n + 2 * m
    rc, a, b = O.fill_incidence_adm(e, nodes)
    assert rc == 0
    zl2, zt2 = zl.copy(), zt.copy()
    yla, ytb = O.calculate_yla_ytb(zl2, zt2, a, b, m, n)
    yn = O.fill_impedance_adm2(yla, ytb, a, b, m, n)
    ie = np.zeros(n, complex)
    ie[0] = 1.0
    info, u = O.solve_admittance(yn, ie, n)
    assert info == 0
    Y = O._sym_lower(yn, n)
    assert np.abs(Y @ u - ie).max() <= 1e-13 * np.abs(Y).max() * np.abs(u).max()
    ie_w = np.zeros(N, complex)
    ie_w[0] = 1.0
    info, x = O.solve_immittance(wg, ie_w, m, n)
    assert info == 0
    scale = np.abs(x[:n]).max()
    assert np.abs(u - x[:n]).max() <= 1e-9 * scale
    il = O._cm(yla, m, n) @ u
    it = O._cm(ytb, m, n) @ u
    assert np.abs(il - x[n:n + m]).max() <= 1e-9 * np.abs(x[n:n + m]).max()
    assert np.abs(it - x[n + m:]).max() <= 1e-9 * np.abs(x[n + m:]).max()
    # KCL: the leakage currents add up to the injection
    assert abs(it.sum() - 1.0) <= 1e-9


def test_inverse_written_to_lower_triangle():
    e, nodes, m, n, wg, zl, zt = _case(1e4)
    rc, a, b = O.fill_incidence_adm(e, nodes)
    zt = zt.copy()
    up = np.triu_indices(m, 1)
    zt_view = O._cm(zt, m, m)
    zt_view[up] = 7.0 + 3.0j  # marker in the unused strict upper triangle
    Z = O._sym_lower(zt, m)
    zl2 = zl.copy()
    O.fill_impedance_adm(zl2, zt, a, b, m, n)
    assert np.all(O._cm(zt, m, m)[up] == 7.0 + 3.0j)
    Zi = O._sym_lower(zt, m)
    assert np.abs(Z @ Zi - np.eye(m)).max() <= 1e-10


def test_adm2_equals_adm():
    e, nodes, m, n, wg, zl, zt = _case(2e5)
    rc, a, b = O.fill_incidence_adm(e, nodes)
    yn1 = O.fill_impedance_adm(zl.copy(), zt.copy(), a, b, m, n)
    yla, ytb = O.calculate_yla_ytb(zl.copy(), zt.copy(), a, b, m, n)
    yn2 = O.fill_impedance_adm2(yla, ytb, a, b, m, n)
    assert np.abs(yn1 - yn2).max() <= 1e-13 * np.abs(yn1).max()
    Y = O._cm(yn1, n, n)
    assert np.abs(Y - Y.T).max() <= 1e-10 * np.abs(Y).max()
