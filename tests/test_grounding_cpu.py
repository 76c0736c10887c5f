"""The mixed form's grounded factor, restated in numpy on the CPU.

The node voltages U (last n unknowns of the [I_L; I_T; U] order) are replaced
by U_g = U[0] and V_i = U_i - U_g: W' = T^T W T with T = I plus ones in
column g on the node rows. The complex64 factor of W' is taken without row
exchanges (Z_L block, Z_T block, node block, each on its own diagonal), and
FP64 iterative refinement runs on W itself: x += T W'^-1 T^T (b - W x).

These tests pin the identities the GPU kernels rely on (only row and column g
change; the incidence rows sum to 0 and 1) and that the refinement converges to
LAPACK's solution on W assembled by the oracle (reference pipeline)."""
import numpy as np
import pytest
import scipy.linalg as sla

from oracle import oracle as O


def _system(freq):
    e, nodes = O.grid(5, 5, 8.0, 8.0, 2, 2, 5e-3, -0.6)
    m, n = len(e), len(nodes)
    N = n + 2 * m
    sc = O.soil_scalars(freq, 1 / 500.0, 20.0)
    wg, _, _ = O.assemble(e, nodes, O.images_of(e), sc, rel_tol=1e-6, threads=8)
    Wr = np.asarray(wg).reshape(N, N).T  # column-major, lower triangle filled
    Wr = np.tril(Wr) + np.tril(Wr, -1).T
    order = np.concatenate([n + np.arange(2 * m), np.arange(n)])  # [I_L; I_T; U]
    return Wr[np.ix_(order, order)], m, n


def _tmat(N, g):
    T = np.eye(N, dtype=complex)
    T[g:, g] = 1.0
    return T


def _lu_nopiv(a):
    a = a.copy()
    for k in range(a.shape[0] - 1):
        a[k + 1:, k] /= a[k, k]
        a[k + 1:, k + 1:] -= np.outer(a[k + 1:, k], a[k, k + 1:])
    return a


def test_grounding_identities():
    W, m, n = _system(1e4)
    N = W.shape[0]
    g = 2 * m
    Wg = _tmat(N, g).T @ W @ _tmat(N, g)
    keep = np.ones(N, bool)
    keep[g] = False
    # only row and column g change
    assert np.array_equal(Wg[np.ix_(keep, keep)], W[np.ix_(keep, keep)])
    np.testing.assert_allclose(Wg[:, g][keep], W[keep][:, g:].sum(1), rtol=0, atol=1e-12)
    np.testing.assert_allclose(Wg[g, :][keep], W[g:, :][:, keep].sum(0), rtol=0, atol=1e-12)
    assert Wg[g, g] == 0  # the node block is zero (Y_E = 0)
    # column g: A 1 = 0 on the Z_L rows, B 1 = 1 on the Z_T rows, 0 on node rows
    assert np.array_equal(Wg[:m, g], np.zeros(m))
    assert np.array_equal(Wg[m:2 * m, g], np.ones(m))
    assert np.array_equal(Wg[g + 1:, g], np.zeros(n - 1))


@pytest.mark.parametrize("freq", [100.0, 1e6])
def test_unpivoted_grounded_refinement_converges(freq):
    W, m, n = _system(freq)
    N = W.shape[0]
    g = 2 * m
    T = _tmat(N, g)
    F = _lu_nopiv((T.T @ W @ T).astype(np.complex64))
    assert np.all(np.isfinite(F))
    growth = np.abs(np.triu(F)).max() / np.abs(W).max()
    assert growth < 1e3
    b = np.zeros(N, complex)
    b[g + 3] = 1.0
    x = np.zeros(N, complex)
    r = b.copy()
    wn = np.abs(W).sum(1).max()
    for it in range(1, 9):
        y = sla.solve_triangular(F, (T.T @ r).astype(np.complex64), lower=True,
                                 unit_diagonal=True)
        y = sla.solve_triangular(F, y, lower=False)
        x = x + T @ y.astype(np.complex128)
        r = b - W @ x
        berr = np.abs(r).max() / (wn * np.abs(x).max())
        if berr <= 1e-15:
            break
    assert berr <= 1e-15 and it <= 6, (it, berr)
    x0 = np.linalg.solve(W, b)
    assert np.abs(x - x0).max() <= 1e-10 * np.abs(x0).max()
