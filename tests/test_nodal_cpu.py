"""The nodal plan's block elimination, restated in numpy on the CPU.

With the unknowns stored as [I_L; I_T; U], W = [[-Z_L, 0, A], [0, -Z_T, B],
[A^T, B^T, 0]]. The nodal plan factors Z_L and Z_T separately, forms
X = Z^-1 A for each, sets the node block to its Schur complement
S = W_UU - (A^T X_L + B^T X_T) from the node rows' non-zeros only, and solves
through the blocks. These tests pin what the GPU kernels rely on:
- the node rows are exactly the incidence (A = +-1, B = 0.5 on the segments
  touching the node, nothing else), so their non-zeros are few;
- W_UU is zero and Z_L, Z_T are uncoupled;
- block elimination reproduces LAPACK's solution of W (oracle W, several
  frequencies);
- Z_L and Z_T factored without row exchanges (as the hybrid plan's complex64
  factor does) stay stable: small growth, as for complex symmetric matrices
  with a definite real or imaginary part."""
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
    Wr = np.asarray(wg).reshape(N, N).T
    Wr = np.tril(Wr) + np.tril(Wr, -1).T
    order = np.concatenate([n + np.arange(2 * m), np.arange(n)])  # [I_L; I_T; U]
    return Wr[np.ix_(order, order)], m, n


def _lu_nopiv(a):
    a = a.copy()
    for k in range(a.shape[0] - 1):
        a[k + 1:, k] /= a[k, k]
        a[k + 1:, k + 1:] -= np.outer(a[k + 1:, k], a[k, k + 1:])
    return a


def test_node_rows_are_the_incidence():
    W, m, n = _system(1e4)
    cur = W[2 * m:, :2 * m]
    vals = np.unique(np.round(cur[cur != 0], 12))
    assert set(np.real(vals)) <= {-1.0, 0.5, 1.0} and np.all(np.imag(vals) == 0)
    # every segment touches exactly two nodes, in both its current columns
    assert np.array_equal((cur[:, :m] != 0).sum(0), np.full(m, 2))
    assert np.array_equal((cur[:, m:] != 0).sum(0), np.full(m, 2))
    # a node row's non-zeros: its segments' I_L and I_T columns, the same set
    assert np.array_equal(cur[:, :m] != 0, cur[:, m:] != 0)
    assert np.count_nonzero(W[2 * m:, 2 * m:]) == 0
    assert np.count_nonzero(W[:m, m:2 * m]) == 0 and np.count_nonzero(W[m:2 * m, :m]) == 0


@pytest.mark.parametrize("freq", [100.0, 1e4, 1e6])
def test_block_elimination_matches_lapack(freq):
    W, m, n = _system(freq)
    N = W.shape[0]
    b = np.zeros(N, complex)
    b[2 * m + 3] = 1.0
    x0 = np.linalg.solve(W, b)
    Z = [W[:m, :m], W[m:2 * m, m:2 * m]]
    C = [W[:m, 2 * m:], W[m:2 * m, 2 * m:]]      # A, B (current rows, node columns)
    R = [W[2 * m:, :m], W[2 * m:, m:2 * m]]      # A^T, B^T (node rows)
    f = [sla.lu_factor(z) for z in Z]
    X = [sla.lu_solve(fi, c) for fi, c in zip(f, C)]
    # the Schur complement from each node row's non-zeros only
    S = W[2 * m:, 2 * m:].copy()
    for i in range(n):
        for r, x in zip(R, X):
            nz = np.nonzero(r[i])[0]
            S[i] -= r[i, nz] @ x[nz]
    z = [sla.lu_solve(fi, b[k * m:(k + 1) * m]) for k, fi in enumerate(f)]
    xu = np.linalg.solve(S, b[2 * m:] - R[0] @ z[0] - R[1] @ z[1])
    xi = [zi - xk @ xu for zi, xk in zip(z, X)]
    x = np.concatenate([xi[0], xi[1], xu])
    # both solutions carry forward errors of about cond(W) u (cond(W) ~ 4e6 at
    # 100 Hz here, so ~4e-10); a wrong elimination step is off by O(1)
    assert np.abs(x - x0).max() <= 1e-8 * np.abs(x0).max()


@pytest.mark.parametrize("freq", [100.0, 1e5, 5e6])
def test_current_blocks_factor_without_row_exchanges(freq):
    W, m, n = _system(freq)
    for z in (W[:m, :m], W[m:2 * m, m:2 * m]):
        lu = _lu_nopiv(z)
        growth = np.abs(np.triu(lu)).max() / np.abs(z).max()
        assert growth <= 2.0, growth
        rel = np.abs(np.tril(lu, -1) + np.eye(m)) @ np.abs(np.triu(lu))
        assert rel.max() <= 4.0 * np.abs(z).max()
