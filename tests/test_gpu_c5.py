"""GPU parity on the lightning-transient grid at its stated size: the 30 x 30 m
grid with 5 m mesh (11 + 11 conductors, 1 m segments: m = 1320, n = 1221,
N = 3861) in the Alipio-Visacro "mean" soil (sigma0 = 5e-4 S/m,
h = 1.26 (1000 sigma0)^-0.73, g = 0.54, eps_inf'/eps0 = 12), injection at the
centre node, at the complex frequencies s_k = c + j 2 pi k / tmax of a
numerical Laplace transform with tmax = 100 us and nt = 2^14 (8193 samples;
examples/three_grids.c:49-87 and src/auxiliary.c:240-262 set the grid up this
way).

* The assembled W against the oracle's own entries (calculate_impedances +
  impedances_images at rel tol 1e-11) on sampled pairs, the touching ones
  included, at the lowest and a high |s|: 1e-9.
* The solve against LAPACK on the same (GPU-assembled) W, LAPACK's solution
  refined with residuals in extended precision, at the lowest |s| (where W is
  worst conditioned: cond_1 ~ 6e8) and across the band: FP64 within 1e-10,
  mixed within 1e-9 (both far inside the 1e-8 bar).
* Every 16th sample of the full grid of s (513 samples) in both precisions:
  mixed against FP64 within 1e-8, with no sample left above the acceptance
  rule's forward-error estimate."""
import numpy as np
import pytest

import tags_pkg

T = tags_pkg.load()
O = tags_pkg.oracle()

pytestmark = pytest.mark.gpu

NT, TMAX = 2 ** 14, 1e-4
SIGMA0 = 5e-4
H = 1.26 * (1000 * SIGMA0) ** -0.73


def _case():
    _, e, nodes = T.electrode_grid(11, 11, 30.0, 30.0, 6, 6, 5e-3, -0.5)
    inj = int(np.argmin(np.sum((nodes - [15.0, 15.0, -0.5]) ** 2, axis=1)))
    return e, nodes, inj


def _s(k):
    c = np.log(NT ** 2) / TMAX
    return c + 2j * np.pi * np.asarray(k, float) / TMAX


def _kappa(s):
    out = np.zeros(len(s), complex)
    for i, sk in enumerate(s):
        sg, er = T.alipio_soil(SIGMA0, sk, H, 0.54, 12.0)
        out[i] = sg + sk * er * T.EPS0
    return out


def _scalars(sk, kap):
    return dict(s=sk, kappa=kap, gamma=complex(np.sqrt(sk * O.MU0 * kap + 0j)), ref_l=1.0 + 0j,
                ref_t=(kap - sk * O.EPS0) / (kap + sk * O.EPS0), mur=1.0)


def _lapack_refined(W, b, steps=3):
    """LAPACK's LU solution, refined with residuals in extended precision."""
    from scipy.linalg import lapack
    lu, piv, x, info = lapack.zgesv(W, b.copy())
    assert info == 0
    wr, wi = W.real.astype(np.longdouble), W.imag.astype(np.longdouble)
    for _ in range(steps):
        xr, xi = x.real.astype(np.longdouble), x.imag.astype(np.longdouble)
        rr = b.real - (wr @ xr - wi @ xi)
        ri = b.imag - (wr @ xi + wi @ xr)
        d, info = lapack.zgetrs(lu, piv, rr.astype(float) + 1j * ri.astype(float))
        x = x + d
    return x


@pytest.mark.parametrize("k", [0, 4096])
def test_c5_assembly_vs_oracle(k):
    e, nodes, _ = _case()
    m, n = len(e), len(nodes)
    sk = _s([k])
    kap = _kappa(sk)
    print("c5 assembly: k=%d |s|/2pi=%.4g Hz" % (k, abs(sk[0]) / (2 * np.pi)), flush=True)
    rng = np.random.default_rng(k + 1)
    # random pairs, the diagonal, and every pair touching segment 660 (grid crossings)
    ends = np.concatenate([e["start_point"], e["end_point"]])
    s0 = 660
    touch = [j for j in range(m) if j != s0 and
             min(np.abs(ends[[j, j + m]][:, None, :] -
                        ends[[s0, s0 + m]][None, :, :]).sum(-1).ravel()) < 1e-9]
    pairs = np.vstack([rng.integers(0, m, (1500, 2)),
                       np.column_stack([np.arange(0, m, 7), np.arange(0, m, 7)]),
                       np.column_stack([np.full(len(touch), s0), touch])])
    assert len(touch) >= 2
    sc = _scalars(sk[0], kap[0])
    zl, zt = O.z_entries(e, O.images_of(e), pairs, sc, rel_tol=1e-11, threads=16)
    sw = T.Sweep(e, nodes)
    wl = sw.assemble_entries(sk[0], kap[0], n + pairs[:, 0], n + pairs[:, 1])
    wt = sw.assemble_entries(sk[0], kap[0], n + m + pairs[:, 0], n + m + pairs[:, 1])
    # incidence entries: node rows of the segment columns are exact
    inc = sw.assemble_entries(sk[0], kap[0], np.arange(n), n + np.zeros(n, int))
    sw.close()
    for got, ref in ((-wl, zl), (-wt, zt)):
        err = np.abs(got - ref)
        assert np.all(err <= 1e-9 * np.abs(ref) + 1e-14 * np.abs(ref).max()), err.max()
    a = O.fill_incidence_imm(e, nodes)[1].reshape(n + 2 * m, n + 2 * m).T
    assert np.array_equal(inc, a[np.arange(n), n])


@pytest.mark.parametrize("k", [0, 1, 3, 64, 2048, 8192])
def test_c5_solve_vs_lapack(k):
    e, nodes, inj = _case()
    m, n = len(e), len(nodes)
    N = n + 2 * m
    sk = _s([k])
    kap = _kappa(sk)
    print("c5 solve: k=%d |s|/2pi=%.4g Hz" % (k, abs(sk[0]) / (2 * np.pi)), flush=True)
    sw = T.Sweep(e, nodes)
    W = sw.assemble(sk[0], kap[0]).reshape(N, N).T.copy()
    sw.set_verify(True)
    u64, st64 = sw.solve(sk, kap, inj, T.PREC_FP64)
    umx, stmx = sw.solve(sk, kap, inj, T.PREC_MIXED)
    sw.close()
    b = np.zeros(N, complex)
    b[inj] = 1.0
    x = _lapack_refined(W, b)
    z = x[inj]
    assert abs(u64[0] - z) <= 1e-10 * abs(z), (u64[0], z)
    assert abs(umx[0] - z) <= 1e-9 * abs(z), (umx[0], z, stmx)
    assert 0 <= st64[0]["backward_err"] <= 1e-14
    q = stmx[0]
    if q["precision_used"] == 1:
        assert q["backward_err"] <= 1e-15 and 0 <= q["forward_est"] <= 1e-10


def test_c5_sweep_mixed_vs_fp64_every_16th():
    e, nodes, inj = _case()
    k = np.arange(0, NT // 2 + 1, 16)
    s = _s(k)
    kap = _kappa(s)
    sw = T.Sweep(e, nodes)
    u64, _ = sw.solve(s, kap, inj, T.PREC_FP64)
    umx, st = sw.solve(s, kap, inj, T.PREC_MIXED)
    sw.close()
    d = np.abs(umx - u64) / np.abs(u64)
    print("c5 every 16th: %d samples, max rel diff %.3g at k=%d, fallbacks %d, steps mean %.2f"
          % (len(s), d.max(), k[int(np.argmax(d))],
             sum(q["precision_used"] == 0 for q in st),
             np.mean([q["refine_iters"] for q in st])), flush=True)
    assert d.max() <= 1e-8
    for q in st:
        if q["precision_used"] == 1:
            assert q["backward_err"] <= 1e-15 and q["forward_est"] <= 1e-10, q
    assert np.all(u64.real > 0)
