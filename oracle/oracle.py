"""CPU oracle for the per-frequency HEM solve -- TEST INFRASTRUCTURE ONLY.

ctypes front end of ``tags_oracle.c`` (the C restatement of the reference's
integrals, assembly and geometry) plus the LAPACK solve the reference calls.
Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this module; the product path never does.

The solve step follows ``solve_immittance`` (src/linalg.c:113-139): LAPACK
``zsysv`` with ``uplo='L'`` on the lower triangle of W. Here it is scipy's
bundled OpenBLAS/LAPACK (the same routine the reference links from the system
LAPACK, src/lapack.h:24-25).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

ELECTRODE_DTYPE = np.dtype(
    [("start_point", "<f8", 3), ("end_point", "<f8", 3), ("middle_point", "<f8", 3),
     ("length", "<f8"), ("radius", "<f8")], align=False)
assert ELECTRODE_DTYPE.itemsize == 88

INTG_NONE, INTG_DOUBLE, INTG_SINGLE, INTG_MHEM = 1, 2, 3, 4

# src/auxiliary.c:17-21
PI = 3.1415926535897932384626433832795029
TWO_PI = 6.283185307179586476925286766559
FOUR_PI = 12.5663706143591729538505735331
MU0 = 1.256637061435917e-6
EPS0 = 8.854187817620e-12


class _Grid(ctypes.Structure):
    _fields_ = [("vertices_x", ctypes.c_uint), ("vertices_y", ctypes.c_uint),
                ("length_x", ctypes.c_double), ("length_y", ctypes.c_double),
                ("edge_segments_x", ctypes.c_uint), ("edge_segments_y", ctypes.c_uint),
                ("radius", ctypes.c_double), ("depth", ctypes.c_double)]


_lib = None


def build() -> str:
    """Compile liboracle.so with the committed Makefile (gcc, no GPU)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        D = ctypes.c_double
        S = ctypes.c_size_t
        I = ctypes.c_int
        L.or_calculate_impedances.argtypes = [P, P, P, S, P, P, D, P, S, D, D, I, I]
        L.or_calculate_impedances.restype = I
        L.or_impedances_images.argtypes = [P, P, P, P, S, P, P, D, P, P, P, S, D, D, I, I]
        L.or_impedances_images.restype = I
        L.or_fill_incidence_imm.argtypes = [P, P, S, P, S]
        L.or_fill_incidence_imm.restype = ctypes.c_long
        L.or_fill_incidence_adm.argtypes = [P, P, P, S, P, S]
        L.or_fill_incidence_adm.restype = ctypes.c_long
        L.or_fill_impedance_imm.argtypes = [P, P, P, S, S]
        L.or_fill_impedance_imm.restype = I
        L.or_self_integral.argtypes = [P]
        L.or_self_integral.restype = D
        L.or_integral.argtypes = [P, P, P, S, D, D, I, P, P]
        L.or_integral.restype = I
        L.or_gk15_rule.argtypes = [P, P, P]
        L.or_gk15_rule.restype = None
        L.or_electric_potential.argtypes = [P, P, S, P, P, P, S, D, D, P]
        L.or_electric_potential.restype = None
        L.or_magnetic_potential.argtypes = [P, P, S, P, P, D, S, D, D, P]
        L.or_magnetic_potential.restype = None
        L.or_electric_field.argtypes = [P, P, S, P, P, P, P, D, P, S, D, D, P]
        L.or_electric_field.restype = I
        L.or_voltage.argtypes = [P, P, P, S, P, P, P, P, D, P, S, D, D, P]
        L.or_voltage.restype = None
        L.or_populate_electrode.argtypes = [P, P, P, D]
        L.or_populate_electrode.restype = I
        L.or_segment_electrode.argtypes = [P, P, S, P, P, D]
        L.or_segment_electrode.restype = I
        L.or_number_segments.argtypes = [_Grid]
        L.or_number_segments.restype = I
        L.or_number_nodes.argtypes = [_Grid]
        L.or_number_nodes.restype = I
        L.or_electrode_grid.argtypes = [_Grid, P, P]
        L.or_electrode_grid.restype = I
        L.or_z_entries.argtypes = [P, P, S, P, S, P, P, D, P, P, P, D, P, P, I]
        L.or_z_entries.restype = I
        L.or_pair_sample.argtypes = [P, P, P, S, P, D, P, I]
        L.or_pair_sample.restype = I
        L.or_logspace.argtypes = [D, D, S, P]
        L.or_logspace.restype = I
        L.or_linspace.argtypes = [D, D, S, P]
        L.or_linspace.restype = I
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def _c2(z) -> np.ndarray:
    z = complex(z)
    return np.array([z.real, z.imag], dtype=np.float64)


# ----------------------------------------------------------------------------
# geometry
# ----------------------------------------------------------------------------
def populate_electrode(start, end, radius):
    e = np.zeros(1, dtype=ELECTRODE_DTYPE)
    s = np.ascontiguousarray(start, dtype=np.float64)
    t = np.ascontiguousarray(end, dtype=np.float64)
    err = lib().or_populate_electrode(_p(e), _p(s), _p(t), float(radius))
    return err, e


def segment_electrode(start, end, num_segments, radius):
    e = np.zeros(num_segments, dtype=ELECTRODE_DTYPE)
    nodes = np.zeros((num_segments + 1, 3))
    s = np.ascontiguousarray(start, dtype=np.float64)
    t = np.ascontiguousarray(end, dtype=np.float64)
    err = lib().or_segment_electrode(_p(e), _p(nodes), num_segments, _p(s), _p(t),
                                      float(radius))
    if err:
        raise ValueError("segment_electrode failed: %d" % err)
    return e, nodes


def grid(vertices_x, vertices_y, length_x, length_y, edge_segments_x, edge_segments_y,
         radius, depth):
    g = _Grid(vertices_x, vertices_y, length_x, length_y, edge_segments_x,
              edge_segments_y, radius, depth)
    m = lib().or_number_segments(g)
    n = lib().or_number_nodes(g)
    e = np.zeros(m, dtype=ELECTRODE_DTYPE)
    nodes = np.zeros((n, 3))
    lib().or_electrode_grid(g, _p(e), _p(nodes))
    return e, nodes


def images_of(e: np.ndarray) -> np.ndarray:
    """Image electrodes through the soil-air interface z = 0 (caller pattern of
    examples/harmonic_impedance.c:58-65)."""
    im = e.copy()
    im["start_point"][:, 2] *= -1
    im["end_point"][:, 2] *= -1
    im["middle_point"][:, 2] *= -1
    return im


def logspace(a, b, n):
    u = np.zeros(n)
    lib().or_logspace(float(a), float(b), n, _p(u))
    return u


# ----------------------------------------------------------------------------
# per-frequency scalars (examples/harmonic_impedance.c:116-123)
# ----------------------------------------------------------------------------
def soil_scalars(freq, sigma, epsr, mur=1.0):
    s = complex(0.0, TWO_PI * freq)
    kappa = sigma + s * epsr * EPS0
    gamma = np.sqrt(s * mur * MU0 * kappa + 0j)
    ref_t = (kappa - s * EPS0) / (kappa + s * EPS0)
    return dict(s=s, kappa=kappa, gamma=complex(gamma), ref_l=1.0 + 0j, ref_t=ref_t,
                mur=mur)


# ----------------------------------------------------------------------------
# integrals and matrices
# ----------------------------------------------------------------------------
def self_integral(e) -> float:
    e = np.ascontiguousarray(e, dtype=ELECTRODE_DTYPE).reshape(1)
    return lib().or_self_integral(_p(e))


def gk15_rule():
    """The Gauss-Kronrod (7, 15) abscissae and weights on [-1, 1] that the 1-D
    adaptive rule uses: (x[8], kronrod weights[8], gauss weights[4])."""
    x, wk, wg = np.zeros(8), np.zeros(8), np.zeros(4)
    lib().or_gk15_rule(_p(x), _p(wk), _p(wg))
    return x, wk, wg


def integral(sender, receiver, gamma, rel_tol=1e-6, abs_tol=0.0, max_eval=0,
             integration_type=INTG_DOUBLE):
    s = np.ascontiguousarray(sender, dtype=ELECTRODE_DTYPE).reshape(1)
    r = np.ascontiguousarray(receiver, dtype=ELECTRODE_DTYPE).reshape(1)
    g = _c2(gamma)
    res = np.zeros(2)
    err = np.zeros(2)
    rc = lib().or_integral(_p(s), _p(r), _p(g), max_eval, abs_tol, rel_tol,
                           integration_type, _p(res), _p(err))
    return rc, complex(res[0], res[1]), complex(err[0], err[1])


def calculate_impedances(e, gamma, s, mur, kappa, rel_tol=1e-6, abs_tol=0.0, max_eval=0,
                         integration_type=INTG_DOUBLE, threads=1, zl=None, zt=None):
    """Returns (rc, zl, zt) as m*m complex flat arrays in the reference's
    lower-half convention (entry (i, k >= i) at i*m + k)."""
    e = np.ascontiguousarray(e, dtype=ELECTRODE_DTYPE)
    m = len(e)
    zl = np.zeros(m * m, np.complex128) if zl is None else zl
    zt = np.zeros(m * m, np.complex128) if zt is None else zt
    rc = lib().or_calculate_impedances(_p(zl), _p(zt), _p(e), m, _p(_c2(gamma)),
                                       _p(_c2(s)), float(mur), _p(_c2(kappa)), max_eval,
                                       abs_tol, rel_tol, integration_type, threads)
    return rc, zl, zt


def impedances_images(zl, zt, e, images, gamma, s, mur, kappa, ref_l, ref_t, rel_tol=1e-6,
                      abs_tol=0.0, max_eval=0, integration_type=INTG_DOUBLE, threads=1):
    e = np.ascontiguousarray(e, dtype=ELECTRODE_DTYPE)
    images = np.ascontiguousarray(images, dtype=ELECTRODE_DTYPE)
    m = len(e)
    return lib().or_impedances_images(_p(zl), _p(zt), _p(e), _p(images), m, _p(_c2(gamma)),
                                      _p(_c2(s)), float(mur), _p(_c2(kappa)),
                                      _p(_c2(ref_l)), _p(_c2(ref_t)), max_eval, abs_tol,
                                      rel_tol, integration_type, threads)


def fill_incidence_imm(e, nodes, wg=None):
    e = np.ascontiguousarray(e, dtype=ELECTRODE_DTYPE)
    nodes = np.ascontiguousarray(nodes, dtype=np.float64)
    m, n = len(e), len(nodes)
    N = n + 2 * m
    wg = np.zeros(N * N, np.complex128) if wg is None else wg
    rc = lib().or_fill_incidence_imm(_p(wg), _p(e), m, _p(nodes), n)
    return rc, wg


def fill_impedance_imm(wg, zl, zt, m, n):
    return lib().or_fill_impedance_imm(_p(wg), _p(zl), _p(zt), m, n)


def solve_immittance(wg, ie, m, n):
    """zsysv('L') on the column-major W (src/linalg.c:113-139). Returns (info, x)."""
    from scipy.linalg import lapack
    N = n + 2 * m
    W = wg.reshape(N, N).T  # column-major view -> row/col indexable
    _, _, x, info = lapack.zsysv(np.ascontiguousarray(W), ie.reshape(N, 1), lower=1)
    return info, x.reshape(N)


# ---------------------------------------------------------------------------
# nodal admittance formulation (src/linalg.c:142-350)
# ---------------------------------------------------------------------------
def fill_incidence_adm(e, nodes, a=None, b=None):
    """src/linalg.c:142-199 (C restatement). a, b: m*n complex flat arrays,
    column-major m x n (entry (segment s, node k) at m*k + s). b may be a
    (aliasing). Returns (rc, a, b)."""
    e = np.ascontiguousarray(e, dtype=ELECTRODE_DTYPE)
    nodes = np.ascontiguousarray(nodes, dtype=np.float64)
    m, n = len(e), len(nodes)
    a = np.zeros(m * n, np.complex128) if a is None else a
    b = np.zeros(m * n, np.complex128) if b is None else b
    rc = lib().or_fill_incidence_adm(_p(a), _p(b), _p(e), m, _p(nodes), n)
    return rc, a, b


def _cm(x, rows, cols):
    """(rows x cols) view of a column-major flat array."""
    return x.reshape(cols, rows).T


def _sym_lower(z, m):
    """The symmetric matrix whose lower triangle the column-major z holds."""
    lo = np.tril(_cm(z, m, m))
    return lo + np.tril(lo, -1).T


def _invert_lower(z, m):
    """zsytrf + zsytri with uplo='L' (src/linalg.c:226-236): the lower triangle
    of z (in place) becomes that of Z^-1; the strict upper triangle is not
    touched. Z^-1 here is LAPACK's zgesv on the identity (numpy), the same
    matrix up to rounding; zsytri's Bunch-Kaufman route is not in scipy."""
    inv = np.linalg.inv(_sym_lower(z, m))
    v = _cm(z, m, m)
    il = np.tril_indices(m)
    v[il] = inv[il]


def calculate_yla_ytb(zl, zt, a, b, m, n):
    """src/linalg.c:263-305: zl, zt replaced (lower triangles) by the inverses;
    yla = symm(inv(Z_L)) A, ytb = symm(inv(Z_T)) B (zsymm, side 'L', uplo 'L').
    Returns (yla, ytb) as column-major flat arrays."""
    _invert_lower(zl, m)
    _invert_lower(zt, m)
    yla = _sym_lower(zl, m) @ _cm(a, m, n)
    ytb = _sym_lower(zt, m) @ _cm(b, m, n)
    return yla.ravel(order="F"), ytb.ravel(order="F")


def fill_impedance_adm2(yla, ytb, a, b, m, n):
    """src/linalg.c:307-330: yn = a^T yla + b^T ytb (zgemm 'T','N'), all of yn,
    column-major flat."""
    yn = _cm(a, m, n).T @ _cm(yla, m, n) + _cm(b, m, n).T @ _cm(ytb, m, n)
    return yn.ravel(order="F")


def fill_impedance_adm(zl, zt, a, b, m, n):
    """src/linalg.c:201-261: the inverses into zl, zt as calculate_yla_ytb, then
    yn = a^T (symm(inv Z_L) a) + b^T (symm(inv Z_T) b). Returns yn."""
    yla, ytb = calculate_yla_ytb(zl, zt, a, b, m, n)
    return fill_impedance_adm2(yla, ytb, a, b, m, n)


def solve_admittance(yn, ie, n):
    """src/linalg.c:332-350: zsysv('L') on the column-major yn. Returns (info, x)."""
    from scipy.linalg import lapack
    Y = np.ascontiguousarray(_cm(yn, n, n))
    _, _, x, info = lapack.zsysv(Y, np.asarray(ie, np.complex128).reshape(n, 1), lower=1)
    return info, x.reshape(n)


# ---------------------------------------------------------------------------
# modified HEM (examples/harmonic_impedance.c:66-164)
# ---------------------------------------------------------------------------
def mhem_potentials(e, images, rel_tol=1e-6, threads=1):
    """The gamma-free potentials of examples/harmonic_impedance.c:78-85:
    calculate_impedances and impedances_images with INTG_MHEM (unit scalars).
    Returns (potzl, potzt, potzli, potzti), flat m*m, lower-half convention."""
    m = len(e)
    rc, pl, pt = calculate_impedances(e, 0.0, 0.0, 0.0, 0.0, rel_tol=rel_tol,
                                      integration_type=INTG_MHEM, threads=threads)
    if rc:
        raise RuntimeError("calculate_impedances (mHEM) failed: %d" % rc)
    pli = np.zeros(m * m, np.complex128)
    pti = np.zeros(m * m, np.complex128)
    rc = impedances_images(pli, pti, e, images, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, rel_tol=rel_tol,
                           integration_type=INTG_MHEM, threads=threads)
    if rc:
        raise RuntimeError("impedances_images (mHEM) failed: %d" % rc)
    return pl, pt, pli, pti


def mhem_impedances(pots, e, images, sc):
    """examples/harmonic_impedance.c:116-140: for k >= m (flat m*ne + k),
    zl = (e^{-gamma rbar} potzl + ref_l e^{-gamma rbar_i} potzli) iwu_4pi and
    zt likewise with ref_t and one_4pik; rbar = |mid(e[k]) - mid(e[m])|,
    rbar_i = |mid(e[k]) - mid(images[m])|. Returns (zl, zt)."""
    pl, pt, pli, pti = pots
    e = np.ascontiguousarray(e, dtype=ELECTRODE_DTYPE)
    images = np.ascontiguousarray(images, dtype=ELECTRODE_DTYPE)
    ne = len(e)
    rows, cols = np.triu_indices(ne)  # reference loop m (rows) <= k (cols)
    flat = rows * ne + cols
    mid, imid = e["middle_point"], images["middle_point"]
    rbar = np.sqrt(((mid[cols] - mid[rows]) ** 2).sum(1))
    rbi = np.sqrt(((mid[cols] - imid[rows]) ** 2).sum(1))
    g = sc["gamma"]
    iwu_4pi = sc["s"] * sc["mur"] * MU0 / FOUR_PI
    one_4pik = 1.0 / (FOUR_PI * sc["kappa"])
    zl = np.zeros(ne * ne, np.complex128)
    zt = np.zeros(ne * ne, np.complex128)
    ed, ei = np.exp(-g * rbar), np.exp(-g * rbi)
    zl[flat] = (ed * pl[flat] + sc["ref_l"] * ei * pli[flat]) * iwu_4pi
    zt[flat] = (ed * pt[flat] + sc["ref_t"] * ei * pti[flat]) * one_4pik
    return zl, zt


def mhem_harmonic_impedance(e, nodes, freqs, sigma, epsr, inj_node=0, mur=1.0, rel_tol=1e-6,
                            threads=1, pots=None):
    """Z_h(f) = U[inj] for a unit current by the examples' route: mHEM
    impedances, fill_impedance_adm, zsysv on Y_N (harmonic_impedance.c:141-164)."""
    images = images_of(e)
    m, n = len(e), len(nodes)
    rc, a, b = fill_incidence_adm(e, nodes)
    if rc:
        raise RuntimeError("fill_incidence_adm failed: %d" % rc)
    if pots is None:
        pots = mhem_potentials(e, images, rel_tol=rel_tol, threads=threads)
    zh = np.zeros(len(freqs), np.complex128)
    for i, f in enumerate(freqs):
        sc = soil_scalars(f, sigma, epsr, mur)
        zl, zt = mhem_impedances(pots, e, images, sc)
        yn = fill_impedance_adm(zl, zt, a, b, m, n)
        ie = np.zeros(n, np.complex128)
        ie[inj_node] = 1.0
        info, u = solve_admittance(yn, ie, n)
        if info != 0:
            raise RuntimeError("zsysv info=%d at f=%g" % (info, f))
        zh[i] = u[inj_node]
    return zh


# ---------------------------------------------------------------------------
# transient wrap (src/auxiliary.c:61-93, :240-329)
# ---------------------------------------------------------------------------
def alipio_soil(sigma0, s, h, g, eps_ratio):
    """src/auxiliary.c:61-70 (complex f = s / (2 pi j), principal branches)."""
    f = complex(s) / (TWO_PI * 1j)
    sigma = sigma0 + sigma0 * h * (f / 1e6) ** g
    t = np.tan(PI * g / 2) / (TWO_PI * EPS0 * 1e6 ** g)
    epsr = eps_ratio + t * sigma0 * h * f ** (g - 1.0)
    return complex(sigma), complex(epsr)


def smith_longmire_soil(sigma0, s, erinf):
    """src/auxiliary.c:72-93."""
    a = [3.4e6, 2.74e5, 2.58e4, 3.38e3, 5.26e2, 1.33e2, 2.72e1, 1.25e1, 4.8e0, 2.17e0, 9.8e-1,
         3.92e-1, 1.73e-1]
    fdc = complex(125.0 * sigma0) ** 0.8312
    se = 0j
    ss = 0j
    for i, ai in enumerate(a):
        F = fdc * 10.0 ** i
        r2 = (complex(s) / (1j * TWO_PI * F)) ** 2.0
        den = 1.0 + r2
        se += ai / den
        ss += ai * F * (r2 / den)
    return complex(sigma0 + TWO_PI * EPS0 * ss), complex(erinf + se)


def laplace_trans(f, tmax):
    """src/auxiliary.c:240-263 with numpy's real FFT for FFTW's r2c plan:
    rows of f -> (g, s)."""
    f = np.asarray(f, np.float64)
    nt = f.shape[-1]
    ns = nt // 2 + 1
    c = np.log(float(nt) ** 2.0) / tmax
    dt = tmax / (nt - 1)
    dw = TWO_PI / tmax
    s = c + 1j * dw * np.arange(ns)
    k = np.arange(nt)
    g = np.fft.rfft(f * (dt * np.exp(-c * k * dt)), axis=-1)
    return g, s


def inlt_window(filter, w, omega):
    """The data windows of src/auxiliary.c:275-305."""
    with np.errstate(invalid="ignore", divide="ignore"):
        if filter == 1:
            a = PI * w / omega
            return 0.42 + 0.5 * np.cos(a) + 0.08 * np.cos(2 * a)
        if filter == 2:
            return (1 + np.cos(PI * w / omega)) / 2
        if filter == 3:
            a = PI * w / omega
            return np.sin(a) / a
        if filter == 4:
            return 1.0 - np.abs(w / omega) ** 2.0
    return np.ones_like(w)


def inv_laplace_trans(g, s, tmax, nt, filter=0):
    """src/auxiliary.c:266-329 with numpy's irfft for FFTW's c2r plan (which is
    unnormalised: irfft times nt)."""
    g = np.asarray(g, np.complex128)
    s = np.asarray(s, np.complex128)
    ns = nt // 2 + 1
    dt = tmax / (nt - 1)
    sig = inlt_window(filter, s.imag, s[ns - 1].imag)
    out = np.fft.irfft(sig * g, n=nt, axis=-1) * nt
    c = s[0].real
    k = np.arange(nt)
    return out * np.exp(c * k * dt) / (nt * dt)


# ---------------------------------------------------------------------------
# potentials, fields, voltages (src/electrode.c:433-668)
# ---------------------------------------------------------------------------
def _cur(c, m):
    return np.ascontiguousarray(np.zeros(m) if c is None else c, np.complex128)


def electric_potential(point, e, it, gamma, kappa, rel_tol=1e-10, abs_tol=0.0, max_eval=0):
    e = np.ascontiguousarray(e, dtype=ELECTRODE_DTYPE)
    out = np.zeros(2)
    lib().or_electric_potential(_p(np.ascontiguousarray(point, np.float64)), _p(e), len(e),
                                _p(_cur(it, len(e))), _p(_c2(gamma)), _p(_c2(kappa)), max_eval,
                                abs_tol, rel_tol, _p(out))
    return complex(out[0], out[1])


def magnetic_potential(point, e, il, gamma, mur=1.0, rel_tol=1e-10, abs_tol=0.0, max_eval=0):
    e = np.ascontiguousarray(e, dtype=ELECTRODE_DTYPE)
    va = np.zeros(6)
    lib().or_magnetic_potential(_p(np.ascontiguousarray(point, np.float64)), _p(e), len(e),
                                _p(_cur(il, len(e))), _p(_c2(gamma)), float(mur), max_eval,
                                abs_tol, rel_tol, _p(va))
    return va[0::2] + 1j * va[1::2]


def electric_field(point, e, il, it, gamma, s, kappa, mur=1.0, rel_tol=1e-10, abs_tol=0.0,
                   max_eval=0):
    """ve of a zeroed input vector (src/electrode.c:517-558)."""
    e = np.ascontiguousarray(e, dtype=ELECTRODE_DTYPE)
    ve = np.zeros(6)
    lib().or_electric_field(_p(np.ascontiguousarray(point, np.float64)), _p(e), len(e),
                            _p(_cur(il, len(e))), _p(_cur(it, len(e))), _p(_c2(gamma)),
                            _p(_c2(s)), float(mur), _p(_c2(kappa)), max_eval, abs_tol, rel_tol,
                            _p(ve))
    return ve[0::2] + 1j * ve[1::2]


def voltage(p1, p2, e, il, it, gamma, s, kappa, mur=1.0, rel_tol=1e-10, abs_tol=0.0,
            max_eval=0):
    e = np.ascontiguousarray(e, dtype=ELECTRODE_DTYPE)
    out = np.zeros(2)
    lib().or_voltage(_p(np.ascontiguousarray(p1, np.float64)),
                     _p(np.ascontiguousarray(p2, np.float64)), _p(e), len(e),
                     _p(_cur(il, len(e))), _p(_cur(it, len(e))), _p(_c2(gamma)), _p(_c2(s)),
                     float(mur), _p(_c2(kappa)), max_eval, abs_tol, rel_tol, _p(out))
    return complex(out[0], out[1])


def mhem_gpd(points, e, it, gamma, kappa):
    """examples/visacro57emc01.c:418-430 for one right-hand side: sum_m
    one_4pik it[m] e^{-gamma rbar} log(r0) / L_m at each point."""
    e = np.ascontiguousarray(e, dtype=ELECTRODE_DTYPE)
    pts = np.asarray(points, np.float64).reshape(-1, 3)
    one_4pik = 1.0 / (FOUR_PI * kappa)
    out = np.zeros(len(pts), np.complex128)
    for p, P in enumerate(pts):
        rbar = np.sqrt(((P - e["middle_point"]) ** 2).sum(1))
        r1 = np.sqrt(((P - e["start_point"]) ** 2).sum(1))
        r2 = np.sqrt(((P - e["end_point"]) ** 2).sum(1))
        r0 = (r1 + r2 + e["length"]) / (r1 + r2 - e["length"])
        out[p] = np.sum(one_4pik * np.asarray(it) * np.exp(-gamma * rbar) * np.log(r0) /
                        e["length"])
    return out


def pair_sample(e, images, pairs, gamma, rel_tol=1e-6, threads=1):
    e = np.ascontiguousarray(e, dtype=ELECTRODE_DTYPE)
    images = np.ascontiguousarray(images, dtype=ELECTRODE_DTYPE)
    pairs = np.ascontiguousarray(pairs, dtype=np.int64)
    out = np.zeros((len(pairs), 4))
    lib().or_pair_sample(_p(e), _p(images), _p(pairs), len(pairs), _p(_c2(gamma)), rel_tol,
                         _p(out), threads)
    return out[:, 0] + 1j * out[:, 1], out[:, 2] + 1j * out[:, 3]


def z_entries(e, images, pairs, sc, rel_tol=1e-6, threads=1):
    """Z_L(i, k), Z_T(i, k) after calculate_impedances + impedances_images for
    each pair (i, k) (symmetric: the lower-half entry of min, max), with the
    same per-entry arithmetic as the full-matrix routines."""
    e = np.ascontiguousarray(e, dtype=ELECTRODE_DTYPE)
    images = np.ascontiguousarray(images, dtype=ELECTRODE_DTYPE)
    pairs = np.ascontiguousarray(pairs, dtype=np.int64).reshape(-1, 2)
    zl = np.zeros(len(pairs), np.complex128)
    zt = np.zeros(len(pairs), np.complex128)
    rc = lib().or_z_entries(_p(e), _p(images), len(e), _p(pairs), len(pairs),
                            _p(_c2(sc["gamma"])), _p(_c2(sc["s"])), sc["mur"],
                            _p(_c2(sc["kappa"])), _p(_c2(sc["ref_l"])), _p(_c2(sc["ref_t"])),
                            rel_tol, _p(zl), _p(zt), threads)
    if rc:
        raise RuntimeError("or_z_entries failed: %d" % rc)
    return zl, zt


def assemble(e, nodes, images, sc, rel_tol=1e-6, threads=1):
    """Full W (column-major flat) for one frequency, reference pipeline:
    calculate_impedances -> impedances_images -> fill_incidence_imm ->
    fill_impedance_imm (caller pattern of examples/harmonic_impedance.c)."""
    m, n = len(e), len(nodes)
    rc, zl, zt = calculate_impedances(e, sc["gamma"], sc["s"], sc["mur"], sc["kappa"],
                                      rel_tol=rel_tol, threads=threads)
    if rc:
        raise RuntimeError("calculate_impedances failed: %d" % rc)
    rc = impedances_images(zl, zt, e, images, sc["gamma"], sc["s"], sc["mur"], sc["kappa"],
                           sc["ref_l"], sc["ref_t"], rel_tol=rel_tol, threads=threads)
    if rc:
        raise RuntimeError("impedances_images failed: %d" % rc)
    rc, wg = fill_incidence_imm(e, nodes)
    if rc:
        raise RuntimeError("fill_incidence_imm failed: %d" % rc)
    fill_impedance_imm(wg, zl, zt, m, n)
    return wg, zl, zt


def harmonic_impedance(e, nodes, freqs, sigma, epsr, inj_node=0, mur=1.0, rel_tol=1e-6,
                       threads=1, return_x=False):
    """Z_h(f) = U[inj] for a unit current injected at `inj_node`."""
    images = images_of(e)
    m, n = len(e), len(nodes)
    N = n + 2 * m
    zh = np.zeros(len(freqs), np.complex128)
    xs = []
    for i, f in enumerate(freqs):
        sc = soil_scalars(f, sigma, epsr, mur)
        wg, _, _ = assemble(e, nodes, images, sc, rel_tol=rel_tol, threads=threads)
        ie = np.zeros(N, np.complex128)
        ie[inj_node] = 1.0
        info, x = solve_immittance(wg, ie, m, n)
        if info != 0:
            raise RuntimeError("zsysv info=%d at f=%g" % (info, f))
        zh[i] = x[inj_node]
        if return_x:
            xs.append(x)
    return (zh, xs) if return_x else zh
