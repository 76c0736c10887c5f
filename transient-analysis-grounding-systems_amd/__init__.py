"""Host-side mirror of the reference's C API for the per-frequency HEM solve,
backed by ``libtags_hip.so`` (HIP kernels for gfx950).

The functions keep the reference names, argument meaning and return codes
(src/electrode.h:247-317, src/linalg.h:38-81, src/grid.h:55-71) and operate on
numpy arrays laid out exactly as the reference expects:

* ``Electrode`` arrays: numpy structured arrays of ``ELECTRODE_DTYPE`` (88 bytes);
* complex matrices: flat ``complex128`` arrays (``zl``/``zt`` m*m in the
  reference's lower-half convention, ``W`` (2m+n)^2 column-major).

There is no CPU fallback: loading fails loudly when the shared library is
missing, and every call runs the HIP path.
"""
from __future__ import annotations

import atexit
import ctypes
import os
import subprocess
import sys
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB_PATH = os.path.join(HERE, "libtags_hip.so")
CSRC = os.path.join(HERE, "csrc")
SOURCES = ["fill.hip", "gemm.hip", "lu.hip", "tslu.hip", "gmres.hip", "api.hip", "nlt.hip",
           "io.hip", "field.hip"]

ELECTRODE_DTYPE = np.dtype(
    [("start_point", "<f8", 3), ("end_point", "<f8", 3), ("middle_point", "<f8", 3),
     ("length", "<f8"), ("radius", "<f8")], align=False)
assert ELECTRODE_DTYPE.itemsize == 88

INTG_NONE, INTG_DOUBLE, INTG_SINGLE, INTG_MHEM = 1, 2, 3, 4
PREC_FP64, PREC_MIXED = 0, 1
PLAN_DENSE, PLAN_STRUCTURED, PLAN_ZL_FIRST, PLAN_BLOCK_PIVOT, PLAN_GROUNDED = 0, 1, 2, 3, 4
PLAN_GROUNDED_PIVOTED = 5
PLAN_NODAL = 6
PLAN_HYBRID = 7
FORM_HEM, FORM_MHEM = 0, 1  # tags_formulation

# src/auxiliary.c:17-21
TWO_PI = 6.283185307179586476925286766559
FOUR_PI = 12.5663706143591729538505735331
MU0 = 1.256637061435917e-6
EPS0 = 8.854187817620e-12


HIPCC_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
               "-Wno-unused-result", "-Wno-unused-value"]
OBJ_DIR = os.path.join(HERE, "_build")


def hipcc_command(out=LIB_PATH):
    """One hipcc line for the whole library (what build() runs per source)."""
    return (["hipcc"] + HIPCC_FLAGS + ["-shared", "-I" + os.path.join(ROOT, "include"),
                                       "-I" + CSRC, "-o", out] +
            [os.path.join(CSRC, s) for s in SOURCES])


def build(verbose=False) -> str:
    """Compile libtags_hip.so for gfx950 in-tree (cross-compiles without a GPU):
    every source to an object in parallel, then one link."""
    from concurrent.futures import ThreadPoolExecutor
    os.makedirs(OBJ_DIR, exist_ok=True)
    inc = ["-I" + os.path.join(ROOT, "include"), "-I" + CSRC]

    def compile_one(src):
        obj = os.path.join(OBJ_DIR, os.path.splitext(src)[0] + ".o")
        r = subprocess.run(["hipcc"] + HIPCC_FLAGS + inc + ["-c", os.path.join(CSRC, src),
                                                            "-o", obj],
                           capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("hipcc failed on %s:\n%s%s" % (src, r.stdout, r.stderr))
        return obj, r.stderr

    workers = max(1, min(len(SOURCES), os.cpu_count() or 1, 8))
    with ThreadPoolExecutor(workers) as ex:
        res = list(ex.map(compile_one, SOURCES))
    tmp = LIB_PATH + ".tmp"
    r = subprocess.run(["hipcc", "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] +
                       [o for o, _ in res], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc link failed:\n" + r.stdout + r.stderr)
    if verbose:
        print("".join(e for _, e in res) + r.stderr)
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


class CDouble(ctypes.Structure):
    """x86-64 SysV passes `_Complex double` exactly like this struct."""
    _fields_ = [("re", ctypes.c_double), ("im", ctypes.c_double)]

    @staticmethod
    def of(z):
        z = complex(z)
        return CDouble(z.real, z.imag)

    def value(self):
        return complex(self.re, self.im)


class Grid(ctypes.Structure):
    _fields_ = [("vertices_x", ctypes.c_uint), ("vertices_y", ctypes.c_uint),
                ("length_x", ctypes.c_double), ("length_y", ctypes.c_double),
                ("edge_segments_x", ctypes.c_uint), ("edge_segments_y", ctypes.c_uint),
                ("radius", ctypes.c_double), ("depth", ctypes.c_double)]


class ItemStats(ctypes.Structure):
    _fields_ = [("precision_used", ctypes.c_int), ("refine_iters", ctypes.c_int),
                ("info", ctypes.c_int), ("backward_err", ctypes.c_double),
                ("refine_hist", ctypes.c_double * 8), ("forward_est", ctypes.c_double)]


class Tuning(ctypes.Structure):
    """tags_tuning: tile shapes and schedule of one sweep and precision."""
    _fields_ = [("gemm_big", ctypes.c_int), ("gemm_small", ctypes.c_int),
                ("step_width", ctypes.c_int), ("update_chunks", ctypes.c_int),
                ("solve_rows", ctypes.c_int), ("inflight", ctypes.c_int),
                ("batch", ctypes.c_int), ("nodal_above", ctypes.c_double),
                ("gmres_steps", ctypes.c_int),
                ("ms_per_sample", ctypes.c_double), ("from_cache", ctypes.c_int),
                ("evaluations", ctypes.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class CaseDesc(ctypes.Structure):
    """tags_case: one grounding case of a multi-case solve."""
    _fields_ = [("electrodes", ctypes.c_void_p), ("images", ctypes.c_void_p),
                ("num_electrodes", ctypes.c_size_t), ("nodes", ctypes.c_void_p),
                ("num_nodes", ctypes.c_size_t), ("inj_node", ctypes.c_size_t),
                ("nfreq", ctypes.c_size_t), ("s", ctypes.c_void_p), ("kappa", ctypes.c_void_p),
                ("mur", ctypes.c_double), ("u_inj", ctypes.c_void_p), ("stats", ctypes.c_void_p)]


class Timers(ctypes.Structure):
    _fields_ = [("ms_fill", ctypes.c_double), ("ms_panel", ctypes.c_double),
                ("ms_gemm", ctypes.c_double), ("ms_other", ctypes.c_double),
                ("ms_solve", ctypes.c_double), ("ms_residual", ctypes.c_double),
                ("gemm_flops", ctypes.c_double), ("trailing_ms", ctypes.c_double),
                ("trailing_flops", ctypes.c_double), ("trailing_launches", ctypes.c_long)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


EXPORTED = ["populate_electrode", "segment_electrode", "number_segments", "number_nodes",
            "electrode_grid", "self_integral", "integral", "calculate_impedances",
            "impedances_images", "fill_incidence_imm", "fill_impedance_imm",
            "solve_immittance", "tags_sweep_create", "tags_sweep_solve",
            "tags_sweep_set_profiling", "tags_sweep_get_timers", "tags_sweep_destroy",
            "tags_dense_solve", "tags_version", "tags_device_synchronize", "tags_gemm_bench",
            "tags_gemm_variants", "tags_gemm_select", "tags_autotune", "tags_sweep_set_plan",
            "tags_sweep_assemble", "tags_immittance_solve", "tags_sweep_set_inflight",
            "tags_gemm_run", "tags_set_step_width", "tags_set_solve_rows", "tags_sweep_set_nodal_above",
            "fill_incidence_adm", "fill_impedance_adm", "calculate_yla_ytb",
            "fill_impedance_adm2", "solve_admittance", "tags_sweep_set_formulation",
            "alipio_soil", "smith_longmire_soil", "laplace_trans", "inv_laplace_trans",
            "tags_laplace_batch", "tags_inv_laplace_batch", "electrodes_file", "nodes_file",
            "tags_count_lines", "tags_frequencies_file", "tags_write_complex_file",
            "electric_potential", "magnetic_potential", "electric_field", "voltage",
            "tags_fields_batch", "tags_mhem_gpd", "tags_sweep_set_verify",
            "tags_sweep_assemble_entries", "tags_sweep_set_batch", "tags_cases_create",
            "tags_cases_solve", "tags_cases_set_plan", "tags_cases_set_lanes",
            "tags_cases_set_profiling", "tags_cases_get_timers", "tags_cases_destroy",
            "tags_sweep_autotune", "tags_sweep_set_tuning", "tags_sweep_get_tuning",
            "tags_cases_autotune",
            "tags_sweep_set_gmres"]

_lib = None


def lib():
    """Loads libtags_hip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libtags_hip.so is missing; run __graft_entry__.build() "
                           "(hipcc for gfx950) first")
    L = ctypes.CDLL(LIB_PATH)
    P, D, S, I = ctypes.c_void_p, ctypes.c_double, ctypes.c_size_t, ctypes.c_int
    C = CDouble
    sig = {
        "populate_electrode": ([P, P, P, D], I),
        "segment_electrode": ([P, P, S, P, P, D], I),
        "number_segments": ([Grid], I),
        "number_nodes": ([Grid], I),
        "electrode_grid": ([Grid, P, P], I),
        "self_integral": ([P], C),
        "integral": ([P, P, C, S, D, D, I, P, P], I),
        "calculate_impedances": ([P, P, P, S, C, C, D, C, S, D, D, I], I),
        "impedances_images": ([P, P, P, P, S, C, C, D, C, C, C, S, D, D, I], I),
        "fill_incidence_imm": ([P, P, S, P, S], I),
        "fill_impedance_imm": ([P, P, P, S, S], I),
        "solve_immittance": ([P, P, S, S], I),
        "fill_incidence_adm": ([P, P, P, S, P, S], I),
        "fill_impedance_adm": ([P, P, P, P, P, S, S], I),
        "calculate_yla_ytb": ([P, P, P, P, P, P, S, S], I),
        "fill_impedance_adm2": ([P, P, P, P, P, S, S], I),
        "solve_admittance": ([P, P, S], I),
        "tags_sweep_set_formulation": ([P, I], I),
        "alipio_soil": ([P, P, D, C, D, D, D], I),
        "smith_longmire_soil": ([P, P, D, C, D], I),
        "laplace_trans": ([P, P, P, D, S], I),
        "inv_laplace_trans": ([P, P, P, D, S, I], I),
        "tags_laplace_batch": ([P, P, P, D, S, S], I),
        "tags_inv_laplace_batch": ([P, P, P, D, S, I, S], I),
        "electrodes_file": ([ctypes.c_char_p, P, S], I),
        "nodes_file": ([ctypes.c_char_p, P, S], I),
        "tags_count_lines": ([ctypes.c_char_p], ctypes.c_long),
        "tags_frequencies_file": ([ctypes.c_char_p, P, S], I),
        "tags_write_complex_file": ([ctypes.c_char_p, P, S], I),
        "electric_potential": ([P, P, S, P, C, C, S, D, D], C),
        "magnetic_potential": ([P, P, S, P, C, D, S, D, D, P], I),
        "electric_field": ([P, P, S, P, P, C, C, D, C, S, D, D, P], I),
        "voltage": ([P, P, P, S, P, P, C, C, D, C, S, D, D], C),
        "tags_fields_batch": ([P, S, P, S, P, P, S, C, C, D, C, P, P, P], I),
        "tags_mhem_gpd": ([P, S, P, S, P, S, C, C, P], I),
        "tags_sweep_create": ([ctypes.POINTER(P), P, P, S, P, S], I),
        "tags_sweep_solve": ([P, S, P, P, D, P, P, P, S, I, P, P, P], I),
        "tags_sweep_set_profiling": ([P, I], I),
        "tags_sweep_get_timers": ([P, ctypes.POINTER(Timers)], I),
        "tags_sweep_destroy": ([P], None),
        "tags_dense_solve": ([P, P, S, I, P], I),
        "tags_version": ([], ctypes.c_char_p),
        "tags_device_synchronize": ([], I),
        "tags_gemm_bench": ([I, I, I, I, I, I, ctypes.POINTER(D)], I),
        "tags_gemm_variants": ([I], I),
        "tags_gemm_select": ([I, I, I], I),
        "tags_gemm_run": ([I, I, I, I, I, P, P, P, P, P, P, I], I),
        "tags_set_panel": ([I], I),
        "tags_set_lookahead": ([I], I),
        "tags_set_update_chunks": ([I], I),
        "tags_set_step_width": ([I], I),
        "tags_set_solve_rows": ([I], I),
        "tags_sweep_autotune": ([P, I, S, P, P, D, S, ctypes.c_char_p, I, P], I),
        "tags_sweep_set_tuning": ([P, I, P], I),
        "tags_sweep_get_tuning": ([P, I, P], I),
        "tags_cases_autotune": ([P, I, ctypes.c_char_p, I, P, P], I),
        "tags_sweep_set_nodal_above": ([P, D], I),
        "tags_autotune": ([I, ctypes.POINTER(I), ctypes.POINTER(I)], I),
        "tags_sweep_set_plan": ([P, I], I),
        "tags_sweep_set_inflight": ([P, I], I),
        "tags_sweep_set_verify": ([P, I], I),
        "tags_sweep_set_batch": ([P, I], I),
        "tags_sweep_set_gmres": ([P, D, I], I),
        "tags_cases_create": ([P, P, S], I),
        "tags_cases_solve": ([P, I], I),
        "tags_cases_set_plan": ([P, I], I),
        "tags_cases_set_lanes": ([P, I], I),
        "tags_cases_set_profiling": ([P, I], I),
        "tags_cases_get_timers": ([P, P], I),
        "tags_cases_destroy": ([P], None),
        "tags_sweep_assemble_entries": ([P, C, C, D, S, P, P, P], I),
        "tags_sweep_assemble": ([P, C, C, D, P], I),
        "tags_immittance_solve": ([P, P, S, S, I, I, P], I),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _electrodes(e):
    e = np.ascontiguousarray(e)
    if e.dtype != ELECTRODE_DTYPE:
        raise TypeError("electrodes must use ELECTRODE_DTYPE")
    return e


# ---------------------------------------------------------------------------
# reference API (src/electrode.h, src/grid.h, src/linalg.h)
# ---------------------------------------------------------------------------
def populate_electrode(start, end, radius):
    """Returns (code, electrode[1]); code 1: equal points, 2: radius <= 0."""
    e = np.zeros(1, ELECTRODE_DTYPE)
    s = np.ascontiguousarray(start, np.float64)
    t = np.ascontiguousarray(end, np.float64)
    return lib().populate_electrode(_p(e), _p(s), _p(t), float(radius)), e


def segment_electrode(start, end, num_segments, radius):
    e = np.zeros(max(num_segments, 1), ELECTRODE_DTYPE)
    nodes = np.zeros((max(num_segments, 1) + 1, 3))
    s = np.ascontiguousarray(start, np.float64)
    t = np.ascontiguousarray(end, np.float64)
    rc = lib().segment_electrode(_p(e), _p(nodes), num_segments, _p(s), _p(t), float(radius))
    return rc, e[:num_segments], nodes[:num_segments + 1]


def electrode_grid(vertices_x, vertices_y, length_x, length_y, edge_segments_x,
                   edge_segments_y, radius, depth):
    g = Grid(vertices_x, vertices_y, length_x, length_y, edge_segments_x, edge_segments_y,
             radius, depth)
    m, n = lib().number_segments(g), lib().number_nodes(g)
    e = np.zeros(m, ELECTRODE_DTYPE)
    nodes = np.zeros((n, 3))
    rc = lib().electrode_grid(g, _p(e), _p(nodes))
    return rc, e, nodes


def images_of(e):
    """Image electrodes through the soil-air interface z = 0 (caller pattern of
    examples/harmonic_impedance.c:58-65)."""
    im = np.array(e, copy=True)
    for f in ("start_point", "end_point", "middle_point"):
        im[f][:, 2] *= -1
    return im


def self_integral(e):
    e = _electrodes(e).reshape(-1)[:1]
    return lib().self_integral(_p(e)).value()


def integral(sender, receiver, gamma, max_eval=0, req_abs_error=0.0, req_rel_error=1e-6,
             integration_type=INTG_DOUBLE):
    s = _electrodes(sender).reshape(-1)[:1]
    r = _electrodes(receiver).reshape(-1)[:1]
    res = np.zeros(2)
    err = np.zeros(2)
    rc = lib().integral(_p(s), _p(r), CDouble.of(gamma), max_eval, req_abs_error,
                        req_rel_error, integration_type, _p(res), _p(err))
    return rc, complex(res[0], res[1]), complex(err[0], err[1])


def calculate_impedances(zl, zt, electrodes, gamma, s, mur, kappa, max_eval=0,
                         req_abs_error=0.0, req_rel_error=1e-6, integration_type=INTG_DOUBLE):
    """In place on zl, zt (m*m complex128); returns the reference's code."""
    e = _electrodes(electrodes)
    return lib().calculate_impedances(_p(zl), _p(zt), _p(e), len(e), CDouble.of(gamma),
                                      CDouble.of(s), float(mur), CDouble.of(kappa), max_eval,
                                      req_abs_error, req_rel_error, integration_type)


def impedances_images(zl, zt, electrodes, images, gamma, s, mur, kappa, ref_l, ref_t,
                      max_eval=0, req_abs_error=0.0, req_rel_error=1e-6,
                      integration_type=INTG_DOUBLE):
    e = _electrodes(electrodes)
    im = _electrodes(images)
    return lib().impedances_images(_p(zl), _p(zt), _p(e), _p(im), len(e), CDouble.of(gamma),
                                   CDouble.of(s), float(mur), CDouble.of(kappa),
                                   CDouble.of(ref_l), CDouble.of(ref_t), max_eval,
                                   req_abs_error, req_rel_error, integration_type)


def fill_incidence_imm(wg, electrodes, nodes):
    e = _electrodes(electrodes)
    nodes = np.ascontiguousarray(nodes, np.float64)
    return lib().fill_incidence_imm(_p(wg), _p(e), len(e), _p(nodes), len(nodes))


def fill_impedance_imm(wg, zl, zt, num_electrodes, num_nodes):
    return lib().fill_impedance_imm(_p(wg), _p(zl), _p(zt), num_electrodes, num_nodes)


def solve_immittance(wg, ie, num_electrodes, num_nodes):
    """ie (complex128, N) is overwritten by [U; I_L; I_T]; returns info."""
    return lib().solve_immittance(_p(wg), _p(ie), num_electrodes, num_nodes)


def fill_incidence_adm(a, b, electrodes, nodes):
    """a, b: complex128 arrays of m*n (column-major m x n), written in place."""
    e = _electrodes(electrodes)
    nodes = np.ascontiguousarray(nodes, np.float64)
    return lib().fill_incidence_adm(_p(a), _p(b), _p(e), len(e), _p(nodes), len(nodes))


def fill_impedance_adm(yn, zl, zt, a, b, num_electrodes, num_nodes):
    """yn (n*n) written; the lower triangles of zl, zt replaced by the inverses'."""
    return lib().fill_impedance_adm(_p(yn), _p(zl), _p(zt), _p(a), _p(b), num_electrodes,
                                    num_nodes)


def calculate_yla_ytb(yla, ytb, zl, zt, a, b, num_electrodes, num_nodes):
    return lib().calculate_yla_ytb(_p(yla), _p(ytb), _p(zl), _p(zt), _p(a), _p(b),
                                   num_electrodes, num_nodes)


def fill_impedance_adm2(yn, yla, ytb, a, b, num_electrodes, num_nodes):
    return lib().fill_impedance_adm2(_p(yn), _p(yla), _p(ytb), _p(a), _p(b), num_electrodes,
                                     num_nodes)


def solve_admittance(yn, ic, num_nodes):
    """ic (complex128, n) is overwritten by the node potentials; returns info."""
    return lib().solve_admittance(_p(yn), _p(ic), num_nodes)


def fields(points, electrodes, il, it, gamma, s, kappa, mur=1.0):
    """Batched potentials and fields: points (np, 3); il, it (nrhs, m) or None.
    Returns (pot (np, nrhs), va (np, nrhs, 3), ve (np, nrhs, 3))."""
    e = _electrodes(electrodes)
    pts = np.ascontiguousarray(points, np.float64).reshape(-1, 3)
    cur = [None if c is None else np.ascontiguousarray(np.atleast_2d(c), np.complex128)
           for c in (il, it)]
    nrhs = next(c.shape[0] for c in cur if c is not None)
    np_ = len(pts)
    pot = np.zeros((np_, nrhs), np.complex128)
    va = np.zeros((np_, nrhs, 3), np.complex128)
    ve = np.zeros((np_, nrhs, 3), np.complex128)
    rc = lib().tags_fields_batch(_p(pts), np_, _p(e), len(e), _p(cur[0]), _p(cur[1]), nrhs,
                                 CDouble.of(gamma), CDouble.of(s), float(mur), CDouble.of(kappa),
                                 _p(pot), _p(va), _p(ve))
    if rc:
        raise RuntimeError("tags_fields_batch failed: %d" % rc)
    return pot, va, ve


def mhem_gpd(points, electrodes, it, gamma, kappa):
    e = _electrodes(electrodes)
    pts = np.ascontiguousarray(points, np.float64).reshape(-1, 3)
    it = np.ascontiguousarray(np.atleast_2d(it), np.complex128)
    out = np.zeros((len(pts), it.shape[0]), np.complex128)
    rc = lib().tags_mhem_gpd(_p(pts), len(pts), _p(e), len(e), _p(it), it.shape[0],
                             CDouble.of(gamma), CDouble.of(kappa), _p(out))
    if rc:
        raise RuntimeError("tags_mhem_gpd failed: %d" % rc)
    return out


def voltage(p1, p2, electrodes, il, it, gamma, s, kappa, mur=1.0):
    e = _electrodes(electrodes)
    p1 = np.ascontiguousarray(p1, np.float64)
    p2 = np.ascontiguousarray(p2, np.float64)
    il = np.ascontiguousarray(il, np.complex128)
    it = np.ascontiguousarray(it, np.complex128)
    v = lib().voltage(_p(p1), _p(p2), _p(e), len(e), _p(il), _p(it), CDouble.of(gamma),
                      CDouble.of(s), float(mur), CDouble.of(kappa), 0, 0.0, 1e-6)
    return v.value()


def read_case_files(electrodes_csv, nodes_csv):
    """Electrodes and nodes from the reference's case files, sized as the
    examples size them (line counts). Returns (rc, electrodes, nodes)."""
    L = lib()
    ne = L.tags_count_lines(os.fsencode(electrodes_csv))
    nn = L.tags_count_lines(os.fsencode(nodes_csv))
    if ne < 0 or nn < 0:
        return -10, None, None
    e = np.zeros(ne, ELECTRODE_DTYPE)
    nodes = np.zeros((nn, 3))
    rc = L.electrodes_file(os.fsencode(electrodes_csv), _p(e), ne)
    if rc == 0:
        rc = L.nodes_file(os.fsencode(nodes_csv), _p(nodes), nn)
    return rc, e, nodes


FILTER_NONE, FILTER_BLACKMAN, FILTER_HANNING, FILTER_LANCZOS, FILTER_RIESZ = 0, 1, 2, 3, 4


def alipio_soil(sigma0, s, h, g, eps_ratio):
    """(sigma(s), eps_r(s)) of the Alipio-Visacro soil model."""
    sg, er = CDouble(), CDouble()
    lib().alipio_soil(ctypes.byref(sg), ctypes.byref(er), float(sigma0), CDouble.of(s),
                      float(h), float(g), float(eps_ratio))
    return complex(sg.re, sg.im), complex(er.re, er.im)


def smith_longmire_soil(sigma0, s, erinf):
    sg, er = CDouble(), CDouble()
    lib().smith_longmire_soil(ctypes.byref(sg), ctypes.byref(er), float(sigma0), CDouble.of(s),
                              float(erinf))
    return complex(sg.re, sg.im), complex(er.re, er.im)


def laplace_trans(f, tmax):
    """Numerical Laplace transform of the rows of f (nt samples on [0, tmax]):
    returns (g, s), g of shape (..., nt // 2 + 1)."""
    f = np.ascontiguousarray(f, np.float64)
    nt = f.shape[-1]
    ns = nt // 2 + 1
    batch = int(np.prod(f.shape[:-1], dtype=np.int64)) if f.ndim > 1 else 1
    g = np.zeros(f.shape[:-1] + (ns,), np.complex128)
    s = np.zeros(ns, np.complex128)
    rc = lib().tags_laplace_batch(_p(f), _p(g), _p(s), float(tmax), nt, batch)
    if rc:
        raise RuntimeError("tags_laplace_batch failed: %d" % rc)
    return g, s


def inv_laplace_trans(g, s, tmax, nt, filter=FILTER_NONE):
    g = np.ascontiguousarray(g, np.complex128)
    s = np.ascontiguousarray(s, np.complex128)
    batch = int(np.prod(g.shape[:-1], dtype=np.int64)) if g.ndim > 1 else 1
    f = np.zeros(g.shape[:-1] + (nt,), np.float64)
    rc = lib().tags_inv_laplace_batch(_p(f), _p(g), _p(s), float(tmax), nt, int(filter), batch)
    if rc:
        raise RuntimeError("tags_inv_laplace_batch failed: %d" % rc)
    return f


def dense_solve(a, b, precision=PREC_FP64):
    """Solves a x = b (a: n x n complex128 row-major). Returns (code, x, stats)."""
    a = np.ascontiguousarray(a, np.complex128)
    x = np.array(b, dtype=np.complex128, copy=True)
    st = ItemStats()
    rc = lib().tags_dense_solve(_p(a), _p(x), len(x), precision, ctypes.byref(st))
    return rc, x, st


def gemm_bench(M, N, K, precision, variant, iters=5):
    """Mean time (ms) of one C -= A B launch of the given MFMA tile variant."""
    ms = ctypes.c_double()
    rc = lib().tags_gemm_bench(M, N, K, precision, variant, iters, ctypes.byref(ms))
    if rc:
        raise RuntimeError("tags_gemm_bench failed: %d" % rc)
    return ms.value


PANEL_TOURNAMENT, PANEL_PARTIAL = 0, 1


def gemm_run(a, b, c, precision, variant, arow=None, brow=None, crow=None, mode=1):
    """C = A B (mode 0) or C -= A B (mode 1) through one GEMM variant; returns C."""
    a = np.ascontiguousarray(a, np.complex128)
    b = np.ascontiguousarray(b, np.complex128)
    c = np.array(c, dtype=np.complex128, copy=True)
    M, K = a.shape
    N = b.shape[1]
    maps = [None if r is None else np.ascontiguousarray(r, np.int32) for r in (arow, brow, crow)]
    rc = lib().tags_gemm_run(M, N, K, precision, variant, _p(a), _p(b), _p(c), _p(maps[0]),
                             _p(maps[1]), _p(maps[2]), mode)
    if rc:
        raise RuntimeError("tags_gemm_run failed: %d" % rc)
    return c


def set_panel(kind):
    """Panel algorithm of every LU in the process (PANEL_TOURNAMENT default)."""
    rc = lib().tags_set_panel(kind)
    if rc:
        raise RuntimeError("tags_set_panel failed: %d" % rc)


def set_lookahead(on=True):
    """Look-ahead of every later LU in the process (on by default)."""
    lib().tags_set_lookahead(1 if on else 0)


def set_step_width(columns):
    rc = lib().tags_set_step_width(columns)
    if rc:
        raise RuntimeError("tags_set_step_width failed: %d" % rc)


def set_solve_rows(rows):
    """Triangular solves of every later LU: 1 or 2 (64 or 128 rows per
    work-group) for one launch per range and direction, 0 for the per-step
    launches."""
    rc = lib().tags_set_solve_rows(int(rows))
    if rc:
        raise RuntimeError("tags_set_solve_rows failed: %d" % rc)


def set_update_chunks(n):
    """Launches each step's look-ahead trailing update is split into."""
    rc = lib().tags_set_update_chunks(int(n))
    if rc:
        raise RuntimeError("tags_set_update_chunks failed: %d" % rc)


def autotune(precision):
    """Times every GEMM variant and installs the fastest (big, small) pair."""
    b, s = ctypes.c_int(), ctypes.c_int()
    rc = lib().tags_autotune(precision, ctypes.byref(b), ctypes.byref(s))
    if rc:
        raise RuntimeError("tags_autotune failed: %d" % rc)
    return b.value, s.value


def immittance_solve(a, b, m, n, precision=PREC_FP64, plan=PLAN_STRUCTURED):
    """Solves the immittance system stored row-major in the plan's order:
    [U; I_L; I_T] for PLAN_DENSE / PLAN_STRUCTURED, [I_L; I_T; U] otherwise."""
    a = np.ascontiguousarray(a, np.complex128)
    x = np.array(b, dtype=np.complex128, copy=True)
    st = ItemStats()
    rc = lib().tags_immittance_solve(_p(a), _p(x), m, n, precision, plan, ctypes.byref(st))
    return rc, x, st


# ---------------------------------------------------------------------------
# per-frequency scalars of the reference caller (examples/harmonic_impedance.c:116-123)
# ---------------------------------------------------------------------------
def soil_scalars(freq, sigma, epsr, mur=1.0):
    s = complex(0.0, TWO_PI * freq)
    kappa = sigma + s * epsr * EPS0
    gamma = complex(np.sqrt(s * mur * MU0 * kappa + 0j))
    ref_t = (kappa - s * EPS0) / (kappa + s * EPS0)
    return dict(s=s, kappa=kappa, gamma=gamma, ref_l=1.0 + 0j, ref_t=ref_t, mur=mur)


# ---------------------------------------------------------------------------
# device-resident sweep
# ---------------------------------------------------------------------------
_live = weakref.WeakSet()


@atexit.register
def _close_live():
    """Destroys the sweeps still open at interpreter exit, while the HIP
    runtime is still up (atexit hooks run before module teardown)."""
    for obj in list(_live):
        try:
            obj.close()
        except Exception:
            pass


class Sweep:
    """Device-resident frequency sweep of one grounding geometry."""

    def __init__(self, electrodes, nodes, images=None):
        e = _electrodes(electrodes)
        nodes = np.ascontiguousarray(nodes, np.float64)
        im = None if images is None else _electrodes(images)
        self._h = ctypes.c_void_p()
        rc = lib().tags_sweep_create(ctypes.byref(self._h), _p(e), _p(im), len(e), _p(nodes),
                                     len(nodes))
        if rc != 0:
            raise RuntimeError("tags_sweep_create failed: %d" % rc)
        self.m, self.n = len(e), len(nodes)
        self.N = self.n + 2 * self.m
        _live.add(self)

    def solve(self, s, kappa, inj_node=0, precision=PREC_FP64, mur=1.0, gamma=None,
              ref_l=None, ref_t=None, want_x=False):
        s = np.ascontiguousarray(s, np.complex128)
        kappa = np.ascontiguousarray(kappa, np.complex128)
        nf = len(s)
        u = np.zeros(nf, np.complex128)
        x = np.zeros((nf, self.N), np.complex128) if want_x else None
        stats = (ItemStats * nf)()
        conv = [None if a is None else np.ascontiguousarray(a, np.complex128)
                for a in (gamma, ref_l, ref_t)]
        rc = lib().tags_sweep_solve(self._h, nf, _p(s), _p(kappa), float(mur), _p(conv[0]),
                                    _p(conv[1]), _p(conv[2]), inj_node, precision, _p(u),
                                    _p(x), ctypes.cast(stats, ctypes.c_void_p))
        if rc != 0:
            raise RuntimeError("tags_sweep_solve failed: %d" % rc)
        st = [dict(precision_used=q.precision_used, refine_iters=q.refine_iters, info=q.info,
                   backward_err=q.backward_err, forward_est=q.forward_est,
                   refine_hist=[v for v in q.refine_hist if v >= 0]) for q in stats]
        return (u, st, x) if want_x else (u, st)

    def assemble(self, s, kappa, mur=1.0):
        """W of one sample as assembled on the GPU (column-major, reference order)."""
        w = np.zeros(self.N * self.N, np.complex128)
        rc = lib().tags_sweep_assemble(self._h, CDouble.of(s), CDouble.of(kappa), float(mur), _p(w))
        if rc:
            raise RuntimeError("tags_sweep_assemble failed: %d" % rc)
        return w

    def assemble_entries(self, s, kappa, rows, cols, mur=1.0):
        """W(rows[i], cols[i]) of one sample as assembled on the GPU (reference
        unknown order)."""
        rows = np.ascontiguousarray(rows, np.int64)
        cols = np.ascontiguousarray(cols, np.int64)
        out = np.zeros(len(rows), np.complex128)
        rc = lib().tags_sweep_assemble_entries(self._h, CDouble.of(s), CDouble.of(kappa),
                                               float(mur), len(rows), _p(rows), _p(cols),
                                               _p(out))
        if rc:
            raise RuntimeError("tags_sweep_assemble_entries failed: %d" % rc)
        return out

    def set_plan(self, plan):
        rc = lib().tags_sweep_set_plan(self._h, plan)
        if rc:
            raise RuntimeError("tags_sweep_set_plan failed: %d" % rc)

    def set_nodal_above(self, hz):
        """Switch frequency of PLAN_HYBRID's mixed form (nodal factor at and above)."""
        rc = lib().tags_sweep_set_nodal_above(self._h, float(hz))
        if rc:
            raise RuntimeError("tags_sweep_set_nodal_above failed: %d" % rc)

    def set_formulation(self, formulation):
        """FORM_HEM (pair integrals per sample) or FORM_MHEM (gamma-free
        potentials once, exp(-gamma rbar) scaling per sample)."""
        rc = lib().tags_sweep_set_formulation(self._h, int(formulation))
        if rc:
            raise RuntimeError("tags_sweep_set_formulation failed: %d" % rc)

    def set_verify(self, on=True):
        """FP64 form: report each sample's backward error (W assembled again)."""
        rc = lib().tags_sweep_set_verify(self._h, 1 if on else 0)
        if rc:
            raise RuntimeError("tags_sweep_set_verify failed: %d" % rc)

    def set_gmres(self, below_hz=2e4, steps=6):
        """Nodal complex64 factors below below_hz refine by FGMRES (steps Arnoldi
        steps per refinement step)."""
        rc = lib().tags_sweep_set_gmres(self._h, float(below_hz), int(steps))
        if rc:
            raise RuntimeError("tags_sweep_set_gmres failed: %d" % rc)

    def set_batch(self, samples):
        """Frequency samples per group (0: chosen from N and memory)."""
        rc = lib().tags_sweep_set_batch(self._h, int(samples))
        if rc:
            raise RuntimeError("tags_sweep_set_batch failed: %d" % rc)

    def set_inflight(self, samples):
        rc = lib().tags_sweep_set_inflight(self._h, samples)
        if rc:
            raise RuntimeError("tags_sweep_set_inflight failed: %d" % rc)

    def autotune(self, precision, s, kappa, inj_node=0, mur=1.0, cache=None, force=False):
        """Picks and installs this sweep's tile shapes and schedule for
        `precision` (tags_sweep_autotune) from timed solves of the probe
        samples s, kappa; with `cache` (a file path) a stored entry for the
        same device and system is used instead unless force, and the result is
        stored. Returns the values as a dict."""
        s = np.ascontiguousarray(s, np.complex128)
        kappa = np.ascontiguousarray(kappa, np.complex128)
        t = Tuning()
        rc = lib().tags_sweep_autotune(self._h, int(precision), len(s), _p(s), _p(kappa),
                                       float(mur), int(inj_node),
                                       None if cache is None else os.fsencode(cache),
                                       1 if force else 0, ctypes.byref(t))
        if rc:
            raise RuntimeError("tags_sweep_autotune failed: %d" % rc)
        return t.as_dict()

    def set_tuning(self, precision, values=None):
        """Installs tuned values (a dict as autotune returns; None removes)."""
        if values is None:
            rc = lib().tags_sweep_set_tuning(self._h, int(precision), None)
        else:
            t = Tuning(**{k: values[k] for k, _ in Tuning._fields_ if k in values})
            rc = lib().tags_sweep_set_tuning(self._h, int(precision), ctypes.byref(t))
        if rc:
            raise RuntimeError("tags_sweep_set_tuning failed: %d" % rc)

    def tuning(self, precision):
        """The installed tuning of `precision` as a dict, or None."""
        t = Tuning()
        rc = lib().tags_sweep_get_tuning(self._h, int(precision), ctypes.byref(t))
        return t.as_dict() if rc == 0 else None

    def set_profiling(self, on=True):
        lib().tags_sweep_set_profiling(self._h, 1 if on else 0)

    def timers(self):
        t = Timers()
        lib().tags_sweep_get_timers(self._h, ctypes.byref(t))
        return t.as_dict()

    def close(self):
        if self._h:
            lib().tags_sweep_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        # no HIP calls during interpreter finalization (the runtime's exit
        # handlers may already have run); _close_live runs before that
        if sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:
            pass


class Cases:
    """Many grounding cases solved together (tags_cases_*): `cases` is a list
    of dicts with electrodes, nodes, inj_node, s, kappa (and optionally
    images, mur). solve() returns one array of U[inj] per case and the
    per-sample statistics."""

    def __init__(self, cases):
        self._keep = []
        self._u = []
        self._stats = []
        descs = (CaseDesc * len(cases))()
        for i, c in enumerate(cases):
            e = _electrodes(c["electrodes"])
            nodes = np.ascontiguousarray(c["nodes"], np.float64)
            im = None if c.get("images") is None else _electrodes(c["images"])
            s = np.ascontiguousarray(c["s"], np.complex128)
            kap = np.ascontiguousarray(c["kappa"], np.complex128)
            u = np.zeros(len(s), np.complex128)
            st = (ItemStats * max(1, len(s)))()
            self._keep += [e, nodes, im, s, kap]
            self._u.append(u)
            self._stats.append(st)
            d = descs[i]
            d.electrodes = e.ctypes.data
            d.images = None if im is None else im.ctypes.data
            d.num_electrodes = len(e)
            d.nodes = nodes.ctypes.data
            d.num_nodes = len(nodes)
            d.inj_node = int(c.get("inj_node", 0))
            d.nfreq = len(s)
            d.s = s.ctypes.data
            d.kappa = kap.ctypes.data
            d.mur = float(c.get("mur", 1.0))
            d.u_inj = u.ctypes.data
            d.stats = ctypes.cast(st, ctypes.c_void_p)
        self._descs = descs
        self._h = ctypes.c_void_p()
        rc = lib().tags_cases_create(ctypes.byref(self._h), ctypes.cast(descs, ctypes.c_void_p),
                                     len(cases))
        if rc != 0:
            raise RuntimeError("tags_cases_create failed: %d" % rc)
        _live.add(self)

    def solve(self, precision=PREC_MIXED):
        rc = lib().tags_cases_solve(self._h, precision)
        if rc != 0:
            raise RuntimeError("tags_cases_solve failed: %d" % rc)
        us = [u.copy() for u in self._u]
        sts = [[dict(precision_used=q.precision_used, refine_iters=q.refine_iters, info=q.info,
                     backward_err=q.backward_err, forward_est=q.forward_est)
                for q in st[:len(u)]] for st, u in zip(self._stats, self._u)]
        return us, sts

    def set_plan(self, plan):
        rc = lib().tags_cases_set_plan(self._h, plan)
        if rc:
            raise RuntimeError("tags_cases_set_plan failed: %d" % rc)

    def set_lanes(self, lanes):
        rc = lib().tags_cases_set_lanes(self._h, lanes)
        if rc:
            raise RuntimeError("tags_cases_set_lanes failed: %d" % rc)

    def autotune(self, precision, cache=None, force=False):
        """Tunes one case per size class and installs its choice on the class
        (tags_cases_autotune); returns the classes' values as dicts."""
        out = (Tuning * 16)()
        nk = ctypes.c_int(0)
        rc = lib().tags_cases_autotune(self._h, int(precision),
                                       None if cache is None else os.fsencode(cache),
                                       1 if force else 0, ctypes.cast(out, ctypes.c_void_p),
                                       ctypes.byref(nk))
        if rc:
            raise RuntimeError("tags_cases_autotune failed: %d" % rc)
        return [out[i].as_dict() for i in range(min(16, nk.value))]

    def set_profiling(self, on=True):
        lib().tags_cases_set_profiling(self._h, 1 if on else 0)

    def timers(self):
        t = Timers()
        lib().tags_cases_get_timers(self._h, ctypes.byref(t))
        return t.as_dict()

    def close(self):
        if self._h:
            lib().tags_cases_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        if sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:
            pass

