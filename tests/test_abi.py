"""C-ABI checks that need no GPU: libtags_hip.so loads, exports every function
include/tags_hip.h declares, keeps the reference's struct layouts, and its
host-side geometry code (src/electrode.c:12-32,103-127; src/grid.c:9-91) agrees
with the oracle exactly."""
import ctypes
import os
import re

import numpy as np
import pytest

import tags_pkg

T = tags_pkg.load()
O = tags_pkg.oracle()
HEADER = os.path.join(tags_pkg.ROOT, "include", "tags_hip.h")


def _declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", text)
    skip = {"if", "defined", "sizeof"}
    return sorted({n for n in names if n not in skip and not n.isupper()})


def test_library_exports_all_declared_symbols():
    L = T.lib()
    decl = _declared_functions()
    assert len(decl) >= 19
    missing = [n for n in decl if not hasattr(L, n)]
    assert not missing, missing
    assert set(T.EXPORTED) <= set(decl)


def test_struct_layouts():
    assert T.ELECTRODE_DTYPE.itemsize == 88
    assert ctypes.sizeof(T.Grid) == 48  # 2 uint, 2 double, 2 uint, 2 double
    assert ctypes.sizeof(T.CDouble) == 16


def test_version_string():
    assert b"gfx950" in T.lib().tags_version()


@pytest.mark.parametrize("g", [(5, 5, 20., 20., 10, 10, 6e-3, -0.5),
                               (31, 31, 60., 60., 4, 4, 5e-3, -0.6),
                               (11, 11, 30., 30., 6, 6, 5e-3, -0.5),
                               (2, 3, 10., 12., 3, 2, 1e-2, -1.0)])
def test_grid_matches_oracle(g):
    rc, e, nodes = T.electrode_grid(*g)
    assert rc == 0
    eo, no = O.grid(*g)
    assert e.tobytes() == eo.tobytes()
    assert np.array_equal(nodes, no)


def test_segment_and_populate_match_oracle():
    rc, e, nodes = T.segment_electrode([0, 0, -0.5], [15, 0, -0.5], 30, 7e-3)
    assert rc == 0
    eo, no = O.segment_electrode([0, 0, -0.5], [15, 0, -0.5], 30, 7e-3)
    assert e.tobytes() == eo.tobytes() and np.array_equal(nodes, no)
    assert T.populate_electrode([1, 2, 3], [1, 2, 3], 0.1)[0] == 1
    assert T.populate_electrode([1, 2, 3], [1, 2, 4], -0.1)[0] == 2
    assert T.segment_electrode([0, 0, 0], [1, 0, 0], 0, 0.1)[0] == 1


def test_self_integral_host_code():
    for L, a in [(0.5, 7e-3), (2.0, 5e-3), (0.1, 0.01)]:
        _, e = T.populate_electrode([0, 0, 0], [L, 0, 0], a)
        v = T.self_integral(e)
        assert v.imag == 0
        assert abs(v.real - O.self_integral(e)) <= 1e-15 * v.real


def test_integral_none_and_unknown_types_host():
    _, a = T.populate_electrode([0, 0, -0.5], [0.5, 0, -0.5], 5e-3)
    _, b = T.populate_electrode([2, 0, -0.5], [2.5, 0, -0.5], 5e-3)
    rc, v, _ = T.integral(a, b, 0.1 + 0.2j, integration_type=T.INTG_NONE)
    assert rc == 0 and v == 2.0
    assert T.integral(a, b, 0.1, integration_type=17)[0] == -10
    zl = np.zeros(4, complex)
    assert T.calculate_impedances(zl, zl, np.concatenate([a, b]), 0.1, 1j, 1.0, 1.0,
                                  integration_type=17) == -10


def test_soil_scalars_match_oracle():
    for f in (100.0, 1e4, 1e6, 5e6):
        a = T.soil_scalars(f, 2e-3, 20.0)
        b = O.soil_scalars(f, 2e-3, 20.0)
        for k in ("s", "kappa", "gamma", "ref_t"):
            assert abs(a[k] - b[k]) <= 1e-15 * abs(b[k])


def test_images_of():
    e, _ = O.grid(2, 2, 4.0, 4.0, 2, 2, 6e-3, -0.5)
    im = T.images_of(e)
    assert np.all(im["start_point"][:, 2] == 0.5) and np.all(im["length"] == e["length"])
    assert np.array_equal(im.tobytes(), O.images_of(e).tobytes())


def test_c_driver_compiles_and_links(tmp_path):
    """A plain-C11 driver written against include/tags_hip.h, making the
    reference caller's calls, compiles without warnings and links against
    libtags_hip.so (every symbol it uses resolves)."""
    out = str(tmp_path / "harmonic_c1")
    tags_pkg.build_c_driver(out)
    assert os.path.getsize(out) > 0
    out = str(tmp_path / "admittance_c1")
    tags_pkg.build_c_driver(out, os.path.join(tags_pkg.C_DRIVER_DIR, "admittance_c1.c"))
    assert os.path.getsize(out) > 0


def test_argument_checks_without_gpu():
    """Entry points that reject their arguments (or have nothing to do) before
    any device call, as the reference returns its codes."""
    L = T.lib()
    P = T._p
    f = np.zeros(8)
    g = np.zeros(5, complex)
    s = np.zeros(5, complex)
    assert L.laplace_trans(P(f), P(g), P(s), 1e-5, 1) == -12          # nt < 2
    assert L.inv_laplace_trans(P(f), P(g), P(s), 0.0, 8, 0) == -12     # tmax <= 0
    assert L.tags_fields_batch(None, 3, None, 1, None, None, 1, T.CDouble.of(0),
                               T.CDouble.of(0), 1.0, T.CDouble.of(1), None, None, None) == -12
    assert L.tags_fields_batch(None, 0, None, 1, None, None, 1, T.CDouble.of(0),
                               T.CDouble.of(0), 1.0, T.CDouble.of(1), None, None, None) == 0
    assert L.tags_mhem_gpd(None, 2, None, 1, None, 1, T.CDouble.of(0), T.CDouble.of(1),
                           None) == -12
    e, nodes = O.segment_electrode([0, 0, -0.5], [3, 0, -0.5], 3, 5e-3)
    a = np.zeros(3 * 4, complex)
    # no nodes: sum(B) = 0 != m -> -m; no segments: node 0 unconnected -> 1
    assert T.fill_incidence_adm(a, a, e, nodes[:0]) == -3
    assert T.fill_incidence_adm(a, a, e[:0], nodes) == 1
    assert L.solve_admittance(None, None, 0) == 0


def test_tuning_entry_points_check_arguments():
    """The autotuner's and the solve knob's entry points reject bad arguments
    before any device call."""
    L = T.lib()
    t = T.Tuning()
    assert L.tags_sweep_autotune(None, 1, 1, None, None, 1.0, 0, None, 0, None) == -12
    assert L.tags_sweep_set_tuning(None, 1, ctypes.byref(t)) == -12
    assert L.tags_sweep_get_tuning(None, 0, ctypes.byref(t)) == -12
    assert L.tags_set_solve_rows(3) == -12
    assert L.tags_set_solve_rows(-1) == -12
    assert L.tags_set_solve_rows(2) == 0
    # 7 ints, padding, nodal_above, gmres_steps + padding, ms_per_sample, 2 ints
    assert ctypes.sizeof(T.Tuning) == 7 * 4 + 4 + 8 + 8 + 8 + 2 * 4
