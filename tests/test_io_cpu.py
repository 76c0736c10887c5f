"""Case files (src/electrode.c:45-101; examples/harmonic_impedance.c:178-257)
through the library's host code, against the oracle's geometry and a plain
Python reading of the same formats: round trips of grids written in the
reference's formats, error codes, and the result format."""
import os

import numpy as np

import tags_pkg

T = tags_pkg.load()
O = tags_pkg.oracle()


def _write_case(tmp, e, nodes, sep=", "):
    pe = os.path.join(tmp, "electrodes.csv")
    pn = os.path.join(tmp, "nodes.csv")
    with open(pe, "w") as fh:
        fh.write("\n".join(sep.join("%.17g" % v for v in (*x["start_point"], *x["end_point"],
                                                          x["radius"])) for x in e))
    with open(pn, "w") as fh:
        fh.write("\n".join(sep.join("%.17g" % v for v in p) for p in nodes))
    return pe, pn


def test_case_files_round_trip(tmp_path):
    e, nodes = O.grid(4, 3, 12.0, 8.0, 2, 3, 5e-3, -0.5)
    pe, pn = _write_case(str(tmp_path), e, nodes)
    rc, e2, n2 = T.read_case_files(pe, pn)
    assert rc == 0 and len(e2) == len(e) and len(n2) == len(nodes)
    for k in ("start_point", "end_point", "middle_point", "length", "radius"):
        assert np.array_equal(e2[k], e[k]), k
    assert np.array_equal(n2, nodes)
    # "%lf, %lf" also takes records without the blank after the comma
    pe, pn = _write_case(str(tmp_path), e, nodes, sep=",")
    rc, e3, n3 = T.read_case_files(pe, pn)
    assert rc == 0 and np.array_equal(e3["length"], e["length"]) and np.array_equal(n3, nodes)


def test_case_file_errors(tmp_path):
    L = T.lib()
    e = np.zeros(3, T.ELECTRODE_DTYPE)
    assert L.electrodes_file(os.fsencode(str(tmp_path / "missing.csv")), T._p(e), 3) == -10
    p = tmp_path / "short.csv"
    p.write_text("0, 0, -0.5, 1, 0, -0.5, 0.01\n1, 0, -0.5, 2, 0\n")
    assert L.electrodes_file(os.fsencode(str(p)), T._p(e), 3) == 5 - 7
    p.write_text("0, 0, -0.5, 0, 0, -0.5, 0.01\n")  # equal points: populate_electrode -> 1
    assert L.electrodes_file(os.fsencode(str(p)), T._p(e), 1) == 1
    p.write_text("0, 0, -0.5, 1, 0, -0.5, -1\n")  # radius <= 0 -> 2
    assert L.electrodes_file(os.fsencode(str(p)), T._p(e), 1) == 2
    nodes = np.zeros((2, 3))
    p.write_text("1, 2, 3\n4, 5\n")
    assert L.nodes_file(os.fsencode(str(p)), T._p(nodes), 2) == 2 - 3
    assert np.array_equal(nodes[0], [1, 2, 3])


def test_frequencies_and_results(tmp_path):
    L = T.lib()
    f = O.logspace(2, 7, 50)
    p = tmp_path / "frequencies.csv"
    p.write_text("\n".join("%.17g" % x for x in f))
    n = L.tags_count_lines(os.fsencode(str(p)))
    assert n == 50
    out = np.zeros(n)
    assert L.tags_frequencies_file(os.fsencode(str(p)), T._p(out), n) == 0
    assert np.array_equal(out, f)
    z = np.array([1.5 - 2.25j, -3e-7 + 4e5j, 0j])
    q = tmp_path / "zh.csv"
    assert L.tags_write_complex_file(os.fsencode(str(q)), T._p(z), len(z)) == 0
    lines = q.read_text().splitlines()
    assert lines == ["%e%+eim" % (v.real, v.imag) for v in z]
    assert lines[0] == "1.500000e+00-2.250000e+00im"
    # the examples' own parse of their result format (Julia-style "a+bim")
    back = [complex(s.replace("im", "j")) for s in lines]
    assert np.allclose(back, z, rtol=1e-6)
