"""The mixed-case workload generator (CPU): 512 seeded cases of vertical rods,
horizontal counterpoises and rectangular grids with the stated ranges of
size, radius, soil and frequency count, built through the library's
reference-order geometry code."""
import numpy as np

import tags_pkg

T = tags_pkg.load()
C = tags_pkg.cases()
O = tags_pkg.oracle()


def test_family_ranges_and_determinism():
    ps = C.case_params()
    assert len(ps) == 512
    for a, b in zip(ps, C.case_params()):
        assert {k: v for k, v in a.items() if k != "freqs"} == \
            {k: v for k, v in b.items() if k != "freqs"}
        assert np.array_equal(a["freqs"], b["freqs"])
    kinds = [p["kind"] for p in ps]
    for k in ("rod", "counterpoise", "grid"):
        assert 120 <= kinds.count(k) <= 220, (k, kinds.count(k))
    for p in ps:
        assert 20 <= p["m"] <= 1500
        assert 5e-3 <= p["radius"] <= 12e-3
        assert 50.0 <= p["rho"] <= 5000.0 and 5.0 <= p["epsr"] <= 40.0
        assert 64 <= p["nf"] <= 512 and len(p["freqs"]) == p["nf"]
        assert p["freqs"][0] == 100.0 and abs(p["freqs"][-1] - 1e7) <= 1e-6
        if p["kind"] == "grid":
            assert 10.0 <= p["lx"] <= 40.0 and 10.0 <= p["ly"] <= 40.0
            assert 2.0 <= p["mesh"] <= 8.0
        elif p["kind"] == "rod":
            assert 1.0 <= p["length"] <= 30.0
        else:
            assert 10.0 <= p["length"] <= 100.0
    N = np.array([p["N"] for p in ps])
    assert N.min() == 61 and N.max() < 4700
    assert sum(p["nf"] for p in ps) == 145424


def test_family_geometry_matches_counts_and_oracle():
    """Every case's geometry has the predicted segment and node counts and
    matches the oracle's restatement of the reference geometry code byte for
    byte; node 0 is the injection point the workload names."""
    for p in C.case_params():
        e, nodes = C.build_case(T, p)
        assert len(e) == p["m"] and len(nodes) == p["n"], p["index"]
        if p["kind"] == "grid":
            eo, no = O.grid(p["vx"], p["vy"], p["lx"], p["ly"], p["ex"], p["ey"], p["radius"],
                            -C.DEPTH)
            assert np.allclose(nodes[0], [0.0, 0.0, -C.DEPTH])
        elif p["kind"] == "rod":
            eo, no = O.segment_electrode([0, 0, -C.DEPTH], [0, 0, -C.DEPTH - p["length"]], p["m"],
                                         p["radius"])
            assert np.allclose(nodes[0], [0.0, 0.0, -C.DEPTH])
        else:
            eo, no = O.segment_electrode([0, 0, -C.DEPTH], [p["length"], 0, -C.DEPTH], p["m"],
                                         p["radius"])
        assert e.tobytes() == eo.tobytes() and np.array_equal(nodes, no)
