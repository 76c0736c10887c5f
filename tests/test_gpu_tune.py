"""GPU tests of the autotuner (tags_sweep_autotune): on the 20 x 20 m grid
(N = 1185) it times the MFMA tile variants on the sweep's own update shapes
and the schedule knobs on probe solves, installs the choice, writes it to the
cache file and reads it back for the same system. A tuned sweep answers as an
untuned one to rounding (the tile shapes, step width, launch split, solve rows
and grouping change no sample's arithmetic beyond rounding): FP64 within
1e-12, the mixed form within 1e-9, both within 1e-8 of each other."""
import numpy as np
import pytest

import tags_pkg

T = tags_pkg.load()
O = tags_pkg.oracle()

pytestmark = pytest.mark.gpu


def _c2():
    rc, e, nodes = T.electrode_grid(5, 5, 20.0, 20.0, 10, 10, 6e-3, -0.5)
    f = O.logspace(2, 7, 400)[::23]  # 18 samples
    s = 2j * np.pi * f
    kap = 1e-3 + s * 15.0 * T.EPS0
    return e, nodes, s, kap


@pytest.mark.parametrize("precision", [0, 1])
def test_autotune_installs_and_caches(precision, tmp_path):
    e, nodes, s, kap = _c2()
    cache = str(tmp_path / "tuning.txt")
    sw = T.Sweep(e, nodes)
    u0, _ = sw.solve(s, kap, 0, precision)
    t = sw.autotune(precision, s[::3], kap[::3], 0, cache=cache)
    print("tuned:", t, flush=True)
    assert t["from_cache"] == 0 and t["evaluations"] >= 5
    assert t["step_width"] in (128, 256) and t["inflight"] in (1, 2, 3, 4)
    assert 0 < t["ms_per_sample"] < 1e4
    assert sw.tuning(precision)["gemm_big"] == t["gemm_big"]
    u1, st1 = sw.solve(s, kap, 0, precision)
    d = np.max(np.abs(u1 - u0) / np.abs(u0))
    print("tuned vs default: %.2e" % d, flush=True)
    assert d <= (1e-12 if precision == 0 else 1e-9)
    if precision == 1:
        assert all(q["precision_used"] == 1 for q in st1)
    sw.close()
    with open(cache) as f:
        lines = [l for l in f if " | " in l]
    assert len(lines) == 1 and ("prec=%s" % ("mixed" if precision else "fp64")) in lines[0]
    # a second sweep of the same system reads the entry instead of timing
    sw2 = T.Sweep(e, nodes)
    t2 = sw2.autotune(precision, s[::3], kap[::3], 0, cache=cache)
    assert t2["from_cache"] == 1 and t2["evaluations"] == 0
    for k in ("gemm_big", "gemm_small", "step_width", "update_chunks", "solve_rows",
              "inflight", "batch"):
        assert t2[k] == t[k], k
    u2, _ = sw2.solve(s, kap, 0, precision)
    assert np.max(np.abs(u2 - u1) / np.abs(u1)) <= 1e-12
    sw2.set_tuning(precision, None)
    assert sw2.tuning(precision) is None
    sw2.close()


def test_tuning_rejects_bad_values():
    e, nodes, s, kap = _c2()
    sw = T.Sweep(e, nodes)
    bad = dict(gemm_big=999, gemm_small=0, step_width=256, update_chunks=-1, solve_rows=-1,
               inflight=0, batch=-1, nodal_above=-1.0)
    with pytest.raises(RuntimeError):
        sw.set_tuning(1, bad)
    bad.update(gemm_big=0, step_width=100)
    with pytest.raises(RuntimeError):
        sw.set_tuning(1, bad)
    sw.close()


def test_cases_autotune_by_size_class(tmp_path):
    """tags_cases_autotune tunes the largest case of each size class and
    installs its choice on the class; the answers agree with the untuned
    solve to rounding."""
    C = tags_pkg.cases()
    fam = C.case_family(T, freq_step=16)
    Ns = np.array([c["N"] for c in fam])
    pick = sorted({int(np.argmin(np.abs(Ns - q))) for q in (70, 200, 300, 600, 900, 1400)})
    sub = [fam[i] for i in pick]
    cache = str(tmp_path / "tuning.txt")
    cs = T.Cases(sub)
    u0, _ = cs.solve(T.PREC_MIXED)
    cl = cs.autotune(T.PREC_MIXED, cache=cache)
    print("classes:", [(t["gemm_big"], t["step_width"], t["nodal_above"], t["gmres_steps"])
                       for t in cl], flush=True)
    assert 3 <= len(cl) <= len(sub)
    assert all(t["from_cache"] == 0 for t in cl)
    u1, st1 = cs.solve(T.PREC_MIXED)
    for a, b, st in zip(u1, u0, st1):
        assert np.max(np.abs(a - b) / np.abs(b)) <= 1e-9
        assert all(q["precision_used"] == 1 for q in st)
    # a second autotune reads every class from the cache
    cl2 = cs.autotune(T.PREC_MIXED, cache=cache)
    assert [t["from_cache"] for t in cl2] == [1] * len(cl)
    cs.close()


def test_c_driver_sweep_with_autotune(tmp_path):
    """The sweep and its autotuner from plain C (tests/c_driver/sweep_c2.c,
    built by build()): 40 frequencies of the 20 x 20 m grid in both
    precisions agree within 1e-8 and with the Python wrapper's FP64 sweep to
    rounding; a second run reads both tunings from the cache it wrote."""
    import subprocess
    exe = tags_pkg.C_SWEEP_DRIVER_BIN
    cache = str(tmp_path / "tuning.txt")
    outs = []
    for run in range(2):
        r = subprocess.run([exe, cache, "40"], capture_output=True, text=True, timeout=240)
        assert r.returncode == 0, (r.returncode, r.stdout[-2000:], r.stderr[-2000:])
        outs.append(r.stdout.splitlines())
    for run, lines in enumerate(outs):
        tl = [l.split() for l in lines if l.startswith("tuning")]
        assert len(tl) == 2 and all(int(t[5]) == run for t in tl), tl
    rows = np.array([[float(v) for v in l.split()] for l in outs[1] if not l.startswith("tuning")])
    f = rows[:, 0]
    zm = rows[:, 1] + 1j * rows[:, 2]
    z64 = rows[:, 3] + 1j * rows[:, 4]
    assert np.all(rows[:, 5] == 1)
    d = np.max(np.abs(zm - z64) / np.abs(z64))
    rc, e, nodes = T.electrode_grid(5, 5, 20.0, 20.0, 10, 10, 6e-3, -0.5)
    s = 2j * np.pi * f
    sw = T.Sweep(e, nodes)
    u64, _ = sw.solve(s, 1e-3 + s * 15.0 * T.EPS0, 0, T.PREC_FP64)
    sw.close()
    dp = np.max(np.abs(z64 - u64) / np.abs(u64))
    print("C driver: mixed vs FP64 %.2e, FP64 vs Python %.2e" % (d, dp), flush=True)
    assert d <= 1e-8 and dp <= 1e-12


def test_split_bf16_products_answer_as_fp32_tiles():
    """The large complex64 products on the bf16 MFMA cores (split-bf16
    variant 24: trailing updates and the node columns' products of the nodal
    plan) answer as the fp32 three-multiply tiles (variant 19) to rounding:
    the 60 x 60 m grid (N = 21,421) in the mixed form at three frequencies
    (one below the 20 kHz FGMRES switch), Z_h within 1e-9 of each other and
    within 1e-8 of the FP64 form, no FP64 fallback."""
    _, e, nodes = T.electrode_grid(31, 31, 60.0, 60.0, 4, 4, 5e-3, -0.6)
    f = np.array([100.0, 3e5, 5e6])
    s = 2j * np.pi * f
    kap = 2e-3 + s * 20.0 * T.EPS0
    base = dict(gemm_small=17, step_width=256, update_chunks=-1, solve_rows=-1, inflight=0,
                batch=-1, nodal_above=-1.0, gmres_steps=0)
    sw = T.Sweep(e, nodes)
    u64, _ = sw.solve(s, kap, 0, T.PREC_FP64)
    out = {}
    for big in (19, 24):
        sw.set_tuning(T.PREC_MIXED, dict(base, gemm_big=big))
        u, st = sw.solve(s, kap, 0, T.PREC_MIXED)
        assert all(q["precision_used"] == 1 for q in st), (big, st)
        out[big] = u
        print("variant %d vs FP64: %.2e" % (big, float(np.max(np.abs(u - u64) / np.abs(u64)))),
              flush=True)
        assert np.max(np.abs(u - u64) / np.abs(u64)) <= 1e-8
    sw.close()
    d = float(np.max(np.abs(out[24] - out[19]) / np.abs(out[19])))
    print("split-bf16 vs fp32 tiles: %.2e" % d, flush=True)
    assert d <= 1e-9
