"""Frequency-sample throughput of the per-frequency HEM solve on one MI355X.

Workload: the 60 x 60 m grid with 2 m mesh (31 + 31 conductors, 0.5 m
segments, radius 5 mm, depth 0.6 m): m = 7,440 segments, n = 6,541 nodes,
immittance order N = 21,421; soil rho = 500 Ohm m, eps_r = 20; 256 log-spaced
frequencies 100 Hz - 5 MHz; unit current injected at corner node 0.

A step is one frequency sample end to end on the GPU: pair-integral fill and
immittance assembly in HBM, LU factorisation, triangular solves (and, for the
mixed form, FP64 iterative refinement). The timed steps walk frequencies spread
over the whole sweep; the sweep keeps --inflight samples in flight (default 4),
as a frequency sweep runs. `value` is the mixed form (complex64 MFMA LU + FP64
refinement); the plain FP64 form is timed on the same samples the same way and
reported beside it together with the largest relative difference of Z_h
between the two. The per-kernel figures (roofline, timers) come from a second
timed pass over the same samples with one in flight, so that each launch's
HIP-event duration is its own.

The same line carries the other configurations of the workload family as
keys, each timed over all its samples in both precisions (the library groups
samples per launch), with the trailing update's roofline from a profiled pass
and the oracle on the host as CPU baseline:
  c1  the 15 m conductor (N = 91), 200 frequencies;
  c2  the 20 x 20 m grid (N = 1185), 400 frequencies;
  c5  the 30 x 30 m grid in Alipio-Visacro soil (N = 3861), every 8th of the
      8193 numerical-Laplace samples;
  c4  512 seeded cases (N = 61 .. 4246), every 8th of their frequencies, in
      one multi-case solve per precision.

Prints one JSON line. Single GPU only. `--rows` instead times the calls next
to the path (nodal admittance, the Laplace transform pair, fields), and
`--transient` runs the lightning transient on the 30 x 30 m grid end to end;
each prints its own JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "tests"))

METRIC = "freq samples/s, 60×60 m grid: FP64 vs FP32+refine at ≤1e-8 rel GPR err"
PEAK = {"fp64": 78.6, "fp32": 157.3, "bf16": 2500.0}  # dense MFMA TFLOP/s, MI355X
# complex64 GEMM variants that run on the bf16 MFMA cores (every fp32 operand
# as three bf16 parts, six bf16 part products per real product), and the
# library's default large-product variants
X6_VARIANTS = range(22, 33)
X6_PARTS_PRODUCTS = 6
DEFAULT_BIG = {"mixed": 24, "fp64": 13}
GRID = (31, 31, 60.0, 60.0, 4, 4, 5e-3, -0.6)
SIGMA, EPSR = 1.0 / 500.0, 20.0
NFREQ = 256
PLAN_NAMES = ["dense", "structured", "zl_first", "block_pivot", "grounded", "grounded_pivoted",
              "nodal", "hybrid"]


def lu_flops(N):
    """(8/3) N^3 + 8 N^2 real flops: complex LU plus two triangular solves."""
    return 8.0 / 3.0 * N ** 3 + 8.0 * N ** 2


TUNING_TABLE = os.path.join(ROOT, "transient-analysis-grounding-systems_amd", "tuning_gfx950.txt")


def apply_tuning(T, sw, prec, s, kap, inj, mode, cache):
    """Installs the sweep's tile shapes and schedule for `prec`: from the
    tuning table (tags_sweep_autotune with the table as cache: an entry for
    this device and system is read, a missing one is measured on the probe
    samples s, kap), measured afresh (mode "measure", written to `cache`), or
    none (mode "off", the library defaults). Returns what was installed."""
    if mode == "off":
        return "defaults"
    t = sw.autotune(prec, s, kap, inj, cache=cache, force=(mode == "measure"))
    t["source"] = "table" if t["from_cache"] else "measured"
    return t


def note(msg):
    """A progress line on stderr (the JSON line alone goes to stdout)."""
    print("bench: " + msg, file=sys.stderr, flush=True)


def load_pkg():
    import tags_pkg
    return tags_pkg.load()


def cpu_baseline(T, e, nodes, N, threads, lu_n=4000, seed=7, freq=1e5):
    """The CPU restatement of the reference (oracle) on a bounded sample of one
    frequency sample: the full pair fill (calculate_impedances then
    impedances_images, adaptive Genz-Malik cubature at rel tol 1e-6, OpenMP
    over `threads`) timed whole, and LAPACK zsysv (scipy's OpenBLAS) of order
    `lu_n` scaled by (N / lu_n)^3 to the immittance order N."""
    import tags_pkg
    O = tags_pkg.oracle()
    from scipy.linalg import lapack
    m = len(e)
    sc = O.soil_scalars(freq, SIGMA, EPSR)
    img = O.images_of(e)
    t0 = time.perf_counter()
    rc, zl, zt = O.calculate_impedances(e, sc["gamma"], sc["s"], 1.0, sc["kappa"],
                                        threads=threads)
    O.impedances_images(zl, zt, e, img, sc["gamma"], sc["s"], 1.0, sc["kappa"],
                        sc["ref_l"], sc["ref_t"], threads=threads)
    t_fill = time.perf_counter() - t0
    del zl, zt
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((lu_n, lu_n)) + 1j * rng.standard_normal((lu_n, lu_n))
    a = a + a.T
    b = np.ones((lu_n, 1), complex)
    lapack.zsysv(a[:500, :500], b[:500], lower=1)  # warm the library
    t0 = time.perf_counter()
    lapack.zsysv(a, b, lower=1)
    t_lu_n = time.perf_counter() - t0
    t_lu = t_lu_n * (N / lu_n) ** 3
    per_sample = t_fill + t_lu
    return {"value": 1.0 / per_sample, "unit": "freq samples/s", "cores": threads,
            "kind": "port",
            "sample": ("oracle on the host, %d threads, f = %g Hz: full pair fill of m = %d "
                       "segments (%d direct + image pairs, cubature rel tol 1e-6) in %.1f s; "
                       "LAPACK zsysv of order %d in %.2f s scaled by (%d/%d)^3 to %.0f s; "
                       "%.0f s per sample"
                       % (threads, freq, m, m * (m + 1) // 2, t_fill, lu_n, t_lu_n, N, lu_n,
                          t_lu, per_sample))}


def copy_calibration(mib):
    """Device-to-device copy of known size: reads and writes mib MiB per launch."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_per_launch": n * 4, "launches": 5}))


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "trailing_update_f32_grid60_traffic.txt")
TRAFFIC_FILE_X6 = os.path.join(ROOT, "profiles", "trailing_update_x6_grid60_traffic.txt")


def measured_traffic(x6=False):
    """HBM bytes per launch of the trailing complex64 GEMM (x6: the split-bf16
    product, its two operand packs included), from the committed rocprofv3
    FETCH_SIZE + WRITE_SIZE summary (scaled by a copy kernel), or None when
    that summary is absent."""
    try:
        with open(TRAFFIC_FILE_X6 if x6 else TRAFFIC_FILE) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch"):
                    return float(line.split("=")[1])
    except OSError:
        pass
    return None


def rows_bench(T):
    """Timings of the rows next to the path, one JSON line: every figure is a
    host-pointer call end to end (PCIe copies included) unless named device."""
    out = {"mode": "rows", "n_gpus": 1}
    L = T.lib()
    rng = np.random.default_rng(1)

    def timed(fn, reps=3):
        fn()
        L.tags_device_synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        L.tags_device_synchronize()
        return (time.perf_counter() - t0) / reps

    # nodal admittance at the 30 x 30 m grid, 5 m mesh (m = 1320, n = 1221),
    # synthetic complex symmetric Z (the LU's work does not depend on values)
    _, e, nodes = T.electrode_grid(11, 11, 30.0, 30.0, 6, 6, 5e-3, -0.5)
    m, n = len(e), len(nodes)
    a = np.zeros(m * n, complex)
    b = np.zeros(m * n, complex)
    T.fill_incidence_adm(a, b, e, nodes)
    zs = []
    for _ in range(2):
        r = rng.standard_normal((m, m)) + 1j * rng.standard_normal((m, m))
        zs.append(((r + r.T) / 2 + 4 * np.sqrt(m) * np.eye(m)).T.ravel().copy())
    yn = np.zeros(n * n, complex)
    dt = timed(lambda: T.fill_impedance_adm(yn, zs[0].copy(), zs[1].copy(), a, b, m, n))
    fl = 2 * (8 / 3 * m ** 3 + 12.0 * m * m * (n + m)) + 2 * 8.0 * m * n * n
    out["fill_impedance_adm"] = {"m": m, "n": n, "ms": 1e3 * dt,
                                 "tflops_incl_pcie": fl / dt / 1e12,
                                 "flops": fl}
    ie = np.zeros(n, complex)
    ie[0] = 1
    dt = timed(lambda: T.solve_admittance(yn, ie.copy(), n))
    out["solve_admittance"] = {"n": n, "ms": 1e3 * dt}
    # numerical Laplace transform pair: nt = 2^14, 512 signals
    nt, B, tmax = 2 ** 14, 512, 1e-4
    f = rng.standard_normal((B, nt))
    g, sgrid = T.laplace_trans(f, tmax)
    dt_f = timed(lambda: T.laplace_trans(f, tmax))
    dt_i = timed(lambda: T.inv_laplace_trans(g, sgrid, tmax, nt))
    passes = int(np.log2(nt))
    out["laplace_batch"] = {"nt": nt, "signals": B, "ms_forward": 1e3 * dt_f,
                            "ms_inverse": 1e3 * dt_i,
                            "device_pass_bytes": passes * 2 * 16 * nt * B}
    # fields: 256 ground points around the 60 x 60 m grid (m = 7440), 1 rhs
    _, e3, _ = T.electrode_grid(*GRID)
    m3 = len(e3)
    pts = np.column_stack([rng.uniform(-5, 65, 256), rng.uniform(-5, 65, 256), np.zeros(256)])
    it = rng.standard_normal((1, m3)) + 1j * rng.standard_normal((1, m3))
    il = rng.standard_normal((1, m3)) + 1j * rng.standard_normal((1, m3))
    sc = T.soil_scalars(1e5, SIGMA, EPSR)
    dt = timed(lambda: T.fields(pts, e3, il, it, sc["gamma"], sc["s"], sc["kappa"]))
    out["fields_batch"] = {"points": 256, "segments": m3, "ms": 1e3 * dt,
                           "pairs_per_s": 256 * m3 / dt}
    dt = timed(lambda: T.mhem_gpd(pts, e3, it, sc["gamma"], sc["kappa"]))
    out["mhem_gpd"] = {"points": 256, "segments": m3, "ms": 1e3 * dt,
                       "pairs_per_s": 256 * m3 / dt}
    print(json.dumps(out))


def transient_bench(T, nt=2 ** 14):
    """The lightning-transient case end to end, one JSON line: the 30 x 30 m
    grid with 5 m mesh (11 + 11 conductors, 1 m segments: m = 1320, n = 1221,
    N = 3861) in an Alipio-Visacro "mean" soil (sigma0 = 5e-4 S/m,
    h = 1.26 (1000 sigma0)^-0.73, g = 0.54, eps_inf'/eps0 = 12), a 1 kA
    double-exponential current (alpha = 1.4659e4 /s, beta = 2.4689e6 /s,
    k = 1.0373) injected at the centre node, tmax = 100 us, nt samples: NLT of
    the current, the sweep over the nt/2 + 1 complex frequencies in both
    precisions, INLT (Blackman window) of Z_h I. Reports the sweeps' rates and
    the GPR waveform difference between them before 0.9 tmax."""
    _, e, nodes = T.electrode_grid(11, 11, 30.0, 30.0, 6, 6, 5e-3, -0.5)
    inj = int(np.argmin(np.sum((nodes - [15.0, 15.0, -0.5]) ** 2, axis=1)))
    tmax = 1e-4
    t = np.linspace(0.0, tmax, nt)
    cur = 1e3 * 1.0373 * (np.exp(-1.4659e4 * t) - np.exp(-2.4689e6 * t))
    ig, s = T.laplace_trans(cur, tmax)
    sigma0 = 5e-4
    h = 1.26 * (1000 * sigma0) ** -0.73
    kap = np.array([sg + sk * er * T.EPS0 for sk in s
                    for sg, er in [T.alipio_soil(sigma0, sk, h, 0.54, 12.0)]])
    sw = T.Sweep(e, nodes)
    sw.set_plan(T.PLAN_HYBRID)
    L = T.lib()
    out = {"mode": "transient", "n_gpus": 1, "m": len(e), "n": len(nodes),
           "N": len(nodes) + 2 * len(e), "inj_node": inj, "nt": nt, "samples": len(s)}
    res = {}
    for name, prec in (("fp64", T.PREC_FP64), ("mixed", T.PREC_MIXED)):
        sw.set_plan(T.PLAN_HYBRID if prec == T.PREC_MIXED else 2)
        sw.solve(s[:4], kap[:4], inj, prec)
        L.tags_device_synchronize()
        t0 = time.perf_counter()
        zh, st = sw.solve(s, kap, inj, prec)
        L.tags_device_synchronize()
        dt = time.perf_counter() - t0
        v = T.inv_laplace_trans(zh * ig, s, tmax, nt, T.FILTER_BLACKMAN)
        res[name] = (zh, v)
        out[name] = {"samples_per_s": len(s) / dt, "s": dt}
        if prec == T.PREC_MIXED:
            out[name]["fp64_fallbacks"] = int(sum(q["precision_used"] == 0 for q in st))
            out[name]["refine_iters_mean"] = float(np.mean([q["refine_iters"] for q in st]))
    sw.close()
    # the INLT multiplies by e^{c t}, which reaches nt^2 at tmax: a relative
    # difference d of the spectra shows as up to d nt^(2 t / tmax) in the
    # waveform, so the waveforms are compared over the first half of the
    # window and, over all of it, with that factor taken out
    v64, vmx = res["fp64"][1], res["mixed"][1]
    half = slice(0, nt // 2)
    damp = np.exp(-s[0].real * t)
    out["gpr_peak_v"] = float(np.max(np.abs(v64[half])))
    out["gpr_max_rel_diff_first_half"] = float(np.max(np.abs(vmx - v64)[half]) /
                                               np.max(np.abs(v64[half])))
    out["gpr_max_rel_diff_damped"] = float(np.max(np.abs((vmx - v64) * damp)) /
                                           np.max(np.abs(v64 * damp)))
    dz = np.abs(res["mixed"][0] - res["fp64"][0]) / np.abs(res["fp64"][0])
    out["zh_rel_diff_mixed_vs_fp64"] = {"max": float(dz.max()), "median": float(np.median(dz)),
                                        "worst_sample_hz": float(abs(s[int(np.argmax(dz))]) /
                                                                 (2 * np.pi))}
    print(json.dumps(out))


def _sync(T):
    T.lib().tags_device_synchronize()


def _timed_solve(T, obj, prec, *args):
    _sync(T)
    t0 = time.perf_counter()
    r = obj.solve(*args, precision=prec) if args else obj.solve(prec)
    _sync(T)
    return r, time.perf_counter() - t0


def gemm_basis(mixed, tuning):
    """(kernel label, peak TFLOP/s, what the peak is) of the trailing update's
    large products under a tuning (a dict with gemm_big, or the defaults)."""
    big = tuning.get("gemm_big", -1) if isinstance(tuning, dict) else -1
    if big < 0:
        big = DEFAULT_BIG["mixed" if mixed else "fp64"]
    if mixed and big in X6_VARIANTS:
        return ("k_trailing_update_x6 (complex64 as bf16 parts, 24 bf16 MFMAs per complex "
                "product), variant %d" % big, PEAK["bf16"] / X6_PARTS_PRODUCTS,
                "bf16 dense MFMA peak / 6 (bf16 part products per fp32-accurate real product)")
    if mixed:
        return ("k_trailing_update<float, 3-multiply>, variant %d" % big, PEAK["fp32"],
                "fp32 dense MFMA peak")
    return "k_trailing_update<double>, variant %d" % big, PEAK["fp64"], "fp64 dense MFMA peak"


def _roofline(tm, basis):
    """Trailing-update GEMM of a profiled pass: algorithmic 8MNK flops over the
    HIP-event time of its launches, against the peak of the arithmetic the
    large products run on (gemm_basis)."""
    if not tm or tm["trailing_launches"] <= 0 or tm["trailing_ms"] <= 0:
        return None
    kernel, peak, what = basis
    ach = tm["trailing_flops"] / tm["trailing_ms"] / 1e9
    r = {"bound": "mfma", "achieved": ach, "peak": peak, "unit": "TFLOP/s",
         "frac": ach / peak, "traffic": None, "kernel": kernel, "peak_basis": what,
         "launches": tm["trailing_launches"],
         "avg_launch_ms": tm["trailing_ms"] / tm["trailing_launches"]}
    if peak != PEAK["fp32"] and "bf16" in what:
        # the same rate against the fp32 MFMA peak the complex64 product had
        r["frac_of_fp32_mfma_peak"] = ach / PEAK["fp32"]
    return r


def probe_step(N):
    """Every how many samples of a sweep the autotuner's probe takes: all of a
    small system's (its groups of samples and samples in flight only show on
    the whole sweep), fewer as the order grows."""
    return 1 if N < 2000 else (4 if N < 8000 else 16)


def sweep_config(T, e, nodes, inj, s, kap, plan=7, tuning=("table", TUNING_TABLE)):
    """Both precisions of one sweep over all its samples (groups of samples per
    launch as the library chooses), one unprofiled timed pass each, then a
    profiled pass for the trailing-update figures. The tile shapes and
    schedule come from the tuning table (probe: probe_step) unless `tuning`
    says otherwise."""
    sw = T.Sweep(e, nodes)
    sw.set_plan(plan)
    out = {}
    us = {}
    ps = probe_step(sw.N)
    for name, prec in (("mixed", T.PREC_MIXED), ("fp64", T.PREC_FP64)):
        tu = apply_tuning(T, sw, prec, s[::ps], kap[::ps], inj, tuning[0], tuning[1])
        sw.solve(s[:2], kap[:2], inj, prec)  # allocations
        (u, st), dt = _timed_solve(T, sw, prec, s, kap, inj)
        sw.set_profiling(True)
        sw.solve(s, kap, inj, prec)
        tm = sw.timers()
        sw.set_profiling(False)
        us[name] = u
        r = {"samples_per_s": len(s) / dt, "s": dt,
             "roofline": _roofline(tm, gemm_basis(prec == T.PREC_MIXED, tu)),
             "executed_gemm_flops_per_sample": tm["gemm_flops"] / len(s),
             "tuning": tu}
        if prec == T.PREC_MIXED:
            r["fp64_fallbacks"] = int(sum(q["precision_used"] == 0 for q in st))
            r["refine_iters_mean"] = float(np.mean([q["refine_iters"] for q in st]))
            r["forward_est_max"] = float(max(q["forward_est"] for q in st))
        out[name] = r
    sw.close()
    d = np.abs(us["mixed"] - us["fp64"]) / np.abs(us["fp64"])
    out["mixed_over_fp64"] = out["mixed"]["samples_per_s"] / out["fp64"]["samples_per_s"]
    out["zh_rel_diff_mixed_vs_fp64"] = {"max": float(d.max()), "median": float(np.median(d))}
    return out, us


def cpu_oracle_samples(e, nodes, freqs, sigma, epsr, threads, inj=0, soil=None):
    """The oracle's pipeline (Genz-Malik cubature at the reference's rel tol 1e-6,
    LAPACK zsysv) over the given samples, timed end to end; soil(s) -> kappa
    overrides the constant soil."""
    import tags_pkg
    O = tags_pkg.oracle()
    images = O.images_of(e)
    m, n = len(e), len(nodes)
    t0 = time.perf_counter()
    for f in freqs:
        if soil is None:
            sc = O.soil_scalars(f, sigma, epsr)
        else:
            sk = complex(f)
            kk = soil(sk)
            sc = dict(s=sk, kappa=kk, gamma=complex(np.sqrt(sk * O.MU0 * kk + 0j)),
                      ref_l=1.0 + 0j, ref_t=(kk - sk * O.EPS0) / (kk + sk * O.EPS0), mur=1.0)
        wg, _, _ = O.assemble(e, nodes, images, sc, rel_tol=1e-6, threads=threads)
        ie = np.zeros(n + 2 * m, complex)
        ie[inj] = 1.0
        O.solve_immittance(wg, ie, m, n)
    return time.perf_counter() - t0


def more_configs(T, threads, tuning):
    """The other configurations of the workload family, each one key."""
    res = {}
    note("c1 (15 m conductor)")
    # c1: one 15 m conductor, 30 segments, 200 frequencies 100 Hz - 10 MHz
    rc, e, nodes = T.segment_electrode([0, 0, -0.5], [15, 0, -0.5], 30, 7e-3)
    f = np.logspace(2, 7, 200)
    s = 2j * np.pi * f
    kap = 0.01 + s * 10.0 * T.EPS0
    r, _ = sweep_config(T, e, nodes, 0, s, kap, tuning=tuning)
    t = cpu_oracle_samples(e, nodes, f, 0.01, 10.0, threads)
    r["cpu_baseline"] = {"value": len(f) / t, "unit": "freq samples/s", "cores": threads,
                         "kind": "port",
                         "sample": "oracle end to end over all 200 frequencies (cubature rel tol "
                                   "1e-6, LAPACK zsysv), %d threads: %.2f s" % (threads, t)}
    r["workload"] = "15 m conductor, m=30, N=91, rho=100, eps_r=10, 200 freqs 100 Hz-10 MHz"
    res["c1"] = r
    note("c2 (20 x 20 m grid)")
    # c2: 20 x 20 m grid, 5 m mesh, N = 1185, 400 frequencies
    rc, e, nodes = T.electrode_grid(5, 5, 20.0, 20.0, 10, 10, 6e-3, -0.5)
    f = np.logspace(2, 7, 400)
    s = 2j * np.pi * f
    kap = 1e-3 + s * 15.0 * T.EPS0
    r, _ = sweep_config(T, e, nodes, 0, s, kap, tuning=tuning)
    fs = f[[0, 133, 266, 399]]
    t = cpu_oracle_samples(e, nodes, fs, 1e-3, 15.0, threads)
    r["cpu_baseline"] = {"value": len(fs) / t, "unit": "freq samples/s", "cores": threads,
                         "kind": "port",
                         "sample": "oracle end to end at 4 of the 400 frequencies, %d threads: "
                                   "%.2f s" % (threads, t)}
    r["workload"] = ("20x20 m grid, 5 m mesh (m=400, n=385, N=1185), rho=1000, eps_r=15, "
                     "400 freqs 100 Hz-10 MHz, corner injection")
    res["c2"] = r
    note("c5 (30 x 30 m grid, Alipio-Visacro soil)")
    # c5: 30 x 30 m grid in Alipio-Visacro soil, every 8th of the 8193 s-samples
    _, e, nodes = T.electrode_grid(11, 11, 30.0, 30.0, 6, 6, 5e-3, -0.5)
    inj = int(np.argmin(np.sum((nodes - [15.0, 15.0, -0.5]) ** 2, axis=1)))
    nt, tmax = 2 ** 14, 1e-4
    k = np.arange(0, nt // 2 + 1, 8)
    s = np.log(nt ** 2) / tmax + 2j * np.pi * k / tmax
    sigma0 = 5e-4
    h = 1.26 * (1000 * sigma0) ** -0.73

    def soil(sk):
        sg, er = T.alipio_soil(sigma0, sk, h, 0.54, 12.0)
        return sg + sk * er * T.EPS0
    kap = np.array([soil(x) for x in s])
    r, _ = sweep_config(T, e, nodes, inj, s, kap, tuning=tuning)
    t = cpu_oracle_samples(e, nodes, s[:1], None, None, threads, inj=inj, soil=soil)
    r["cpu_baseline"] = {"value": 1.0 / t, "unit": "freq samples/s", "cores": threads,
                         "kind": "port",
                         "sample": "oracle end to end at the lowest |s| sample, %d threads: "
                                   "%.2f s" % (threads, t)}
    r["workload"] = ("30x30 m grid, 5 m mesh (m=1320, n=1221, N=3861), Alipio-Visacro mean soil, "
                     "every 8th of the 8193 Laplace samples (nt=2^14, tmax=100 us), centre "
                     "injection")
    res["c5"] = r
    note("c4 (512 seeded cases)")
    res["c4"] = cases_config(T, threads, tuning=tuning)
    return res


def cases_config(T, threads, step=8, tuning=("table", TUNING_TABLE)):
    """512 seeded cases (rods, counterpoises, grids; N = 61 .. 4246), every
    step-th frequency of each, one multi-case solve per precision, with the
    tile shapes and schedule of each size class from the tuning table (one
    case per class tuned, tags_cases_autotune) unless `tuning` says
    otherwise."""
    import tags_pkg
    C = tags_pkg.cases()
    fam = C.case_family(T, freq_step=step)
    n = sum(len(c["s"]) for c in fam)
    cs = T.Cases(fam)
    out = {"workload": "512 seeded cases (seed 2024), every %dth of their 64-512 frequencies: "
                       "%d samples" % (step, n)}
    us = {}
    for name, prec in (("mixed", T.PREC_MIXED), ("fp64", T.PREC_FP64)):
        tu = "defaults"
        if tuning[0] != "off":
            tu = cs.autotune(prec, cache=tuning[1], force=(tuning[0] == "measure"))
        cs.solve(prec)  # allocations
        (u, st), dt = _timed_solve(T, cs, prec)
        cs.set_profiling(True)
        cs.solve(prec)
        tm = cs.timers()
        cs.set_profiling(False)
        us[name] = u
        r = {"samples_per_s": n / dt, "s": dt,
             "roofline": _roofline(tm, gemm_basis(prec == T.PREC_MIXED,
                                                  tu[-1] if isinstance(tu, list) and tu
                                                  else tu)),
             "dense_equivalent_tflops": sum(len(c["s"]) * lu_flops(c["N"]) for c in fam) / dt
             / 1e12,
             "tuning_classes": tu}
        if prec == T.PREC_MIXED:
            r["fp64_fallbacks"] = int(sum(q["precision_used"] == 0 for s_ in st for q in s_))
        out[name] = r
    cs.close()
    out["mixed_over_fp64"] = out["mixed"]["samples_per_s"] / out["fp64"]["samples_per_s"]
    out["zh_rel_diff_mixed_vs_fp64_max"] = max(float(np.max(np.abs(a - b) / np.abs(b)))
                                               for a, b in zip(us["mixed"], us["fp64"]))
    # CPU: one sample of each of 6 cases spread over the size range, timed; the
    # whole workload's CPU time modelled per case as t(N) interpolated in
    # log-log between those measurements
    Ns = np.array([c["N"] for c in fam])
    pick = [int(np.argmin(np.abs(Ns - q))) for q in (61, 181, 401, 801, 1501, 2201)]
    tn = []
    for i in pick:
        c = fam[i]
        t = cpu_oracle_samples(c["electrodes"], c["nodes"], c["f"][len(c["f"]) // 2:][:1],
                               c["sigma"], c["epsr"], threads)
        tn.append((c["N"], t))
    tn.sort()
    lx, ly = np.log([a for a, _ in tn]), np.log([b for _, b in tn])
    slope = (ly[-1] - ly[-2]) / (lx[-1] - lx[-2])

    def tmodel(N):
        x = np.log(N)
        return float(np.exp(np.interp(x, lx, ly) if x <= lx[-1] else ly[-1] + slope * (x - lx[-1])))
    total = sum(len(c["s"]) * tmodel(c["N"]) for c in fam)
    out["cpu_baseline"] = {"value": n / total, "unit": "freq samples/s", "cores": threads,
                           "kind": "port",
                           "sample": "oracle end to end, one sample each of 6 cases (N = %s: "
                                     "%s s), %d threads; the workload's CPU time modelled as "
                                     "t(N) interpolated in log-log (extrapolated as the last "
                                     "slope above N = %d): %.0f s"
                                     % ([a for a, _ in tn], ["%.3g" % b for _, b in tn],
                                        threads, tn[-1][0], total)}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=24)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-kernel-figures", action="store_true",
                    help="skip the profiled one-at-a-time pass (roofline and timers are "
                         "then missing; for quick parameter sweeps)")
    ap.add_argument("--headline-only", action="store_true",
                    help="only the 60 x 60 m grid line (no c1, c2, c5, c4 keys)")
    ap.add_argument("--plan-mixed", type=int, default=7,
                    help="LU plan of the mixed form: 0 dense, 1 structured, 2 Z_L first, "
                         "3 Z_L-block pivoting, 4 grounded without row exchanges, 5 "
                         "grounded with pivoting, 6 nodal, 7 hybrid (default: nodal "
                         "complex64 factor from --nodal-above, Z_L first below)")
    ap.add_argument("--nodal-above", type=float, default=0.0,
                    help="switch frequency of the hybrid plan (Hz, default 0: the nodal "
                         "complex64 factor at every frequency)")
    ap.add_argument("--gmres-below", type=float, default=2e4,
                    help="nodal factors below this frequency refine by FGMRES (Hz, "
                         "default 2e4)")
    ap.add_argument("--gmres-steps", type=int, default=6,
                    help="Arnoldi steps per FGMRES refinement step (default 6)")
    ap.add_argument("--plan-fp64", type=int, default=2,
                    help="LU plan of the FP64 form (default 2, Z_L first)")
    ap.add_argument("--update-chunks", type=int, default=None,
                    help="launches per step for the trailing update beside the panel")
    ap.add_argument("--step-width-mixed", type=int, default=256, choices=(128, 256),
                    help="columns per factorisation step of the mixed form (the trailing "
                         "update's K; default 256)")
    ap.add_argument("--step-width-fp64", type=int, default=128, choices=(128, 256),
                    help="columns per factorisation step of the FP64 form (default 128)")
    ap.add_argument("--solve-rows", type=int, default=2, choices=(0, 1, 2),
                    help="triangular solves: one launch per range and direction with work-"
                         "groups of 64 (1) or 128 (2, default) rows, or the per-step "
                         "launches (0)")
    ap.add_argument("--inflight", type=int, default=4,
                    help="frequency samples in flight in the timed sweep (default 4)")
    ap.add_argument("--formulation", choices=("hem", "mhem"), default="hem",
                    help="hem: pair integrals of e^{-gamma r}/r every sample (the headline); "
                         "mhem: the examples' modified HEM, gamma-free potentials once per "
                         "geometry (in the warm-up call) and exp(-gamma rbar) scaling per sample")
    ap.add_argument("--transient", action="store_true",
                    help="the lightning transient on the 30 x 30 m grid end to end (NLT, "
                         "both sweeps, INLT); prints its own JSON line")
    ap.add_argument("--rows", action="store_true",
                    help="time the rows next to the path (admittance, Laplace transform "
                         "pair, fields) instead of the sweep; prints its own JSON line")
    ap.add_argument("--no-lookahead", action="store_true",
                    help="factor the steps strictly in order on one stream")
    ap.add_argument("--gemm-mixed", type=int, default=None, metavar="VARIANT",
                    help="complex64 GEMM variant for large products in place of the tuned "
                         "one (22-24: split-bf16 MFMA)")
    ap.add_argument("--tuning", choices=("table", "measure", "off"), default="table",
                    help="tile shapes and schedule per sweep and precision: from the "
                         "tuning table shipped with the package (default; a missing entry "
                         "is measured), measured afresh by the autotuner (written to "
                         "--tuning-cache), or the library defaults")
    ap.add_argument("--tuning-cache", default=None,
                    help="cache file of --tuning measure (default: a file in the temporary "
                         "directory)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB five times (the "
                         "reference kernel for scaling FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: measures one GPU only; frequency samples are independent, so "
              "more GPUs would each take a share of the samples", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    T = load_pkg()
    if args.rows:
        rows_bench(T)
        return
    if args.transient:
        transient_bench(T)
        return
    rc, e, nodes = T.electrode_grid(*GRID)
    m, n = len(e), len(nodes)
    N = n + 2 * m
    freqs = np.logspace(2, np.log10(5e6), NFREQ)
    K, W = args.steps, args.warmup
    idx = np.unique(np.linspace(0, NFREQ - 1, K).round().astype(int))
    K = len(idx)
    s = 2j * np.pi * freqs[idx]
    kap = SIGMA + s * EPSR * T.EPS0
    sw = T.Sweep(e, nodes)
    if args.formulation == "mhem":
        sw.set_formulation(T.FORM_MHEM)
    L = T.lib()
    if args.no_lookahead:
        T.set_lookahead(False)
    if args.update_chunks is not None:
        T.set_update_chunks(args.update_chunks)
    T.set_solve_rows(args.solve_rows)
    if args.tuning == "measure":
        import tempfile
        cache = args.tuning_cache or os.path.join(tempfile.gettempdir(), "tags_tuning.txt")
    else:
        cache = TUNING_TABLE
    tuning = (args.tuning, cache)
    # probe samples of the autotuner: a spread of the sweep (16 samples, on
    # which it also compares groups of up to four samples, four in flight)
    probe = np.arange(8, NFREQ, 16)
    s_probe = 2j * np.pi * freqs[probe]
    kap_probe = SIGMA + s_probe * EPSR * T.EPS0
    tuned = {}

    def run(prec, one_at_a_time, profile):
        """W warm-up samples, then the K timed samples, as many in flight as
        the tuning (or --inflight) sets, or one at a time; with `profile`
        every kernel launch is bracketed by HIP events on the stream it runs
        on. Returns the results, the wall time, the timers and the number of
        samples in flight."""
        mixed = prec == T.PREC_MIXED
        sw.set_plan(args.plan_mixed if mixed else args.plan_fp64)
        sw.set_nodal_above(args.nodal_above)
        sw.set_gmres(args.gmres_below, args.gmres_steps)
        T.set_step_width(args.step_width_mixed if mixed else args.step_width_fp64)
        sw.set_inflight(args.inflight)
        if prec not in tuned:
            tuned[prec] = apply_tuning(T, sw, prec, s_probe, kap_probe, 0, args.tuning, cache)
        if mixed and args.gemm_mixed is not None:
            base = tuned[prec] if isinstance(tuned[prec], dict) else sw.tuning(prec)
            if base is None:
                raise SystemExit("--gemm-mixed needs a tuning (table or measure)")
            tuned[prec] = dict(base, gemm_big=args.gemm_mixed)
        tu = tuned[prec]
        inflight = args.inflight
        group = 1
        if isinstance(tu, dict):
            inflight = tu["inflight"] if tu["inflight"] > 0 else args.inflight
            group = max(1, tu["batch"])
            if one_at_a_time:
                sw.set_tuning(prec, dict(tu, inflight=1))
            else:
                sw.set_tuning(prec, tu)
        if one_at_a_time:
            inflight = 1
            sw.set_inflight(1)
        if W:
            # the warm-up sets up every slot: `inflight` groups of `group`
            # samples
            nw = max(W, inflight * group)
            sw.solve(s[:nw], kap[:nw], 0, prec)
        sw.set_profiling(profile)
        L.tags_device_synchronize()
        t0 = time.perf_counter()
        u, st = sw.solve(s, kap, 0, prec)
        L.tags_device_synchronize()
        dt = time.perf_counter() - t0
        tm = sw.timers()
        sw.set_profiling(False)
        return u, st, dt, tm, inflight

    def effective(key, default, lowest):
        """The mixed form's value of a setting: the tuned one where the tuning
        sets it (>= lowest), else the command line's."""
        tu = tuned.get(T.PREC_MIXED)
        if isinstance(tu, dict) and tu.get(key, -1) >= lowest:
            return tu[key]
        return default

    def width_of(prec):
        tu = tuned.get(prec)
        if isinstance(tu, dict) and tu["step_width"] > 0:
            return tu["step_width"]
        return args.step_width_mixed if prec == T.PREC_MIXED else args.step_width_fp64

    # headline: the sweep as it runs (the tuned number of samples in flight);
    # kernel figures: the same samples one at a time, so that each launch's
    # HIP-event duration is its own (with --inflight 1 and no tuning the
    # headline pass is that pass)
    one = args.inflight == 1 and args.tuning == "off"
    note("60 x 60 m grid, mixed")
    u_mx, st_mx, dt_mx, tm_mx, infl_mx = run(T.PREC_MIXED, one, one)
    u_m1, dt_m1 = u_mx, dt_mx
    if not one and not args.no_kernel_figures:
        u_m1, _, dt_m1, tm_mx, _ = run(T.PREC_MIXED, True, True)
    out = {
        "metric": METRIC,
        "value": K / dt_mx,
        "unit": "freq samples/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": 1e3 * dt_mx / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 LU (MFMA) + f64 refinement",
        "data": "synthetic",
        "config": {"workload": "60x60 m grid, 2 m mesh, 0.5 m segments (m=7440, n=6541, "
                               "N=21421), rho=500, eps_r=20, 256 freqs 100 Hz-5 MHz, unit "
                               "current at corner node",
                   "samples_timed": [float(f) for f in freqs[idx]],
                   "formulation": args.formulation,
                   "plan": PLAN_NAMES[args.plan_mixed],
                   "step_width": width_of(T.PREC_MIXED),
                   "nodal_above_hz": effective("nodal_above", args.nodal_above, 0.0),
                   "gmres_below_hz": args.gmres_below,
                   "gmres_steps": effective("gmres_steps", args.gmres_steps, 1),
                   "solve_rows": effective("solve_rows", args.solve_rows, 0),
                   "tuning": tuned[T.PREC_MIXED]},
        "mixed": {"refine_iters": [q["refine_iters"] for q in st_mx],
                  "precision_used": [q["precision_used"] for q in st_mx],
                  "backward_err_max": max(q["backward_err"] for q in st_mx),
                  "inflight": infl_mx,
                  "ms_per_step_one_in_flight": 1e3 * dt_m1 / K,
                  "max_rel_diff_one_in_flight": float(np.max(np.abs(u_m1 - u_mx) /
                                                             np.abs(u_mx))),
                  "timers_ms_one_in_flight": {k: v for k, v in tm_mx.items()
                                              if k.startswith("ms_")}},
    }
    if not args.no_fp64:
        note("60 x 60 m grid, FP64")
        u64, st64, dt64, tm64, infl64 = run(T.PREC_FP64, one, one)
        dt641 = dt64
        if not one and not args.no_kernel_figures:
            _, _, dt641, tm64, _ = run(T.PREC_FP64, True, True)
        rel = float(np.max(np.abs(u_mx - u64) / np.abs(u64)))
        out["fp64"] = {"value": K / dt64, "ms_per_step": 1e3 * dt64 / K,
                       "executed_gemm_flops_per_sample": tm64["gemm_flops"] / K,
                       "plan": PLAN_NAMES[args.plan_fp64],
                       "step_width": width_of(T.PREC_FP64),
                       "tuning": tuned[T.PREC_FP64],
                       "inflight": infl64,
                       "ms_per_step_one_in_flight": 1e3 * dt641 / K,
                       "timers_ms_one_in_flight": {k: v for k, v in tm64.items()
                                                   if k.startswith("ms_")},
                       "trailing_gemm_tflops": (tm64["trailing_flops"] / tm64["trailing_ms"] / 1e9
                                                if tm64["trailing_ms"] > 0 else None),
                       "dense_equivalent_tflops": lu_flops(N) * K / dt64 / 1e12}
        out["mixed_over_fp64"] = (K / dt_mx) / (K / dt64)
        out["max_rel_err_zh_vs_fp64"] = rel
        # the error of every timed sample, not only the largest
        out["rel_err_zh_vs_fp64_per_sample"] = [float(v) for v in
                                                np.abs(u_mx - u64) / np.abs(u64)]
    ach = tm_mx["trailing_flops"] / tm_mx["trailing_ms"] / 1e9 if tm_mx["trailing_ms"] > 0 else 0.0
    kernel, peak, what = gemm_basis(True, tuned[T.PREC_MIXED])
    x6 = "bf16" in what
    out["roofline"] = {"bound": "mfma", "achieved": ach, "peak": peak,
                       "unit": "TFLOP/s", "frac": ach / peak, "peak_basis": what,
                       "traffic": measured_traffic(x6),
                       "kernel": "%s (A22 -= L21 U12, K=%d); products under 512 tiles: the "
                                 "small variant" % (kernel, width_of(T.PREC_MIXED)),
                       "launches": tm_mx["trailing_launches"],
                       "avg_launch_ms": tm_mx["trailing_ms"] / max(1, tm_mx["trailing_launches"]),
                       "measured_in": ("the timed samples" if one else
                                       "the same %d samples timed one at a time" % K)}
    if x6:
        out["roofline"]["frac_of_fp32_mfma_peak"] = ach / PEAK["fp32"]
    # F(N) = (8/3) N^3 + 8 N^2 per sample counts the dense LU, including the
    # exactly-zero blocks the plans skip: a dense-equivalent rate, not a
    # hardware rate; the executed one counts the products the plans ran
    out["dense_equivalent_tflops"] = lu_flops(N) * K / dt_mx / 1e12
    out["mixed"]["executed_gemm_flops_per_sample"] = tm_mx["gemm_flops"] / K
    out["mixed"]["executed_gemm_tflops"] = tm_mx["gemm_flops"] / K * (K / dt_mx) / 1e12
    out["mixed"]["forward_est_max"] = max(q["forward_est"] for q in st_mx)
    threads = min(16, os.cpu_count() or 1)
    if not args.no_cpu_baseline:
        note("CPU baseline (oracle, %d threads)" % threads)
        out["cpu_baseline"] = cpu_baseline(T, e, nodes, N, threads)
    sw.close()
    if not args.headline_only:
        out.update(more_configs(T, threads, tuning))
    print(json.dumps(out))


if __name__ == "__main__":
    main()
