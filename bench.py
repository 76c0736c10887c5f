"""Frequency-sample throughput of the per-frequency HEM solve on one MI355X.

Workload: the 60 x 60 m grid with 2 m mesh (31 + 31 conductors, 0.5 m
segments, radius 5 mm, depth 0.6 m): m = 7,440 segments, n = 6,541 nodes,
immittance order N = 21,421; soil rho = 500 Ohm m, eps_r = 20; 256 log-spaced
frequencies 100 Hz - 5 MHz; unit current injected at corner node 0.

A step is one frequency sample end to end on the GPU: pair-integral fill and
immittance assembly in HBM, LU factorisation, triangular solves (and, for the
mixed form, FP64 iterative refinement). The timed steps walk frequencies spread
over the whole sweep. `value` is the mixed form (complex64 MFMA LU + FP64
refinement); the plain FP64 form is timed on the same samples and reported
beside it together with the largest relative difference of Z_h between the two.

Prints one JSON line. Single GPU only.
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
PEAK = {"fp64": 78.6, "fp32": 157.3}  # dense MFMA TFLOP/s, MI355X
GRID = (31, 31, 60.0, 60.0, 4, 4, 5e-3, -0.6)
SIGMA, EPSR = 1.0 / 500.0, 20.0
NFREQ = 256


def lu_flops(N):
    """(8/3) N^3 + 8 N^2 real flops: complex LU plus two triangular solves."""
    return 8.0 / 3.0 * N ** 3 + 8.0 * N ** 2


def load_pkg():
    import tags_pkg
    return tags_pkg.load()


def cpu_baseline(T, e, nodes, N, threads, pair_sample=2000000, lu_n=4000, seed=7):
    """The CPU restatement (oracle: adaptive Genz-Malik cubature at rel tol 1e-6
    plus LAPACK zsysv from scipy's OpenBLAS) on a bounded sample: `pair_sample`
    random segment pairs (direct + image integrals) scaled to all m(m+1)/2
    pairs, and zsysv of order `lu_n` scaled by (N / lu_n)^3."""
    import tags_pkg
    O = tags_pkg.oracle()
    from scipy.linalg import lapack
    m = len(e)
    rng = np.random.default_rng(seed)
    i = rng.integers(0, m, pair_sample)
    k = rng.integers(0, m, pair_sample)
    pairs = np.stack([np.minimum(i, k), np.maximum(i, k)], 1)
    sc = T.soil_scalars(1e5, SIGMA, EPSR)
    img = O.images_of(e)
    t0 = time.perf_counter()
    O.pair_sample(e, img, pairs, sc["gamma"], rel_tol=1e-6, threads=threads)
    t_pairs = time.perf_counter() - t0
    t_fill = t_pairs * (m * (m + 1) / 2) / pair_sample
    a = rng.standard_normal((lu_n, lu_n)) + 1j * rng.standard_normal((lu_n, lu_n))
    a = a + a.T
    b = np.ones((lu_n, 1), complex)
    lapack.zsysv(a, b, lower=1)  # warm
    t0 = time.perf_counter()
    lapack.zsysv(a, b, lower=1)
    t_lu = (time.perf_counter() - t0) * (N / lu_n) ** 3
    per_sample = t_fill + t_lu
    return {"value": 1.0 / per_sample, "unit": "freq samples/s", "cores": threads,
            "kind": "port",
            "sample": ("oracle on the box host: %d random segment pairs (direct + image, "
                       "cubature rel tol 1e-6) in %.1f s scaled to %d pairs, and LAPACK zsysv "
                       "of order %d in %.2f s scaled by (%d/%d)^3; extrapolated %.0f s per "
                       "sample (fill %.0f s, solve %.0f s)"
                       % (pair_sample, t_pairs, m * (m + 1) // 2, lu_n, t_lu / (N / lu_n) ** 3,
                          N, lu_n, per_sample, t_fill, t_lu))}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=6)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--plan", type=int, default=1,
                    help="0 dense, 1 structured (default), 2 structured with Z_L-block pivoting")
    ap.add_argument("--no-autotune", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU; every frequency sample is "
              "independent, see the sharding paragraph of DESIGN.md", file=sys.stderr)
        sys.exit(2)
    T = load_pkg()
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
    sw.set_plan(args.plan)
    L = T.lib()
    tuned = {}
    if not args.no_autotune:
        tuned = {"fp32": T.autotune(T.PREC_MIXED), "fp64": T.autotune(T.PREC_FP64)}

    def run(prec, profile):
        if W:
            sw.solve(s[:W], kap[:W], 0, prec)
        sw.set_profiling(profile)
        L.tags_device_synchronize()
        t0 = time.perf_counter()
        u, st = sw.solve(s, kap, 0, prec)
        L.tags_device_synchronize()
        dt = time.perf_counter() - t0
        tm = sw.timers()
        sw.set_profiling(False)
        return u, st, dt, tm

    u_mx, st_mx, dt_mx, tm_mx = run(T.PREC_MIXED, True)
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
                   "plan": ["dense", "structured", "block_pivot"][args.plan],
                   "gemm_variants": tuned},
        "mixed": {"refine_iters": [q["refine_iters"] for q in st_mx],
                  "precision_used": [q["precision_used"] for q in st_mx],
                  "backward_err_max": max(q["backward_err"] for q in st_mx),
                  "timers_ms": {k: v for k, v in tm_mx.items() if k.startswith("ms_")}},
    }
    if not args.no_fp64:
        u64, st64, dt64, tm64 = run(T.PREC_FP64, True)
        rel = float(np.max(np.abs(u_mx - u64) / np.abs(u64)))
        out["fp64"] = {"value": K / dt64, "ms_per_step": 1e3 * dt64 / K,
                       "timers_ms": {k: v for k, v in tm64.items() if k.startswith("ms_")},
                       "trailing_gemm_tflops": tm64["trailing_flops"] / tm64["trailing_ms"] / 1e9,
                       "lu_tflops_per_sample": lu_flops(N) * K / dt64 / 1e12}
        out["mixed_over_fp64"] = (K / dt_mx) / (K / dt64)
        out["max_rel_err_zh_vs_fp64"] = rel
    ach = tm_mx["trailing_flops"] / tm_mx["trailing_ms"] / 1e9
    out["roofline"] = {"bound": "mfma", "achieved": ach, "peak": PEAK["fp32"],
                       "unit": "TFLOP/s", "frac": ach / PEAK["fp32"], "traffic": None,
                       "kernel": "k_cgemm<float> trailing update (A22 -= L21 U12, K=128)",
                       "launches": tm_mx["trailing_launches"],
                       "avg_launch_ms": tm_mx["trailing_ms"] / max(1, tm_mx["trailing_launches"])}
    out["lu_algorithmic_tflops"] = lu_flops(N) * K / dt_mx / 1e12
    if not args.no_cpu_baseline:
        threads = min(16, os.cpu_count() or 1)
        out["cpu_baseline"] = cpu_baseline(T, e, nodes, N, threads)
    sw.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
