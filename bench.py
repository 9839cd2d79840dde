"""Benchmark: MRA log-likelihood evaluations per second on one MI355X.

Workload (the metric's configuration): n = 2,000,000 points on [0,1]^2, 60%
uniform + 40% in 20 Gaussian clusters (sigma 0.02); nonstationary exponential
kernel with log sigma^2 = 0.5x - 0.3y, log beta = log 0.05 + 0.8x, nugget
0.05 (planar metric); M = 10 levels, J = 4, r = 64 knots; observations drawn
from the MRA prior itself (mra_sample_prior).  One step = one full
log-likelihood evaluation (coarse prior, finest level, posterior, reduction)
with the inputs resident in HBM.

Prints one JSON line (see the keys below).  --gpus N > 1 is refused: this
build runs on one GPU.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from nonstationary_mra_amd import capi, synth, workmodel  # noqa: E402

PEAK_FP64 = 78.6e12   # MI355X dense FP64 (matrix = vector) FLOP/s
PEAK_HBM = 8.0e12     # bytes/s
SEED_LOC, SEED_FIELD = 20260101, 20260102


def make_workload(n, M, J, r, seed=SEED_LOC):
    rng = np.random.default_rng(seed)
    x, y = synth.clustered_points(n, rng)
    tree = capi.partition_program(x, y, np.zeros(n), None, None, M=M, J=J, r=r, bounds=synth.UNIT_BOUNDS)
    return tree


def group_names(M):
    names = {capi.T_FINEST_CROSS: "finest_cross", capi.T_FINEST_LOCAL: "finest_local"}
    for l in range(M - 1):
        names[capi.T_PRIOR0 + l] = f"prior_l{l}"
        names[capi.T_POST0 + l] = f"post_l{l}"
    return names


KERNEL_OF = {"prior": "k_prior_coarse", "post": "k_posterior", "finest_cross": "k_finest_cross",
             "finest_local": "k_finest_local"}


def cpu_baseline(M, J, r, n, budget_s=30.0):
    """The FP64 oracle (a restatement of the reference's CPU algorithm, OpenMP
    over regions, MKL kernels when libmkl_rt is present) on a bounded sample:
    the same generator at n/16 points and M-2 levels (same points per finest
    region), scaled to the full workload by the ratio of algorithmic flops."""
    from oracle import oracle
    mkl = oracle.lib().orc_set_blas(1)
    ns, Ms = n // 16, M - 2
    tree = make_workload(ns, Ms, J, r, seed=SEED_LOC + 1)
    cov = synth.linear_trend_cov()
    prm = dict(synth.UNIT_PARAMS)
    tree["obs_val"] = np.random.default_rng(3).standard_normal(len(tree["obs_lon"]))
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    times = []
    t_start = time.time()
    while True:
        t0 = time.perf_counter()
        oracle.evaluate(tree, cov, prm, predict=False, nthreads=threads)
        times.append(time.perf_counter() - t0)
        if len(times) >= 3 or time.time() - t_start > budget_s:
            break
    t = float(np.median(times))
    cs = np.diff(tree["obs_offsets"].astype(np.int64))
    w_sample = workmodel.total(Ms, J, r, cs)
    return t, w_sample, threads, mkl, f"n={ns} M={Ms} J={J} r={r} (same generator), {len(times)} runs, median"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=2_000_000)
    ap.add_argument("--M", type=int, default=10)
    ap.add_argument("--r", type=int, default=64)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--graph", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU only (no multi-GPU path); use --gpus 1", file=sys.stderr)
        sys.exit(2)
    J = 4
    t0 = time.time()
    tree = make_workload(args.n, args.M, J, args.r)
    counts = np.diff(tree["obs_offsets"].astype(np.int64))
    print(f"[bench] partition {time.time() - t0:.1f}s n_obs={counts.sum()} max_nF={counts.max()} "
          f"empty={(counts == 0).sum()}", file=sys.stderr, flush=True)
    h = capi.Handle(tree, with_pred=False)
    h.set_covariance(synth.linear_trend_cov())
    h.set_params(**synth.UNIT_PARAMS)
    h.sample_prior(SEED_FIELD, replace=True)
    if args.precision == "mixed":
        h.set_precision(capi.PRECISION_MIXED)
    if args.graph:
        h.use_graph(True)
    print(f"[bench] setup {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        D = h.likelihood()
    tim = np.zeros(capi.T_COUNT)
    t1 = time.perf_counter()
    for _ in range(args.steps):
        D = h.likelihood()
        tim += h.timings()
    elapsed = time.perf_counter() - t1
    tim /= args.steps
    ms_per_step = 1e3 * elapsed / args.steps
    value = args.steps / elapsed
    n = int(counts.sum())
    loglik = -0.5 * (D + n * np.log(2 * 3.14159265359))
    print(f"[bench] {ms_per_step:.2f} ms/eval D={D:.10e}", file=sys.stderr, flush=True)
    # roofline of the dominant kernel (largest average launch time)
    names = group_names(args.M)
    flops = workmodel.groups(args.M, J, args.r, counts)
    gid = max(names, key=lambda g: tim[g])
    gname = names[gid]
    achieved = flops[gname] / (tim[gid] * 1e-3)
    per_group = {names[g]: round(float(tim[g]), 3) for g in sorted(names) if tim[g] > 0.01}
    kern = KERNEL_OF[gname.split("_l")[0]] if gname.startswith(("prior_l", "post_l")) else KERNEL_OF[gname]
    roof = {"bound": "mfma", "kernel": kern, "launch": gname, "achieved": achieved / 1e12, "peak": PEAK_FP64 / 1e12,
            "unit": "TFLOP/s", "frac": achieved / PEAK_FP64, "traffic": None,
            "flops_per_launch": flops[gname], "avg_launch_ms": float(tim[gid])}
    w_total = workmodel.total(args.M, J, args.r, counts)
    out = {
        "metric": "MRA loglik evals/sec at n=2M,M=10,r=64",
        "value": value, "unit": "evals/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32+f64",
        "data": "synthetic (clustered points, field drawn from the MRA prior)",
        "config": {"workload": f"n={args.n} clustered, M={args.M}, J={J}, r={args.r}, "
                               f"nonstationary linear-trend Matern-1/2, planar", "n": n, "M": args.M, "J": J,
                   "r": args.r, "precision": args.precision, "graph": bool(args.graph),
                   "max_obs_per_finest_region": int(counts.max())},
        "loglik": loglik, "D": D,
        "whole_eval": {"flops": w_total, "tflops": w_total / (ms_per_step * 1e-3) / 1e12,
                       "frac_fp64_peak": w_total / (ms_per_step * 1e-3) / PEAK_FP64},
        "launch_ms": per_group,
        "roofline": roof,
    }
    h.close()
    if not args.no_cpu_baseline:
        t_cpu, w_sample, threads, mkl, sample = cpu_baseline(args.M, J, args.r, args.n)
        cpu_evals = (1.0 / t_cpu) * (w_sample / w_total)
        out["cpu_baseline"] = {"value": cpu_evals, "unit": "evals/s", "cores": threads, "kind": "port",
                               "sample": sample + f"; {t_cpu:.2f}s per sample eval, "
                                                  f"{w_sample:.3e} of {w_total:.3e} flops; "
                                                  f"blas={'mkl' if mkl else 'oracle loops'}"}
        out["gpu_vs_cpu"] = value / cpu_evals
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
