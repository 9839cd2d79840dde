"""Benchmark: MRA log-likelihood evaluations per second on one MI355X.

Workload: n = 2,000,000 points on [0,1]^2, 60% uniform + 40% in 20 Gaussian
clusters (sigma 0.02); nonstationary exponential kernel with
log sigma^2 = 0.5x - 0.3y, log beta = log 0.05 + 0.8x, nugget 0.05 (planar
metric); M = 10 levels, J = 4, r = 64 knots; observations drawn from the MRA
prior itself (mra_sample_prior, FP64); 200,000 uniform prediction points.
One step = one full log-likelihood evaluation (coarse prior, finest level,
posterior, reduction) with every input resident in HBM.

The headline `value` is the mixed-precision path (FP32 covariance and
finest-level observation GEMMs, FP64 knot Cholesky / conditioning /
log-determinants) replayed as one captured HIP graph per evaluation.  The
same line reports the pure-FP64 GPU path, |loglik_mixed - loglik_fp64| / n,
the prediction error of the mixed path against FP64 at the prediction
points, per-launch-group times (HIP events on the evaluation stream, from
an eager pass of the same number of steps), the roofline of the dominant
kernel, and the CPU oracle timed on a bounded sample.

Prints one JSON line.  --gpus N > 1 is refused: this build runs on one GPU.
"""
import argparse
import json
import os
import sys
import time

# The C3 batch runs its fields on up to 16 streams, which run side by side
# only on as many hardware queues (HIP's default is 4; --hw-queues N sets
# another count, at most 32); the setting is read when the process first uses
# the GPU, so it is made before anything does.
def _hw_queues(argv):
    for i, a in enumerate(argv):
        if a.startswith("--hw-queues="):
            return a.split("=", 1)[1]
        if a == "--hw-queues" and i + 1 < len(argv):
            return argv[i + 1]
    return "16"


os.environ["GPU_MAX_HW_QUEUES"] = _hw_queues(sys.argv)

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from nonstationary_mra_amd import capi, synth, workloads, workmodel  # noqa: E402

PEAK = {"f64": 78.6e12, "f32": 157.3e12}   # MI355X dense matrix FLOP/s (FP64 matrix = vector)
PEAK_HBM = 8.0e12                         # bytes/s
SEED_LOC, SEED_FIELD = 20260101, 20260102
METRIC = "MRA loglik evals/sec at n=2M,M=10,r=64: mixed vs FP64 GPU, |Δℓ|/n≤1e-6"
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic.json")


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_workload(n, npred, M, J, r, seed=SEED_LOC):
    rng = np.random.default_rng(seed)
    x, y = synth.clustered_points(n, rng)
    px, py = synth.uniform_points(npred, rng) if npred else (None, None)
    return capi.partition_program(x, y, np.zeros(n), px, py, M=M, J=J, r=r, bounds=synth.UNIT_BOUNDS)


def group_names(M):
    names = {capi.T_FINEST_CROSS: "finest_cross", capi.T_FINEST_LOCAL: "finest_local"}
    for l in range(M - 1):
        names[capi.T_PRIOR0 + l] = f"prior_l{l}"
        names[capi.T_POST0 + l] = f"post_l{l}"
    return names


KERNEL_OF = {"prior": "k_prior_coarse", "post": "k_posterior", "finest_cross": "k_finest_cross",
             "finest_local": "k_finest_local"}


def kernel_of(gname):
    return KERNEL_OF[gname.split("_l")[0]] if gname.startswith(("prior_l", "post_l")) else KERNEL_OF[gname]


def timed(h, steps, warmup):
    """(evals/s, ms/step, D) over `steps` evaluations after `warmup`."""
    D = None
    for _ in range(warmup):
        D = h.likelihood()
    t1 = time.perf_counter()
    for _ in range(steps):
        D = h.likelihood()
    el = time.perf_counter() - t1
    return steps / el, 1e3 * el / steps, D


def eager_groups(h, steps):
    """Average per-launch-group milliseconds over `steps` eager evaluations
    (HIP events recorded on the evaluation stream around every group)."""
    h.use_graph(False)
    tim = np.zeros(capi.T_COUNT)
    t1 = time.perf_counter()
    for _ in range(steps):
        h.likelihood()
        tim += h.timings()
    return tim / steps, 1e3 * (time.perf_counter() - t1) / steps


def cpu_per_level(ph, Ms, M, J, r, cs, counts):
    """C2 CPU seconds from the oracle's phase times on the sample (coarse prior
    per level, finest level, posterior per level), level by level: the
    sample's coarse prior levels 0..Ms-2 and posterior levels 0..Ms-3 do the
    same work as C2's levels of the same index (same regions, same ancestor
    counts) and are taken as measured; C2's deepest coarse prior level is
    priced at the flop rate of the sample's deepest, its finest level at the
    sample finest level's rate, its deepest dense posterior level at the
    sample's deepest dense one, and its row-reading (finest-but-one) posterior
    level at the sample's row-reading level.  Returns (seconds, measured
    seconds)."""
    Ls, L = Ms - 1, M - 1
    fs, fc = workmodel.groups(Ms, J, r, cs), workmodel.groups(M, J, r, counts)
    tp, tf, tq = ph[:Ls], ph[Ls], ph[Ls + 1:]

    def rate(f, t):
        return f / max(t, 1e-9)

    total = measured = 0.0
    for l in range(L):
        if l < Ls:
            t = fc[f"prior_l{l}"] / rate(fs[f"prior_l{l}"], tp[l])
            measured += t
        else:
            t = fc[f"prior_l{l}"] / rate(fs[f"prior_l{Ls - 1}"], tp[Ls - 1])
        total += t
    ffs = fs["finest_cross"] + fs["finest_local"]
    total += (fc["finest_cross"] + fc["finest_local"]) / rate(ffs, tf)
    for l in range(L):
        if l < Ls - 1:
            t = fc[f"post_l{l}"] / rate(fs[f"post_l{l}"], tq[l])
            measured += t
        elif l < L - 1:
            t = fc[f"post_l{l}"] / rate(fs[f"post_l{Ls - 2}"], tq[Ls - 2])
        else:
            t = fc[f"post_l{l}"] / rate(fs[f"post_l{Ls - 1}"], tq[Ls - 1])
        total += t
    return total, measured


def cpu_baseline(h_unused, M, J, r, n, counts, budget_s=30.0):
    """The FP64 oracle (a restatement of the reference's CPU algorithm, OpenMP
    over regions, MKL kernels when libmkl_rt is present) on a bounded sample:
    the same generator at n/4 points and M-1 levels (same points per finest
    region; about 4 s per evaluation on 16 cores and 110 GB of host memory,
    where the full size would need about 4x that memory).  The C2 time is
    the oracle's phase times on the sample carried over level by level
    (cpu_per_level); the scaling of the whole sample time by the ratio of
    algorithmic flops is reported beside it.  The FP64 GPU path is evaluated
    on the same sample for its error against the oracle."""
    from oracle import oracle
    mkl = oracle.lib().orc_set_blas(1)
    ns, Ms = n // 4, M - 1
    tree = make_workload(ns, 0, Ms, J, r, seed=SEED_LOC + 1)
    cov = synth.linear_trend_cov()
    prm = dict(synth.UNIT_PARAMS)
    tree["obs_val"] = np.random.default_rng(3).standard_normal(len(tree["obs_lon"]))
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    times, phases = [], []
    t_start = time.time()
    Do = None
    while True:
        t0 = time.perf_counter()
        Do, ph = oracle.evaluate_timed(tree, cov, prm, nthreads=threads)
        times.append(time.perf_counter() - t0)
        phases.append(ph)
        if len(times) >= 3 or time.time() - t_start > budget_s:
            break
    t = float(np.median(times))
    ph = np.median(np.array(phases), axis=0)
    cs = np.diff(tree["obs_offsets"].astype(np.int64))
    w_sample = workmodel.total(Ms, J, r, cs)
    t_c2, t_meas = cpu_per_level(ph, Ms, M, J, r, cs, counts)
    g = capi.Handle(tree, with_pred=False)
    try:
        g.set_covariance(cov)
        g.set_params(**prm)
        Dg = g.likelihood()
    finally:
        g.close()
    return dict(t=t, w=w_sample, t_c2=t_c2, t_measured=t_meas, threads=threads, mkl=mkl, n_runs=len(times),
                sample=f"n={ns} M={Ms} J={J} r={r} (same generator)", rel=abs(Dg - Do) / abs(Do))


def load_traffic(launch):
    """HBM bytes per launch of `launch` from the committed PMC summary
    (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration), or None."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
        return t["per_launch_bytes"].get(launch)
    except (OSError, ValueError, KeyError):
        return None


def calibrate_copy(mib):
    """A plain copy kernel of known size: reads and writes mib MiB once each
    (three launches), for scaling the FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    a = torch.ones(mib * (1 << 20) // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": mib << 20, "copy_bytes_written": mib << 20, "launches": 3}), flush=True)


def cov_kernel_rate(n1, n2, precision, steps, seed=SEED_LOC + 5):
    """The pairwise covariance kernel alone (mra_cov_tile_device): an n1 x n2
    row-major matrix of the C2 kernel (clustered points, linear-trend fields,
    planar) written to HBM.  Returns (ms per launch from HIP events on the
    launch stream, bytes written per launch)."""
    rng = np.random.default_rng(seed)
    cov = synth.linear_trend_cov()

    def pts(n):
        x, y = synth.clustered_points(n, rng)
        s = cov["sigmasq"]
        b = cov["beta"]
        ls = s["intercept"] + s["slope_x"] * x + s["slope_y"] * y
        lb = b["intercept"] + b["slope_x"] * x + b["slope_y"] * y
        return (capi.DeviceBuffer.from_array(np.concatenate([x, y, np.zeros(n)])),
                capi.DeviceBuffer.from_array(np.concatenate([np.exp(0.5 * ls), np.exp(lb)])))

    xyz1, sb1 = pts(n1)
    xyz2, sb2 = pts(n2)
    esz = 4 if precision == capi.PRECISION_MIXED else 8
    out = capi.DeviceBuffer(n1 * n2 * esz)
    try:
        ms = capi.cov_matrix_timed(precision, synth.METRIC["planar"], xyz1.ptr, sb1.ptr, n1, xyz2.ptr, sb2.ptr, n2,
                                   out.ptr, steps)
    finally:
        for b in (out, xyz1, sb1, xyz2, sb2):
            b.free()
    return ms, n1 * n2 * esz


def run_cov(args):
    """The covariance kernel on its own: a 65,536 x 65,536 matrix (4.29e9
    pairs, about the pair count of one C2 evaluation) in FP32 and FP64."""
    n1 = n2 = args.cov_n
    res = {}
    for name, prec in (("f32", capi.PRECISION_MIXED), ("f64", capi.PRECISION_FP64)):
        ms, nbytes = cov_kernel_rate(n1, n2, prec, args.steps)
        res[name] = {"ms_per_launch": ms, "bytes_written": nbytes, "GB_per_s": nbytes / (ms * 1e-3) / 1e9,
                     "pairs_per_s": n1 * n2 / (ms * 1e-3)}
        log(f"cov {name}: {ms:.3f} ms, {res[name]['GB_per_s']:.0f} GB/s")
    f32 = res["f32"]
    out = {"metric": f"covariance evaluation GB/s ({n1}x{n2} pairs, FP32 output)", "value": f32["GB_per_s"],
           "unit": "GB/s", "n_gpus": 1, "steps": args.steps, "warmup": 1, "ms_per_step": f32["ms_per_launch"],
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic (clustered points, linear-trend nonstationary fields, planar)",
           "config": {"workload": f"{n1} x {n2} nonstationary Matern-1/2 covariance matrix, row-major", "n1": n1,
                      "n2": n2},
           "roofline": {"bound": "hbm", "kernel": "k_cov_tile_f32", "achieved": f32["GB_per_s"],
                        "peak": PEAK_HBM / 1e9, "unit": "GB/s", "frac": f32["GB_per_s"] * 1e9 / PEAK_HBM,
                        "traffic": load_traffic("cov_f32") if n1 == n2 == 65536 else None,
                        "avg_launch_ms": f32["ms_per_launch"]},
           "fp64": res["f64"]}
    print(json.dumps(out), flush=True)


def c3_measure(n_fields, steps, warmup):
    """Configuration of the uneven batch: independent fields with n drawn
    log-uniform in [5,000, 200,000] (uniform locations), per-field kernel
    parameters (log range in [log 0.02, log 0.2], range and variance slopes
    in [-1, 1], nugget in [0.01, 0.2]), M chosen so that a finest region holds
    at most 256 points on average, r = 32, mixed precision, fields drawn from
    the MRA prior (workloads.c3_specs / c3_field); all likelihoods evaluated
    as one batch (mra_batch_*: one captured graph for all fields) whose launch
    setting comes from the batch autotuner.  One step = one evaluation of
    every field."""
    t0 = time.time()
    specs = workloads.c3_specs(n_fields)
    hs = []
    for i, spec in enumerate(specs):
        hs.append(workloads.c3_field(spec, i))
        if i % 32 == 31:
            log(f"c3: {i + 1} fields ready ({time.time() - t0:.1f}s)")
    n_tot = sum(s[0] for s in specs)
    b = capi.Batch(hs)
    streams, graph, tune_ms = b.autotune()
    log(f"c3 batch autotune: streams={streams} graph={graph} {tune_ms:.1f} ms")
    setup_s = time.time() - t0
    for _ in range(warmup):
        D, st = b.likelihood()
    t1 = time.perf_counter()
    for _ in range(steps):
        D, st = b.likelihood()
    el = time.perf_counter() - t1
    # the same fields one after another (each synchronised), for comparison
    t2 = time.perf_counter()
    Ds = np.array([h.likelihood() for h in hs])
    seq_ms = 1e3 * (time.perf_counter() - t2)
    res = {"fields": n_fields, "total_points": n_tot, "ms_per_batch": 1e3 * el / steps,
           "field_evals_per_s": n_fields * steps / el, "sequential_ms_per_batch": seq_ms,
           "batch_matches_sequential": bool(np.array_equal(D, Ds)), "not_pd_fields": int((st != 0).sum()),
           "streams": streams, "graph": graph, "setup_s": setup_s}
    b.close()
    for h in hs:
        h.close()
    return res


def run_c3(args):
    """--config c3: the batch of independent fields on its own line (see
    c3_measure)."""
    r = c3_measure(args.c3_fields, args.steps, args.warmup)
    out = {"metric": "C3 batch loglik evals/sec (256 uneven fields, r=32, mixed)", "value": r["field_evals_per_s"],
           "unit": "evals/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": r["ms_per_batch"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32+f64", "data": "synthetic (uniform points, fields drawn from the MRA prior)",
           "config": {"workload": "256 fields, n log-uniform in [5e3, 2e5], r=32, M per field, mixed precision",
                      "fields": args.c3_fields, "total_points": r["total_points"], "streams": r["streams"],
                      "graph": r["graph"]},
           "sequential_ms_per_step": r["sequential_ms_per_batch"],
           "batch_matches_sequential": r["batch_matches_sequential"], "not_pd_fields": r["not_pd_fields"]}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=2_000_000)
    ap.add_argument("--npred", type=int, default=200_000)
    ap.add_argument("--M", type=int, default=10)
    ap.add_argument("--r", type=int, default=64)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="mixed")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--tuning", default=None, metavar="X32,LR,FB16",
                    help="fixed work-list shape instead of the autotuner (e.g. 64,128,1)")
    ap.add_argument("--config", choices=["c2", "c3", "cov"], default="c2",
                    help="c2: the headline workload; c3: the uneven batch of independent fields; "
                         "cov: the covariance kernel alone")
    ap.add_argument("--cov-n", type=int, default=65536)
    ap.add_argument("--c3-fields", type=int, default=256)
    ap.add_argument("--finest-wave", type=int, default=None, choices=[0, 1],
                    help="mixed-precision finest conditioning one wave per tile (MRA_OPT_FINEST_WAVE)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-c3", action="store_true", help="skip the C3 batch leg of the default run")
    ap.add_argument("--hw-queues", type=int, default=16, choices=range(1, 33), metavar="N",
                    help="HIP hardware queues of the process (GPU_MAX_HW_QUEUES, read before the GPU is used)")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 comparison leg")
    ap.add_argument("--no-cov-kernel", action="store_true", help="skip the standalone covariance-kernel leg")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB (HBM counter calibration) and exit")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU only (no multi-GPU path); use --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.config == "c3":
        run_c3(args)
        return
    if args.config == "cov":
        run_cov(args)
        return
    J = 4
    graph = not args.no_graph
    t0 = time.time()
    tree = make_workload(args.n, args.npred, args.M, J, args.r)
    counts = np.diff(tree["obs_offsets"].astype(np.int64))
    n = int(counts.sum())
    log(f"partition {time.time() - t0:.1f}s n_obs={n} max_nF={counts.max()} empty={(counts == 0).sum()}")
    h = capi.Handle(tree, with_pred=args.npred > 0)
    h.set_covariance(synth.linear_trend_cov())
    h.set_params(**synth.UNIT_PARAMS)
    h.sample_prior(SEED_FIELD, replace=True)
    log(f"setup {time.time() - t0:.1f}s")
    prec = capi.PRECISION_MIXED if args.precision == "mixed" else capi.PRECISION_FP64
    h.set_precision(prec)
    if args.finest_wave is not None:
        h.set_option(capi.OPT_FINEST_WAVE, args.finest_wave)
    # work-list shape from the library's autotuner (timed on this workload)
    if args.tuning:
        h.set_tuning(*[int(v) for v in args.tuning.split(",")])
        tune_cfg, tune_ms = h.get_tuning(), None
    else:
        tune_cfg, tune_ms = h.autotune() if not args.no_autotune else (h.get_tuning(), None)
    log(f"autotune {tune_cfg} {tune_ms}")
    # headline: the requested precision, graph replay
    h.use_graph(graph)
    value, ms_per_step, D = timed(h, args.steps, args.warmup)
    log(f"{args.precision}{' graph' if graph else ''}: {ms_per_step:.2f} ms/eval D={D:.10e}")
    # per-group times of the same evaluation, eager
    tim, eager_ms = eager_groups(h, args.steps)
    names = group_names(args.M)
    flops = workmodel.groups(args.M, J, args.r, counts)
    gid = max(names, key=lambda g: tim[g])
    gname = names[gid]
    fp32_group = args.precision == "mixed" and gname == "finest_cross"
    peak = PEAK["f32" if fp32_group else "f64"]
    achieved = flops[gname] / (tim[gid] * 1e-3)
    per_group = {names[g]: round(float(tim[g]), 3) for g in sorted(names) if tim[g] > 0.01}
    roof = {"bound": "mfma", "kernel": kernel_of(gname), "launch": gname, "achieved": achieved / 1e12,
            "peak": peak / 1e12, "unit": "TFLOP/s", "frac": achieved / peak, "traffic": load_traffic(gname),
            "flops_per_launch": flops[gname], "avg_launch_ms": float(tim[gid]),
            "arith": "f32" if fp32_group else "f64"}
    w_total = workmodel.total(args.M, J, args.r, counts)
    out = {
        "metric": METRIC,
        "value": value, "unit": "evals/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32+f64",
        "data": "synthetic (clustered points, field drawn from the MRA prior)",
        "config": {"workload": f"n={args.n} clustered, M={args.M}, J={J}, r={args.r}, "
                               f"nonstationary linear-trend Matern-1/2, planar, {args.npred} prediction points",
                   "n": n, "n_pred": args.npred, "M": args.M, "J": J, "r": args.r, "precision": args.precision,
                   "graph": graph, "max_obs_per_finest_region": int(counts.max()),
                   "worklist_shape": dict(zip(("x32max", "lrmax", "fb16"), tune_cfg)),
                   "finest_wave": h.get_option(capi.OPT_FINEST_WAVE),
                   "seeds": {"locations": SEED_LOC, "field": SEED_FIELD}},
        "loglik": -0.5 * (D + n * np.log(2 * 3.14159265359)), "D": D,
        "eager_ms_per_step": eager_ms,
        "whole_eval": {"flops": w_total, "tflops": w_total / (ms_per_step * 1e-3) / 1e12,
                       "frac_fp64_peak": w_total / (ms_per_step * 1e-3) / PEAK["f64"]},
        "launch_ms": per_group,
        "roofline": roof,
    }
    if args.precision == "mixed" and not args.no_fp64:
        h.set_precision(capi.PRECISION_FP64)
        # the FP64 path has its own best work-list shape (tile costs differ)
        tune64 = tune_cfg
        if not args.tuning and not args.no_autotune:
            tune64, _ = h.autotune()
            log(f"autotune fp64 {tune64}")
        h.use_graph(graph)
        v64, ms64, D64 = timed(h, args.steps, args.warmup)
        log(f"fp64: {ms64:.2f} ms/eval D={D64:.10e}")
        out["fp64_gpu"] = {"value": v64, "ms_per_step": ms64, "D": D64,
                           "worklist_shape": dict(zip(("x32max", "lrmax", "fb16"), tune64))}
        out["mixed_vs_fp64"] = {"speedup": value / v64, "abs_dloglik_per_n": abs(D - D64) / (2 * n)}
        if args.npred:
            t1 = time.perf_counter()
            _, m64, s64 = h.predict()
            tp64 = time.perf_counter() - t1
            h.set_precision(capi.PRECISION_MIXED)
            h.set_tuning(*tune_cfg)
            h.predict()
            t1 = time.perf_counter()
            _, mm, sm = h.predict()
            tpm = time.perf_counter() - t1
            out["prediction"] = {"n_pred": len(mm), "ms_mixed": 1e3 * tpm, "ms_fp64": 1e3 * tp64,
                                 "mean_max_abs_err_rel_to_max": float(np.abs(mm - m64).max() / np.abs(m64).max()),
                                 "var_max_rel_err": float(np.max(np.abs(sm - s64) / np.abs(s64)))}
    h.close()
    if not args.no_cov_kernel:
        # the pairwise covariance kernel alone, materialised: HBM-bound
        cn = args.cov_n
        cms, cbytes = cov_kernel_rate(cn, cn, capi.PRECISION_MIXED, 10)
        out["covariance_kernel"] = {"bound": "hbm", "kernel": "k_cov_tile_f32", "n1": cn, "n2": cn, "dtype": "f32",
                                    "ms_per_launch": cms, "achieved": cbytes / (cms * 1e-3) / 1e9,
                                    "peak": PEAK_HBM / 1e9, "unit": "GB/s",
                                    "frac": cbytes / (cms * 1e-3) / PEAK_HBM,
                                    "traffic": load_traffic("cov_f32") if cn == 65536 else None}
    if not args.no_c3:
        # the uneven batch of independent fields (C3), three batch evaluations
        out["c3"] = c3_measure(args.c3_fields, 3, 1)
        log(f"c3: {out['c3']['ms_per_batch']:.1f} ms per batch of {args.c3_fields} fields")
    if not args.no_cpu_baseline:
        cb = cpu_baseline(None, args.M, J, args.r, args.n, counts)
        cpu_evals = 1.0 / cb["t_c2"]
        out["cpu_baseline"] = {"value": cpu_evals, "unit": "evals/s", "cores": cb["threads"], "kind": "port",
                               "sample": cb["sample"] + f", {cb['n_runs']} runs, median {cb['t']:.2f}s per sample "
                                         f"eval; C2 time {cb['t_c2']:.2f}s from the phase times level by level "
                                         f"({cb['t_measured']:.2f}s of it on levels whose work the sample does "
                                         f"identically); blas={'mkl' if cb['mkl'] else 'oracle loops'}",
                               "value_by_total_flops": (1.0 / cb["t"]) * (cb["w"] / w_total)}
        out["gpu_vs_cpu"] = value / cpu_evals
        if "fp64_gpu" in out:
            out["fp64_gpu"]["vs_cpu"] = out["fp64_gpu"]["value"] / cpu_evals
        out["fp64_gpu_vs_oracle_rel_on_sample"] = cb["rel"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
