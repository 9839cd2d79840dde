"""The synthetic workloads of the five benchmark configurations.

All live on the unit square with the planar metric and linear-trend fields
(see synth.py); locations and values come from fixed seeds so that the
benchmark and the tests build identical inputs.

* C0: n = 4,000 uniform points, an exact dense GP draw, M = 4, J = 4, r = 16,
  1,000 prediction points.
* C1: n = 250,000 clustered points (60 % uniform, 40 % in 20 Gaussian
  clusters of sigma 0.02), field drawn from the MRA prior, M = 8, J = 4,
  r = 64, 50,000 prediction points.
* C2: as C1 with n = 2,000,000, M = 10 and 200,000 prediction points.
* C3: 256 independent fields, n log-uniform in [5,000, 200,000], per-field
  kernel parameters, M chosen so a finest region holds at most 256 points on
  average, r = 32, mixed precision.
* C4: fields of n = 50,000 uniform points, M = 8, r = 32, drawn from the MRA
  prior with parameters from the warm-start ranges (warmstart.py).
"""
import numpy as np

from . import capi, synth

SEED_LOC, SEED_FIELD = 20260101, 20260102


def clustered_tree(n, npred, M, r, seed=SEED_LOC, J=4):
    """Clustered locations partitioned by the program rule on [0,1]^2, values zero."""
    rng = np.random.default_rng(seed)
    x, y = synth.clustered_points(n, rng)
    px, py = synth.uniform_points(npred, rng) if npred else (None, None)
    return capi.partition_program(x, y, np.zeros(n), px, py, M=M, J=J, r=r, bounds=synth.UNIT_BOUNDS)


def prior_field(tree, seed=SEED_FIELD, cov=None, with_pred=True):
    """A handle on `tree` whose observations are one draw from the MRA prior of
    the C1/C2 kernel (log sigma^2 = 0.5x - 0.3y, log beta = log 0.05 + 0.8x,
    nugget 0.05); the drawn values are also written into tree["obs_val"]."""
    h = capi.Handle(tree, with_pred=with_pred and tree.get("pred_offsets") is not None)
    h.set_covariance(cov or synth.linear_trend_cov())
    h.set_params(**synth.UNIT_PARAMS)
    tree["obs_val"] = h.sample_prior(seed, replace=True)
    return h


def c1_tree():
    return clustered_tree(250_000, 50_000, M=8, r=64, seed=SEED_LOC + 11)


def c2_tree(n=2_000_000, npred=200_000, M=10, r=64):
    return clustered_tree(n, npred, M=M, r=r, seed=SEED_LOC)


def c3_levels(n):
    """M such that a finest region holds at most 256 points on average."""
    return 1 + max(1, int(np.ceil(np.log(n / 256.0) / np.log(4.0))))


def c3_specs(n_fields=256, seed=SEED_LOC + 3):
    """Per field: (n, M, theta) with theta = (a_s, bx_s, by_s, a_b, bx_b,
    log tau^2): log range in [log 0.02, log 0.2], range and variance slopes in
    [-1, 1], nugget in [0.01, 0.2]."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_fields):
        n = int(np.exp(rng.uniform(np.log(5000), np.log(200000))))
        theta = (0.0, rng.uniform(-1, 1), 0.0, rng.uniform(np.log(0.02), np.log(0.2)), rng.uniform(-1, 1),
                 np.log(rng.uniform(0.01, 0.2)))
        out.append((n, c3_levels(n), theta, int(rng.integers(1 << 62))))
    return out


def trend_cov(theta):
    a_s, bx_s, by_s, a_b, bx_b, lt = (float(v) for v in theta)
    return dict(nonstationary=1, metric=synth.METRIC["planar"], wendland_len=1.0,
                sigmasq=dict(kind=synth.FIELD_LINEAR, intercept=a_s, slope_x=bx_s, slope_y=by_s),
                beta=dict(kind=synth.FIELD_LINEAR, intercept=a_b, slope_x=bx_b, slope_y=0.0),
                tausq=dict(kind=synth.FIELD_CONSTANT, intercept=lt))


def c3_field(spec, index, precision=capi.PRECISION_MIXED, keep_tree=False):
    """One C3 field as a handle (values drawn from the MRA prior with the
    field's own parameters); with keep_tree also returns its tree."""
    n, M, theta, seed = spec
    rng = np.random.default_rng(seed)
    x, y = synth.uniform_points(n, rng)
    tree = capi.partition_program(x, y, np.zeros(n), None, None, M=M, J=4, r=32, bounds=synth.UNIT_BOUNDS)
    h = capi.Handle(tree, with_pred=False)
    h.set_covariance(trend_cov(theta))
    h.set_params(**synth.UNIT_PARAMS)
    tree["obs_val"] = h.sample_prior(SEED_FIELD + index, replace=True)
    h.set_precision(precision)
    return (h, tree) if keep_tree else h
