"""GPU tests of the reference's driver flow through the C++ layer
(include/mra_classes.hpp, mra_approximation.hpp, mra_optimization.hpp):
nonstationary_mra_amd/examples/reference_main.cpp runs the sequence of
src/main.cpp:104-148 on a `user_parameters` file, in each CALCULATION_MODE,
and its outputs are compared with the oracle on the same inputs.  Also the
field, covariance and log-likelihood entry points the classes use."""
import math
import os
import subprocess

import numpy as np
import pytest

from nonstationary_mra_amd import build, capi, formats, synth
from oracle import oracle
from tests.cases import EXAMPLE, EXAMPLE_PARAMS, example_cov
from tests.test_io import example_params, write_data

pytestmark = pytest.mark.gpu

PI_LITERAL = 3.14159265359


def run_main(tmp_path, params_file):
    assert os.path.exists(build.REFMAIN), "build the library first (it also builds the C++ examples)"
    res = subprocess.run([build.REFMAIN, str(params_file)], capture_output=True, text=True, timeout=300,
                         cwd=str(tmp_path))
    print(res.stdout, res.stderr, flush=True)
    assert res.returncode == 0
    out = {}
    for line in res.stdout.splitlines():
        parts = line.split()
        if parts and parts[0] in ("D", "loglik", "mean_sum"):
            out[parts[0]] = float(parts[1])
        if parts and parts[0] == "opt":
            out["opt"] = [float(v) for v in parts[1:4]]
            out["opt_D"] = float(parts[5])
            out["evals"] = int(parts[7])
    return out


def test_example_prediction_mode(tmp_path, example):
    """The reference example (user files, VARIOUS_J, chordal Wendland fields)
    in prediction mode: D, the printed log-likelihood, and the prediction file
    against the oracle."""
    out = run_main(tmp_path, example_params(tmp_path))
    tree, cov, prm = example
    Do, mo, vo = oracle.evaluate(tree, cov, prm)
    n = len(tree["obs_lon"])
    assert abs(out["D"] - Do) <= 1e-10 * abs(Do)
    assert abs(out["loglik"] - (-0.5 * (Do + n * math.log(2 * PI_LITERAL)))) <= 1e-10 * abs(Do)
    pr = formats.read_prediction_result(str(tmp_path / "pred_res.bin"))
    np.testing.assert_array_equal(pr["lon"], tree["pred_lon"])
    np.testing.assert_allclose(pr["mean"], mo, rtol=1e-9, atol=1e-10)
    np.testing.assert_allclose(pr["var"], vo, rtol=1e-9, atol=1e-10)
    assert abs(out["mean_sum"] - mo.sum()) <= 1e-9 * np.abs(mo).sum()


def _synthetic(tmp_path, n=3000, dup=300, seed=3):
    """Planar points in degrees near the equator with exact duplicates and a
    linear trend in the values."""
    rng = np.random.default_rng(seed)
    lon, lat = rng.uniform(-8, 8, n), rng.uniform(-6, 6, n)
    k = rng.integers(0, n, dup)
    lon = np.concatenate([lon, lon[k]])
    lat = np.concatenate([lat, lat[k]])
    val = 2.0 + 0.1 * lon - 0.05 * lat + rng.standard_normal(len(lon))
    write_data(tmp_path / "d.bin", lon, lat, val)
    return lon, lat, val


def _stationary_params(tmp_path, mode, **over):
    kv = dict(DATA_FILE_NAME=str(tmp_path / "d.bin"), ELIMINATION_DUPLICATES_FLAG="true", REGRESSION_FLAG="true",
              NONSTATIONARY_FLAG="false", OFFSET="default", NUM_PARTITIONS_J=4, NUM_KNOTS_r=16,
              NUM_LEVELS_M="default", CALCULATION_MODE=mode, PREDICTION_LOCATION_MODE="A",
              DUMP_PREDICTION_RESULTS_FLAG="false", SIGMASQ=1.3, BETA=400, TAUSQ=0.2, MAX_ITERATIONS=400,
              CHORDAL_DISTANCE_FLAG="false", PROVIDED_KNOTS_FLAG="false", OBS_AS_KNOTS_FLAG="true",
              VARIOUS_J_FINEST_LEVEL_FLAG="false", SIGMASQ_INITIAL_GUESS=1.0, SIGMASQ_LOWER_BOUND=0.1,
              SIGMASQ_UPPER_BOUND=10, BETA_INITIAL_GUESS=300, BETA_LOWER_BOUND=50, BETA_UPPER_BOUND=2000,
              TAUSQ_INITIAL_GUESS=0.5, TAUSQ_LOWER_BOUND=0.01, TAUSQ_UPPER_BOUND=2)
    kv.update(over)
    from tests.test_io import write_params
    return write_params(tmp_path / "user_parameters", **kv)


def test_likelihood_mode_with_preprocessing(tmp_path):
    """Program partition with the default M after duplicate elimination and
    the regression detrend, stationary model, great-circle distance: the
    printed log-likelihood against the oracle on the same preprocessed data
    (duplicates removed and detrended by numpy, partitioned by the oracle's
    restatement of build_partition_by_program)."""
    lon, lat, val = _synthetic(tmp_path)
    out = run_main(tmp_path, _stationary_params(tmp_path, "likelihood"))
    order = np.lexsort((np.arange(len(lon)), lat, lon))
    keep = np.ones(len(lon), bool)
    keep[1:] = (lon[order][1:] != lon[order][:-1]) | (lat[order][1:] != lat[order][:-1])
    idx = order[keep]
    X = np.column_stack([np.ones(len(idx)), lon[idx], lat[idx]])
    coef, *_ = np.linalg.lstsq(X, val[idx], rcond=None)
    resid = val[idx] - X @ coef
    n = len(idx)
    M = capi.default_levels(n, 4, 16)
    tree = oracle.partition_program(lon[idx], lat[idx], resid, None, None, M=M, J=4, r=16)
    cov = dict(nonstationary=0, metric=capi.METRIC["great_circle"], wendland_len=1.0,
               sigmasq=dict(kind=0, intercept=0.0), beta=dict(kind=0, intercept=0.0), tausq=dict(kind=0, intercept=0.0))
    prm = dict(sill=1.3, range=400.0, nugget=0.2, sill_scale=0.0, range_scale=0.0, nugget_scale=0.0)
    Do, _, _ = oracle.evaluate(tree, cov, prm, predict=False)
    print("M", M, "n", n, "D", out["D"], "oracle", Do, flush=True)
    # the residuals differ from numpy's lstsq in the last bits only
    assert abs(out["D"] - Do) <= 1e-9 * abs(Do)
    assert abs(out["loglik"] - (-0.5 * (out["D"] + n * math.log(2 * PI_LITERAL)))) <= 1e-12 * abs(out["loglik"])


def test_build_structure_only_mode(tmp_path):
    lon, lat, val = _synthetic(tmp_path, n=2000, dup=0)
    run_main(tmp_path, _stationary_params(tmp_path, "build_structure_only", ELIMINATION_DUPLICATES_FLAG="false",
                                          REGRESSION_FLAG="false"))
    text = (tmp_path / "structure_information.txt").read_text()
    assert text.startswith(f"The number of levels: {capi.default_levels(2000, 4, 16)}\n")
    assert "The number of knots in each region at the finest level:" in text


def test_optimization_mode_nonstationary(tmp_path, example):
    """The example in optimization mode: the three scales from (1, 1, 1)
    within [1e-2, 1e2] (src/optimization.cpp:158-192); the optimum's D is
    the oracle's D at the returned scales and no larger than at the start."""
    out = run_main(tmp_path, example_params(tmp_path, CALCULATION_MODE="optimization", MAX_ITERATIONS=300))
    tree, cov, _ = example
    s = out["opt"]
    prm = dict(sill=0.0, range=0.0, nugget=0.0, sill_scale=s[0], range_scale=s[1], nugget_scale=s[2])
    Do, _, _ = oracle.evaluate(tree, cov, prm, predict=False)
    D1, _, _ = oracle.evaluate(tree, cov, dict(prm, sill_scale=1.0, range_scale=1.0, nugget_scale=1.0), predict=False)
    print("opt", s, out["opt_D"], Do, D1, out["evals"], flush=True)
    assert abs(out["opt_D"] - Do) <= 1e-10 * abs(Do)
    assert out["opt_D"] <= D1
    assert all(1e-2 <= v <= 1e2 for v in s) and out["evals"] >= 8


def test_fields_match_oracle(example):
    """get_sigmasq / get_beta / get_tausq through mra_fields: chordal
    (x, y, z as given and derived from lon/lat) and great circle."""
    tree, cov, _ = example
    lon, lat = tree["obs_lon"][:300], tree["obs_lat"][:300]
    s, b, t = capi.fields(cov, lon, lat)
    so, bo, to = oracle.fields(cov, lon, lat)
    for a, o in ((s, so), (b, bo), (t, to)):
        np.testing.assert_allclose(a, o, rtol=1e-13, atol=0)
    x = np.cos(np.radians(lat)) * np.cos(np.radians(lon))
    y = np.cos(np.radians(lat)) * np.sin(np.radians(lon))
    z = np.sin(np.radians(lat))
    s2, _, _ = capi.fields(cov, x=x, y=y, z=z)
    np.testing.assert_allclose(s2, so, rtol=1e-13, atol=0)
    gc = dict(cov, metric=capi.METRIC["great_circle"])
    s3, b3, _ = capi.fields(gc, lon, lat)
    so3, bo3, _ = oracle.fields(gc, lon, lat)
    np.testing.assert_allclose(s3, so3, rtol=1e-13, atol=0)
    np.testing.assert_allclose(b3, bo3, rtol=1e-13, atol=0)


def test_covariance_with_given_coordinates(example):
    """evaluate_nonstationary_{cross,variance}_covariance with the caller's
    x, y, z (chordal): equal to the oracle's matrices, whose coordinates are
    computed on the host as the reference's are."""
    tree, cov, prm = example
    lon1, lat1 = tree["obs_lon"][:70], tree["obs_lat"][:70]
    lon2, lat2 = tree["obs_lon"][100:140], tree["obs_lat"][100:140]

    def xyz(lon, lat):
        return (np.cos(np.radians(lat)) * np.cos(np.radians(lon)), np.cos(np.radians(lat)) * np.sin(np.radians(lon)),
                np.sin(np.radians(lat)))
    keep = capi.Keep()
    c = capi.cov_desc(cov, keep)
    out = np.zeros(70 * 40)
    p = [prm[k] for k in ("sill", "range", "nugget", "sill_scale", "range_scale", "nugget_scale")]
    x1, x2 = xyz(lon1, lat1), xyz(lon2, lat2)
    capi.check(capi.lib().mra_cross_covariance_xyz(capi.C.byref(c), *p, keep.d(lon1), keep.d(lat1),
                                                    *[keep.d(a) for a in x1], 70, keep.d(lon2), keep.d(lat2),
                                                    *[keep.d(a) for a in x2], 40, capi._d(out)))
    ref = oracle.cross_covariance(cov, prm, lon1, lat1, lon2, lat2)
    np.testing.assert_allclose(out.reshape(70, 40), ref, rtol=1e-13, atol=1e-15)
    v = np.zeros(70 * 70)
    capi.check(capi.lib().mra_variance_covariance_xyz(capi.C.byref(c), *p, keep.d(lon1), keep.d(lat1),
                                                       *[keep.d(a) for a in x1], 70, capi._d(v)))
    np.testing.assert_allclose(v.reshape(70, 70), oracle.variance_covariance(cov, prm, lon1, lat1), rtol=1e-13,
                               atol=1e-15)


def test_loglikelihood_literal_pi(example):
    tree, cov, prm = example
    h = capi.Handle(tree, with_pred=False)
    try:
        h.set_covariance(cov)
        h.set_params(**prm)
        D, ell = h.loglikelihood()
        D2, ell2 = h.loglikelihood(5000)
    finally:
        h.close()
    n = len(tree["obs_lon"])
    assert D == D2
    assert ell == -0.5 * (D + n * math.log(2 * PI_LITERAL))
    assert ell2 == -0.5 * (D + 5000 * math.log(2 * PI_LITERAL))
