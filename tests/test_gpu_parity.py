"""GPU parity: the HIP FP64 path (through the C ABI) against the CPU oracle.

Tolerances: the pure-FP64 GPU path must match the FP64 reference restatement
to 1e-10 relative in D = -2 loglik - n log(2 pi); predictive means and
variances to 1e-9 relative (absolute floor 1e-10).
"""
import numpy as np
import pytest

from nonstationary_mra_amd import capi, synth
from oracle import oracle
from tests import cases

pytestmark = pytest.mark.gpu

TOL_D = 1e-10
TOL_PRED = 1e-9


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def run_gpu(tree, cov, prm, predict=True, factored=1, gram=1, sum_msgs=1):
    h = capi.Handle(tree, with_pred=predict)
    try:
        h.set_covariance(cov)
        h.set_params(**prm)
        h.set_option(capi.OPT_FACTORED_POSTERIOR, factored)
        h.set_option(capi.OPT_GRAM_KERNEL, gram)
        h.set_option(capi.OPT_SUM_MESSAGES, sum_msgs)
        if predict and h.n_pred:
            return h.predict()
        return h.likelihood(), None, None
    finally:
        h.close()


def check(tree, cov, prm, predict=True, factored=1, gram=1, sum_msgs=1):
    D, mean, var = run_gpu(tree, cov, prm, predict, factored, gram, sum_msgs)
    Do, mo, vo = oracle.evaluate(tree, cov, prm, predict=predict)
    print(f"D gpu {D!r} oracle {Do!r} rel {rel(D, Do):.3e}", flush=True)
    assert rel(D, Do) < TOL_D
    if mo is not None:
        np.testing.assert_allclose(mean, mo, rtol=TOL_PRED, atol=1e-10)
        np.testing.assert_allclose(var, vo, rtol=TOL_PRED, atol=1e-10)


def test_example_dataset(example):
    """The reference's own example (J=2, M=5, r=4, user knots, parent file,
    chordal Wendland fields, finest regions up to 268 observations)."""
    tree, cov, prm = example
    check(tree, cov, prm)


@pytest.mark.parametrize("M,J,r,metric,stationary,clustered,seed", [
    (2, 4, 9, "planar", False, False, 1),
    (3, 4, 16, "planar", False, True, 2),
    (3, 2, 4, "planar", True, False, 3),
    (4, 2, 4, "chordal", False, False, 4),
    (3, 4, 16, "great_circle", False, True, 5),
    (4, 4, 30, "planar", False, True, 6),
    (3, 4, 64, "planar", False, False, 7),
])
def test_synthetic(M, J, r, metric, stationary, clustered, seed):
    tree, cov, prm = cases.synthetic_case(n=600, npred=80, M=M, J=J, r=r, seed=seed, metric=metric,
                                          stationary=stationary, clustered=clustered)
    check(tree, cov, prm)


def test_levels_match_oracle():
    tree, cov, prm = cases.synthetic_case(n=800, npred=0, M=4, J=4, r=16, seed=21, clustered=True)
    h = capi.Handle(tree, with_pred=False)
    h.set_covariance(cov)
    h.set_params(**prm)
    lev = h.likelihood_levels()
    h.close()
    lo = oracle.evaluate_levels(tree, cov, prm)
    for a, b in zip(lev, lo):
        assert abs(a - b) <= 1e-10 * max(abs(b), abs(lo).max() * 1e-3)


def test_config0_small_fp64():
    """n = 4,000 uniform points, exact dense GP draw (log sigma^2 = 0.5x - 0.3y,
    log beta = log 0.05 + 0.8x, nugget 0.05), M=4, J=4, r=16, 1,000
    prediction points."""
    rng = np.random.default_rng(2024)
    n, npred = 4000, 1000
    x, y = synth.uniform_points(n, rng)
    cov = synth.linear_trend_cov()
    val = synth.dense_gp_draw(cov, x, y, rng)
    px, py = synth.uniform_points(npred, rng)
    tree = capi.partition_program(x, y, val, px, py, M=4, J=4, r=16, bounds=synth.UNIT_BOUNDS)
    check(tree, cov, dict(synth.UNIT_PARAMS))


def test_clustered_r64():
    """Clustered points (uneven finest counts, some > 64), r = 64, J = 4."""
    rng = np.random.default_rng(77)
    n, npred = 12000, 600
    x, y = synth.clustered_points(n, rng)
    px, py = synth.uniform_points(npred, rng)
    tree = capi.partition_program(x, y, rng.standard_normal(n), px, py, M=4, J=4, r=64,
                                  bounds=synth.UNIT_BOUNDS)
    counts = np.diff(tree["obs_offsets"].astype(np.int64))
    print("finest counts max", counts.max(), "empty", (counts == 0).sum(), flush=True)
    assert counts.max() > 64
    check(tree, synth.linear_trend_cov(), dict(synth.UNIT_PARAMS))


def test_covariance_entry_points():
    tree, cov, prm = cases.synthetic_case(n=70, npred=0, M=2, J=4, r=4, seed=31, metric="chordal")
    lon, lat = tree["obs_lon"], tree["obs_lat"]
    C = capi.cross_covariance(cov, prm, lon[:40], lat[:40], lon[40:], lat[40:])
    Co = oracle.cross_covariance(cov, prm, lon[:40], lat[:40], lon[40:], lat[40:])
    np.testing.assert_allclose(C, Co, rtol=1e-13)
    V = capi.variance_covariance(cov, prm, lon, lat)
    Vo = oracle.variance_covariance(cov, prm, lon, lat)
    np.testing.assert_allclose(V, Vo, rtol=1e-13)


def test_repeated_evaluations_identical():
    """Device buffers are reused across evaluations; results are bitwise
    repeatable (deterministic reductions)."""
    tree, cov, prm = cases.synthetic_case(n=500, npred=50, M=3, J=4, r=16, seed=41, clustered=True)
    h = capi.Handle(tree)
    h.set_covariance(cov)
    h.set_params(**prm)
    a = h.predict()
    b = h.predict()
    c = h.likelihood()
    h.close()
    assert a[0] == b[0] == c
    np.testing.assert_array_equal(a[1], b[1])
    np.testing.assert_array_equal(a[2], b[2])


def test_not_positive_definite_reports_error():
    """A negative nugget scale makes every variance-covariance diagonal
    sigma^2 - 2 < 0: the first pivot is negative and the library must report
    MRA_E_NOT_PD, not return a number."""
    rng = np.random.default_rng(5)
    x, y = rng.random(200), rng.random(200)
    cov = synth.linear_trend_cov(nugget=1.0)
    prm = dict(synth.UNIT_PARAMS, nugget_scale=-2.0)
    tree = capi.partition_program(x, y, rng.standard_normal(200), None, None, M=2, J=4, r=4,
                                  bounds=synth.UNIT_BOUNDS)
    h = capi.Handle(tree, with_pred=False)
    h.set_covariance(cov)
    h.set_params(**prm)
    with pytest.raises(capi.MraError) as e:
        h.likelihood()
    assert e.value.code == capi.MRA_E_NOT_PD
    h.close()


def test_prior_sampler_matches_dense_covariance():
    """Draws from mra_sample_prior have covariance Sigma_MRA (dense oracle):
    every entry of the empirical covariance of N draws lies within 5.5
    standard errors sqrt((S_ii S_jj + S_ij^2)/N) of Sigma_MRA; the same seed
    gives the same draw."""
    from oracle import dense
    tree, cov, prm = cases.synthetic_case(n=24, npred=0, M=3, J=2, r=4, seed=51)
    h = capi.Handle(tree, with_pred=False)
    h.set_covariance(cov)
    h.set_params(**prm)
    N = 1500
    draws = np.stack([h.sample_prior(seed=1000 + s, replace=False) for s in range(N)])
    again = h.sample_prior(seed=1000, replace=False)
    h.close()
    np.testing.assert_array_equal(again, draws[0])
    sig, n, _ = dense.dense_sigma(tree, cov, prm)
    emp = draws.T @ draws / N
    d = np.diag(sig)
    se = np.sqrt((np.outer(d, d) + sig ** 2) / N)
    z = np.abs(emp - sig) / se
    print("max |z|", z.max(), flush=True)
    assert z.max() < 5.5


@pytest.mark.parametrize("factored", [0, 1])
@pytest.mark.parametrize("case", ["example", "clustered64", "j2"])
def test_dense_and_factored_posterior(case, factored, example):
    """The reference's dense posterior update and the factored form of the
    finest-but-one level both match the oracle (the clustered case has
    regions with more than 64 observations below them)."""
    if case == "example":
        tree, cov, prm = example
    elif case == "clustered64":
        rng = np.random.default_rng(78)
        x, y = synth.clustered_points(9000, rng, n_clusters=3, sigma=0.03)
        px, py = synth.uniform_points(300, rng)
        tree = capi.partition_program(x, y, rng.standard_normal(9000), px, py, M=4, J=4, r=64,
                                      bounds=synth.UNIT_BOUNDS)
        cov, prm = synth.linear_trend_cov(), dict(synth.UNIT_PARAMS)
    else:
        tree, cov, prm = cases.synthetic_case(n=900, npred=60, M=5, J=2, r=9, seed=8, clustered=True)
    h = capi.Handle(tree)
    h.set_covariance(cov)
    h.set_params(**prm)
    h.set_option(capi.OPT_FACTORED_POSTERIOR, factored)
    assert h.get_option(capi.OPT_FACTORED_POSTERIOR) == factored
    h.close()
    check(tree, cov, prm, factored=factored)


@pytest.mark.parametrize("factored", [1, 0])
@pytest.mark.parametrize("gram", [1, 0])
def test_message_block_kernel(gram, factored):
    """Posterior message blocks of 64-knot regions from the one-wave-per-block
    kernel (gram=1) or inside the posterior kernel (gram=0): both match the
    oracle, with dense messages at three coarse levels and, in the factored
    form, a rows-mode level mixing factored and dense children."""
    rng = np.random.default_rng(91)
    x, y = synth.clustered_points(12000, rng, n_clusters=4, sigma=0.02)
    px, py = synth.uniform_points(200, rng)
    tree = capi.partition_program(x, y, rng.standard_normal(12000), px, py, M=5, J=4, r=64,
                                  bounds=synth.UNIT_BOUNDS)
    cov, prm = synth.linear_trend_cov(), dict(synth.UNIT_PARAMS)
    h = capi.Handle(tree)
    h.set_option(capi.OPT_GRAM_KERNEL, gram)
    assert h.get_option(capi.OPT_GRAM_KERNEL) == gram
    h.close()
    check(tree, cov, prm, factored=factored, gram=gram)


@pytest.mark.parametrize("factored", [1, 0])
@pytest.mark.parametrize("sum_msgs", [1, 0])
def test_summed_child_messages(sum_msgs, factored):
    """Message blocks of the rows-reading level stored summed per parent
    (sum_msgs=1: the parent reads one block set) or per region (0), in the
    factored form (summed at the finest-but-two level) and the dense form
    (summed at the finest-but-one level): both match the oracle, and the
    option reports the form in effect."""
    rng = np.random.default_rng(92)
    x, y = synth.clustered_points(12000, rng, n_clusters=4, sigma=0.02)
    px, py = synth.uniform_points(200, rng)
    tree = capi.partition_program(x, y, rng.standard_normal(12000), px, py, M=5, J=4, r=64,
                                  bounds=synth.UNIT_BOUNDS)
    cov, prm = synth.linear_trend_cov(), dict(synth.UNIT_PARAMS)
    h = capi.Handle(tree)
    h.set_option(capi.OPT_SUM_MESSAGES, sum_msgs)
    assert h.get_option(capi.OPT_SUM_MESSAGES) == sum_msgs
    h.set_option(capi.OPT_GRAM_KERNEL, 0)
    assert h.get_option(capi.OPT_SUM_MESSAGES) == 0
    h.close()
    check(tree, cov, prm, factored=factored, sum_msgs=sum_msgs)


@pytest.mark.parametrize("wave", [1, 0])
@pytest.mark.parametrize("case", ["uniform64", "small9"])
def test_small_region_local_kernel(case, wave):
    """Finest regions with at most 16 observations one wave each (wave=1) or
    one work group each (wave=0): D and predictions match the oracle in FP64,
    the mixed path stays within its tolerance; regions of 1 to 16
    observations, plus larger ones in the other buckets."""
    if case == "uniform64":
        rng = np.random.default_rng(97)
        x, y = synth.clustered_points(6000, rng, n_clusters=2, sigma=0.05)
        px, py = synth.uniform_points(120, rng)
        tree = capi.partition_program(x, y, rng.standard_normal(6000), px, py, M=5, J=4, r=64,
                                      bounds=synth.UNIT_BOUNDS)
        cov, prm = synth.linear_trend_cov(), dict(synth.UNIT_PARAMS)
    else:
        tree, cov, prm = cases.synthetic_case(n=1500, npred=80, M=4, J=4, r=9, seed=12, clustered=True)
    counts = np.diff(tree["obs_offsets"].astype(np.int64))
    assert ((counts >= 1) & (counts <= 16)).sum() > 20 and counts.max() > 16
    h = capi.Handle(tree)
    try:
        h.set_covariance(cov)
        h.set_params(**prm)
        h.set_option(capi.OPT_WAVE_LOCAL16, wave)
        assert h.get_option(capi.OPT_WAVE_LOCAL16) == wave
        D, mean, var = h.predict()
        h.set_precision(capi.PRECISION_MIXED)
        Dm = h.likelihood()
    finally:
        h.close()
    Do, mo, vo = oracle.evaluate(tree, cov, prm)
    print(f"{case} wave={wave} D gpu {D!r} oracle {Do!r} rel {rel(D, Do):.3e}", flush=True)
    assert rel(D, Do) < TOL_D
    np.testing.assert_allclose(mean, mo, rtol=TOL_PRED, atol=1e-10)
    np.testing.assert_allclose(var, vo, rtol=TOL_PRED, atol=1e-10)
    assert abs(Dm - D) / (2 * len(tree["obs_lon"])) < 1e-6


@pytest.mark.parametrize("cross16", [1, 0])
def test_small_segment_cross_tiles(cross16):
    """Finest-level segments (observations below one finest-but-one region)
    conditioned with a last tile of 16, 32 or 64 columns, the narrowest that
    holds the rest (cross16=1), or in tiles of one width (0): D and
    predictions match the oracle in FP64, mixed within tolerance."""
    rng = np.random.default_rng(16)
    x, y = synth.clustered_points(1500, rng, n_clusters=3, sigma=0.1)
    px, py = synth.uniform_points(60, rng)
    tree = capi.partition_program(x, y, rng.standard_normal(1500), px, py, M=5, J=4, r=16,
                                  bounds=synth.UNIT_BOUNDS)
    cov, prm = synth.linear_trend_cov(), dict(synth.UNIT_PARAMS)
    seg = np.diff(tree["obs_offsets"].astype(np.int64)).reshape(-1, 4).sum(1)
    assert ((seg >= 1) & (seg <= 16)).sum() > 10 and (seg > 32).sum() > 5
    h = capi.Handle(tree)
    try:
        h.set_covariance(cov)
        h.set_params(**prm)
        h.set_option(capi.OPT_CROSS16, cross16)
        assert h.get_option(capi.OPT_CROSS16) == cross16
        D, mean, var = h.predict()
        h.set_precision(capi.PRECISION_MIXED)
        Dm = h.likelihood()
    finally:
        h.close()
    Do, mo, vo = oracle.evaluate(tree, cov, prm)
    print(f"cross16={cross16} D gpu {D!r} oracle {Do!r} rel {rel(D, Do):.3e}", flush=True)
    assert rel(D, Do) < TOL_D
    np.testing.assert_allclose(mean, mo, rtol=TOL_PRED, atol=1e-10)
    np.testing.assert_allclose(var, vo, rtol=TOL_PRED, atol=1e-10)
    assert abs(Dm - D) / (2 * len(tree["obs_lon"])) < 1e-6


@pytest.mark.parametrize("r", [16, 32])
def test_large_region_local_kernel(r):
    """Finest regions of 65 to 256 observations (padded workspace, blocked
    Cholesky and inverse of up to 16 blocks) and above 256 (generic path):
    D and predictions match the oracle in FP64, mixed within tolerance."""
    rng = np.random.default_rng(40 + r)
    x, y = synth.clustered_points(3600, rng, n_clusters=2, sigma=0.08)
    px, py = synth.uniform_points(60, rng)
    tree = capi.partition_program(x, y, rng.standard_normal(3600), px, py, M=3, J=4, r=r,
                                  bounds=synth.UNIT_BOUNDS)
    cov, prm = synth.linear_trend_cov(log_range=(np.log(0.2), 0.8, 0.0)), dict(synth.UNIT_PARAMS)
    counts = np.diff(tree["obs_offsets"].astype(np.int64))
    print("counts", sorted(counts.tolist()), flush=True)
    assert ((counts > 128) & (counts <= 256)).sum() >= 4
    h = capi.Handle(tree)
    try:
        h.set_covariance(cov)
        h.set_params(**prm)
        D, mean, var = h.predict()
        h.set_precision(capi.PRECISION_MIXED)
        assert h.get_option(capi.OPT_MIXED_FP64_ABOVE) == 128
        Dm = h.likelihood()
        h.set_option(capi.OPT_MIXED_FP64_ABOVE, 256)
        Dm256 = h.likelihood()
        h.set_option(capi.OPT_MIXED_FP64_ABOVE, 64)
        Dm64 = h.likelihood()
        for bad in (63, 257):
            with pytest.raises(capi.MraError):
                h.set_option(capi.OPT_MIXED_FP64_ABOVE, bad)
        assert h.get_option(capi.OPT_MIXED_FP64_ABOVE) == 64
    finally:
        h.close()
    Do, mo, vo = oracle.evaluate(tree, cov, prm)
    print(f"r={r} D gpu {D!r} oracle {Do!r} rel {rel(D, Do):.3e}", flush=True)
    assert rel(D, Do) < TOL_D
    np.testing.assert_allclose(mean, mo, rtol=TOL_PRED, atol=1e-10)
    np.testing.assert_allclose(var, vo, rtol=TOL_PRED, atol=1e-10)
    # mixed: regions of strongly correlated observations (range 0.2) above
    # the threshold (default 128) are conditioned in FP64; with FP32 for the
    # 700-900 observation regions |dl|/n reached 1e-5
    n2 = 2 * len(tree["obs_lon"])
    dl, dl256, dl64 = abs(Dm - D) / n2, abs(Dm256 - D) / n2, abs(Dm64 - D) / n2
    print(f"mixed |dl|/n {dl:.3e} (threshold 256: {dl256:.3e}, 64: {dl64:.3e})", flush=True)
    assert dl <= 1e-6 and dl64 <= 1e-6 and dl256 < 1e-5


@pytest.mark.parametrize("gram", [1, 0])
def test_binary_tree_padded_knots(gram):
    """J = 2 with r = 49 knots padded to 64-wide blocks: the one-wave
    posterior kernels, the factored level, and finest regions in the small,
    middle and large buckets, against the oracle."""
    rng = np.random.default_rng(88)
    x, y = synth.clustered_points(5000, rng, n_clusters=3, sigma=0.04)
    px, py = synth.uniform_points(100, rng)
    tree = capi.partition_program(x, y, rng.standard_normal(5000), px, py, M=6, J=2, r=49,
                                  bounds=synth.UNIT_BOUNDS)
    counts = np.diff(tree["obs_offsets"].astype(np.int64))
    print("counts", sorted(counts.tolist()), flush=True)
    cov, prm = synth.linear_trend_cov(log_range=(np.log(0.1), 0.8, 0.0)), dict(synth.UNIT_PARAMS)
    check(tree, cov, prm, gram=gram)


@pytest.mark.parametrize("r", [64, 50])
def test_deepest_coarse_level(r):
    """J = 2, M = 10: the coarse prior down to level 8, the depth of the
    n = 2e6 benchmark tree (eight ancestor levels, K_R = V - G^T G over
    K = 8 * 64), with r = 64 and r = 50 padded to 64-wide blocks, on a field
    drawn from the model: FP64 against the oracle, mixed precision within
    |dloglik|/n <= 1e-6 of FP64 (white-noise data, far from the model, give
    a larger mixed error, about 4e-7 |D| here: the documented limit of the
    mixed contract)."""
    rng = np.random.default_rng(90 + r)
    n = 6000
    x, y = synth.clustered_points(n, rng, n_clusters=3, sigma=0.05)
    px, py = synth.uniform_points(100, rng)
    tree = capi.partition_program(x, y, np.zeros(n), px, py, M=10, J=2, r=r, bounds=synth.UNIT_BOUNDS)
    cov, prm = synth.linear_trend_cov(log_range=(np.log(0.1), 0.8, 0.0)), dict(synth.UNIT_PARAMS)
    h = capi.Handle(tree, with_pred=False)
    h.set_covariance(cov)
    h.set_params(**prm)
    tree["obs_val"] = h.sample_prior(seed=11, replace=False)
    h.close()
    check(tree, cov, prm)
    h = capi.Handle(tree, with_pred=False)
    h.set_covariance(cov)
    h.set_params(**prm)
    D64 = h.likelihood()
    h.set_precision(capi.PRECISION_MIXED)
    Dm = h.likelihood()
    h.close()
    print(f"mixed {Dm!r} fp64 {D64!r} |dloglik|/n {abs(Dm - D64) / (2 * n):.3e}", flush=True)
    assert abs(Dm - D64) / (2 * n) <= 1e-6


def test_concurrent_streams():
    """Independent launches on forked streams (default) against one stream:
    the same D to the last bit, eager and graph replay, FP64 and mixed, and
    the oracle's D; a tree whose finest-but-one level has both dense and
    factored regions and whose finest level fills every length bucket."""
    rng = np.random.default_rng(89)
    x, y = synth.clustered_points(8000, rng, n_clusters=3, sigma=0.03)
    tree = capi.partition_program(x, y, rng.standard_normal(8000), None, None, M=9, J=2, r=49,
                                  bounds=synth.UNIT_BOUNDS)
    counts = np.diff(tree["obs_offsets"].astype(np.int64))
    assert all(np.histogram(counts, [1, 17, 33, 65, 129, 257, 10**9])[0] > 0)
    rows = counts.reshape(-1, 2).sum(1)
    assert (rows > 128).sum() > 0 and ((rows > 0) & (rows <= 128)).sum() > 0
    cov, prm = synth.linear_trend_cov(log_range=(np.log(0.1), 0.8, 0.0)), dict(synth.UNIT_PARAMS)
    out = {}
    for conc in (1, 0):
        h = capi.Handle(tree, with_pred=False)
        try:
            h.set_covariance(cov)
            h.set_params(**prm)
            h.set_option(capi.OPT_CONCURRENT, conc)
            assert h.get_option(capi.OPT_CONCURRENT) == conc
            for prec in (capi.PRECISION_FP64, capi.PRECISION_MIXED):
                h.set_precision(prec)
                for graph in (False, True):
                    h.use_graph(graph)
                    out[(conc, prec, graph)] = [h.likelihood() for _ in range(2)]
        finally:
            h.close()
    print(out, flush=True)
    for prec in (capi.PRECISION_FP64, capi.PRECISION_MIXED):
        ref = out[(0, prec, False)][0]
        for k, v in out.items():
            if k[1] == prec:
                assert v == [ref, ref], k
    Do = oracle.evaluate(tree, cov, prm, predict=False)[0]
    assert rel(out[(1, capi.PRECISION_FP64, True)][0], Do) < TOL_D


@pytest.mark.parametrize("case", [(1500, 3000, 4, 71), (2000, 400, 5, 72)])
def test_prediction_group_tiles(case):
    """Prediction tiles of 16 points of a group (the finest regions below one
    grandparent): groups of several hundred points, so that segments (the
    points below one parent) span tile boundaries and tiles hold segment
    boundaries, next to groups of a few points; means and variances match
    the oracle in FP64, the mixed path stays within its tolerance."""
    n, npred, M, seed = case
    rng = np.random.default_rng(seed)
    x, y = synth.clustered_points(n, rng, n_clusters=3, sigma=0.08)
    px, py = synth.clustered_points(npred, rng, n_clusters=2, sigma=0.1)
    tree = capi.partition_program(x, y, rng.standard_normal(n), px, py, M=M, J=4, r=16,
                                  bounds=synth.UNIT_BOUNDS)
    cov, prm = synth.linear_trend_cov(), dict(synth.UNIT_PARAMS)
    pc = np.diff(tree["pred_offsets"].astype(np.int64))
    print("largest groups", np.sort(pc.reshape(-1, 16).sum(1))[-4:], flush=True)
    check(tree, cov, prm)
    h = capi.Handle(tree)
    try:
        h.set_covariance(cov)
        h.set_params(**prm)
        h.set_precision(capi.PRECISION_MIXED)
        Dm, mm, vm = h.predict()
    finally:
        h.close()
    Do, mo, vo = oracle.evaluate(tree, cov, prm, predict=True)
    assert abs(Dm - Do) / (2 * n) < TOL_MIXED
    assert np.abs(mm - mo).max() <= 1e-3 * np.abs(mo).max()
    assert np.max(np.abs(vm - vo) / np.abs(vo)) < 1e-3


TOL_MIXED = 1e-6   # |delta loglik| / n for the mixed-precision path


@pytest.mark.parametrize("case", ["clustered64", "example"])
def test_mixed_precision_loglik(case, example):
    """Mixed precision (FP32 covariance evaluation, FP32 MFMA finest
    conditioning, FP32 storage of G_F/B_F, and for likelihood evaluations at
    RP = 64 the row Gram products X_j^T X_k, X_l^T X_j of the posterior on
    FP32 MFMA; FP64 Sigma_F assembly, Cholesky, log-determinants, -P^T P and
    everything else above the finest level) stays within |delta loglik| / n
    <= 1e-6 of the FP64 oracle, and predictions within 1e-4 relative."""
    if case == "example":
        tree, cov, prm = example
    else:
        rng = np.random.default_rng(79)
        x, y = synth.clustered_points(12000, rng)
        px, py = synth.uniform_points(400, rng)
        tree = capi.partition_program(x, y, np.zeros(12000), px, py, M=4, J=4, r=64,
                                      bounds=synth.UNIT_BOUNDS)
        cov, prm = synth.linear_trend_cov(), dict(synth.UNIT_PARAMS)
        h = capi.Handle(tree, with_pred=False)
        h.set_covariance(cov)
        h.set_params(**prm)
        tree["obs_val"] = h.sample_prior(seed=7, replace=False)
        h.close()
    n = len(tree["obs_lon"])
    Do, mo, vo = oracle.evaluate(tree, cov, prm, predict=True)
    h = capi.Handle(tree)
    try:
        h.set_covariance(cov)
        h.set_params(**prm)
        h.set_precision(capi.PRECISION_MIXED)
        D, mean, var = h.predict()
        Dl = h.likelihood()
    finally:
        h.close()
    err = abs(D - Do) / (2 * n)
    print(f"mixed D {D!r} oracle {Do!r} |dl|/n {err:.3e}", flush=True)
    assert err <= TOL_MIXED
    # likelihood() takes the FP32 row products, predict() the FP64 ones: both
    # D within the mixed tolerance of the oracle
    assert abs(Dl - Do) / (2 * n) <= TOL_MIXED
    np.testing.assert_allclose(mean, mo, rtol=1e-4, atol=1e-4 * np.abs(mo).max())
    np.testing.assert_allclose(var, vo, rtol=1e-4, atol=1e-6)


def test_graph_replay_matches_eager():
    """A captured evaluation replayed many times, with eager evaluations,
    parameter changes and precision changes in between, gives bitwise the
    eager results (same kernels, deterministic reductions)."""
    tree, cov, prm = cases.synthetic_case(n=3000, npred=200, M=5, J=4, r=16, seed=61, clustered=True)
    prm2 = dict(prm, sill_scale=prm["sill_scale"] * 1.3, range_scale=prm["range_scale"] * 0.7,
                nugget_scale=prm["nugget_scale"] * 1.5)
    eager = capi.Handle(tree)
    graph = capi.Handle(tree)
    try:
        for h in (eager, graph):
            h.set_covariance(cov)
        graph.use_graph(True)
        for rep in range(3):
            for p in (prm, prm2):
                for precision in (capi.PRECISION_FP64, capi.PRECISION_MIXED):
                    for h in (eager, graph):
                        h.set_params(**p)
                        h.set_precision(precision)
                    a = eager.likelihood()
                    b = graph.likelihood()
                    assert a == b, (rep, precision, a, b)
                    pa = eager.predict()
                    pb = graph.predict()
                    assert pa[0] == pb[0]
                    np.testing.assert_array_equal(pa[1], pb[1])
                    np.testing.assert_array_equal(pa[2], pb[2])
            assert graph.timings()[capi.T_GRAPH] > 0
        Do, _, _ = oracle.evaluate(tree, cov, prm2, predict=False)
        graph.set_params(**prm2)
        graph.set_precision(capi.PRECISION_FP64)
        assert rel(graph.likelihood(), Do) < TOL_D
    finally:
        eager.close()
        graph.close()


def test_xcd_item_order():
    """Regions handed out in XCD-contiguous ranges (MRA_OPT_XCD_MAP, the
    default) or in launch order: the same D and predictions to the last bit,
    eager and graph, FP64 and mixed, and the oracle's D."""
    rng = np.random.default_rng(90)
    x, y = synth.clustered_points(6000, rng, n_clusters=3, sigma=0.03)
    px, py = synth.uniform_points(300, rng)
    tree = capi.partition_program(x, y, rng.standard_normal(6000), px, py, M=5, J=4, r=64,
                                  bounds=synth.UNIT_BOUNDS)
    cov, prm = synth.linear_trend_cov(log_range=(np.log(0.1), 0.8, 0.0)), dict(synth.UNIT_PARAMS)
    h = capi.Handle(tree)
    out = {}
    try:
        h.set_covariance(cov)
        h.set_params(**prm)
        for prec in (capi.PRECISION_FP64, capi.PRECISION_MIXED):
            h.set_precision(prec)
            for xcd in (1, 0):
                h.set_option(capi.OPT_XCD_MAP, xcd)
                assert h.get_option(capi.OPT_XCD_MAP) == xcd
                for graph in (0, 1):
                    h.use_graph(graph)
                    out[(prec, xcd, graph)] = h.predict()
    finally:
        h.close()
    for prec in (capi.PRECISION_FP64, capi.PRECISION_MIXED):
        ref = out[(prec, 1, 0)]
        for k, v in out.items():
            if k[0] == prec:
                assert v[0] == ref[0]
                np.testing.assert_array_equal(v[1], ref[1])
                np.testing.assert_array_equal(v[2], ref[2])
    Do, _, _ = oracle.evaluate(tree, cov, prm, predict=False)
    assert rel(out[(capi.PRECISION_FP64, 1, 0)][0], Do) < TOL_D


@pytest.mark.parametrize("case", ["buckets", "large"])
def test_persistent_finest_launches(case):
    """Finest-level kernels on a resident grid walking the length-bucketed
    work lists (MRA_OPT_PERSISTENT = 1) or one work group per item (0): the
    same D and predictions to the last bit, eager and graph, FP64 and mixed
    (which also runs the FP64 large-region path), and the oracle's D.  The
    trees cover every bucket: <= 16 (one wave each), 17-32, 33-64, 65-128,
    129-256 and above 256 observations."""
    rng = np.random.default_rng(91 if case == "buckets" else 56)
    if case == "buckets":
        x, y = synth.clustered_points(12000, rng, n_clusters=4, sigma=0.02)
        ux, uy = synth.uniform_points(2000, rng)
        x, y = np.concatenate([x, ux]), np.concatenate([y, uy])
        px, py = synth.uniform_points(300, rng)
        M, r = 6, 64
    else:
        x, y = synth.clustered_points(3600, rng, n_clusters=2, sigma=0.08)
        px, py = synth.uniform_points(60, rng)
        M, r = 3, 32
    tree = capi.partition_program(x, y, rng.standard_normal(len(x)), px, py, M=M, J=4, r=r,
                                  bounds=synth.UNIT_BOUNDS)
    counts = np.diff(tree["obs_offsets"].astype(np.int64))
    sizes = [int(((counts > lo) & (counts <= hi)).sum())
             for lo, hi in ((0, 16), (16, 32), (32, 64), (64, 128), (128, 256), (256, 1 << 30))]
    print(case, "bucket sizes", sizes, flush=True)
    assert all(sizes) if case == "buckets" else sizes[3] and sizes[4] and sizes[5]
    cov, prm = synth.linear_trend_cov(log_range=(np.log(0.2), 0.8, 0.0)), dict(synth.UNIT_PARAMS)
    h = capi.Handle(tree)
    out = {}
    try:
        h.set_covariance(cov)
        h.set_params(**prm)
        # the work-group form of the mixed conditioning in both (persistent
        # launches have no one-wave form)
        h.set_option(capi.OPT_FINEST_WAVE, 0)
        for prec in (capi.PRECISION_FP64, capi.PRECISION_MIXED):
            h.set_precision(prec)
            for pers in (0, 1):
                h.set_option(capi.OPT_PERSISTENT, pers)
                assert h.get_option(capi.OPT_PERSISTENT) == pers
                for graph in (0, 1):
                    h.use_graph(graph)
                    out[(prec, pers, graph)] = h.predict()
                    out[(prec, pers, graph, "l")] = h.likelihood()
    finally:
        h.close()
    for prec in (capi.PRECISION_FP64, capi.PRECISION_MIXED):
        ref = out[(prec, 0, 0)]
        for k, v in out.items():
            if k[0] != prec:
                continue
            if len(k) == 4:
                assert v == out[(prec, 0, 0, "l")]
                continue
            assert v[0] == ref[0]
            np.testing.assert_array_equal(v[1], ref[1])
            np.testing.assert_array_equal(v[2], ref[2])
    Do, _, _ = oracle.evaluate(tree, cov, prm, predict=False)
    assert rel(out[(capi.PRECISION_FP64, 1, 0)][0], Do) < TOL_D


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["buckets", "large", "r16"])
def test_local_left_looking(case):
    """Finest local step of the regions above 64 observations as the
    left-looking blocked Cholesky (MRA_OPT_LOCAL_LL = 2: every such region;
    1, the default: all but the 65-128-observation regions with more than 256
    ancestor knot columns) against the workspace and unblocked forms (0).  The same arithmetic in another
    order: FP64 D within 1e-12 relative and predictions within TOL_PRED, the
    mixed-precision D within 1e-10 (both forms run this step in FP64 from the
    same G); graph replay equal to eager to the last bit; the oracle's D."""
    rng = np.random.default_rng({"buckets": 91, "large": 56, "r16": 17}[case])
    if case == "buckets":
        x, y = synth.clustered_points(12000, rng, n_clusters=4, sigma=0.02)
        ux, uy = synth.uniform_points(2000, rng)
        x, y = np.concatenate([x, ux]), np.concatenate([y, uy])
        px, py = synth.uniform_points(300, rng)
        M, r = 6, 64
    else:
        x, y = synth.clustered_points(3600 if case == "large" else 3000, rng, n_clusters=2, sigma=0.08)
        px, py = synth.uniform_points(60, rng)
        M, r = 3, 32 if case == "large" else 16
    tree = capi.partition_program(x, y, rng.standard_normal(len(x)), px, py, M=M, J=4, r=r,
                                  bounds=synth.UNIT_BOUNDS)
    counts = np.diff(tree["obs_offsets"].astype(np.int64))
    sizes = [int(((counts > lo) & (counts <= hi)).sum()) for lo, hi in ((64, 128), (128, 256), (256, 1 << 30))]
    print(case, "regions in (64,128], (128,256], >256:", sizes, "max", int(counts.max()), flush=True)
    assert sizes[0] and sizes[1] and sizes[2]
    cov, prm = synth.linear_trend_cov(log_range=(np.log(0.2), 0.8, 0.0)), dict(synth.UNIT_PARAMS)
    h = capi.Handle(tree)
    out = {}
    try:
        h.set_covariance(cov)
        h.set_params(**prm)
        assert h.get_option(capi.OPT_LOCAL_LL) == 1
        for prec in (capi.PRECISION_FP64, capi.PRECISION_MIXED):
            h.set_precision(prec)
            for ll in (2, 1, 0):
                h.set_option(capi.OPT_LOCAL_LL, ll)
                assert h.get_option(capi.OPT_LOCAL_LL) == ll
                for graph in (0, 1):
                    h.use_graph(graph)
                    out[(prec, ll, graph)] = h.predict()
                    out[(prec, ll, graph, "l")] = h.likelihood()
        h.set_option(capi.OPT_LOCAL_LL, 1)
    finally:
        h.close()
    for prec in (capi.PRECISION_FP64, capi.PRECISION_MIXED):
        for ll in (0, 1, 2):
            # graph replay equal to eager; in FP64 the prediction's D is the
            # likelihood's (mixed likelihoods use FP32 Gram products, predictions not)
            ref = out[(prec, ll, 0)]
            assert out[(prec, ll, 1, "l")] == out[(prec, ll, 0, "l")]
            if prec == capi.PRECISION_FP64:
                assert out[(prec, ll, 0, "l")] == ref[0]
            assert out[(prec, ll, 1)][0] == ref[0]
            np.testing.assert_array_equal(out[(prec, ll, 1)][1], ref[1])
            np.testing.assert_array_equal(out[(prec, ll, 1)][2], ref[2])
        for ll in (1, 2):
            a, b = out[(prec, ll, 0)], out[(prec, 0, 0)]
            tol = 1e-12 if prec == capi.PRECISION_FP64 else 1e-10
            print(case, prec, ll, "D", a[0], b[0], "rel", rel(a[0], b[0]), flush=True)
            assert rel(a[0], b[0]) < tol
            la, lb = out[(prec, ll, 0, "l")], out[(prec, 0, 0, "l")]
            print(case, prec, ll, "likelihood D", la, lb, "rel", rel(la, lb), flush=True)
            assert rel(la, lb) < tol
            if prec == capi.PRECISION_FP64:
                for q in (1, 2):
                    err = np.max(np.abs(a[q] - b[q])) / np.max(np.abs(b[q]))
                    print(case, ll, "prediction output", q, "max rel diff", err, flush=True)
                    assert err < TOL_PRED
    Do, _, _ = oracle.evaluate(tree, cov, prm, predict=False)
    assert rel(out[(capi.PRECISION_FP64, 1, 0)][0], Do) < TOL_D


@pytest.mark.gpu
@pytest.mark.parametrize("r", [16, 32, 64])
@pytest.mark.parametrize("sizes", [(65, 96, 129, 257), (97, 128, 160, 320), (100, 200, 300, 415)])
def test_local_left_looking_region_sizes(sizes, r):
    """The left-looking local step at chosen region sizes (one finest region
    per quadrant: block-column edges at 64, 96, 128, 160, 256, 320 and sizes
    just past them), 16/32/64 knots per region: FP64 D against the earlier
    forms (1e-12) and the oracle (TOL_D), mixed D against the earlier forms
    (1e-10)."""
    rng = np.random.default_rng(sum(sizes) + r)
    xs, ys = [], []
    for k, m in enumerate(sizes):
        xs.append(0.5 * (k % 2) + 0.02 + 0.46 * rng.random(m))
        ys.append(0.5 * (k // 2) + 0.02 + 0.46 * rng.random(m))
    x, y = np.concatenate(xs), np.concatenate(ys)
    tree = capi.partition_program(x, y, rng.standard_normal(len(x)), None, None, M=2, J=4, r=r,
                                  bounds=synth.UNIT_BOUNDS)
    assert tuple(np.diff(tree["obs_offsets"].astype(np.int64))) == tuple(sizes)
    cov, prm = synth.linear_trend_cov(log_range=(np.log(0.2), 0.8, 0.0)), dict(synth.UNIT_PARAMS)
    h = capi.Handle(tree, with_pred=False)
    out = {}
    try:
        h.set_covariance(cov)
        h.set_params(**prm)
        for prec in (capi.PRECISION_FP64, capi.PRECISION_MIXED):
            h.set_precision(prec)
            for ll in (2, 0):
                h.set_option(capi.OPT_LOCAL_LL, ll)
                out[(prec, ll)] = h.likelihood()
        h.set_option(capi.OPT_LOCAL_LL, 1)
    finally:
        h.close()
    Do, _, _ = oracle.evaluate(tree, cov, prm, predict=False)
    print(sizes, r, out, "oracle", Do, flush=True)
    assert rel(out[(capi.PRECISION_FP64, 2)], out[(capi.PRECISION_FP64, 0)]) < 1e-12
    assert rel(out[(capi.PRECISION_FP64, 2)], Do) < TOL_D
    assert rel(out[(capi.PRECISION_MIXED, 2)], out[(capi.PRECISION_MIXED, 0)]) < 1e-10


def test_prior_one_wave_per_region_refused():
    """The one-wave-per-region coarse prior is switched off: setting the
    option is refused with MRA_E_UNSUPPORTED and the default form stays."""
    tree, cov, prm = cases.synthetic_case(n=300, npred=0, M=3, J=4, r=64, seed=3, partition="program")
    h = capi.Handle(tree, with_pred=False)
    try:
        with pytest.raises(capi.MraError) as e:
            h.set_option(capi.OPT_PRIOR_WAVE, 1)
        assert e.value.code == capi.MRA_E_UNSUPPORTED
        assert h.get_option(capi.OPT_PRIOR_WAVE) == 0
    finally:
        h.close()


@pytest.mark.skip(reason="MRA_OPT_PRIOR_WAVE is refused: a GPU run of this form (mixed precision and graph "
                         "capture, first used in test_xcd_item_order) stopped making progress; cause not established")
@pytest.mark.parametrize("case", ["planar64", "chordal40", "gc_stationary50", "j2_64"])
def test_prior_one_wave_per_region(case):
    """Coarse prior of 64-knot (padded) regions one wave per region
    (MRA_OPT_PRIOR_WAVE = 1, the default) or one work group per region (0):
    D within 1e-12 of each other and within TOL_D of the oracle, predictions
    within TOL_PRED, FP64 and mixed (FP32 copy of G), eager and graph.  The
    cases cover r = 64 and padded r (40, 50), the three metrics, the
    stationary kernel and J = 2 (one work group then spans two parents)."""
    M, J, r, metric, stationary = {"planar64": (4, 4, 64, "planar", False),
                                   "chordal40": (4, 4, 40, "chordal", False),
                                   "gc_stationary50": (4, 4, 50, "great_circle", True),
                                   "j2_64": (6, 2, 64, "planar", False)}[case]
    tree, cov, prm = cases.synthetic_case(n=6000, npred=200, M=M, J=J, r=r, seed=97, clustered=True,
                                          metric=metric, stationary=stationary, partition="program")
    h = capi.Handle(tree)
    out = {}
    try:
        h.set_covariance(cov)
        h.set_params(**prm)
        for prec in (capi.PRECISION_FP64, capi.PRECISION_MIXED):
            h.set_precision(prec)
            for wave in (1, 0):
                h.set_option(capi.OPT_PRIOR_WAVE, wave)
                assert h.get_option(capi.OPT_PRIOR_WAVE) == wave
                for graph in (0, 1):
                    h.use_graph(graph)
                    out[(prec, wave, graph)] = h.predict()
    finally:
        h.close()
    for prec in (capi.PRECISION_FP64, capi.PRECISION_MIXED):
        ref = out[(prec, 0, 0)]
        for k, v in out.items():
            if k[0] == prec:
                print(case, k, v[0], ref[0], rel(v[0], ref[0]), flush=True)
                assert rel(v[0], ref[0]) < 1e-12
                np.testing.assert_allclose(v[1], ref[1], rtol=TOL_PRED, atol=1e-10)
                np.testing.assert_allclose(v[2], ref[2], rtol=TOL_PRED, atol=1e-10)
        # the graph replays the eager launches bit for bit
        for wave in (1, 0):
            assert out[(prec, wave, 1)][0] == out[(prec, wave, 0)][0]
    Do, mo, vo = oracle.evaluate(tree, cov, prm)
    D, mean, var = out[(capi.PRECISION_FP64, 1, 0)]
    assert rel(D, Do) < TOL_D
    np.testing.assert_allclose(mean, mo, rtol=TOL_PRED, atol=1e-10)
    np.testing.assert_allclose(var, vo, rtol=TOL_PRED, atol=1e-10)


@pytest.mark.parametrize("case", ["clustered64", "padded40", "chordal_stationary", "cross16_off", "clustered32",
                                  "padded24_chordal_stationary"])
def test_finest_one_wave_per_tile(case):
    """Mixed-precision finest conditioning one wave per tile
    (MRA_OPT_FINEST_WAVE = 1) against one work group per tile (0): both
    within the mixed contract (|delta loglik| / n <= 1e-6) of the FP64
    oracle and within 1e-7 of each other per observation (FP32 rounding in
    another order; measured 2.5e-9 against 3.6e-9 and 6.1e-9 from the
    oracle), predictions within 1e-4, graph replay equal to eager bit for bit.
    The cases cover 64-, 32- and 16-column tiles (and 64/32 only), 64- and
    32-knot regions, padded knots (r = 40, 24), the chordal metric and the
    stationary kernel."""
    rng = np.random.default_rng(83)
    if case in ("clustered64", "cross16_off", "clustered32"):
        x, y = synth.clustered_points(12000, rng)
        px, py = synth.uniform_points(300, rng)
        tree = capi.partition_program(x, y, np.zeros(12000), px, py, M=4, J=4, r=32 if case == "clustered32" else 64,
                                      bounds=synth.UNIT_BOUNDS)
        cov, prm = synth.linear_trend_cov(), dict(synth.UNIT_PARAMS)
    elif case == "padded24_chordal_stationary":
        tree, cov, prm = cases.synthetic_case(n=8000, npred=200, M=5, J=4, r=24, seed=86, clustered=True,
                                              metric="chordal", stationary=True, partition="program")
    elif case == "padded40":
        tree, cov, prm = cases.synthetic_case(n=8000, npred=200, M=4, J=4, r=40, seed=84, clustered=True,
                                              partition="program")
    else:
        tree, cov, prm = cases.synthetic_case(n=8000, npred=200, M=4, J=4, r=50, seed=85, clustered=True,
                                              metric="chordal", stationary=True, partition="program")
    # observations drawn from the model (the mixed contract is stated for data
    # the model describes; on white noise the rounding error scales with |D|)
    h = capi.Handle(tree, with_pred=False)
    h.set_covariance(cov)
    h.set_params(**prm)
    tree["obs_val"] = h.sample_prior(seed=8, replace=False)
    h.close()
    n = len(tree["obs_lon"])
    out = {}
    h = capi.Handle(tree)
    try:
        h.set_covariance(cov)
        h.set_params(**prm)
        h.set_precision(capi.PRECISION_MIXED)
        if case == "cross16_off":
            h.set_option(capi.OPT_CROSS16, 0)
        for wave in (0, 1):
            h.set_option(capi.OPT_FINEST_WAVE, wave)
            assert h.get_option(capi.OPT_FINEST_WAVE) == wave
            for graph in (0, 1):
                h.use_graph(graph)
                out[(wave, graph)] = h.predict()
                out[(wave, graph, "l")] = h.likelihood()
    finally:
        h.close()
    Do, mo, vo = oracle.evaluate(tree, cov, prm, predict=True)
    for wave in (0, 1):
        assert out[(wave, 1)][0] == out[(wave, 0)][0]
        assert out[(wave, 1, "l")] == out[(wave, 0, "l")]
        np.testing.assert_array_equal(out[(wave, 1)][1], out[(wave, 0)][1])
        D, mean, var = out[(wave, 0)]
        err = abs(D - Do) / (2 * n)
        print(case, "wave", wave, "D", D, "oracle", Do, "|dl|/n", err, flush=True)
        assert err <= TOL_MIXED
        assert abs(out[(wave, 0, "l")] - Do) / (2 * n) <= TOL_MIXED
        np.testing.assert_allclose(mean, mo, rtol=1e-4, atol=1e-4 * np.abs(mo).max())
        np.testing.assert_allclose(var, vo, rtol=1e-4, atol=1e-6)
    # the two orders of FP32 rounding differ by less than either differs from FP64
    assert abs(out[(1, 0)][0] - out[(0, 0)][0]) / (2 * n) <= 1e-7
