"""GPU tests of the learned warm start: variogram features, the MLP's own
forward / backward kernels against the float64 CPU oracle (oracle/mlp.py,
itself pinned by finite differences), bf16 inference, training, the
six-parameter fit, and the gain of the trained network's start over the
default start."""
import numpy as np
import pytest

from nonstationary_mra_amd import capi, synth, warmstart
from oracle import mlp as om

pytestmark = pytest.mark.gpu


def test_variogram_features_match_oracle():
    tree = warmstart.make_locations(6000, 5, 16, seed=3)
    h = capi.Handle(tree, with_pred=False)
    try:
        h.set_covariance(warmstart.trend_cov(warmstart.DEFAULT_START))
        h.set_params(**synth.UNIT_PARAMS)
        y = h.sample_prior(11, replace=True)
        f = h.variogram_features()
        f2 = h.variogram_features()
    finally:
        h.close()
    fo = om.variogram_features(tree["obs_lon"], tree["obs_lat"], y)
    print("max |df|", np.abs(f - fo).max(), flush=True)
    np.testing.assert_array_equal(f, f2)  # integer accumulation: bitwise repeatable
    np.testing.assert_allclose(f, fo, atol=2e-5)


def bf16_round(x):
    """Round to the nearest bfloat16 (ties to even), through float32."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def test_mlp_forward_and_gradient_match_oracle():
    rng = np.random.default_rng(12)
    m = capi.Mlp(seed=5)
    try:
        p = m.get_params().astype(np.float64)
        X = rng.standard_normal((37, 64)).astype(np.float32)
        T = rng.standard_normal((37, 6)).astype(np.float32)
        Y = m.forward(X)
        Yo, _ = om.forward(p, X)
        print("fwd max err", np.abs(Y - Yo).max(), flush=True)
        np.testing.assert_allclose(Y, Yo, rtol=1e-4, atol=1e-5)
        loss, g = m.loss_grad(X, T)
        lo, go = om.loss_grad(p, X, T)
        print("loss", loss, lo, "grad max err", np.abs(g - go).max(), "max |g|", np.abs(go).max(), flush=True)
        assert abs(loss - lo) <= 1e-5 * abs(lo)
        np.testing.assert_allclose(g, go, rtol=0, atol=2e-5 * np.abs(go).max())
        # bf16 inference: within the bfloat16 rounding of weights and activations
        Yb = m.forward(X, bf16=True)
        rel = np.abs(Yb - Yo).max() / np.abs(Yo).max()
        print("bf16 rel err", rel, flush=True)
        assert 0 < rel < 3e-2
        # ... and the bfloat16 matrix instructions against a float64 restatement of
        # the same roundings (inputs, weights, activations; FP32 accumulation on
        # the GPU): a misplaced fragment would give O(1) errors, while a rare
        # activation rounded the other way changes Y by about 1e-4 of its scale
        w = om.unpack(bf16_round(p))
        A = bf16_round(X)
        for k in (1, 2):
            A = bf16_round(om.gelu(A @ w[f"W{k}"].T + w[f"b{k}"]))
        Ye = A @ w["W3"].T + w["b3"]
        err = np.abs(Yb - Ye).max() / np.abs(Ye).max()
        print("bf16 MFMA vs emulation", err, flush=True)
        assert err < 2e-3
    finally:
        m.close()


def test_training_reduces_loss():
    rng = np.random.default_rng(13)
    A = rng.standard_normal((64, 6)) / 8
    X = rng.standard_normal((512, 64)).astype(np.float32)
    T = np.tanh(X @ A).astype(np.float32)
    m = capi.Mlp(seed=1)
    try:
        l0 = m.loss_grad(X, T)[0]
        for _ in range(150):
            for b in range(0, 512, 128):
                m.train_step(X[b:b + 128], T[b:b + 128], 1e-3)
        l1 = m.loss_grad(X, T)[0]
    finally:
        m.close()
    print("loss", l0, "->", l1, flush=True)
    assert l1 < 0.05 * l0


def test_fit_reaches_the_likelihood_of_the_truth():
    """The maximum-likelihood fit ends at least as high in likelihood as the
    true parameters (D_fit <= D(theta_true)), from the default start."""
    tree = warmstart.make_locations(5000, 5, 16, seed=21)
    h = capi.Handle(tree, with_pred=False)
    try:
        truth = np.array([0.3, 0.6, -0.2, np.log(0.08), 0.5, np.log(0.05)])
        warmstart.simulate(h, truth, 5)
        D_true = h.likelihood()
        th, D, ne, st = warmstart.fit(h, warmstart.DEFAULT_START)
        # the handle is left at the fitted parameters
        assert abs(h.likelihood() - D) <= 1e-9 * abs(D)
    finally:
        h.close()
    print("D_true", D_true, "D_fit", D, "evals", ne, "theta", th, flush=True)
    assert st in (0, capi.MRA_W_MAX_EVALS)
    assert D <= D_true + 1e-6 * abs(D_true)


def test_trained_start_needs_fewer_evaluations():
    """A small network trained on 512 prior draws starts the optimizer closer
    to the optimum: fewer evaluations on average over held-out fields, at an
    equal final likelihood (within the optimizer's resolution)."""
    X, th = warmstart.training_set(512, 3000, 4, 16, seed=31)
    ws, hist = warmstart.train(X, th, epochs=200, batch=128, lr=2e-3)
    print("train/val", hist[-1], flush=True)
    rows = warmstart.compare(ws, 6, 3000, 4, 16, seed=41, max_evals=600)
    s = warmstart.summarize(rows)
    print(s, flush=True)
    assert s["mean_evals_warm"] < s["mean_evals_default"]
    assert s["max_D_warm_minus_default"] < 0.05
