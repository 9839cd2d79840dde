"""CPU checks of the warm-start oracle (oracle/mlp.py): its analytic
gradient against central finite differences of its loss, and the feature
definition on fields with known semivariograms."""
import numpy as np

from oracle import mlp as om


def test_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(4)
    p = rng.standard_normal(om.NPARAM) * 0.05
    X = rng.standard_normal((9, om.NIN))
    T = rng.standard_normal((9, om.NOUT))
    _, g = om.loss_grad(p, X, T)
    idx = np.concatenate([rng.choice(om.NPARAM, 60, replace=False), [om.NPARAM - 1, 0]])
    h = 1e-6
    for i in idx:
        e = np.zeros(om.NPARAM)
        e[i] = h
        fd = (om.loss(p + e, X, T) - om.loss(p - e, X, T)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 + 1e-6 * abs(g[i]), (i, fd, g[i])


def test_feature_definition_constant_and_scaled_fields():
    rng = np.random.default_rng(5)
    x, y = rng.random(3000), rng.random(3000)
    f0 = om.variogram_features(x, y, np.full(3000, 2.5))
    assert np.all(f0 == np.log(1e-12))  # constant field: zero semivariance
    v = rng.standard_normal(3000)
    f1 = om.variogram_features(x, y, v)
    f2 = om.variogram_features(x, y, 3.0 * v)
    np.testing.assert_allclose(f2 - f1, np.log(9.0), atol=1e-12)  # gamma scales with the square
    # white noise of variance 1: semivariance 1 at every lag (sample average)
    assert abs(np.exp(f1).mean() - 1.0) < 0.05


def test_whitening_makes_the_curvature_the_identity():
    """The whitened coordinates of the warm start: theta = theta0 + S z with
    S^T H S = I, on a curvature with the correlations of D (-0.98 between
    intercepts) and scales spread over two decades."""
    from nonstationary_mra_amd import warmstart
    rng = np.random.default_rng(6)
    C = np.full((6, 6), 0.5) + 0.5 * np.eye(6)
    C[0, 3] = C[3, 0] = -0.98
    C[1, 4] = C[4, 1] = -0.97
    C = C + 6 * np.eye(6) * (np.linalg.eigvalsh(C)[0] < 0.05)  # keep it positive definite
    d = np.exp(rng.uniform(np.log(1e2), np.log(1e4), 6))
    H = np.sqrt(np.outer(d, d)) * C
    S = warmstart.whitening(H)
    np.testing.assert_allclose(S.T @ H @ S, np.eye(6), atol=1e-10)
    # a step of length t in z changes the quadratic D by t^2 / 2 in every direction
    z = rng.standard_normal(6)
    z /= np.linalg.norm(z)
    dth = S @ (0.3 * z)
    assert abs(0.5 * dth @ H @ dth - 0.5 * 0.09) < 1e-12
