"""CPU tests of the derivative-free minimiser (mra_bobyqa, host code of the
library; no GPU needed) on functions whose minimisers are known in closed
form, including minimisers on the bounds."""
import numpy as np
import pytest

from nonstationary_mra_amd import capi


def test_quadratic_exact_minimiser():
    """A convex quadratic in 6 variables: the minimiser is A^{-1} b."""
    rng = np.random.default_rng(1)
    Q = rng.standard_normal((6, 6))
    A = Q @ Q.T + 6 * np.eye(6)
    b = rng.standard_normal(6)
    xs = np.linalg.solve(A, b)
    f = lambda x: 0.5 * x @ A @ x - b @ x
    x, fb, ne, rc = capi.bobyqa(f, np.zeros(6), -5 * np.ones(6), 5 * np.ones(6), 0.5, 1e-8, 500)
    print(ne, rc, np.abs(x - xs).max())
    assert rc == 0
    np.testing.assert_allclose(x, xs, atol=1e-6)
    assert ne < 200


def test_rosenbrock():
    f = lambda x: 100 * (x[1] - x[0] ** 2) ** 2 + (1 - x[0]) ** 2
    x, fb, ne, rc = capi.bobyqa(f, np.array([-1.2, 1.0]), np.array([-5.0, -5.0]), np.array([5.0, 5.0]), 0.5, 1e-8,
                                2000)
    print(ne, rc, x, fb)
    assert rc == 0
    np.testing.assert_allclose(x, [1.0, 1.0], atol=1e-5)


def test_active_bounds():
    """The unconstrained minimiser (3, -3, 0.5) lies outside the box: the
    solution is its projection (2, -1, 0.5), since f is separable."""
    c = np.array([3.0, -3.0, 0.5])
    f = lambda x: float(((x - c) ** 2 * np.array([1.0, 2.0, 3.0])).sum())
    x, fb, ne, rc = capi.bobyqa(f, np.zeros(3), np.array([-1.0, -1.0, -1.0]), np.array([2.0, 2.0, 2.0]), 0.3, 1e-9,
                                500)
    print(ne, rc, x)
    assert rc == 0
    np.testing.assert_allclose(x, [2.0, -1.0, 0.5], atol=1e-7)


def test_evaluation_limit_reported():
    f = lambda x: float(((x - 0.3) ** 2).sum() + np.sin(7 * x).sum() ** 2)
    x, fb, ne, rc = capi.bobyqa(f, np.zeros(4), -np.ones(4), np.ones(4), 0.4, 1e-12, 15)
    assert rc == capi.MRA_W_MAX_EVALS
    assert ne <= 16
    assert fb <= f(np.zeros(4))


def test_argument_checks():
    f = lambda x: float((x ** 2).sum())
    with pytest.raises(capi.MraError):
        capi.bobyqa(f, np.zeros(2), np.zeros(2), np.ones(2) * 0.1, 0.5, 1e-6, 100)  # box narrower than 2 rho


@pytest.mark.parametrize("npt", [13, 20, 28])
def test_interpolation_set_sizes(npt):
    """npt from 2n+1 up to the full quadratic (n+1)(n+2)/2: the same
    minimiser of a correlated convex quadratic; with the full set the first
    model is the function itself, so very few evaluations follow it."""
    rng = np.random.default_rng(2)
    Q = rng.standard_normal((6, 6))
    A = Q @ Q.T + 0.5 * np.eye(6)
    b = rng.standard_normal(6)
    xs = np.linalg.solve(A, b)
    f = lambda x: 0.5 * x @ A @ x - b @ x
    x, fb, ne, rc = capi.bobyqa(f, np.zeros(6), -20 * np.ones(6), 20 * np.ones(6), 0.5, 1e-8, 3000, npt=npt)
    print(npt, ne, rc, np.abs(x - xs).max())
    assert rc == 0
    np.testing.assert_allclose(x, xs, atol=1e-5)
    if npt == 28:
        assert ne < 120


def test_rosenbrock_full_quadratic_model():
    f = lambda x: 100 * (x[1] - x[0] ** 2) ** 2 + (1 - x[0]) ** 2
    x, fb, ne, rc = capi.bobyqa(f, np.array([-1.2, 1.0]), np.array([-5.0, -5.0]), np.array([5.0, 5.0]), 0.5, 1e-8,
                                2000, npt=6)
    assert rc == 0
    np.testing.assert_allclose(x, [1.0, 1.0], atol=1e-5)


def test_interpolation_set_size_limits():
    f = lambda x: float((x ** 2).sum())
    for npt in (12, 29):
        with pytest.raises(capi.MraError):
            capi.bobyqa(f, np.ones(6), -2 * np.ones(6), 2 * np.ones(6), 0.5, 1e-6, 100, npt=npt)


def test_seeded_curvature():
    """A starting curvature close to the true Hessian (here within 10 %):
    the 2n+1-point run reaches the same minimiser in far fewer evaluations
    than without it; an exact seed makes the first model the function."""
    rng = np.random.default_rng(3)
    Q = rng.standard_normal((6, 6))
    A = Q @ Q.T + 0.2 * np.eye(6)
    b = rng.standard_normal(6)
    xs = np.linalg.solve(A, b)
    f = lambda x: 0.5 * x @ A @ x - b @ x
    lo, hi = -30 * np.ones(6), 30 * np.ones(6)
    _, _, ne0, rc0 = capi.bobyqa(f, np.zeros(6), lo, hi, 0.5, 1e-8, 5000)
    E = 1 + 0.1 * rng.uniform(-1, 1, (6, 6))
    x1, _, ne1, rc1 = capi.bobyqa(f, np.zeros(6), lo, hi, 0.5, 1e-8, 5000, h0=A * (E + E.T) / 2)
    x2, _, ne2, rc2 = capi.bobyqa(f, np.zeros(6), lo, hi, 0.5, 1e-8, 5000, h0=A)
    print(ne0, ne1, ne2)
    assert rc0 == rc1 == rc2 == 0
    np.testing.assert_allclose(x1, xs, atol=1e-5)
    np.testing.assert_allclose(x2, xs, atol=1e-6)
    assert ne1 < ne0 / 2 and ne2 < 40
