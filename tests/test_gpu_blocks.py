"""GPU unit tests of the work-group building blocks (mra_selftest): the FP64
matrix-core GEMM lane layout, the LDS Cholesky and triangular solve, against
numpy on the same inputs."""
import numpy as np
import pytest

from nonstationary_mra_amd import capi

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("K,Mv,Nv", [(4, 64, 64), (64, 64, 64), (130, 64, 64), (37, 40, 23), (5, 16, 17)])
def test_blocks(K, Mv, Nv):
    rng = np.random.default_rng(K * 100 + Mv)
    A = rng.standard_normal((K, 64))
    B = rng.standard_normal((K, 64))
    C0 = rng.standard_normal((64, 64))
    X = rng.standard_normal((64, 64))
    S = X @ X.T / 64 + np.eye(64)
    Bt = rng.standard_normal((64, 64))
    Co, Lo, Xo, bad = capi.selftest(K, Mv, Nv, A, B, C0, S, Bt)
    ref = C0.copy()
    ref[:Mv, :Nv] -= (A.T @ B)[:Mv, :Nv]
    print("gemm max err", np.abs(Co - ref).max(), flush=True)
    np.testing.assert_allclose(Co, ref, rtol=0, atol=1e-12 * K)
    Lr = np.linalg.cholesky(S)
    assert bad == 0
    np.testing.assert_allclose(np.tril(Lo), Lr, rtol=0, atol=1e-13)
    np.testing.assert_allclose(Xo, np.linalg.solve(Lr, Bt), rtol=0, atol=1e-11)


@pytest.mark.parametrize("K,Mv,Nv", [(4, 64, 64), (64, 64, 64), (130, 64, 64), (37, 40, 23)])
def test_gemm_f32(K, Mv, Nv):
    """FP32 matrix-core GEMM with FP32 accumulation: agrees with the float64
    product to the FP32 error bound K * 2^-24 * max|A||B| (plus rounding)."""
    rng = np.random.default_rng(K + 7 * Mv)
    A = rng.standard_normal((K, 64)).astype(np.float32)
    B = rng.standard_normal((K, 64)).astype(np.float32)
    C0 = rng.standard_normal((64, 64)).astype(np.float32)
    Co = capi.selftest_f32(K, Mv, Nv, A, B, C0)
    ref = C0.astype(np.float64)
    ref[:Mv, :Nv] -= (A.astype(np.float64).T @ B.astype(np.float64))[:Mv, :Nv]
    bound = (K + 2) * 2.0 ** -24 * (np.abs(A).T @ np.abs(B) + np.abs(C0)).max()
    print("f32 gemm max err", np.abs(Co - ref).max(), "bound", bound, flush=True)
    assert np.abs(Co - ref).max() <= bound


@pytest.mark.parametrize("n", [16, 32, 48, 64, 128])
def test_blocked_cholesky_and_inverse(n):
    """Blocked LDS Cholesky (16 x 16 blocks, MFMA panel and trailing updates)
    and in-place triangular inverse against numpy; the strict upper part of
    the inverse is zero; an indefinite matrix raises the pivot flag."""
    rng = np.random.default_rng(n)
    X = rng.standard_normal((n, n))
    S = X @ X.T / n + 0.5 * np.eye(n)
    Lo, Xo, bad = capi.selftest_blk(S)
    Lr = np.linalg.cholesky(S)
    assert bad == 0
    print("blk chol err", np.abs(Lo - Lr).max(), "inv err", np.abs(Xo - np.linalg.inv(Lr)).max(), flush=True)
    np.testing.assert_allclose(Lo, Lr, rtol=0, atol=1e-12)
    np.testing.assert_allclose(Xo, np.linalg.inv(Lr), rtol=0, atol=1e-10 * np.abs(np.linalg.inv(Lr)).max())
    assert np.all(np.triu(Xo, 1) == 0.0)
    S2 = S.copy()
    S2[n - 3, n - 3] = -1.0
    _, _, bad2 = capi.selftest_blk(S2)
    assert bad2 == 1
