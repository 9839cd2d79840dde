"""GPU tests of the work-list autotuner: every candidate shape gives the
oracle's D to 1e-10 and the same predictions; mra_autotune keeps one of
the candidates and the results do not move."""
import itertools

import numpy as np
import pytest

from nonstationary_mra_amd import capi, synth
from oracle import oracle

pytestmark = pytest.mark.gpu


def _case():
    rng = np.random.default_rng(123)
    n, npred = 12000, 300
    x, y = synth.clustered_points(n, rng, n_clusters=4, sigma=0.03)
    px, py = synth.uniform_points(npred, rng)
    tree = capi.partition_program(x, y, rng.standard_normal(n), px, py, M=4, J=4, r=64, bounds=synth.UNIT_BOUNDS)
    return tree, synth.linear_trend_cov(), dict(synth.UNIT_PARAMS)


def test_every_shape_matches_oracle_and_autotune_keeps_results():
    tree, cov, prm = _case()
    Do, mo, vo = oracle.evaluate(tree, cov, prm, predict=True)
    h = capi.Handle(tree)
    try:
        h.set_covariance(cov)
        h.set_params(**prm)
        assert h.get_tuning() == (32, 128, 1)
        D0, m0, v0 = h.predict()
        for x32, lr, fb in itertools.product((0, 32, 64, -1), (64, 128), (0, 1)):
            h.set_tuning(x32, lr, fb)
            D, m, v = h.predict()
            print(x32, lr, fb, abs(D - Do) / abs(Do), flush=True)
            assert abs(D - Do) / abs(Do) < 1e-10
            np.testing.assert_allclose(m, mo, rtol=1e-9, atol=1e-10)
            np.testing.assert_allclose(v, vo, rtol=1e-9, atol=1e-10)
        cfg, ms = h.autotune()
        print("autotune", cfg, ms, flush=True)
        assert cfg[0] in (0, 32, 64, -1) and cfg[1] in (64, 128) and cfg[2] in (0, 1) and ms > 0
        assert h.get_tuning() == cfg
        D1 = h.likelihood()
        assert abs(D1 - D0) / abs(D0) < 1e-12
    finally:
        h.close()
    with pytest.raises(capi.MraError):
        h2 = capi.Handle(tree)
        try:
            h2.set_tuning(16, 128, 1)
        finally:
            h2.close()
