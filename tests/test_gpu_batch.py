"""GPU tests of the batch of independent fields (mra_batch_*): every
handle's D in the batch equals its own evaluation bitwise, for every stream
count and launch mode; a field with a non-positive-definite model is
reported in its status slot without disturbing the others; the autotuner
keeps a valid setting."""
import numpy as np
import pytest

from nonstationary_mra_amd import capi, synth, warmstart

pytestmark = pytest.mark.gpu


def _fields():
    rng = np.random.default_rng(101)
    specs = [(800, 3, 16), (3000, 4, 16), (12000, 5, 32), (2000, 4, 32), (5000, 5, 16), (1500, 3, 16)]
    hs = []
    for i, (n, M, r) in enumerate(specs):
        x, y = synth.clustered_points(n, rng) if i % 2 else synth.uniform_points(n, rng)
        tree = capi.partition_program(x, y, rng.standard_normal(n), None, None, M=M, J=4, r=r,
                                      bounds=synth.UNIT_BOUNDS)
        h = capi.Handle(tree, with_pred=False)
        th = warmstart.draw_theta(rng, 1)[0]
        h.set_covariance(warmstart.trend_cov(th))
        h.set_params(**synth.UNIT_PARAMS)
        h.sample_prior(500 + i, replace=True)
        if i % 3 == 1:
            h.set_precision(capi.PRECISION_MIXED)
        hs.append(h)
    return hs


def test_batch_matches_individual_evaluations():
    hs = _fields()
    ref = np.array([h.likelihood() for h in hs])
    b = capi.Batch(hs)
    try:
        for streams, graph in [(1, 0), (4, 0), (2, 1), (8, 1), (4, 1)]:
            b.set(streams, graph)
            D, st = b.likelihood()
            print(streams, graph, np.abs(D - ref).max(), flush=True)
            assert np.all(st == 0)
            np.testing.assert_array_equal(D, ref)
        s, g, ms = b.autotune()
        print("autotune", s, g, ms, flush=True)
        assert s in (1, 2, 4, 8) and g in (0, 1) and ms > 0
        assert b.get() == (s, g)
        D, st = b.likelihood()
        np.testing.assert_array_equal(D, ref)
    finally:
        b.close()
        for h in hs:
            h.close()


def test_batch_reports_not_positive_definite_per_field():
    hs = _fields()[:3]
    hs[1].set_params(**dict(synth.UNIT_PARAMS, sill_scale=1.0, nugget_scale=-50.0))
    b = capi.Batch(hs)
    try:
        b.set(2, 1)
        D, st = b.likelihood()
    finally:
        b.close()
    ok0, ok2 = hs[0].likelihood(), hs[2].likelihood()
    for h in hs:
        h.close()
    assert st[1] == capi.MRA_E_NOT_PD and st[0] == 0 and st[2] == 0
    assert D[0] == ok0 and D[2] == ok2
