"""GPU tests of the batch of independent fields (mra_batch_*): every
handle's D in the batch equals its own evaluation bitwise, for every stream
count and launch mode, and after changes of parameters, precision and
options; a field with a non-positive-definite model is reported in its
status slot without disturbing the others; the autotuner keeps a valid
setting; repeated handles are refused."""
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
        for streams, graph in [(1, 0), (4, 0), (2, 1), (8, 1), (4, 1), (16, 0), (16, 1)]:
            b.set(streams, graph)
            D, st = b.likelihood()
            print(streams, graph, np.abs(D - ref).max(), flush=True)
            assert np.all(st == 0)
            np.testing.assert_array_equal(D, ref)
        s, g, ms = b.autotune()
        print("autotune", s, g, ms, flush=True)
        assert s in (1, 2, 4, 8, 16) and g in (0, 1) and ms > 0
        with pytest.raises(capi.MraError):
            b.set(17, 0)
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


def test_batch_follows_parameter_and_state_changes():
    """The batch graph is captured again when a handle's state changes: a
    stationary model's sill and range (passed to the kernels by value), a
    handle's precision and its options.  After each change every field's D
    equals the handle's own evaluation bitwise."""
    rng = np.random.default_rng(202)
    hs = []
    for i, (n, M, r) in enumerate([(1500, 3, 16), (4000, 4, 16), (2500, 4, 32)]):
        x, y = synth.uniform_points(n, rng)
        tree = capi.partition_program(x, y, rng.standard_normal(n), None, None, M=M, J=4, r=r,
                                      bounds=synth.UNIT_BOUNDS)
        h = capi.Handle(tree, with_pred=False)
        cov = synth.linear_trend_cov()
        if i != 1:
            cov = dict(cov, nonstationary=0)  # stationary sill * exp(-d / range)
        h.set_covariance(cov)
        h.set_params(sill=1.0, range=0.1, nugget=0.05)
        hs.append(h)
    b = capi.Batch(hs)
    try:
        b.set(2, 1)
        for step, change in enumerate(["none", "params", "precision", "option", "params"]):
            if change == "params":
                for k, h in enumerate(hs):
                    h.set_params(sill=1.0 + 0.3 * (step + k), range=0.05 + 0.02 * step, nugget=0.05)
            elif change == "precision":
                hs[2].set_precision(capi.PRECISION_MIXED)
            elif change == "option":
                hs[0].set_option(capi.OPT_WAVE_LOCAL16, 0)
            D, st = b.likelihood()
            ref = np.array([h.likelihood() for h in hs])
            print(change, D, ref, flush=True)
            assert np.all(st == 0)
            np.testing.assert_array_equal(D, ref)
    finally:
        b.close()
        for h in hs:
            h.close()


def test_batch_rejects_repeated_handles():
    hs = _fields()[:2]
    try:
        with pytest.raises(capi.MraError) as e:
            capi.Batch([hs[0], hs[1], hs[0]])
        assert e.value.code == capi.MRA_E_ARG
        assert "repeated" in str(e.value)
        # the handles keep their own side streams after a batch is gone
        b = capi.Batch(hs)
        b.likelihood()
        b.close()
        assert hs[0].get_option(capi.OPT_CONCURRENT) == 1
    finally:
        for h in hs:
            h.close()
