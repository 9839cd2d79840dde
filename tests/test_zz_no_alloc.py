"""GPU test of the allocation-free entry point mra_cross_covariance_noalloc:
through the C ABI its output equals mra_cross_covariance's bit for bit, on
two calls in a row (the second smaller, reusing the same static device
array), for Wendland fields on the sphere and linear fields in the plane;
inputs beyond its array are refused with MRA_E_ARG."""
import ctypes as C

import numpy as np
import pytest

from nonstationary_mra_amd import capi, synth
from tests.cases import EXAMPLE_PARAMS, example_cov

pytestmark = pytest.mark.gpu


def _call(fn, cov, prm, lon1, lat1, lon2, lat2):
    keep = capi.Keep()
    c = capi.cov_desc(cov, keep)
    out = np.full(len(lon1) * len(lon2), np.nan)
    p = [prm[k] for k in ("sill", "range", "nugget", "sill_scale", "range_scale", "nugget_scale")]
    rc = fn(C.byref(c), *p, keep.d(lon1), keep.d(lat1), len(lon1), keep.d(lon2), keep.d(lat2), len(lon2),
            capi._d(out))
    return rc, out.reshape(len(lon1), len(lon2))


@pytest.mark.parametrize("case", ["sphere", "plane"])
def test_noalloc_matches_allocating_entry_point(case):
    rng = np.random.default_rng(11)
    if case == "sphere":
        cov, prm = example_cov(), dict(EXAMPLE_PARAMS)
        pts = lambda n: (rng.uniform(-9, 7, n), rng.uniform(-9, 5, n))
    else:
        cov, prm = synth.linear_trend_cov(), dict(synth.UNIT_PARAMS)
        pts = lambda n: synth.uniform_points(n, rng)
    lib = capi.lib()
    for n1, n2 in ((700, 900), (130, 17)):  # the second call smaller than the first
        lon1, lat1 = pts(n1)
        lon2, lat2 = pts(n2)
        rc_a, ref = _call(lib.mra_cross_covariance, cov, prm, lon1, lat1, lon2, lat2)
        rc_b, got = _call(lib.mra_cross_covariance_noalloc, cov, prm, lon1, lat1, lon2, lat2)
        assert rc_a == 0 and rc_b == 0
        assert np.all(np.isfinite(got))
        np.testing.assert_array_equal(got, ref)


def test_noalloc_refuses_inputs_beyond_its_array():
    lib = capi.lib()
    limit = lib.mra_cross_covariance_noalloc_limit()
    assert limit == 32 << 20
    n = int(np.sqrt(limit / 8)) + 16  # the n x n output alone exceeds the array
    rng = np.random.default_rng(12)
    x, y = synth.uniform_points(n, rng)
    rc, _ = _call(lib.mra_cross_covariance_noalloc, synth.linear_trend_cov(), dict(synth.UNIT_PARAMS), x, y, x, y)
    assert rc == capi.MRA_E_ARG
    assert b"scratch" in lib.mra_last_error(None)
    # and the array still serves a call that fits
    rc, got = _call(lib.mra_cross_covariance_noalloc, synth.linear_trend_cov(), dict(synth.UNIT_PARAMS), x[:50],
                    y[:50], x[:60], y[:60])
    _, ref = _call(lib.mra_cross_covariance, synth.linear_trend_cov(), dict(synth.UNIT_PARAMS), x[:50], y[:50],
                   x[:60], y[:60])
    assert rc == 0
    np.testing.assert_array_equal(got, ref)
