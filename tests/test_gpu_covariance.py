"""GPU: the device covariance-matrix kernel (mra_cov_tile_device) against the
oracle's evaluate_nonstationary_cross_covariance restatement.

The device entry takes coordinates and the scaled sqrt-sill / range per point;
they are formed here from the oracle's own fields (sigma^2, beta at each
point), so the comparison covers the pair formula, the metric and the output
layout (row-major n1 x n2, src/evaluate_covariance.cpp:155-234).
Tolerances: FP64 output 1e-13 relative; FP32 output (mixed precision) within
2e-5 relative plus 1e-6 of the largest entry (inputs rounded to FP32, hardware
rsqrt and expf).
"""
import numpy as np
import pytest

from nonstationary_mra_amd import capi, synth
from oracle import oracle

pytestmark = pytest.mark.gpu


def device_points(cov, lon, lat, metric):
    s2, b, _ = oracle.fields(cov, lon, lat)
    if metric == "planar":
        x, y, z = lon, lat, np.zeros_like(lon)
    else:  # chordal: unit-sphere coordinates (src/class_partition.cpp:507-517)
        la, lo = np.radians(lat), np.radians(lon)
        x, y, z = np.cos(la) * np.cos(lo), np.cos(la) * np.sin(lo), np.sin(la)
    xyz = capi.DeviceBuffer.from_array(np.concatenate([x, y, z]))
    sb = capi.DeviceBuffer.from_array(np.concatenate([np.sqrt(s2), b]))
    return xyz, sb


@pytest.mark.parametrize("metric,n1,n2", [("planar", 300, 517), ("planar", 129, 512), ("chordal", 70, 1030)])
@pytest.mark.parametrize("precision", [capi.PRECISION_FP64, capi.PRECISION_MIXED])
def test_covariance_matrix_kernel(metric, n1, n2, precision):
    rng = np.random.default_rng(n1 * 7 + n2)
    if metric == "planar":
        cov = synth.linear_trend_cov()
        lon1, lat1 = synth.uniform_points(n1, rng)
        lon2, lat2 = synth.uniform_points(n2, rng)
    else:
        cov = dict(synth.linear_trend_cov(log_sill=(0.0, 0.05, -0.03), log_range=(np.log(300.0), 0.01, 0.0)),
                   metric=capi.METRIC["chordal"])
        lon1, lat1 = rng.uniform(-10, 10, n1), rng.uniform(-8, 8, n1)
        lon2, lat2 = rng.uniform(-10, 10, n2), rng.uniform(-8, 8, n2)
    prm = dict(synth.UNIT_PARAMS)
    expect = oracle.cross_covariance(cov, prm, lon1, lat1, lon2, lat2)
    xyz1, sb1 = device_points(cov, lon1, lat1, metric)
    xyz2, sb2 = device_points(cov, lon2, lat2, metric)
    dt = np.float32 if precision == capi.PRECISION_MIXED else np.float64
    out = capi.DeviceBuffer.from_array(np.full(n1 * n2, np.nan, dtype=dt))
    print(f"cov kernel metric={metric} n1={n1} n2={n2} precision={precision}", flush=True)
    capi.cov_matrix_device(precision, capi.METRIC[metric], xyz1.ptr, sb1.ptr, n1, xyz2.ptr, sb2.ptr, n2, out.ptr)
    got = out.to_array(dt, n1 * n2).astype(np.float64).reshape(n1, n2)
    assert np.isfinite(got).all()
    if precision == capi.PRECISION_FP64:
        np.testing.assert_allclose(got, expect, rtol=1e-13, atol=0)
    else:
        np.testing.assert_allclose(got, expect, rtol=2e-5, atol=1e-6 * np.abs(expect).max())


def test_covariance_matrix_kernel_unaligned_output():
    """An output address that is not 16-byte aligned takes the scalar-store
    path and gives the same values."""
    rng = np.random.default_rng(5)
    cov = synth.linear_trend_cov()
    n1, n2 = 65, 260
    lon1, lat1 = synth.uniform_points(n1, rng)
    lon2, lat2 = synth.uniform_points(n2, rng)
    xyz1, sb1 = device_points(cov, lon1, lat1, "planar")
    xyz2, sb2 = device_points(cov, lon2, lat2, "planar")
    for prec, dt in ((capi.PRECISION_FP64, np.float64), (capi.PRECISION_MIXED, np.float32)):
        esz = np.dtype(dt).itemsize
        buf = capi.DeviceBuffer.from_array(np.zeros(n1 * n2 + 1, dtype=dt))
        aligned = capi.DeviceBuffer.from_array(np.zeros(n1 * n2, dtype=dt))
        for o in (buf.ptr + esz, aligned.ptr):
            capi.cov_matrix_device(prec, capi.METRIC["planar"], xyz1.ptr, sb1.ptr, n1, xyz2.ptr, sb2.ptr, n2, o)
        b = buf.to_array(dt, n1 * n2 + 1)
        assert b[0] == 0.0
        np.testing.assert_array_equal(b[1:], aligned.to_array(dt, n1 * n2))
