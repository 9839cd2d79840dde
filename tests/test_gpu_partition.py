"""GPU tests of the partition built on the device (mra_partition_program_gpu):
identical to the host builder (knots, parents, offsets, locations and their
order, input indices) on uniform, clustered and edge-case inputs, up to the
2,000,000-point benchmark workload."""
import time

import numpy as np
import pytest

from nonstationary_mra_amd import capi, synth

pytestmark = pytest.mark.gpu

KEYS = ("knot_lon", "knot_lat", "finest_parent", "obs_offsets", "obs_lon", "obs_lat", "obs_val", "obs_index",
        "pred_offsets", "pred_lon", "pred_lat", "pred_index")


def same(a, b):
    for k in KEYS:
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)
    assert a["r"] == b["r"] and a["n_finest"] == b["n_finest"]


@pytest.mark.parametrize("n, npred, M, J, r, clustered", [(5000, 700, 4, 4, 16, False), (20000, 0, 6, 2, 30, True),
                                                          (3000, 300, 3, 4, 64, True), (0, 10, 3, 4, 9, False)])
def test_gpu_partition_matches_host(n, npred, M, J, r, clustered):
    rng = np.random.default_rng(n + M)
    x, y = synth.clustered_points(n, rng) if clustered else synth.uniform_points(n, rng)
    px, py = synth.uniform_points(npred, rng)
    v = rng.standard_normal(n)
    b = synth.UNIT_BOUNDS
    same(capi.partition_program_gpu(x, y, v, px, py, M=M, J=J, r=r, bounds=b),
         capi.partition_program(x, y, v, px, py, M=M, J=J, r=r, bounds=b))


def test_gpu_partition_edge_cases():
    """Locations on interior box edges, on the outer edges, outside the
    domain, duplicated, and observations on ancestors' knots."""
    rng = np.random.default_rng(3)
    M, J, r = 4, 4, 16
    b = synth.UNIT_BOUNDS
    x, y = synth.uniform_points(3000, rng)
    probe = capi.partition_program(x, y, np.zeros(3000), None, None, M=M, J=J, r=r, bounds=b)
    ex = np.concatenate([rng.integers(0, 9, 200) / 8.0, [1.0, -0.1, 0.5, 1.2], probe["knot_lon"][::7], x[:50]])
    ey = np.concatenate([rng.integers(0, 9, 200) / 8.0, [0.3, 0.5, 1.0, 0.2], probe["knot_lat"][::7], y[:50]])
    x, y = np.concatenate([x, ex]), np.concatenate([y, ey])
    v = rng.standard_normal(len(x))
    g = capi.partition_program_gpu(x, y, v, ex, ey, M=M, J=J, r=r, bounds=b)
    h = capi.partition_program(x, y, v, ex, ey, M=M, J=J, r=r, bounds=b)
    same(g, h)
    assert len(g["obs_lon"]) < len(x)  # edge, outside and knot locations were dropped


def test_gpu_partition_benchmark_size():
    rng = np.random.default_rng(20260101)
    x, y = synth.clustered_points(2_000_000, rng)
    px, py = synth.uniform_points(200_000, rng)
    z = np.zeros(len(x))
    t0 = time.time()
    h = capi.partition_program(x, y, z, px, py, M=10, J=4, r=64, bounds=synth.UNIT_BOUNDS)
    t1 = time.time()
    capi.partition_program_gpu(x[:1000], y[:1000], z[:1000], None, None, M=3, J=4, r=64, bounds=synth.UNIT_BOUNDS)
    t2 = time.time()
    g = capi.partition_program_gpu(x, y, z, px, py, M=10, J=4, r=64, bounds=synth.UNIT_BOUNDS)
    t3 = time.time()
    print(f"partition n=2e6 M=10: host {t1 - t0:.2f}s, GPU {t3 - t2:.2f}s (incl. transfers)", flush=True)
    same(g, h)
