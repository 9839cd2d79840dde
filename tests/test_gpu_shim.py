"""The reference-style C++ layer (include/mra_approximation.hpp) driven by a
compiled program (nonstationary_mra_amd/examples/shim_example.cpp) in the
reference's prediction mode: Approximation(...).likelihood(), predict(),
dump_prediction_result() and evaluate_nonstationary_variance_covariance,
against the oracle on the same partition."""
import os
import subprocess

import numpy as np
import pytest

from nonstationary_mra_amd import build, capi, formats, synth
from oracle import oracle

pytestmark = pytest.mark.gpu


def _write_xy(path, lon, lat, val=None):
    with open(path, "wb") as fh:
        fh.write(np.asarray([len(lon)], dtype="<u8").tobytes())
        for a in (lon, lat) + ((val,) if val is not None else ()):
            fh.write(np.ascontiguousarray(a, dtype="<f8").tobytes())


@pytest.mark.parametrize("mixed", [False, True])
def test_cpp_approximation_prediction_mode(tmp_path, mixed):
    assert os.path.exists(build.SHIM), "build the library first (it also builds the C++ example)"
    rng = np.random.default_rng(91)
    n, npred, M, J, r = 3000, 300, 4, 4, 16
    x, y = synth.clustered_points(n, rng)
    px, py = synth.uniform_points(npred, rng)
    val = rng.standard_normal(n)
    _write_xy(tmp_path / "data.bin", x, y, val)
    _write_xy(tmp_path / "pred.bin", px, py)
    out = tmp_path / "pred_out.bin"
    cmd = [build.SHIM, str(tmp_path / "data.bin"), str(tmp_path / "pred.bin"), str(out), str(M), str(J), str(r)]
    if mixed:
        cmd.append("mixed")
    res = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    print(res.stdout, res.stderr, flush=True)
    assert res.returncode == 0
    D = float([l for l in res.stdout.splitlines() if l.startswith("D ")][0].split()[1])
    C00 = float([l for l in res.stdout.splitlines() if l.startswith("C00 ")][0].split()[1])
    tree = capi.partition_program(x, y, val, px, py, M=M, J=J, r=r, bounds=synth.UNIT_BOUNDS)
    cov, prm = synth.linear_trend_cov(), dict(synth.UNIT_PARAMS)
    Do, mo, vo = oracle.evaluate(tree, cov, prm, predict=True)
    pr = formats.read_prediction_result(str(out))
    np.testing.assert_array_equal(pr["lon"], tree["pred_lon"])
    if mixed:
        assert abs(D - Do) / (2 * n) < 1e-6
        np.testing.assert_allclose(pr["mean"], mo, rtol=1e-4, atol=1e-4 * np.abs(mo).max())
    else:
        assert abs(D - Do) / abs(Do) < 1e-10
        np.testing.assert_allclose(pr["mean"], mo, rtol=1e-9, atol=1e-10)
        np.testing.assert_allclose(pr["var"], vo, rtol=1e-9, atol=1e-10)
    V = oracle.variance_covariance(cov, prm, x[:4], y[:4])
    assert abs(C00 - V[0, 0]) < 1e-13 * abs(V[0, 0])
