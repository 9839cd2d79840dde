"""CPU tests of the library's host side: the C ABI loads and exports every
declared symbol, and the sort-based partition builder reproduces the
reference's O(n*F) partition (restated in the oracle) exactly."""
import os
import re

import numpy as np
import pytest

from nonstationary_mra_amd import capi, formats
from oracle import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "mra.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(mra_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_declared_symbols():
    lib = capi.lib()
    syms = declared_symbols()
    assert len(syms) >= 15
    for s in syms:
        assert hasattr(lib, s), s
    assert sorted(capi.EXPORTED) == syms
    assert b"gfx950" in lib.mra_version()


@pytest.mark.parametrize("M,J,r,seed", [(3, 4, 16, 0), (4, 2, 9, 1), (5, 2, 4, 2), (4, 4, 64, 3), (3, 4, 30, 4)])
def test_partition_matches_reference_restatement(M, J, r, seed):
    rng = np.random.default_rng(seed)
    n, npred = 3000, 500
    lon, lat = rng.uniform(-3, 5, n), rng.uniform(1, 4, n)
    val = rng.standard_normal(n)
    plon, plat = rng.uniform(-4, 6, npred), rng.uniform(0, 5, npred)  # some outside the domain
    ref = oracle.partition_program(lon, lat, val, plon, plat, M=M, J=J, r=r)
    # put some observations exactly on ancestor knots and on region boundaries
    k = rng.integers(0, len(ref["knot_lon"]), 40)
    lon[:40], lat[:40] = ref["knot_lon"][k], ref["knot_lat"][k]
    b = ref["bounds"]
    lon[40:60] = (b[0] + b[1]) / 2
    lat[60:80] = (b[2] + b[3]) / 2
    ref = oracle.partition_program(lon, lat, val, plon, plat, M=M, J=J, r=r, bounds=b)
    got = capi.partition_program(lon, lat, val, plon, plat, M=M, J=J, r=r, bounds=b)
    assert len(ref["obs_lon"]) < n  # the knot-coincident observations were dropped
    for key in ("knot_lon", "knot_lat", "obs_offsets", "obs_lon", "obs_lat", "obs_val", "obs_index", "pred_offsets",
                "pred_lon", "pred_lat", "pred_index", "finest_parent"):
        np.testing.assert_array_equal(np.asarray(got[key]), np.asarray(ref[key]), err_msg=key)


def test_partition_rejects_bad_arguments():
    with pytest.raises(capi.MraError):
        capi.partition_program(np.zeros(3), np.zeros(3), np.zeros(3), None, None, M=3, J=3, r=4,
                               bounds=np.array([0, 1, 0, 1.0]))


def test_formats_roundtrip(tmp_path):
    p = tmp_path / "pred.bin"
    x = np.linspace(0, 1, 7)
    formats.write_prediction_result(str(p), x, x + 1, x * 2, x * 3)
    d = formats.read_prediction_result(str(p))
    np.testing.assert_array_equal(d["mean"], x * 2)
    with pytest.raises(formats.FormatError):
        open(tmp_path / "short.bin", "wb").write(np.array([10], dtype="<u8").tobytes())
        formats.read_data_file(str(tmp_path / "short.bin"))


def test_example_files_decode():
    ex = os.path.join(ROOT, "tests", "golden", "example")
    d = formats.read_data_file(os.path.join(ex, "dataLonLat.bin"))
    assert len(d["lon"]) == 1001 and "region" in d
    par = formats.read_parent_file(os.path.join(ex, "parentOfRegionsInFinestLevel.bin"))
    assert par["start"] == 15 and par["end"] == 30
    kn = formats.read_knots_file(os.path.join(ex, "knots.bin"))
    assert len(kn["lon"]) == 60
