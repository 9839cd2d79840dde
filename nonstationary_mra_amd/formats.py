"""Readers and writers for the reference program's binary file formats.

All files are little-endian; u64 = unsigned 64-bit integer, f64 = IEEE double.

  data file        u64 n; f64 lon[n]; f64 lat[n]; f64 val[n]; [u64 region[n]]
                   (src/class_data.cpp:47-143; the region array is present when
                   knots are provided by files)
  prediction file  u64 np; f64 lon[np]; f64 lat[np]; [u64 region[np]]
                   (src/class_data.cpp:227-319)
  knots file       u64 nr; f64 lon[nr]; f64 lat[nr]; u64 region[nr]
                   (src/class_partition.cpp:122-263)
  parent file      u64 start; u64 end; u64 parent[end-start+1]
                   (src/class_partition.cpp:685-734)
  basis file       u64 nb; f64 intercept; f64 lon[nb]; f64 lat[nb]; f64 w[nb]
                   (src/class_nonstationary_covariance.cpp:84-170)
  prediction out   u64 n; f64 lon[n]; f64 lat[n]; f64 mean[n]; f64 var[n]
                   (src/class_approximation.cpp:481-538)
"""
import numpy as np


class FormatError(ValueError):
    pass


def _read(path):
    with open(path, "rb") as fh:
        return fh.read()


def _u64(buf, off, n):
    if off + 8 * n > len(buf):
        raise FormatError("file too short")
    return np.frombuffer(buf, dtype="<u8", count=n, offset=off).astype(np.uint64), off + 8 * n


def _f64(buf, off, n):
    if off + 8 * n > len(buf):
        raise FormatError("file too short")
    return np.frombuffer(buf, dtype="<f8", count=n, offset=off).astype(np.float64), off + 8 * n


def read_data_file(path, with_region=None):
    """Returns dict(lon, lat, val[, region]).  with_region=None detects the
    optional region array from the file size."""
    buf = _read(path)
    (n,), off = _u64(buf, 0, 1)
    n = int(n)
    lon, off = _f64(buf, off, n)
    lat, off = _f64(buf, off, n)
    val, off = _f64(buf, off, n)
    out = dict(lon=lon, lat=lat, val=val)
    if with_region is None:
        with_region = len(buf) >= off + 8 * n and n > 0
    if with_region:
        out["region"], off = _u64(buf, off, n)
    return out


def read_prediction_file(path, with_region=None):
    buf = _read(path)
    (n,), off = _u64(buf, 0, 1)
    n = int(n)
    lon, off = _f64(buf, off, n)
    lat, off = _f64(buf, off, n)
    out = dict(lon=lon, lat=lat)
    if with_region is None:
        with_region = len(buf) >= off + 8 * n and n > 0
    if with_region:
        out["region"], off = _u64(buf, off, n)
    return out


def read_knots_file(path):
    buf = _read(path)
    (n,), off = _u64(buf, 0, 1)
    n = int(n)
    lon, off = _f64(buf, off, n)
    lat, off = _f64(buf, off, n)
    region, off = _u64(buf, off, n)
    return dict(lon=lon, lat=lat, region=region)


def read_parent_file(path):
    buf = _read(path)
    (start, end), off = _u64(buf, 0, 2)
    m = int(end) - int(start) + 1
    parent, off = _u64(buf, off, m)
    return dict(start=int(start), end=int(end), parent=parent)


def read_basis_file(path):
    buf = _read(path)
    (nb,), off = _u64(buf, 0, 1)
    nb = int(nb)
    (ic,), off = _f64(buf, off, 1)
    lon, off = _f64(buf, off, nb)
    lat, off = _f64(buf, off, nb)
    w, off = _f64(buf, off, nb)
    return dict(kind=1, intercept=float(ic), lon=lon, lat=lat, w=w)


def write_prediction_result(path, lon, lat, mean, var):
    n = len(lon)
    with open(path, "wb") as fh:
        fh.write(np.asarray([n], dtype="<u8").tobytes())
        for a in (lon, lat, mean, var):
            fh.write(np.ascontiguousarray(a, dtype="<f8").tobytes())


def read_prediction_result(path):
    buf = _read(path)
    (n,), off = _u64(buf, 0, 1)
    n = int(n)
    out = {}
    for k in ("lon", "lat", "mean", "var"):
        out[k], off = _f64(buf, off, n)
    return out
