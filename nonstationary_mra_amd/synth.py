"""Synthetic workloads on the unit square (planar metric, linear-trend fields).

The benchmark configurations place points on [0,1]^2 with log-parameter
fields that are linear in the coordinates:
    log sigma^2 = a_s + b_s x + c_s y,  log beta = a_b + b_b x + c_b y,
    tau^2 = constant nugget.
These are expressed through the library's MRA_FIELD_LINEAR / MRA_FIELD_CONSTANT
fields with the planar metric (distance = Euclidean distance in the plane).
"""
import numpy as np

from .capi import FIELD_CONSTANT, FIELD_LINEAR, METRIC


def uniform_points(n, rng):
    return rng.random(n), rng.random(n)


def clustered_points(n, rng, frac_uniform=0.6, n_clusters=20, sigma=0.02):
    """60% uniform on [0,1]^2 plus 40% from Gaussian clusters (points outside
    the square are redrawn)."""
    nu = int(round(n * frac_uniform))
    x = np.empty(n)
    y = np.empty(n)
    x[:nu] = rng.random(nu)
    y[:nu] = rng.random(nu)
    cx, cy = rng.random(n_clusters), rng.random(n_clusters)
    k = n - nu
    lab = rng.integers(0, n_clusters, k)
    px = cx[lab] + sigma * rng.standard_normal(k)
    py = cy[lab] + sigma * rng.standard_normal(k)
    bad = (px < 0) | (px >= 1) | (py < 0) | (py >= 1)
    while bad.any():
        m = int(bad.sum())
        px[bad] = cx[lab[bad]] + sigma * rng.standard_normal(m)
        py[bad] = cy[lab[bad]] + sigma * rng.standard_normal(m)
        bad = (px < 0) | (px >= 1) | (py < 0) | (py >= 1)
    x[nu:], y[nu:] = px, py
    perm = rng.permutation(n)
    return x[perm], y[perm]


def linear_trend_cov(log_sill=(0.0, 0.5, -0.3), log_range=(np.log(0.05), 0.8, 0.0), nugget=0.05):
    """Nonstationary exponential (Matern 1/2) kernel with linear-trend fields."""
    return dict(nonstationary=1, metric=METRIC["planar"], wendland_len=1.0,
                sigmasq=dict(kind=FIELD_LINEAR, intercept=log_sill[0], slope_x=log_sill[1], slope_y=log_sill[2]),
                beta=dict(kind=FIELD_LINEAR, intercept=log_range[0], slope_x=log_range[1], slope_y=log_range[2]),
                tausq=dict(kind=FIELD_CONSTANT, intercept=float(np.log(nugget))))


UNIT_PARAMS = dict(sill=0.0, range=1.0, nugget=0.0, sill_scale=1.0, range_scale=1.0, nugget_scale=1.0)

UNIT_BOUNDS = np.array([0.0, 1.0, 0.0, 1.0])


def dense_gp_draw(cov, lon, lat, rng, params=UNIT_PARAMS):
    """Exact draw from the nonstationary GP with nugget by dense Cholesky
    (small n only).  Uses the numpy kernel of the dense oracle module's
    formula, restated here to keep the product package self-contained."""
    n = len(lon)
    d = np.sqrt((lon[:, None] - lon[None, :]) ** 2 + (lat[:, None] - lat[None, :]) ** 2)
    s = cov["sigmasq"]
    b = cov["beta"]
    ls = s["intercept"] + s.get("slope_x", 0.0) * lon + s.get("slope_y", 0.0) * lat
    lb = b["intercept"] + b.get("slope_x", 0.0) * lon + b.get("slope_y", 0.0) * lat
    sg = params["sill_scale"] * np.exp(ls)
    bt = params["range_scale"] * np.exp(lb)
    bms = (bt[:, None] ** 2 + bt[None, :] ** 2) / 2
    C = np.sqrt(sg[:, None] * sg[None, :]) * (bt[:, None] * bt[None, :] / bms) ** 1.5 * np.exp(-d / np.sqrt(bms))
    tau = params["nugget_scale"] * np.exp(cov["tausq"]["intercept"])
    C[np.diag_indices(n)] += tau
    Lc = np.linalg.cholesky(C)
    return Lc @ rng.standard_normal(n)
