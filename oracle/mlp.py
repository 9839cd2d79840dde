"""CPU oracle of the warm-start network and its features (test
infrastructure only; nothing in the product imports this module).

* MLP 64 -> 256 -> 256 -> 6 with exact GELU, the library's flat parameter
  layout (W1 256x64, b1, W2 256x256, b2, W3 6x256, b3; row-major out x in),
  mean-squared-error loss; forward and backward in float64 numpy.  Pinned by
  central finite differences of the loss (tests/test_warmstart_cpu.py).
* Variogram features: the same quadrant sampling rule as the library, pairs
  enumerated in numpy (brute force).
"""
import math

import numpy as np
from scipy.special import erf

NIN, NH, NOUT = 64, 256, 6
SIZES = [("W1", (NH, NIN)), ("b1", (NH,)), ("W2", (NH, NH)), ("b2", (NH,)), ("W3", (NOUT, NH)), ("b3", (NOUT,))]
NPARAM = sum(int(np.prod(s)) for _, s in SIZES)


def unpack(p):
    out, o = {}, 0
    for name, shape in SIZES:
        k = int(np.prod(shape))
        out[name] = np.asarray(p[o:o + k], dtype=np.float64).reshape(shape)
        o += k
    return out


def pack(d):
    return np.concatenate([d[name].ravel() for name, _ in SIZES])


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / math.sqrt(2.0)))


def gelu_d(z):
    return 0.5 * (1.0 + erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2.0 * math.pi)


def forward(p, X):
    w = unpack(p)
    X = np.asarray(X, dtype=np.float64)
    Z1 = X @ w["W1"].T + w["b1"]
    A1 = gelu(Z1)
    Z2 = A1 @ w["W2"].T + w["b2"]
    A2 = gelu(Z2)
    Y = A2 @ w["W3"].T + w["b3"]
    return Y, (X, Z1, A1, Z2, A2)


def loss(p, X, T):
    Y, _ = forward(p, X)
    return float(np.mean((Y - np.asarray(T, dtype=np.float64)) ** 2))


def loss_grad(p, X, T):
    w = unpack(p)
    Y, (X, Z1, A1, Z2, A2) = forward(p, X)
    T = np.asarray(T, dtype=np.float64)
    B = X.shape[0]
    dY = 2.0 * (Y - T) / (B * NOUT)
    g = {"W3": dY.T @ A2, "b3": dY.sum(0)}
    dZ2 = (dY @ w["W3"]) * gelu_d(Z2)
    g["W2"] = dZ2.T @ A1
    g["b2"] = dZ2.sum(0)
    dZ1 = (dZ2 @ w["W2"]) * gelu_d(Z1)
    g["W1"] = dZ1.T @ X
    g["b1"] = dZ1.sum(0)
    return float(np.mean((Y - T) ** 2)), pack(g)


def variogram_features(lon, lat, val, S=16384, bins=16):
    """log semivariogram per quadrant of the bounding box and lag bin
    floor(bins * sqrt(d / hmax)), hmax = half the box's larger side; up to S
    points per quadrant taken at storage positions k * n_q // S (all points
    when the quadrant has at most S).  Pairs are enumerated in row blocks."""
    lon, lat, val = (np.asarray(a, dtype=np.float64) for a in (lon, lat, val))
    x0, x1, y0, y1 = lon.min(), lon.max(), lat.min(), lat.max()
    xm, ym = 0.5 * (x0 + x1), 0.5 * (y0 + y1)
    hm = max(0.5 * (x1 - x0), 0.5 * (y1 - y0))
    quad = (lon >= xm).astype(int) + 2 * (lat >= ym).astype(int)
    feats = np.zeros(4 * bins)
    for q in range(4):
        members = np.flatnonzero(quad == q)
        nq = len(members)
        Sq = min(nq, S)
        sel = members[(np.arange(Sq) * nq) // Sq] if Sq else members[:0]
        s = np.zeros(bins)
        c = np.zeros(bins, dtype=np.int64)
        for r0 in range(0, Sq, 256):
            rows = np.arange(r0, min(Sq, r0 + 256))
            i, j = np.nonzero(np.arange(Sq)[None, :] > rows[:, None])
            a, b = sel[rows[i]], sel[j]
            d = np.sqrt((lon[a] - lon[b]) ** 2 + (lat[a] - lat[b]) ** 2)
            keep = d < hm
            binix = np.minimum(bins - 1, (bins * np.sqrt(d[keep] / hm)).astype(int))
            e2 = (val[a] - val[b])[keep] ** 2
            s += np.bincount(binix, weights=e2, minlength=bins)
            c += np.bincount(binix, minlength=bins)
        g = np.where(c > 0, s / np.maximum(2 * c, 1), 0.0)
        have = c > 0
        for bix in range(bins):
            v = 1e-12
            for r in range(bins):
                if bix - r >= 0 and have[bix - r]:
                    v = g[bix - r]
                    break
                if bix + r < bins and have[bix + r]:
                    v = g[bix + r]
                    break
            feats[q * bins + bix] = math.log(max(v, 1e-12))
    return feats
