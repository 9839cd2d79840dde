"""Brute-force dense MRA covariance oracle (numpy).  TEST INFRASTRUCTURE ONLY.

This is structurally independent of the region-by-region algorithm that the
reference implements (and that oracle/mra_oracle.cpp restates): it builds the
MRA covariance matrix Sigma_M of all observation and prediction points directly
from the recursive definition of the multi-resolution approximation, and then
evaluates the Gaussian log-likelihood and the kriging predictor with dense
numpy linear algebra (no block Cholesky factors, no w/A messages).

Definition (Katzfuss 2017, with the reference's nugget-on-knot quirk):
  C_0(a, b)   = C(a, b)                         base cross-covariance (no nugget)
  for level k = 0..M-2 and every region R at level k with knots Q_R:
      K_R     = C_k(Q_R, Q_R) + diag(tau^2(Q_R))   (the reference puts the nugget on
                                                  every knot diagonal: src/class_approximation.cpp:143-152)
      Sigma(a,b) += C_k(a, Q_R) K_R^{-1} C_k(Q_R, b)          for a, b below R
      C_{k+1}(a,b) = C_k(a,b) - C_k(a, Q_R) K_R^{-1} C_k(Q_R, b) (a, b below R)
  finest region F: Sigma(a,b) += C_{M-1}(a,b) + tau^2(a) [a == b]   for a, b in F
Log-likelihood: l = -1/2 (y' Sigma^{-1} y + log det Sigma + n log(2 * 3.14159265359)).
Prediction: mean = Sigma_PS Sigma_SS^{-1} y, var = diag(Sigma_PP) - diag(Sigma_PS Sigma_SS^{-1} Sigma_SP).
"""
import numpy as np

PI_LITERAL = 3.14159265359  # src/class_approximation-posterior-inference.cpp:244


def to_xyz(lon, lat):
    lon = np.asarray(lon, dtype=np.float64)
    lat = np.asarray(lat, dtype=np.float64)
    cl, sl = np.cos(lat * np.pi / 180), np.sin(lat * np.pi / 180)
    co, so = np.cos(lon * np.pi / 180), np.sin(lon * np.pi / 180)
    return np.stack([cl * co, cl * so, sl], axis=-1)


def pair_dist(cov, lon1, lat1, lon2, lat2):
    """Distances between two point sets (matrix n1 x n2)."""
    metric = cov["metric"]
    if metric == 0:
        p1, p2 = to_xyz(lon1, lat1), to_xyz(lon2, lat2)
        d = np.sqrt(((p1[:, None, :] - p2[None, :, :]) ** 2).sum(-1))
        return 6371.0 * d
    if metric == 1:
        la1, la2 = np.radians(lat1)[:, None], np.radians(lat2)[None, :]
        lo1, lo2 = np.radians(lon1)[:, None], np.radians(lon2)[None, :]
        ang = np.sin(la1) * np.sin(la2) + np.cos(la1) * np.cos(la2) * np.cos(lo1 - lo2)
        return 6371.0 * np.arccos(np.clip(ang, -1.0, 1.0))
    dx = np.asarray(lon1)[:, None] - np.asarray(lon2)[None, :]
    dy = np.asarray(lat1)[:, None] - np.asarray(lat2)[None, :]
    return np.sqrt(dx * dx + dy * dy)


def field(cov, which, lon, lat):
    """exp of the log-parameter field at the points (unscaled)."""
    f = cov[which]
    lon = np.asarray(lon, dtype=np.float64)
    lat = np.asarray(lat, dtype=np.float64)
    if f["kind"] == 0:
        return np.full(lon.shape, np.exp(f["intercept"]))
    if f["kind"] == 2:
        return np.exp(f["intercept"] + f["slope_x"] * lon + f["slope_y"] * lat)
    blon, blat, w = (np.asarray(f[k], dtype=np.float64) for k in ("lon", "lat", "w"))
    if cov["metric"] == 0:
        d = pair_dist(cov, lon, lat, blon, blat) / 6371.0
    elif cov["metric"] == 1:
        d = pair_dist(cov, lon, lat, blon, blat) / 6371.0
    else:
        d = pair_dist(cov, lon, lat, blon, blat)
    d = d / cov["wendland_len"]
    psi = np.where(d < 1, (1 - np.minimum(d, 1)) ** 6 * (35 * d * d + 18 * d + 3) / 3, 0.0)
    return np.exp(f["intercept"] + psi @ w)


def cross_cov(cov, prm, lon1, lat1, lon2, lat2):
    d = pair_dist(cov, lon1, lat1, lon2, lat2)
    if not cov["nonstationary"]:
        return prm["sill"] * np.exp(-d / prm["range"])
    s1 = prm["sill_scale"] * field(cov, "sigmasq", lon1, lat1)
    s2 = prm["sill_scale"] * field(cov, "sigmasq", lon2, lat2)
    b1 = prm["range_scale"] * field(cov, "beta", lon1, lat1)
    b2 = prm["range_scale"] * field(cov, "beta", lon2, lat2)
    bms = (b1[:, None] ** 2 + b2[None, :] ** 2) / 2
    pre = np.sqrt(s1[:, None] * s2[None, :]) * (b1[:, None] * b2[None, :] / bms) ** 1.5
    return pre * np.exp(-d / np.sqrt(bms))


def nugget(cov, prm, lon, lat):
    if not cov["nonstationary"]:
        return np.full(np.asarray(lon).shape, prm["nugget"])
    return prm["nugget_scale"] * field(cov, "tausq", lon, lat)


def dense_sigma(tree, cov, prm):
    """Sigma_M over [obs; pred] (n+np square), built from the definition."""
    M, J, r = tree["M"], tree["J"], tree["r"]
    L = M - 1
    nC = tree["n_coarse"]
    nF = tree["n_finest"]
    klon, klat = np.asarray(tree["knot_lon"]), np.asarray(tree["knot_lat"])
    oo = np.asarray(tree["obs_offsets"], dtype=np.int64)
    olon, olat = np.asarray(tree["obs_lon"]), np.asarray(tree["obs_lat"])
    po = tree.get("pred_offsets")
    if po is None:
        po = np.zeros(nF + 1, dtype=np.int64)
        plon = plat = np.zeros(0)
    else:
        po = np.asarray(po, dtype=np.int64)
        plon, plat = np.asarray(tree["pred_lon"]), np.asarray(tree["pred_lat"])
    n, npred = len(olon), len(plon)
    lon = np.concatenate([klon, olon, plon])
    lat = np.concatenate([klat, olat, plat])
    N = len(lon)
    nk = nC * r
    # region of every point at every level (-1 = not below any region there)
    lvl_start = [(J ** l - 1) // (J - 1) for l in range(L)]
    reg = -np.ones((L + 1, N), dtype=np.int64)
    for R in range(nC):
        l = max(i for i in range(L) if R >= lvl_start[i])
        a = R
        for k in range(l, -1, -1):
            reg[k, R * r:(R + 1) * r] = a
            a = (a - 1) // J
    parent = np.asarray(tree["finest_parent"], dtype=np.int64)
    for f in range(nF):
        idx = np.concatenate([nk + np.arange(oo[f], oo[f + 1]), nk + n + np.arange(po[f], po[f + 1])])
        reg[L, idx] = f
        a = parent[f]
        for k in range(L - 1, -1, -1):
            reg[k, idx] = a
            a = (a - 1) // J
    cur = cross_cov(cov, prm, lon, lat, lon, lat)
    # C(x, x) = sigma^2(x) exactly (the variance-covariance matrices of the
    # reference set their diagonal directly, src/evaluate_covariance.cpp:307;
    # an acos-based great-circle distance of a point to itself is not 0)
    if cov["nonstationary"]:
        cur[np.diag_indices(N)] = prm["sill_scale"] * field(cov, "sigmasq", lon, lat)
    else:
        cur[np.diag_indices(N)] = prm["sill"]
    tau = nugget(cov, prm, lon, lat)
    sig = np.zeros((n + npred, n + npred))
    for k in range(L):
        for R in range(lvl_start[k], lvl_start[k] + J ** k):
            q = np.arange(R * r, (R + 1) * r)
            below = np.nonzero((reg[k] == R))[0]
            below = below[~np.isin(below, q)]
            if below.size == 0:
                continue
            K = cur[np.ix_(q, q)] + np.diag(tau[q])
            Cbq = cur[np.ix_(below, q)]
            T = Cbq @ np.linalg.solve(K, Cbq.T)
            data = below[below >= nk]
            sel = np.isin(below, data)
            sig[np.ix_(data - nk, data - nk)] += T[np.ix_(sel, sel)]
            cur[np.ix_(below, below)] -= T
    for f in range(nF):
        idx = np.nonzero(reg[L] == f)[0]
        sig[np.ix_(idx - nk, idx - nk)] += cur[np.ix_(idx, idx)]
        sig[idx - nk, idx - nk] += tau[idx]
    return sig, n, npred


def dense_evaluate(tree, cov, prm):
    """Return (D, loglik, mean, var): D = -2 loglik - n log(2*pi_literal)."""
    sig, n, npred = dense_sigma(tree, cov, prm)
    y = np.asarray(tree["obs_val"], dtype=np.float64)
    Sss = sig[:n, :n]
    Lc = np.linalg.cholesky(Sss)
    alpha = np.linalg.solve(Lc, y)
    D = float(alpha @ alpha + 2 * np.log(np.diag(Lc)).sum())
    ll = -0.5 * (D + n * np.log(2 * PI_LITERAL))
    mean = var = None
    if npred:
        Sps = sig[n:, :n]
        W = np.linalg.solve(Lc, Sps.T)
        mean = W.T @ alpha
        var = np.diag(sig[n:, n:]) - (W * W).sum(0)
    return D, ll, mean, var
