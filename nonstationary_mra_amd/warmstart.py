"""Learned warm start for the six-parameter maximum-likelihood fit.

The parameters are theta = (a_s, bx_s, by_s, a_b, bx_b, log_tau2) of the
planar linear-trend model (mra_fit_linear_trend):
    log sigma^2 = a_s + bx_s x + by_s y,  log beta = a_b + bx_b x,  tau^2 = exp(log_tau2).

Training fields are drawn from the MRA prior itself (mra_sample_prior) at
fixed locations, with theta drawn uniformly from PRIOR_LO..PRIOR_HI; the
features are the 64 log-semivariogram values of mra_variogram_features; the
network (mra_mlp_*, 64 -> 256 -> 256 -> 6, GELU) is trained on the GPU by the
library's own kernels on standardised features and targets.  At inference
the predicted theta (clipped into the optimizer's bounds) replaces the
default start of the optimizer.

The network's start comes with a learned mean curvature: the average
Hessian H of D at the optimum over fields of the training distribution
(central differences).  The six parameters are strongly correlated in D
(about -0.98 between the intercepts of log sigma^2 and log beta, the same on
every field), so from the network's start the minimiser runs in whitened
coordinates z, theta = theta0 + S z with S^T H S = I (S = L^-T, H = L L^T),
where the trust region is round in the metric of D.

Command line (one GPU):
    python -m nonstationary_mra_amd.warmstart train --fields 2048 --n 50000 --M 8 --r 32 --out FILE.npz
    python -m nonstationary_mra_amd.warmstart curvature --model FILE.npz --fields 64 --n 50000 --M 8 --r 32
    python -m nonstationary_mra_amd.warmstart evaluate --model FILE.npz --fields 64 --n 50000 --M 8 --r 32
"""
import argparse
import json
import os
import sys
import time

import numpy as np

from . import capi, synth

NTHETA = 6
NAMES = ("a_s", "bx_s", "by_s", "a_b", "bx_b", "log_tau2")
# sampling ranges of the training fields (log range in [log 0.02, log 0.2],
# range and variance slopes in [-1, 1], nugget in [0.01, 0.2])
PRIOR_LO = np.array([-0.5, -1.0, -0.5, np.log(0.02), -1.0, np.log(0.01)])
PRIOR_HI = np.array([0.5, 1.0, 0.5, np.log(0.2), 1.0, np.log(0.2)])
# optimizer bounds and the plain default start
BOUND_LO = np.array([-3.0, -3.0, -3.0, np.log(0.005), -3.0, np.log(1e-3)])
BOUND_HI = np.array([3.0, 3.0, 3.0, np.log(1.0), 3.0, np.log(1.0)])
DEFAULT_START = np.array([0.0, 0.0, 0.0, np.log(0.05), 0.0, np.log(0.05)])
RHO_BEGIN, RHO_END, MAX_EVALS = 0.3, 1e-4, 800
# interpolation points of the minimiser's quadratic models: the full
# quadratic in six parameters, for both starts (2n+1 = 13 needs about 2.6x
# more evaluations on these fields from either start)
NPT = 28
# initial trust radius from the network's start: of the order of its
# held-out prediction error in the standardised parameters, so the optimizer
# does not spend its first evaluations re-exploring the whole box
RHO_BEGIN_WARM = 0.03
# trust radii of the whitened minimisation from the network's start (z
# units: a step of length t changes D by about t^2/2 along any direction);
# 1e-3 reaches at least the final D of the default start's rho_end = 1e-4
RHO_BEGIN_Z, RHO_END_Z = 3.0, 1e-3
# box of the whitened coordinates (theta is also clipped into the bounds)
Z_BOX = 200.0
# central-difference step of the curvature estimate (theta units)
HESS_STEP = 0.02


def trend_cov(theta):
    a_s, bx_s, by_s, a_b, bx_b, lt = (float(v) for v in theta)
    return dict(nonstationary=1, metric=synth.METRIC["planar"], wendland_len=1.0,
                sigmasq=dict(kind=synth.FIELD_LINEAR, intercept=a_s, slope_x=bx_s, slope_y=by_s),
                beta=dict(kind=synth.FIELD_LINEAR, intercept=a_b, slope_x=bx_b, slope_y=0.0),
                tausq=dict(kind=synth.FIELD_CONSTANT, intercept=lt))


def make_locations(n, M, r, seed, J=4):
    rng = np.random.default_rng(seed)
    x, y = synth.uniform_points(n, rng)
    return capi.partition_program(x, y, np.zeros(n), None, None, M=M, J=J, r=r, bounds=synth.UNIT_BOUNDS)


def draw_theta(rng, k):
    return PRIOR_LO + (PRIOR_HI - PRIOR_LO) * rng.random((k, NTHETA))


def simulate(handle, theta, seed):
    """Draw a field from the MRA prior with parameters theta into the handle's
    observations; returns its features."""
    handle.set_covariance(trend_cov(theta))
    handle.set_params(**synth.UNIT_PARAMS)
    handle.sample_prior(seed, replace=True)
    return handle.variogram_features()


def training_set(n_fields, n, M, r, seed=2026, log=None):
    tree = make_locations(n, M, r, seed)
    h = capi.Handle(tree, with_pred=False)
    rng = np.random.default_rng(seed + 1)
    th = draw_theta(rng, n_fields)
    X = np.zeros((n_fields, 64), dtype=np.float32)
    try:
        for i in range(n_fields):
            X[i] = simulate(h, th[i], seed * 1000 + i)
            if log and (i + 1) % 1024 == 0:
                log(f"{i + 1} fields simulated")
    finally:
        h.close()
    return X, th


def trend_D(handle, theta):
    """D at theta (clipped into the optimizer's bounds); inf where a
    factorisation fails, as the library's own objective."""
    th = np.clip(np.asarray(theta, dtype=np.float64), BOUND_LO, BOUND_HI)
    handle.set_covariance(trend_cov(th))
    handle.set_params(**synth.UNIT_PARAMS)
    try:
        return handle.likelihood()
    except capi.MraError as e:
        if e.code == capi.MRA_E_NOT_PD:
            return float("inf")
        raise


def hessian(handle, theta, step=HESS_STEP):
    """Central-difference Hessian of D at theta (1 + 2k + 2k(k-1) evaluations)."""
    k = len(theta)
    E = np.eye(k) * step
    D0 = trend_D(handle, theta)
    H = np.zeros((k, k))
    for i in range(k):
        H[i, i] = (trend_D(handle, theta + E[i]) - 2.0 * D0 + trend_D(handle, theta - E[i])) / step ** 2
        for j in range(i):
            H[i, j] = H[j, i] = (trend_D(handle, theta + E[i] + E[j]) - trend_D(handle, theta + E[i] - E[j])
                                 - trend_D(handle, theta - E[i] + E[j])
                                 + trend_D(handle, theta - E[i] - E[j])) / (4.0 * step ** 2)
    return H


def whitening(H):
    """S with S^T H S = I for a symmetric positive definite H (S = L^-T)."""
    L = np.linalg.cholesky(0.5 * (H + H.T))
    return np.linalg.inv(L).T


class WarmStart:
    """Trained network plus the standardisation of features and targets, and
    (optional) the mean curvature of D at the optimum."""

    def __init__(self, params, fmean, fstd, tmean, tstd, hess=None):
        self.params = np.asarray(params, dtype=np.float32)
        self.fmean, self.fstd = np.asarray(fmean, np.float32), np.asarray(fstd, np.float32)
        self.tmean, self.tstd = np.asarray(tmean, np.float64), np.asarray(tstd, np.float64)
        self.hess = None if hess is None else np.asarray(hess, np.float64)
        self.mlp = capi.Mlp(0)
        self.mlp.set_params(self.params)

    def predict(self, features, bf16=True):
        Xs = (np.atleast_2d(features) - self.fmean) / self.fstd
        Y = self.mlp.forward(Xs.astype(np.float32), bf16=bf16).astype(np.float64)
        th = Y * self.tstd + self.tmean
        return np.clip(th, BOUND_LO + 1e-6, BOUND_HI - 1e-6)

    def save(self, path):
        extra = {} if self.hess is None else dict(hess=self.hess)
        np.savez(path, params=self.params, fmean=self.fmean, fstd=self.fstd, tmean=self.tmean, tstd=self.tstd,
                 **extra)

    @staticmethod
    def load(path):
        z = np.load(path, allow_pickle=False)
        return WarmStart(z["params"], z["fmean"], z["fstd"], z["tmean"], z["tstd"],
                         z["hess"] if "hess" in z.files else None)

    def fit(self, handle, features, max_evals=MAX_EVALS):
        """The optimizer from the network's start: in whitened coordinates
        when the model carries a curvature, else from theta0 in the plain
        ones.  Returns (theta, D, n_evals, status)."""
        t0 = self.predict(features)[0]
        if self.hess is None:
            return fit(handle, t0, max_evals, RHO_BEGIN_WARM)
        return fit_whitened(handle, t0, whitening(self.hess), max_evals)


def train(X, theta, epochs=300, batch=256, lr=1e-3, seed=0, val_frac=0.1, log=None):
    """Adam on standardised features / targets; returns (WarmStart, history)."""
    rng = np.random.default_rng(seed)
    n = X.shape[0]
    perm = rng.permutation(n)
    nv = max(1, int(round(val_frac * n))) if n > 10 else 0
    vi, ti = perm[:nv], perm[nv:]
    fmean, fstd = X[ti].mean(0), X[ti].std(0) + 1e-6
    tmean, tstd = theta[ti].mean(0), theta[ti].std(0) + 1e-12
    Xs = ((X - fmean) / fstd).astype(np.float32)
    Ts = ((theta - tmean) / tstd).astype(np.float32)
    m = capi.Mlp(seed)
    hist = []
    best = (float("inf"), None)  # early stopping on the held-out loss
    for ep in range(epochs):
        order = rng.permutation(ti)
        tot = 0.0
        for b0 in range(0, len(order), batch):
            bi = order[b0:b0 + batch]
            tot += m.train_step(Xs[bi], Ts[bi], lr) * len(bi)
        tl = tot / len(ti)
        vl = float(np.mean((m.forward(Xs[vi]) - Ts[vi]) ** 2)) if nv else float("nan")
        hist.append((tl, vl))
        if nv and vl < best[0]:
            best = (vl, m.get_params())
        if log and (ep % 25 == 0 or ep == epochs - 1):
            log(f"epoch {ep} train {tl:.4f} val {vl:.4f}")
    ws = WarmStart(best[1] if best[1] is not None else m.get_params(), fmean, fstd, tmean, tstd)
    m.close()
    return ws, hist


def fit(handle, theta0, max_evals=MAX_EVALS, rho_begin=RHO_BEGIN, npt=NPT):
    """The optimizer from theta0; returns (theta, D, n_evals, status)."""
    return handle.fit_linear_trend(theta0, BOUND_LO, BOUND_HI, rho_begin, RHO_END, max_evals, npt=npt)


def fit_whitened(handle, theta0, S, max_evals=MAX_EVALS, rho_begin=RHO_BEGIN_Z, rho_end=RHO_END_Z, npt=NPT):
    """The optimizer (mra_bobyqa) over z with theta = theta0 + S z, from z = 0;
    leaves the handle at the best theta.  Returns (theta, D, n_evals, status)."""
    theta0 = np.asarray(theta0, dtype=np.float64)
    S = np.asarray(S, dtype=np.float64)
    box = np.full(NTHETA, Z_BOX)
    z, D, ne, st = capi.bobyqa(lambda z: trend_D(handle, theta0 + S @ z), np.zeros(NTHETA), -box, box, rho_begin,
                               rho_end, max_evals, npt=npt)
    th = np.clip(theta0 + S @ z, BOUND_LO, BOUND_HI)
    trend_D(handle, th)
    return th, D, ne, st


def curvature(ws, n_fields, n, M, r, seed=31, log=None):
    """Mean Hessian of D at the optimum over fields of the training
    distribution (each fitted from the network's start); returns (H, the
    per-field Hessians)."""
    tree = make_locations(n, M, r, seed)
    h = capi.Handle(tree, with_pred=False)
    rng = np.random.default_rng(seed + 1)
    th = draw_theta(rng, n_fields)
    Hs = []
    try:
        for i in range(n_fields):
            f = simulate(h, th[i], seed * 1000 + i)
            ts, _, ne, _ = fit(h, ws.predict(f)[0], rho_begin=RHO_BEGIN_WARM)
            Hs.append(hessian(h, ts))
            if log:
                ev = np.linalg.eigvalsh(Hs[-1])
                log(json.dumps(dict(field=i, evals=ne, eig_min=float(ev[0]), eig_max=float(ev[-1]))))
    finally:
        h.close()
    Hs = np.array(Hs)
    return Hs.mean(axis=0), Hs


def compare(ws, n_fields, n, M, r, seed=77, log=None, max_evals=MAX_EVALS, rho_warm=RHO_BEGIN_WARM):
    """Held-out fields: optimizer evaluations and final D from the network's
    start (whitened when the model carries a curvature) vs the default
    start."""
    tree = make_locations(n, M, r, seed)
    h = capi.Handle(tree, with_pred=False)
    rng = np.random.default_rng(seed + 1)
    th = draw_theta(rng, n_fields)
    rows = []
    try:
        for i in range(n_fields):
            # the fits change only the covariance model, never the observations
            f = simulate(h, th[i], seed * 1000 + i)
            t0 = ws.predict(f)[0] if ws is not None else None
            _, Dd, ned, sd = fit(h, DEFAULT_START, max_evals)
            row = dict(field=i, theta=th[i].tolist(), D_default=Dd, evals_default=ned, status_default=sd)
            if t0 is not None and ws.hess is not None:
                _, Dw, new, sw = fit_whitened(h, t0, whitening(ws.hess), max_evals)
                row.update(D_warm=Dw, evals_warm=new, status_warm=sw, theta0=t0.tolist())
            elif t0 is not None:
                _, Dw, new, sw = fit(h, t0, max_evals, rho_warm)
                row.update(D_warm=Dw, evals_warm=new, status_warm=sw, theta0=t0.tolist())
            rows.append(row)
            if log:
                log(json.dumps({k: (round(v, 4) if isinstance(v, float) else v) for k, v in row.items()
                                if k not in ("theta", "theta0")}))
    finally:
        h.close()
    return rows


def summarize(rows):
    ed = np.array([r["evals_default"] for r in rows], dtype=float)
    out = dict(fields=len(rows), mean_evals_default=float(ed.mean()))
    if "evals_warm" in rows[0]:
        ew = np.array([r["evals_warm"] for r in rows], dtype=float)
        dD = np.array([r["D_warm"] - r["D_default"] for r in rows])
        out.update(mean_evals_warm=float(ew.mean()), evals_ratio=float(ed.mean() / ew.mean()),
                   median_evals_ratio=float(np.median(ed / ew)),
                   mean_D_warm_minus_default=float(dD.mean()), max_D_warm_minus_default=float(dD.max()))
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("cmd", choices=["train", "curvature", "evaluate"])
    ap.add_argument("--fields", type=int, default=2048)
    ap.add_argument("--n", type=int, default=50000)
    ap.add_argument("--M", type=int, default=8)
    ap.add_argument("--r", type=int, default=32)
    ap.add_argument("--epochs", type=int, default=300)
    ap.add_argument("--out", default=None)
    ap.add_argument("--model", default=None)
    ap.add_argument("--max-evals", type=int, default=MAX_EVALS)
    ap.add_argument("--rho-warm", type=float, default=RHO_BEGIN_WARM)
    ap.add_argument("--seed", type=int, default=2026)
    a = ap.parse_args(argv)
    log = lambda s: print(f"[warmstart] {s}", file=sys.stderr, flush=True)
    if a.cmd == "train":
        t0 = time.time()
        X, th = training_set(a.fields, a.n, a.M, a.r, seed=a.seed, log=log)
        log(f"{a.fields} fields simulated in {time.time() - t0:.1f}s")
        ws, hist = train(X, th, epochs=a.epochs, log=log)
        log(f"trained in {time.time() - t0:.1f}s; final train {hist[-1][0]:.4f} val {hist[-1][1]:.4f}")
        if a.out:
            ws.save(a.out)
        print(json.dumps(dict(fields=a.fields, n=a.n, M=a.M, r=a.r, epochs=a.epochs, train_loss=hist[-1][0],
                              val_loss=hist[-1][1], best_val_loss=min(v for _, v in hist),
                              seconds=time.time() - t0)), flush=True)
    elif a.cmd == "curvature":
        # the mean curvature at the optimum, stored in the model file
        t0 = time.time()
        ws = WarmStart.load(a.model)
        H, Hs = curvature(ws, a.fields, a.n, a.M, a.r, seed=a.seed, log=log)
        Si = np.linalg.cholesky(H)
        conds = [float(np.linalg.cond(np.linalg.solve(Si, np.linalg.solve(Si, Hf).T))) for Hf in Hs]
        ws.hess = H
        ws.save(a.out or a.model)
        print(json.dumps(dict(fields=a.fields, n=a.n, M=a.M, r=a.r, seconds=time.time() - t0,
                              mean_cond=float(np.mean([np.linalg.cond(Hf) for Hf in Hs])),
                              mean_cond_whitened=float(np.mean(conds)), max_cond_whitened=float(np.max(conds)))),
              flush=True)
    else:
        ws = WarmStart.load(a.model) if a.model else None
        rows = compare(ws, a.fields, a.n, a.M, a.r, log=log, max_evals=a.max_evals, rho_warm=a.rho_warm)
        print(json.dumps(summarize(rows)), flush=True)


if __name__ == "__main__":
    main()
