// Bound-constrained derivative-free minimisation for the MRA objective
// (src/optimization.cpp:107-192 drives dlib::find_min_bobyqa with npt = 2n+1,
// rho_begin 0.1, rho_end 1e-6 and a cap on evaluations; the objective is D).
//
// Method: a trust-region method with quadratic interpolation models in the
// manner of Powell's BOBYQA (M.J.D. Powell, "The BOBYQA algorithm for bound
// constrained optimization without derivatives", Cambridge DAMTP 2009/NA06):
//  * npt interpolation points (2n+1 by default) inside the bounds;
//  * each new model Q_new = Q_old + D, where D interpolates the residuals
//    f_i - Q_old(y_i) with the least Frobenius norm of its Hessian (Powell's
//    symmetric Broyden update); the KKT system of that problem is solved
//    directly (n is small), and its inverse also gives the Lagrange functions
//    used to choose which point a new one replaces;
//  * the step minimises the model in the intersection of the trust region
//    and the bounds by projected truncated conjugate gradients;
//  * the trust-region radius follows the ratio of actual to predicted
//    reduction; the resolution rho decreases from rho_begin to rho_end
//    (x0.1 while large, geometric mean near the end), with a geometry step
//    that moves the farthest point when the model is not trusted.
// Host code only; it reaches the GPU through the library's C ABI.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <limits>
#include <vector>

#include "../../include/mra.h"

namespace {

typedef std::vector<double> Vec;

// Solve A x = b (dense, partial pivoting); A is k x k row-major, destroyed.
bool lu_solve(std::vector<double>& A, std::vector<double>& B, int k, int nrhs) {
    for (int c = 0; c < k; c++) {
        int p = c;
        for (int r = c + 1; r < k; r++)
            if (std::fabs(A[r * k + c]) > std::fabs(A[p * k + c])) p = r;
        if (std::fabs(A[p * k + c]) < 1e-300) return false;
        if (p != c) {
            for (int j = 0; j < k; j++) std::swap(A[c * k + j], A[p * k + j]);
            for (int j = 0; j < nrhs; j++) std::swap(B[c * nrhs + j], B[p * nrhs + j]);
        }
        for (int r = c + 1; r < k; r++) {
            const double f = A[r * k + c] / A[c * k + c];
            if (f == 0.0) continue;
            for (int j = c; j < k; j++) A[r * k + j] -= f * A[c * k + j];
            for (int j = 0; j < nrhs; j++) B[r * nrhs + j] -= f * B[c * nrhs + j];
        }
    }
    for (int c = k - 1; c >= 0; c--)
        for (int j = 0; j < nrhs; j++) {
            double s = B[c * nrhs + j];
            for (int q = c + 1; q < k; q++) s -= A[c * k + q] * B[q * nrhs + j];
            B[c * nrhs + j] = s / A[c * k + c];
        }
    return true;
}

struct Bobyqa {
    int n, m;
    std::function<double(const double*)> f;
    Vec lo, hi;
    std::vector<Vec> Y;  // interpolation points (absolute)
    Vec F;
    int kopt = 0;
    // model around Y[kopt]: Q(x) = F[kopt] + g.(x - xopt) + 1/2 (x - xopt)' H (x - xopt)
    Vec g, H;
    // KKT inverse (Lagrange functions), base point Y[kopt]
    std::vector<double> Winv;
    int nev = 0, maxev = 0;

    double dist2(const Vec& a, const Vec& b) const {
        double s = 0;
        for (int i = 0; i < n; i++) s += (a[i] - b[i]) * (a[i] - b[i]);
        return s;
    }
    double model(const Vec& x) const {  // Q(x) - F[kopt]
        const Vec& b = Y[kopt];
        double s = 0;
        for (int i = 0; i < n; i++) {
            const double di = x[i] - b[i];
            s += g[i] * di;
            for (int j = 0; j < n; j++) s += 0.5 * di * H[i * n + j] * (x[j] - b[j]);
        }
        return s;
    }
    // KKT matrix of the least-Frobenius-norm interpolation around Y[kopt]
    bool build_kkt(std::vector<double>& W) const {
        const int k = m + n + 1;
        W.assign((size_t)k * k, 0.0);
        const Vec& b = Y[kopt];
        for (int i = 0; i < m; i++) {
            for (int j = 0; j < m; j++) {
                double s = 0;
                for (int q = 0; q < n; q++) s += (Y[i][q] - b[q]) * (Y[j][q] - b[q]);
                W[i * k + j] = 0.5 * s * s;
            }
            W[i * k + m] = 1.0;
            W[m * k + i] = 1.0;
            for (int q = 0; q < n; q++) {
                W[i * k + m + 1 + q] = Y[i][q] - b[q];
                W[(m + 1 + q) * k + i] = Y[i][q] - b[q];
            }
        }
        return true;
    }
    // refresh Winv (identity right-hand sides) around the current Y[kopt]
    bool refresh_inverse() {
        const int k = m + n + 1;
        std::vector<double> W;
        build_kkt(W);
        Winv.assign((size_t)k * k, 0.0);
        for (int i = 0; i < k; i++) Winv[i * k + i] = 1.0;
        return lu_solve(W, Winv, k, k);
    }
    // Lagrange function values l_j(x), j < m (columns of Winv)
    void lagrange(const Vec& x, Vec& l) const {
        const int k = m + n + 1;
        const Vec& b = Y[kopt];
        Vec w(k, 0.0);
        for (int i = 0; i < m; i++) {
            double s = 0;
            for (int q = 0; q < n; q++) s += (Y[i][q] - b[q]) * (x[q] - b[q]);
            w[i] = 0.5 * s * s;
        }
        w[m] = 1.0;
        for (int q = 0; q < n; q++) w[m + 1 + q] = x[q] - b[q];
        l.assign(m, 0.0);
        // l_j(x) = sum_r Winv[r][j] w[r] (Winv symmetric)
        for (int j = 0; j < m; j++) {
            double s = 0;
            for (int r = 0; r < k; r++) s += Winv[r * k + j] * w[r];
            l[j] = s;
        }
    }
    // Q_new = Q_old + D (least Frobenius norm change), around Y[kopt]
    void update_model(const Vec& gold_at_base, const Vec& Hold, double cold_at_base) {
        const int k = m + n + 1;
        const Vec& b = Y[kopt];
        Vec r(k, 0.0);
        for (int i = 0; i < m; i++) {
            // residual f_i - Q_old(y_i), Q_old expanded around the base
            double q = cold_at_base;
            for (int a = 0; a < n; a++) {
                const double da = Y[i][a] - b[a];
                q += gold_at_base[a] * da;
                for (int c = 0; c < n; c++) q += 0.5 * da * Hold[a * n + c] * (Y[i][c] - b[c]);
            }
            r[i] = F[i] - q;
        }
        Vec sol(k, 0.0);
        for (int i = 0; i < k; i++) {
            double s = 0;
            for (int j = 0; j < k; j++) s += Winv[i * k + j] * r[j];
            sol[i] = s;
        }
        g = gold_at_base;
        H = Hold;
        for (int a = 0; a < n; a++) g[a] += sol[m + 1 + a];
        for (int i = 0; i < m; i++)
            for (int a = 0; a < n; a++)
                for (int c = 0; c < n; c++) H[a * n + c] += sol[i] * (Y[i][a] - b[a]) * (Y[i][c] - b[c]);
        // the constant: Q_new(base) = cold + sol[m]; the model is kept relative
        // to F[kopt], which interpolation makes equal to Q_new(base)
    }
    // trust-region step within the bounds: projected truncated CG
    Vec tr_step(double delta) const {
        const Vec& x = Y[kopt];
        Vec d(n, 0.0), gr = g;
        std::vector<char> fixed(n, 0);
        for (int outer = 0; outer <= n; outer++) {
            for (int i = 0; i < n; i++)
                if ((x[i] + d[i] <= lo[i] + 1e-15 && gr[i] > 0) || (x[i] + d[i] >= hi[i] - 1e-15 && gr[i] < 0))
                    fixed[i] = 1;
            Vec s(n, 0.0);
            double gg = 0;
            for (int i = 0; i < n; i++)
                if (!fixed[i]) {
                    s[i] = -gr[i];
                    gg += gr[i] * gr[i];
                }
            if (gg <= 1e-30) break;
            bool restart = false;
            for (int it = 0; it < n && !restart; it++) {
                Vec Hs(n, 0.0);
                double sHs = 0, gs = 0, ss = 0, ds = 0, dd = 0;
                for (int i = 0; i < n; i++) {
                    for (int j = 0; j < n; j++) Hs[i] += H[i * n + j] * s[j];
                    sHs += s[i] * Hs[i];
                    gs += gr[i] * s[i];
                    ss += s[i] * s[i];
                    ds += d[i] * s[i];
                    dd += d[i] * d[i];
                }
                if (ss <= 1e-30) break;
                // largest alpha inside the trust region
                const double disc = ds * ds + ss * (delta * delta - dd);
                const double atr = disc > 0 ? (-ds + std::sqrt(disc)) / ss : 0.0;
                // largest alpha inside the bounds
                double abd = std::numeric_limits<double>::infinity();
                int ib = -1;
                for (int i = 0; i < n; i++) {
                    if (s[i] > 0) {
                        const double a = (hi[i] - x[i] - d[i]) / s[i];
                        if (a < abd) abd = a, ib = i;
                    } else if (s[i] < 0) {
                        const double a = (lo[i] - x[i] - d[i]) / s[i];
                        if (a < abd) abd = a, ib = i;
                    }
                }
                const double acg = sHs > 0 ? -gs / sHs : std::numeric_limits<double>::infinity();
                double a = std::min(acg, std::min(atr, abd));
                if (a < 0) a = 0;
                for (int i = 0; i < n; i++) {
                    d[i] += a * s[i];
                    gr[i] += a * Hs[i];
                }
                if (a == atr) return d;  // on the trust-region boundary
                if (a == abd && ib >= 0) {
                    fixed[ib] = 1;
                    d[ib] = (s[ib] > 0 ? hi[ib] : lo[ib]) - x[ib];
                    restart = true;
                    break;
                }
                // conjugate direction
                double gn = 0;
                for (int i = 0; i < n; i++)
                    if (!fixed[i]) gn += gr[i] * gr[i];
                if (gn <= 1e-24 * (gg + 1e-300)) return d;
                const double beta = gn / gg;
                gg = gn;
                for (int i = 0; i < n; i++) s[i] = fixed[i] ? 0.0 : -gr[i] + beta * s[i];
            }
            if (!restart) break;
        }
        return d;
    }
    double eval(const Vec& x) {
        nev++;
        double v = f(x.data());
        if (!std::isfinite(v)) v = std::numeric_limits<double>::max() / 4;
        return v;
    }
    void set_opt() {
        kopt = 0;
        for (int i = 1; i < m; i++)
            if (F[i] < F[kopt]) kopt = i;
    }
};

}  // namespace

// Generic entry point (C callback objective); see mra.h.  h0 (n x n, may be
// NULL): the Hessian the first quadratic model starts from; the first model
// is then the interpolant closest to it in the Frobenius norm.
extern "C" int mra_bobyqa_h0(int n, int npt, mra_objective_fn fn, void* ctx, double* x, const double* lo,
                             const double* hi, double rho_begin, double rho_end, int max_evals, const double* h0,
                             double* f_best, int* n_evals) {
    if (n < 1 || n > 64 || !fn || !x || !lo || !hi || !(rho_begin > 0) || !(rho_end > 0) || rho_end > rho_begin ||
        max_evals < 2 * n + 2)
        return MRA_E_ARG;
    // 2n+1 as the reference (src/optimization.cpp:80,169), up to the
    // (n+1)(n+2)/2 points that determine a full quadratic
    if (npt <= 0) npt = 2 * n + 1;
    if (npt < 2 * n + 1 || npt > (n + 1) * (n + 2) / 2) return MRA_E_UNSUPPORTED;
    for (int i = 0; i < n; i++)
        if (!(hi[i] - lo[i] >= 2 * rho_begin)) return MRA_E_ARG;
    Bobyqa B;
    B.n = n;
    B.m = npt;
    B.f = [&](const double* p) { return fn(p, ctx); };
    B.lo.assign(lo, lo + n);
    B.hi.assign(hi, hi + n);
    B.maxev = max_evals;
    double rho = rho_begin, delta = rho_begin;
    // initial points: x0 (moved inside [lo + rho, hi - rho]) and x0 +- rho e_i
    Vec x0(x, x + n);
    for (int i = 0; i < n; i++) x0[i] = std::min(std::max(x0[i], lo[i] + rho), hi[i] - rho);
    B.Y.assign(B.m, x0);
    for (int i = 0; i < n; i++) {
        B.Y[1 + i][i] += rho;
        B.Y[1 + n + i][i] -= rho;
    }
    B.F.assign(B.m, 0.0);
    for (int i = 0; i < 2 * n + 1; i++) B.F[i] = B.eval(B.Y[i]);
    // beyond 2n+1: x0 + s_i rho e_i + s_j rho e_j for the pairs i < j in
    // order, s_i the side of the lower value along coordinate i (Powell's
    // initial points for npt > 2n+1)
    for (int k = 2 * n + 1, i = 0, j = 1; k < B.m; k++) {
        Vec y = x0;
        const double si = B.F[1 + i] <= B.F[1 + n + i] ? rho : -rho;
        const double sj = B.F[1 + j] <= B.F[1 + n + j] ? rho : -rho;
        y[i] += si;
        y[j] += sj;
        B.Y[k] = y;
        B.F[k] = B.eval(y);
        if (++j == n) j = ++i + 1;
    }
    B.set_opt();
    B.g.assign(n, 0.0);
    B.H.assign((size_t)n * n, 0.0);
    if (!B.refresh_inverse()) return MRA_E_STATE;
    Vec H0((size_t)n * n, 0.0);
    if (h0)
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) H0[i * n + j] = 0.5 * (h0[i * n + j] + h0[j * n + i]);
    B.update_model(Vec(n, 0.0), H0, B.F[B.kopt]);
    int status = MRA_OK;
    int short_steps = 0;
    while (true) {
        if (B.nev >= max_evals) {
            status = MRA_W_MAX_EVALS;
            break;
        }
        const Vec d = B.tr_step(delta);
        double dn = 0;
        for (int i = 0; i < n; i++) dn += d[i] * d[i];
        dn = std::sqrt(dn);
        // farthest point from the best one
        int kfar = -1;
        double dfar = 0;
        for (int i = 0; i < B.m; i++) {
            const double dd = std::sqrt(B.dist2(B.Y[i], B.Y[B.kopt]));
            if (i != B.kopt && dd > dfar) dfar = dd, kfar = i;
        }
        if (dn < 0.5 * rho) {
            // short step: improve the geometry if some point is far, else refine rho
            if (dfar > 2.0 * delta && short_steps < 2) {
                short_steps++;
                // move the farthest point to where its Lagrange function is largest
                // among the coordinate steps of length delta (within the bounds)
                Vec best = B.Y[B.kopt];
                double lbest = -1;
                Vec lv;
                for (int i = 0; i < n; i++)
                    for (int sgn = -1; sgn <= 1; sgn += 2) {
                        Vec cand = B.Y[B.kopt];
                        cand[i] = std::min(std::max(cand[i] + sgn * delta, lo[i]), hi[i]);
                        if (cand[i] == B.Y[B.kopt][i]) continue;
                        B.lagrange(cand, lv);
                        if (std::fabs(lv[kfar]) > lbest) lbest = std::fabs(lv[kfar]), best = cand;
                    }
                if (lbest < 0) {
                    short_steps = 2;
                    continue;
                }
                const Vec gold = B.g, Hold = B.H;
                const double fo = B.F[B.kopt];
                const Vec xo = B.Y[B.kopt];
                B.Y[kfar] = best;
                B.F[kfar] = B.eval(best);
                B.set_opt();
                // re-expand the old model around the (possibly new) base
                Vec gb = gold;
                const double cb = fo + [&] {
                    double s = 0;
                    for (int a = 0; a < n; a++) {
                        const double da = B.Y[B.kopt][a] - xo[a];
                        s += gold[a] * da;
                        for (int c = 0; c < n; c++) s += 0.5 * da * Hold[a * n + c] * (B.Y[B.kopt][c] - xo[c]);
                    }
                    return s;
                }();
                for (int a = 0; a < n; a++)
                    for (int c = 0; c < n; c++) gb[a] += Hold[a * n + c] * (B.Y[B.kopt][c] - xo[c]);
                if (!B.refresh_inverse()) return MRA_E_STATE;
                B.update_model(gb, Hold, cb);
                continue;
            }
            short_steps = 0;
            if (rho <= rho_end) break;
            // reduce the resolution (Powell's schedule)
            const double ratio = rho / rho_end;
            rho = ratio <= 16.0 ? rho_end : (ratio <= 250.0 ? std::sqrt(ratio) * rho_end : 0.1 * rho);
            delta = std::max(0.5 * delta, rho);
            continue;
        }
        short_steps = 0;
        Vec xn = B.Y[B.kopt];
        for (int i = 0; i < n; i++) xn[i] = std::min(std::max(xn[i] + d[i], lo[i]), hi[i]);
        const double pred = -B.model(xn);
        const double fn_new = B.eval(xn);
        const double fopt = B.F[B.kopt];
        const double ratio = pred > 0 ? (fopt - fn_new) / pred : -1.0;
        if (ratio <= 0.1)
            delta = 0.5 * delta;
        else if (ratio <= 0.7)
            delta = std::max(0.5 * delta, dn);
        else
            delta = std::max(0.5 * delta, 2.0 * dn);
        if (delta <= 1.5 * rho) delta = rho;
        // choose the point the new one replaces: largest |l_k(xn)| weighted by
        // the distance from the best point
        Vec lv;
        B.lagrange(xn, lv);
        int krep = -1;
        double wbest = -1;
        for (int i = 0; i < B.m; i++) {
            if (i == B.kopt && fn_new >= fopt) continue;
            const double dd = B.dist2(B.Y[i], B.Y[B.kopt]) / (delta * delta);
            const double w = std::fabs(lv[i]) * std::max(1.0, dd * dd);
            if (w > wbest) wbest = w, krep = i;
        }
        if (krep < 0) krep = kfar >= 0 ? kfar : 0;
        const Vec gold = B.g, Hold = B.H;
        const Vec xo = B.Y[B.kopt];
        const double fo = B.F[B.kopt];
        B.Y[krep] = xn;
        B.F[krep] = fn_new;
        B.set_opt();
        Vec gb = gold;
        double cb = fo;
        for (int a = 0; a < n; a++) {
            const double da = B.Y[B.kopt][a] - xo[a];
            cb += gold[a] * da;
            for (int c = 0; c < n; c++) {
                cb += 0.5 * da * Hold[a * n + c] * (B.Y[B.kopt][c] - xo[c]);
                gb[a] += Hold[a * n + c] * (B.Y[B.kopt][c] - xo[c]);
            }
        }
        if (!B.refresh_inverse()) return MRA_E_STATE;
        B.update_model(gb, Hold, cb);
        // no progress at the smallest radius: with the points close by, the
        // model is as good as this resolution allows, so refine rho
        if (ratio <= 0.1 && delta <= rho) {
            double dmax = 0;
            for (int i = 0; i < B.m; i++) dmax = std::max(dmax, std::sqrt(B.dist2(B.Y[i], B.Y[B.kopt])));
            if (dmax <= 2.0 * delta) {
                if (rho <= rho_end) break;
                const double rr = rho / rho_end;
                rho = rr <= 16.0 ? rho_end : (rr <= 250.0 ? std::sqrt(rr) * rho_end : 0.1 * rho);
                delta = std::max(0.5 * delta, rho);
            }
        }
    }
    for (int i = 0; i < n; i++) x[i] = B.Y[B.kopt][i];
    if (f_best) *f_best = B.F[B.kopt];
    if (n_evals) *n_evals = B.nev;
    return status;
}

extern "C" int mra_bobyqa(int n, int npt, mra_objective_fn fn, void* ctx, double* x, const double* lo,
                          const double* hi, double rho_begin, double rho_end, int max_evals, double* f_best,
                          int* n_evals) {
    return mra_bobyqa_h0(n, npt, fn, ctx, x, lo, hi, rho_begin, rho_end, max_evals, nullptr, f_best, n_evals);
}

// ------------------------------------------------------------------------
// The MRA objective: D as a function of the six parameters of the planar
// linear-trend model (log sigma^2 = a_s + bx_s x + by_s y, log beta = a_b +
// bx_b x, tau^2 = exp(log_tau2)), the extension of objective_nonstationary
// (src/optimization.cpp:107-156) from three scales to six field parameters.
namespace {
struct TrendCtx {
    mra_handle* h;
    int metric;
    int failed;
};

double trend_objective(const double* th, void* p) {
    TrendCtx* c = (TrendCtx*)p;
    mra_cov_desc cov;
    std::memset(&cov, 0, sizeof(cov));
    cov.nonstationary = 1;
    cov.metric = c->metric;
    cov.wendland_len = 1.0;
    cov.sigmasq.kind = MRA_FIELD_LINEAR;
    cov.sigmasq.intercept = th[0];
    cov.sigmasq.slope_x = th[1];
    cov.sigmasq.slope_y = th[2];
    cov.beta.kind = MRA_FIELD_LINEAR;
    cov.beta.intercept = th[3];
    cov.beta.slope_x = th[4];
    cov.tausq.kind = MRA_FIELD_CONSTANT;
    cov.tausq.intercept = th[5];
    double D = 0.0;
    if (mra_set_covariance(c->h, &cov) != MRA_OK || mra_set_params(c->h, 0.0, 1.0, 0.0, 1.0, 1.0, 1.0) != MRA_OK) {
        c->failed = 1;
        return std::numeric_limits<double>::infinity();
    }
    const int rc = mra_likelihood(c->h, &D);
    if (rc == MRA_E_NOT_PD) return std::numeric_limits<double>::infinity();
    if (rc != MRA_OK) {
        c->failed = 1;
        return std::numeric_limits<double>::infinity();
    }
    return D;
}
}  // namespace

extern "C" int mra_fit_linear_trend_npt(mra_handle* h, int metric, double* theta, const double* lo,
                                        const double* hi, double rho_begin, double rho_end, int max_evals, int npt,
                                        const double* h0, double* D_best, int* n_evals) {
    if (!h || !theta || !lo || !hi) return MRA_E_ARG;
    TrendCtx c{h, metric, 0};
    int rc = mra_bobyqa_h0(MRA_TREND_NPARAM, npt, trend_objective, &c, theta, lo, hi, rho_begin, rho_end, max_evals,
                           h0, D_best, n_evals);
    if (c.failed) return MRA_E_STATE;
    // leave the handle at the best parameters
    (void)trend_objective(theta, &c);
    return rc;
}

extern "C" int mra_fit_linear_trend(mra_handle* h, int metric, double* theta, const double* lo, const double* hi,
                                    double rho_begin, double rho_end, int max_evals, double* D_best, int* n_evals) {
    return mra_fit_linear_trend_npt(h, metric, theta, lo, hi, rho_begin, rho_end, max_evals, 0, nullptr, D_best,
                                    n_evals);
}
