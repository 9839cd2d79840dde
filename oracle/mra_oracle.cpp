// CPU FP64 oracle: restatement of the reference MRA likelihood / prediction
// path.  TEST INFRASTRUCTURE ONLY (the checker, never the product).
//
// Follows, operation by operation:
//   src/evaluate_covariance.cpp            (covariances, distances)
//   src/class_nonstationary_covariance.cpp (Wendland parameter fields)
//   src/class_approximation.cpp            (coarse prior, aggregation, posterior w/A update)
//   src/class_approximation-loop-regions-at-finest-level-in-prior.cpp
//   src/class_approximation-posterior-inference.cpp
//   src/class_approximation-predict.cpp
//   src/class_partition.cpp                (program-placed partition)
// Dense kernels are plain loops (column-major, the BLAS/LAPACK semantics the
// reference calls) or, for timing only, MKL loaded at run time.
#include "mra_oracle.h"

#include <cmath>
#include <cstring>
#include <cstdio>
#include <vector>
#include <algorithm>
#include <dlfcn.h>
#include <omp.h>

namespace {

const double kPi = M_PI;

// ---------------------------------------------------------------- BLAS layer
// Column-major.  Semantics of cblas_dgemm(ColMajor,Trans,NoTrans,...),
// dtrtrs_('L','N','N',...), dpotrf_('L',...).
typedef void (*cblas_dgemm_t)(int, int, int, int, int, int, double, const double*, int,
                              const double*, int, double, double*, int);
typedef void (*dtrtrs_t)(const char*, const char*, const char*, const int*, const int*,
                         const double*, const int*, double*, const int*, int*);
typedef void (*dpotrf_t)(const char*, const int*, double*, const int*, int*);
typedef void (*mkl_set_t)(int);

struct Blas {
    int mkl = 0;
    cblas_dgemm_t dgemm = nullptr;
    dtrtrs_t dtrtrs = nullptr;
    dpotrf_t dpotrf = nullptr;
} g_blas;

// C(MxN) += alpha * A^T * B, A is KxM (lda), B is KxN (ldb).
void gemm_tn(int M, int N, int K, double alpha, const double* A, int lda, const double* B, int ldb,
             double* C, int ldc) {
    if (M <= 0 || N <= 0 || K <= 0) return;
    if (g_blas.mkl) {
        g_blas.dgemm(102 /*ColMajor*/, 112 /*Trans*/, 111 /*NoTrans*/, M, N, K, alpha, A, lda, B, ldb,
                     1.0, C, ldc);
        return;
    }
    for (int j = 0; j < N; j++)
        for (int i = 0; i < M; i++) {
            const double* a = A + (size_t)i * lda;
            const double* b = B + (size_t)j * ldb;
            double s = 0;
            for (int k = 0; k < K; k++) s += a[k] * b[k];
            C[i + (size_t)j * ldc] += alpha * s;
        }
}

// B <- L^{-1} B, L lower n x n (ldl), B n x m (ldb).
void trsm_lower(int n, int m, const double* L, int ldl, double* B, int ldb) {
    if (n <= 0 || m <= 0) return;
    if (g_blas.mkl) {
        int info = 0;
        const char l = 'L', nn = 'N';
        g_blas.dtrtrs(&l, &nn, &nn, &n, &m, L, &ldl, B, &ldb, &info);
        return;
    }
    for (int c = 0; c < m; c++) {
        double* b = B + (size_t)c * ldb;
        for (int i = 0; i < n; i++) {
            double s = b[i];
            for (int k = 0; k < i; k++) s -= L[i + (size_t)k * ldl] * b[k];
            b[i] = s / L[i + (size_t)i * ldl];
        }
    }
}

// In-place lower Cholesky; upper triangle untouched.  Returns 0 or i+1 of the
// first non-positive pivot.
int potrf_lower(int n, double* A, int lda) {
    if (n <= 0) return 0;
    if (g_blas.mkl) {
        int info = 0;
        const char l = 'L';
        g_blas.dpotrf(&l, &n, A, &lda, &info);
        return info;
    }
    for (int j = 0; j < n; j++) {
        double d = A[j + (size_t)j * lda];
        for (int k = 0; k < j; k++) d -= A[j + (size_t)k * lda] * A[j + (size_t)k * lda];
        if (!(d > 0)) return j + 1;
        d = std::sqrt(d);
        A[j + (size_t)j * lda] = d;
        for (int i = j + 1; i < n; i++) {
            double s = A[i + (size_t)j * lda];
            for (int k = 0; k < j; k++) s -= A[i + (size_t)k * lda] * A[j + (size_t)k * lda];
            A[i + (size_t)j * lda] = s / d;
        }
    }
    return 0;
}

// Armadillo chol(X,"lower"): lower factor with a zeroed upper triangle.
int chol_lower_clean(int n, double* A) {
    int info = potrf_lower(n, A, n);
    for (int j = 1; j < n; j++)
        for (int i = 0; i < j; i++) A[i + (size_t)j * n] = 0.0;
    return info;
}

// ------------------------------------------------------------ geometry/model
// src/evaluate_covariance.cpp:28-37
inline double euclidean_distance(double x1, double y1, double z1, double x2, double y2, double z2) {
    double dx = x2 - x1, dy = y2 - y1, dz = z2 - z1;
    return std::sqrt(dx * dx + dy * dy + dz * dz);
}
// src/evaluate_covariance.cpp:9-25
inline double great_circle_dist(double lon1, double lon2, double lat1, double lat2) {
    lon1 *= kPi / 180;
    lon2 *= kPi / 180;
    lat1 *= kPi / 180;
    lat2 *= kPi / 180;
    double angle = std::sin(lat1) * std::sin(lat2) + std::cos(lat1) * std::cos(lat2) * std::cos(lon1 - lon2);
    if (angle > 1.0) angle = 1.0;
    if (angle < -1.0) angle = -1.0;
    return std::acos(angle);
}
// src/class_partition.cpp:507-517 (lon/lat degrees -> unit sphere)
inline void to_xyz(double lon, double lat, double& x, double& y, double& z) {
    double cosLat = std::cos(lat * kPi / 180);
    double cosLon = std::cos(lon * kPi / 180);
    double sinLat = std::sin(lat * kPi / 180);
    double sinLon = std::sin(lon * kPi / 180);
    x = cosLat * cosLon;
    y = cosLat * sinLon;
    z = sinLat;
}

struct Pt {
    double lon, lat, x, y, z;
};

struct Model {
    const orc_cov* c;
    orc_params p;
    std::vector<double> bx[3], by[3], bz[3];
    void init(const orc_cov* cov, const orc_params* prm) {
        c = cov;
        p = *prm;
        const orc_field* f[3] = {&cov->sigmasq, &cov->beta, &cov->tausq};
        for (int q = 0; q < 3; q++) {
            bx[q].assign(f[q]->n_basis, 0);
            by[q].assign(f[q]->n_basis, 0);
            bz[q].assign(f[q]->n_basis, 0);
            if (f[q]->kind == 1)
                for (uint64_t i = 0; i < f[q]->n_basis; i++)
                    to_xyz(f[q]->basis_lon[i], f[q]->basis_lat[i], bx[q][i], by[q][i], bz[q][i]);
        }
    }
    // src/class_nonstationary_covariance.cpp:479-573 (Wendland fields);
    // kinds 0/2 (constant, linear trend) are this build's additions.
    double field(int q, const Pt& s) const {
        const orc_field* f = q == 0 ? &c->sigmasq : (q == 1 ? &c->beta : &c->tausq);
        double v = f->intercept;
        if (f->kind == 2) return std::exp(f->intercept + f->slope_x * s.lon + f->slope_y * s.lat);
        if (f->kind == 1) {
            for (uint64_t i = 0; i < f->n_basis; i++) {
                double d;
                if (c->metric == 0)
                    d = euclidean_distance(s.x, s.y, s.z, bx[q][i], by[q][i], bz[q][i]);
                else if (c->metric == 1)
                    d = great_circle_dist(s.lon, f->basis_lon[i], s.lat, f->basis_lat[i]);
                else
                    d = euclidean_distance(s.lon, s.lat, 0.0, f->basis_lon[i], f->basis_lat[i], 0.0);
                d /= c->wendland_len;
                if (d >= 1) continue;
                v += std::pow(1 - d, 6) * (35 * d * d + 18 * d + 3) * f->basis_weight[i] / 3;
            }
        }
        return std::exp(v);
    }
    // src/evaluate_covariance.cpp:72-77, 207-216
    double dist(const Pt& a, const Pt& b) const {
        if (c->metric == 0) return 6371 * euclidean_distance(a.x, a.y, a.z, b.x, b.y, b.z);
        if (c->metric == 1) return 6371 * great_circle_dist(a.lon, b.lon, a.lat, b.lat);
        return euclidean_distance(a.lon, a.lat, 0.0, b.lon, b.lat, 0.0);
    }
    // Cross covariance, out[i1*n2+i2] (src/evaluate_covariance.cpp:56-84,155-234)
    void cross(const Pt* s1, int n1, const Pt* s2, int n2, double* out) const {
        if (!c->nonstationary) {
            for (int i1 = 0; i1 < n1; i1++)
                for (int i2 = 0; i2 < n2; i2++)
                    out[(size_t)i1 * n2 + i2] = p.sill * std::exp(-dist(s1[i1], s2[i2]) / p.range);
            return;
        }
        std::vector<double> sg1(n1), b1(n1), sg2(n2), b2(n2);
        for (int i = 0; i < n1; i++) {
            sg1[i] = p.sill_scale * field(0, s1[i]);
            b1[i] = p.range_scale * field(1, s1[i]);
        }
        for (int i = 0; i < n2; i++) {
            sg2[i] = p.sill_scale * field(0, s2[i]);
            b2[i] = p.range_scale * field(1, s2[i]);
        }
        for (int i1 = 0; i1 < n1; i1++)
            for (int i2 = 0; i2 < n2; i2++) {
                double distance = dist(s1[i1], s2[i2]);
                double bms = (b1[i1] * b1[i1] + b2[i2] * b2[i2]) / 2;
                double bp = b1[i1] * b2[i2];
                double covar = std::sqrt(sg1[i1] * sg2[i2]) * std::pow(bp / bms, 1.5);
                covar *= std::exp(-distance / std::sqrt(bms));
                out[(size_t)i1 * n2 + i2] = covar;
            }
    }
    // Variance-covariance with the nugget on the diagonal
    // (src/evaluate_covariance.cpp:100-132,248-313)
    void var(const Pt* s, int n, double* out) const {
        if (!c->nonstationary) {
            double diagonal = p.sill + p.nugget;
            for (int i1 = 0; i1 < n; i1++) {
                for (int i2 = 0; i2 < i1; i2++) {
                    double off = p.sill * std::exp(-dist(s[i1], s[i2]) / p.range);
                    out[(size_t)i1 * n + i2] = off;
                    out[(size_t)i2 * n + i1] = off;
                }
                out[(size_t)i1 * n + i1] = diagonal;
            }
            return;
        }
        std::vector<double> sg(n), b(n), t(n);
        for (int i = 0; i < n; i++) {
            sg[i] = p.sill_scale * field(0, s[i]);
            b[i] = p.range_scale * field(1, s[i]);
            t[i] = p.nugget_scale * field(2, s[i]);
        }
        for (int i1 = 0; i1 < n; i1++) {
            for (int i2 = 0; i2 < i1; i2++) {
                double distance = dist(s[i1], s[i2]);
                double bms = (b[i1] * b[i1] + b[i2] * b[i2]) / 2;
                double bp = b[i1] * b[i2];
                double covar = std::sqrt(sg[i1] * sg[i2]) * std::pow(bp / bms, 1.5);
                covar *= std::exp(-distance / std::sqrt(bms));
                out[(size_t)i1 * n + i2] = covar;
                out[(size_t)i2 * n + i1] = covar;
            }
            out[(size_t)i1 * n + i1] = sg[i1] + t[i1];
        }
    }
};

std::vector<Pt> make_points(const double* lon, const double* lat, size_t n) {
    std::vector<Pt> v(n);
    for (size_t i = 0; i < n; i++) {
        v[i].lon = lon[i];
        v[i].lat = lat[i];
        to_xyz(lon[i], lat[i], v[i].x, v[i].y, v[i].z);
    }
    return v;
}

// -------------------------------------------------------------- the engine
struct Engine {
    const orc_tree* t;
    Model m;
    int M, J, r, L;
    uint64_t nC, nF;
    std::vector<uint64_t> lvlStart;  // first global id of coarse level l (l=0..L-1)
    std::vector<Pt> knots, obs, pred;
    std::vector<uint64_t> chS, chE;  // children range of level L-1 regions (global ids; 0 = none)
    std::vector<char> working;
    bool predicting = false;

    // prior
    std::vector<std::vector<double>> KCw;       // coarse: l blocks r*r ; finest: G_{F,j} (r x n) * L
    std::vector<std::vector<double>> RChol;     // coarse r*r (prior, then posterior)
    // finest
    std::vector<std::vector<double>> LF, Sicy, SicB;  // SicB: H_j (n x r) * L, concatenated
    std::vector<std::vector<double>> wT, AT;        // wTilde (r x l), ATilde (r*r*l(l+1)/2)
    std::vector<std::vector<double>> KCcw, KCcA;    // posterior v (r) and P_j (r*r*l)
    std::vector<std::vector<double>> BT;            // BTilde per finest: L blocks r x p
    std::vector<double> predMean, predVar;          // finest-region order
    std::vector<double> levelTerm;                  // D partial sums
    // wall seconds per phase (measurement only): coarse prior per level,
    // finest level, posterior per level
    std::vector<double> tPrior, tPost;
    double tFinest = 0;
    int err = 0;

    int level_of(uint64_t id) const {
        for (int l = L - 1; l >= 0; l--)
            if (id >= lvlStart[l]) return l;
        return 0;
    }
    // src/class_approximation.cpp:36-43 and the VARIOUS_J variant
    // (…finest-level-in-prior.cpp:93-99): ancestors a_0..a_{L-1} of finest f.
    void finest_ancestors(uint64_t f, std::vector<uint64_t>& anc) const {
        anc.assign(L, 0);
        uint64_t a = t->finest_parent[f];
        anc[L - 1] = a;
        for (int l = L - 2; l >= 0; l--) {
            a = (a - 1) / J;
            anc[l] = a;
        }
    }
    void coarse_ancestors(uint64_t R, int l, std::vector<uint64_t>& anc) const {
        anc.assign(l, 0);
        uint64_t a = R;
        for (int i = 0; i < l; i++) {
            a = (a - 1) / J;
            anc[l - i - 1] = a;
        }
    }

    int setup(const orc_tree* tree, const orc_cov* cov, const orc_params* prm) {
        t = tree;
        M = t->num_levels;
        J = t->num_partitions;
        r = t->num_knots;
        L = M - 1;
        nC = t->n_coarse;
        nF = t->n_finest;
        if (M < 2 || (J != 2 && J != 4) || r < 1) return -1;
        m.init(cov, prm);
        lvlStart.assign(L, 0);
        uint64_t s = 0, cnt = 1;
        for (int l = 0; l < L; l++) {
            lvlStart[l] = s;
            s += cnt;
            cnt *= J;
        }
        if (s != nC) return -2;
        knots = make_points(t->knot_lon, t->knot_lat, nC * r);
        obs = make_points(t->obs_lon, t->obs_lat, t->n_obs);
        predicting = t->pred_offsets != nullptr && t->n_pred > 0;
        if (predicting) pred = make_points(t->pred_lon, t->pred_lat, t->n_pred);
        // children ranges (src/class_partition.cpp:746-769) and working flags
        // (src/class_partition.cpp:880-904): a coarse region works iff some
        // finest region lies below it.
        uint64_t nSecond = (L >= 1) ? (nC - lvlStart[L - 1]) : 0;
        chS.assign(nSecond, 0);
        chE.assign(nSecond, 0);
        working.assign(nC + nF, 0);
        for (uint64_t f = 0; f < nF; f++) {
            uint64_t gid = nC + f;
            uint64_t pr = t->finest_parent[f];
            if (pr < lvlStart[L - 1] || pr >= nC) return -3;
            uint64_t k = pr - lvlStart[L - 1];
            if (chS[k] == 0 || chS[k] > gid) chS[k] = gid;
            if (chE[k] == 0 || chE[k] < gid) chE[k] = gid;
            working[gid] = 1;
            uint64_t a = pr;
            working[a] = 1;
            while (a > 0) {
                a = (a - 1) / J;
                working[a] = 1;
            }
        }
        KCw.assign(nC + nF, {});
        RChol.assign(nC, {});
        LF.assign(nF, {});
        Sicy.assign(nF, {});
        SicB.assign(nF, {});
        wT.assign(nC + nF, {});
        AT.assign(nC + nF, {});
        KCcw.assign(nC, {});
        KCcA.assign(nC, {});
        if (predicting) {
            BT.assign(nF, {});
            predMean.assign(t->n_pred, 0);
            predVar.assign(t->n_pred, 0);
        }
        levelTerm.assign(M, 0);
        return 0;
    }

    // src/class_approximation.cpp:84-169
    void prior_coarse() {
        tPrior.assign(L, 0);
        for (int l = 0; l < L; l++) {
            const double t0 = omp_get_wtime();
            uint64_t n_l = (l + 1 < L ? lvlStart[l + 1] : nC) - lvlStart[l];
#pragma omp parallel for schedule(dynamic, 1)
            for (uint64_t i = 0; i < n_l; i++) {
                uint64_t R = lvlStart[l] + i;
                std::vector<uint64_t> anc;
                coarse_ancestors(R, l, anc);
                std::vector<double>& G = KCw[R];
                G.assign((size_t)l * r * r, 0);
                for (int j = 0; j < l; j++) {
                    double* Gj = G.data() + (size_t)j * r * r;
                    m.cross(&knots[R * r], r, &knots[anc[j] * r], r, Gj);
                    const std::vector<double>& Ga = KCw[anc[j]];
                    for (int k = 0; k < j; k++)
                        gemm_tn(r, r, r, -1.0, Ga.data() + (size_t)k * r * r, r, G.data() + (size_t)k * r * r, r, Gj, r);
                    trsm_lower(r, r, RChol[anc[j]].data(), r, Gj, r);
                }
                std::vector<double>& K = RChol[R];
                K.assign((size_t)r * r, 0);
                m.var(&knots[R * r], r, K.data());
                for (int k = 0; k < l; k++)
                    gemm_tn(r, r, r, -1.0, G.data() + (size_t)k * r * r, r, G.data() + (size_t)k * r * r, r, K.data(), r);
                int info = chol_lower_clean(r, K.data());
                if (info) {
#pragma omp critical
                    err = -10;
                }
            }
            tPrior[l] = omp_get_wtime() - t0;
        }
    }

    // src/class_approximation-loop-regions-at-finest-level-in-prior.cpp:15-409
    void prior_finest() {
        const double t0 = omp_get_wtime();
        std::vector<double> term(nF, 0);
#pragma omp parallel for schedule(dynamic, 1)
        for (uint64_t f = 0; f < nF; f++) {
            uint64_t F = nC + f;
            int n = (int)(t->obs_offsets[f + 1] - t->obs_offsets[f]);
            const Pt* S = obs.data() + t->obs_offsets[f];
            const double* y = t->obs_val + t->obs_offsets[f];
            std::vector<uint64_t> anc;
            finest_ancestors(f, anc);
            int nslices = L * (L + 1) / 2;
            wT[F].assign((size_t)r * L, 0);
            AT[F].assign((size_t)r * r * nslices, 0);
            if (n > 0) {
                std::vector<double>& G = KCw[F];
                G.assign((size_t)L * r * n, 0);
                std::vector<double> B((size_t)L * r * n);
                std::vector<double>& V = LF[f];
                V.assign((size_t)n * n, 0);
                for (int j = 0; j < L; j++) m.cross(S, n, &knots[anc[j] * r], r, G.data() + (size_t)j * r * n);
                m.var(S, n, V.data());
                for (int j = 0; j < L; j++) {
                    double* Gj = G.data() + (size_t)j * r * n;
                    const std::vector<double>& Ga = KCw[anc[j]];
                    for (int k = 0; k < j; k++)
                        gemm_tn(r, n, r, -1.0, Ga.data() + (size_t)k * r * r, r, G.data() + (size_t)k * r * n, r, Gj, r);
                    std::memcpy(B.data() + (size_t)j * r * n, Gj, sizeof(double) * r * n);
                    trsm_lower(r, n, RChol[anc[j]].data(), r, Gj, r);
                }
                for (int k = 0; k < L; k++)
                    gemm_tn(n, n, r, -1.0, G.data() + (size_t)k * r * n, r, G.data() + (size_t)k * r * n, r, V.data(), n);
                int info = potrf_lower(n, V.data(), n);
                if (info) {
#pragma omp critical
                    err = -11;
                }
                std::vector<double>& u = Sicy[f];
                u.assign(y, y + n);
                trsm_lower(n, 1, V.data(), n, u.data(), n);
                // SicB[j] = L_F^{-1} B_j^T  (n x r), stored j-major
                std::vector<double>& H = SicB[f];
                H.assign((size_t)L * n * r, 0);
                for (int j = 0; j < L; j++) {
                    double* Hj = H.data() + (size_t)j * n * r;
                    const double* Bj = B.data() + (size_t)j * r * n;
                    for (int c = 0; c < n; c++)
                        for (int q = 0; q < r; q++) Hj[c + (size_t)q * n] = Bj[q + (size_t)c * r];
                    trsm_lower(n, r, V.data(), n, Hj, n);
                }
                // src/class_approximation.cpp:52-82
                int count = 0;
                for (int j = 0; j < L; j++) {
                    const double* Hj = H.data() + (size_t)j * n * r;
                    double* wcol = wT[F].data() + (size_t)j * r;
                    for (int q = 0; q < r; q++) {
                        double s = 0;
                        for (int c = 0; c < n; c++) s += Hj[c + (size_t)q * n] * u[c];
                        wcol[q] = s;
                    }
                    for (int k = j; k < L; k++) {
                        double* slice = AT[F].data() + (size_t)(count++) * r * r;
                        gemm_tn(r, r, n, 1.0, Hj, n, H.data() + (size_t)k * n * r, n, slice, r);
                    }
                }
                double add = 0;
                int pos = 0;
                for (int i = 0; i < n; i++) {
                    add += u[i] * u[i] + 2 * std::log(V[pos]);
                    pos += n + 1;
                }
                term[f] = add;
            }
            if (predicting) predict_prepass(f, n, S, anc);
        }
        double s = 0;
        for (uint64_t f = 0; f < nF; f++) s += term[f];
        levelTerm[0] = s;
        tFinest = omp_get_wtime() - t0;
    }

    // src/class_approximation-loop-regions-at-finest-level-in-prior.cpp:271-392
    void predict_prepass(uint64_t f, int n, const Pt* S, const std::vector<uint64_t>& anc) {
        uint64_t F = nC + f;
        int p = (int)(t->pred_offsets[f + 1] - t->pred_offsets[f]);
        if (p == 0) return;
        const Pt* P = pred.data() + t->pred_offsets[f];
        std::vector<double> cov((size_t)L * r * p), pk((size_t)L * r * p), pL;
        for (int j = 0; j < L; j++) {
            double* Cj = cov.data() + (size_t)j * r * p;
            m.cross(P, p, &knots[anc[j] * r], r, Cj);
            const std::vector<double>& Ga = KCw[anc[j]];
            for (int k = 0; k < j; k++)
                gemm_tn(r, p, r, -1.0, Ga.data() + (size_t)k * r * r, r, pk.data() + (size_t)k * r * p, r, Cj, r);
            std::memcpy(pk.data() + (size_t)j * r * p, Cj, sizeof(double) * r * p);
            trsm_lower(r, p, RChol[anc[j]].data(), r, pk.data() + (size_t)j * r * p, r);
        }
        if (n > 0) {
            pL.assign((size_t)n * p, 0);
            m.cross(P, p, S, n, pL.data());
            for (int k = 0; k < L; k++)
                gemm_tn(n, p, r, -1.0, KCw[F].data() + (size_t)k * r * n, r, pk.data() + (size_t)k * r * p, r, pL.data(), n);
            trsm_lower(n, p, LF[f].data(), n, pL.data(), n);
        }
        std::vector<double> V((size_t)p * p);
        m.var(P, p, V.data());
        for (int k = 0; k < L; k++)
            gemm_tn(p, p, r, -1.0, pk.data() + (size_t)k * r * p, r, pk.data() + (size_t)k * r * p, r, V.data(), p);
        double* mean = predMean.data() + t->pred_offsets[f];
        double* var = predVar.data() + t->pred_offsets[f];
        if (n > 0) {
            gemm_tn(p, p, n, -1.0, pL.data(), n, pL.data(), n, V.data(), p);
            for (int i = 0; i < p; i++) {
                double s = 0;
                for (int c = 0; c < n; c++) s += pL[c + (size_t)i * n] * Sicy[f][c];
                mean[i] = s;
            }
        } else {
            for (int i = 0; i < p; i++) mean[i] = 0;
        }
        for (int i = 0; i < p; i++) var[i] = V[i + (size_t)i * p];
        BT[f] = cov;
        if (n > 0)
            for (int j = 0; j < L; j++)
                gemm_tn(r, p, n, -1.0, SicB[f].data() + (size_t)j * n * r, n, pL.data(), n, BT[f].data() + (size_t)j * r * p, r);
    }

    // src/class_approximation.cpp:186-264 (single worker: every child that
    // works is summed in index order)
    void aggregate(uint64_t cs, uint64_t ce, int col, int slice, std::vector<double>* w, std::vector<double>& A) {
        bool first = true;
        for (uint64_t c = cs; c <= ce && cs != 0; c++) {
            if (!working[c]) continue;
            const double* s = AT[c].data() + (size_t)slice * r * r;
            if (first) {
                A.assign(s, s + (size_t)r * r);
                if (w) w->assign(wT[c].data() + (size_t)col * r, wT[c].data() + (size_t)(col + 1) * r);
                first = false;
            } else {
                for (size_t i = 0; i < (size_t)r * r; i++) A[i] += s[i];
                if (w)
                    for (int i = 0; i < r; i++) (*w)[i] += wT[c][(size_t)col * r + i];
            }
        }
    }

    // src/class_approximation-posterior-inference.cpp:14-246 and
    // src/class_approximation.cpp:266-306
    void posterior() {
        tPost.assign(L, 0);
        for (int l = L - 1; l >= 0; l--) {
            const double t0 = omp_get_wtime();
            uint64_t n_l = (l + 1 < L ? lvlStart[l + 1] : nC) - lvlStart[l];
            std::vector<double> term(n_l, 0);
#pragma omp parallel for schedule(dynamic, 1)
            for (uint64_t i = 0; i < n_l; i++) {
                uint64_t R = lvlStart[l] + i;
                if (!working[R]) continue;
                uint64_t cs, ce;
                if (l == L - 1) {
                    cs = chS[R - lvlStart[L - 1]];
                    ce = chE[R - lvlStart[L - 1]];
                } else {
                    cs = R * J + 1;
                    ce = R * J + J;
                }
                std::vector<double> w, A;
                aggregate(cs, ce, l, (l + 1) * (l + 2) / 2 - 1, &w, A);
                std::vector<double>& Lr = RChol[R];
                double add = 0;
                for (int q = 0; q < r; q++) add += std::log(Lr[q + (size_t)q * r]);
                add = -2 * add;
                // chol(L L^T + A)
                std::vector<double> K((size_t)r * r);
                for (int c = 0; c < r; c++)
                    for (int q = 0; q < r; q++) {
                        double s = 0;
                        for (int k = 0; k < r; k++) s += Lr[q + (size_t)k * r] * Lr[c + (size_t)k * r];
                        K[q + (size_t)c * r] = s + A[q + (size_t)c * r];
                    }
                if (chol_lower_clean(r, K.data())) {
#pragma omp critical
                    err = -12;
                }
                Lr = K;
                std::vector<double>& v = KCcw[R];
                v = w;
                trsm_lower(r, 1, Lr.data(), r, v.data(), r);
                std::vector<double>& P = KCcA[R];
                P.assign((size_t)r * r * l, 0);
                if (l > 0) {
                    wT[R].assign((size_t)r * l, 0);
                    AT[R].assign((size_t)r * r * (l + 1) * l / 2, 0);
                    int idxChild = (l + 1) * (l + 2) / 2 - 1;
                    int idxCur = l * (l + 1) / 2;
                    for (int j = l - 1; j >= 0; j--) {
                        aggregate(cs, ce, j, --idxChild, &w, A);
                        double* Pj = P.data() + (size_t)j * r * r;
                        for (int c = 0; c < r; c++)
                            for (int q = 0; q < r; q++) Pj[q + (size_t)c * r] = A[c + (size_t)q * r];
                        trsm_lower(r, r, Lr.data(), r, Pj, r);
                        double* wcol = wT[R].data() + (size_t)j * r;
                        for (int q = 0; q < r; q++) {
                            double s = 0;
                            for (int k = 0; k < r; k++) s += Pj[k + (size_t)q * r] * v[k];
                            wcol[q] = w[q] - s;
                        }
                        for (int k = l - 1; k >= j; k--) {
                            aggregate(cs, ce, 0, --idxChild, nullptr, A);
                            double* out = AT[R].data() + (size_t)(--idxCur) * r * r;
                            std::memcpy(out, A.data(), sizeof(double) * r * r);
                            gemm_tn(r, r, r, -1.0, Pj, r, P.data() + (size_t)k * r * r, r, out, r);
                        }
                    }
                }
                double s2 = 0, vv = 0;
                for (int q = 0; q < r; q++) s2 += std::log(Lr[q + (size_t)q * r]);
                for (int q = 0; q < r; q++) vv += v[q] * v[q];
                add += 2 * s2 - vv;
                term[i] = add;
            }
            // free the children's messages (src/class_approximation.cpp:365-374)
            double s = 0;
            for (uint64_t i = 0; i < n_l; i++) s += term[i];
            levelTerm[L - l] = s;
            tPost[l] = omp_get_wtime() - t0;
        }
    }

    // src/class_approximation-predict.cpp:105-189
    void predict() {
#pragma omp parallel for schedule(dynamic, 1)
        for (uint64_t f = 0; f < nF; f++) {
            int p = (int)(t->pred_offsets[f + 1] - t->pred_offsets[f]);
            if (p == 0) continue;
            std::vector<uint64_t> anc;
            finest_ancestors(f, anc);
            double* mean = predMean.data() + t->pred_offsets[f];
            double* var = predVar.data() + t->pred_offsets[f];
            std::vector<double>& BTf = BT[f];
            int jl = L - 1;
            std::vector<double> KcB(BTf.data() + (size_t)jl * r * p, BTf.data() + (size_t)(jl + 1) * r * p);
            trsm_lower(r, p, RChol[anc[jl]].data(), r, KcB.data(), r);
            for (int k = 0; k < jl; k++)
                gemm_tn(r, p, r, -1.0, KCcA[anc[jl]].data() + (size_t)k * r * r, r, KcB.data(), r,
                        BTf.data() + (size_t)k * r * p, r);
            for (int j = L - 2; j > 0; j--) {
                KcB.assign(BTf.data() + (size_t)j * r * p, BTf.data() + (size_t)(j + 1) * r * p);
                trsm_lower(r, p, RChol[anc[j]].data(), r, KcB.data(), r);
                for (int k = 0; k < j; k++)
                    gemm_tn(r, p, r, -1.0, KCcA[anc[j]].data() + (size_t)k * r * r, r, KcB.data(), r,
                            BTf.data() + (size_t)k * r * p, r);
            }
            for (int j = 0; j <= L - 1; j++) {
                double* Z = BTf.data() + (size_t)j * r * p;
                trsm_lower(r, p, RChol[anc[j]].data(), r, Z, r);
                const std::vector<double>& v = KCcw[anc[j]];
                for (int i = 0; i < p; i++) {
                    double s = 0;
                    for (int q = 0; q < r; q++) s += Z[q + (size_t)i * r] * v[q];
                    mean[i] += s;
                }
                for (int i = 0; i < p; i++) {
                    double s = 0;
                    for (int q = 0; q < r; q++) s += Z[q + (size_t)i * r] * Z[q + (size_t)i * r];
                    var[i] += s;
                }
            }
        }
    }
};

}  // namespace

extern "C" {

int orc_set_blas(int use_mkl) {
    if (!use_mkl) {
        g_blas.mkl = 0;
        return 0;
    }
    static void* h = nullptr;
    if (!h) h = dlopen("libmkl_rt.so", RTLD_NOW | RTLD_GLOBAL);
    if (!h) h = dlopen("/opt/conda/lib/libmkl_rt.so", RTLD_NOW | RTLD_GLOBAL);
    if (!h) return 0;
    g_blas.dgemm = (cblas_dgemm_t)dlsym(h, "cblas_dgemm");
    g_blas.dtrtrs = (dtrtrs_t)dlsym(h, "dtrtrs_");
    g_blas.dpotrf = (dpotrf_t)dlsym(h, "dpotrf_");
    mkl_set_t set1 = (mkl_set_t)dlsym(h, "MKL_Set_Num_Threads");
    if (set1) set1(1);
    g_blas.mkl = (g_blas.dgemm && g_blas.dtrtrs && g_blas.dpotrf) ? 1 : 0;
    return g_blas.mkl;
}

int orc_fields(const orc_cov* cov, const double* lon, const double* lat, uint64_t n, double* sig2, double* beta,
               double* tau2) {
    Model m;
    orc_params p{0, 0, 0, 1, 1, 1};
    m.init(cov, &p);
    std::vector<Pt> pts = make_points(lon, lat, n);
    for (uint64_t i = 0; i < n; i++) {
        sig2[i] = m.field(0, pts[i]);
        beta[i] = m.field(1, pts[i]);
        tau2[i] = m.field(2, pts[i]);
    }
    return 0;
}

int orc_cross_covariance(const orc_cov* cov, const orc_params* p, const double* lon1, const double* lat1,
                         uint64_t n1, const double* lon2, const double* lat2, uint64_t n2, double* out) {
    Model m;
    m.init(cov, p);
    std::vector<Pt> a = make_points(lon1, lat1, n1), b = make_points(lon2, lat2, n2);
    m.cross(a.data(), (int)n1, b.data(), (int)n2, out);
    return 0;
}

int orc_variance_covariance(const orc_cov* cov, const orc_params* p, const double* lon, const double* lat,
                            uint64_t n, double* out) {
    Model m;
    m.init(cov, p);
    std::vector<Pt> a = make_points(lon, lat, n);
    m.var(a.data(), (int)n, out);
    return 0;
}

static int run(const orc_tree* t, const orc_cov* cov, const orc_params* p, double* D, double* mean, double* var,
               double* levels, int nthreads, double* phase_s = nullptr) {
    if (nthreads > 0) omp_set_num_threads(nthreads);
    Engine e;
    int rc = e.setup(t, cov, p);
    if (rc) return rc;
    e.prior_coarse();
    if (e.err) return e.err;
    e.prior_finest();
    if (e.err) return e.err;
    e.posterior();
    if (e.err) return e.err;
    // D accumulation order: finest level, then levels M-2 .. 0
    // (…finest-level-in-prior.cpp:402, …posterior-inference.cpp:136,219)
    double d = 0;
    for (int i = 0; i < e.M; i++) d += e.levelTerm[i];
    if (D) *D = d;
    if (levels)
        for (int i = 0; i < e.M; i++) levels[i] = e.levelTerm[i];
    if (phase_s) {
        for (int l = 0; l < e.L; l++) phase_s[l] = e.tPrior[l];
        phase_s[e.L] = e.tFinest;
        for (int l = 0; l < e.L; l++) phase_s[e.L + 1 + l] = e.tPost[l];
    }
    if (e.predicting && mean && var) {
        e.predict();
        std::memcpy(mean, e.predMean.data(), sizeof(double) * t->n_pred);
        std::memcpy(var, e.predVar.data(), sizeof(double) * t->n_pred);
    }
    return 0;
}

int orc_evaluate(const orc_tree* t, const orc_cov* cov, const orc_params* p, double* D, double* mean, double* var,
                 int nthreads) {
    return run(t, cov, p, D, mean, var, nullptr, nthreads);
}

int orc_evaluate_timed(const orc_tree* t, const orc_cov* cov, const orc_params* p, double* D, double* phase_s,
                       int nthreads) {
    orc_tree t2 = *t;
    t2.pred_offsets = nullptr;
    t2.n_pred = 0;
    return run(&t2, cov, p, D, nullptr, nullptr, nullptr, nthreads, phase_s);
}

int orc_evaluate_levels(const orc_tree* t, const orc_cov* cov, const orc_params* p, double* level_terms,
                        int nthreads) {
    orc_tree t2 = *t;
    t2.pred_offsets = nullptr;
    t2.n_pred = 0;
    return run(&t2, cov, p, nullptr, nullptr, nullptr, level_terms, nthreads);
}

// ------------------------------------------------------- partition (program)
// src/class_partition.cpp:24-45 (create_knots), :49-89 (create_partitions),
// :459-654 (build_partition_by_program), :778-797,820-826 (region counts).
namespace {
struct Boxes {
    std::vector<double> lonMin, lonMax, latMin, latMax;
};
void build_boxes(int M, int J, const double* bounds, uint64_t nTotal, Boxes& b) {
    b.lonMin.assign(nTotal, 0);
    b.lonMax.assign(nTotal, 0);
    b.latMin.assign(nTotal, 0);
    b.latMax.assign(nTotal, 0);
    b.lonMin[0] = bounds[0];
    b.lonMax[0] = bounds[1];
    b.latMin[0] = bounds[2];
    b.latMax[0] = bounds[3];
    uint64_t start = 0, cnt = 1;
    for (int l = 0; l < M - 1; l++) {
        for (uint64_t R = start; R < start + cnt; R++) {
            double lonMin = b.lonMin[R], lonMax = b.lonMax[R], latMin = b.latMin[R], latMax = b.latMax[R];
            if (J == 2) {
                if ((lonMax - lonMin) >= (latMax - latMin)) {
                    double lonMid = (lonMax + lonMin) / 2;
                    b.lonMin[R * 2 + 1] = lonMin; b.lonMin[R * 2 + 2] = lonMid;
                    b.lonMax[R * 2 + 1] = lonMid; b.lonMax[R * 2 + 2] = lonMax;
                    b.latMin[R * 2 + 1] = latMin; b.latMin[R * 2 + 2] = latMin;
                    b.latMax[R * 2 + 1] = latMax; b.latMax[R * 2 + 2] = latMax;
                } else {
                    double latMid = (latMax + latMin) / 2;
                    b.lonMin[R * 2 + 1] = lonMin; b.lonMin[R * 2 + 2] = lonMin;
                    b.lonMax[R * 2 + 1] = lonMax; b.lonMax[R * 2 + 2] = lonMax;
                    b.latMin[R * 2 + 1] = latMin; b.latMin[R * 2 + 2] = latMid;
                    b.latMax[R * 2 + 1] = latMid; b.latMax[R * 2 + 2] = latMax;
                }
            } else {
                double lonMid = (lonMax + lonMin) / 2, latMid = (latMax + latMin) / 2;
                b.lonMin[R * 4 + 1] = lonMin; b.lonMin[R * 4 + 2] = lonMid; b.lonMin[R * 4 + 3] = lonMin; b.lonMin[R * 4 + 4] = lonMid;
                b.lonMax[R * 4 + 1] = lonMid; b.lonMax[R * 4 + 2] = lonMax; b.lonMax[R * 4 + 3] = lonMid; b.lonMax[R * 4 + 4] = lonMax;
                b.latMin[R * 4 + 1] = latMin; b.latMin[R * 4 + 2] = latMin; b.latMin[R * 4 + 3] = latMid; b.latMin[R * 4 + 4] = latMid;
                b.latMax[R * 4 + 1] = latMid; b.latMax[R * 4 + 2] = latMid; b.latMax[R * 4 + 3] = latMax; b.latMax[R * 4 + 4] = latMax;
            }
        }
        start += cnt;
        cnt *= J;
    }
}
void create_knots(double* kl, double* kt, double lonMin, double lonMax, int nLon, double latMin, double latMax, int nLat,
                  double OFFSET) {
    double offLon = (lonMax - lonMin) * OFFSET, offLat = (latMax - latMin) * OFFSET;
    double lonStart = lonMin + offLon, lonEnd = lonMax - offLon;
    double latStart = latMin + offLat, latEnd = latMax - offLat;
    double incLon = (nLon == 1 ? 0 : (lonEnd - lonStart) / (nLon - 1));
    double incLat = (nLat == 1 ? 0 : (latEnd - latStart) / (nLat - 1));
    for (int i = 0; i < nLon; i++)
        for (int j = 0; j < nLat; j++) {
            kl[i * nLat + j] = lonEnd - (nLon - i - 1) * incLon;
            kt[i * nLat + j] = latEnd - (nLat - j - 1) * incLat;
        }
}
struct Layout {
    uint64_t nCoarse, nFinest, nTotal;
    int nKnots, nLon, nLat;
};
Layout layout(int M, int J, int r) {
    Layout o;
    o.nLon = (int)std::ceil(std::sqrt((double)r));
    o.nLat = r / o.nLon;
    o.nKnots = o.nLon * o.nLat;
    uint64_t tot = 1, cnt = 1;
    for (int l = 1; l < M - 1; l++) {
        cnt *= J;
        tot += cnt;
    }
    o.nCoarse = tot;
    o.nFinest = cnt * J;
    o.nTotal = tot + o.nFinest;
    return o;
}
}  // namespace

int orc_partition_program_sizes(const double* lon, const double* lat, uint64_t n, const double* plon, const double* plat,
                                uint64_t np, int M, int J, int r, double offset, const double* bounds, uint64_t* sizes) {
    Layout o = layout(M, J, r);
    std::vector<double> kl(o.nCoarse * o.nKnots), kt(o.nCoarse * o.nKnots);
    std::vector<uint64_t> oo(o.nFinest + 1), po(o.nFinest + 1);
    // Count by running the full assignment with sized-to-max buffers.
    std::vector<double> a(n), b(n), c(n), pa(np), pb(np);
    std::vector<uint64_t> oi(n), pi(np);
    std::vector<double> val(n, 0.0);
    int rc = orc_partition_program(lon, lat, val.data(), n, plon, plat, np, M, J, r, offset, bounds, kl.data(), kt.data(),
                                   oo.data(), a.data(), b.data(), c.data(), oi.data(), po.data(), pa.data(), pb.data(),
                                   pi.data());
    sizes[0] = o.nCoarse;
    sizes[1] = o.nFinest;
    sizes[2] = oo[o.nFinest];
    sizes[3] = po[o.nFinest];
    return rc;
}

int orc_partition_program(const double* lon, const double* lat, const double* val, uint64_t n, const double* plon,
                          const double* plat, uint64_t np, int M, int J, int r, double offset, const double* bounds,
                          double* knot_lon, double* knot_lat, uint64_t* obs_offsets, double* obs_lon, double* obs_lat,
                          double* obs_val, uint64_t* obs_index, uint64_t* pred_offsets, double* pred_lon,
                          double* pred_lat, uint64_t* pred_index) {
    Layout o = layout(M, J, r);
    Boxes b;
    build_boxes(M, J, bounds, o.nTotal, b);
    for (uint64_t R = 0; R < o.nCoarse; R++)
        create_knots(knot_lon + R * o.nKnots, knot_lat + R * o.nKnots, b.lonMin[R], b.lonMax[R], o.nLon, b.latMin[R],
                     b.latMax[R], o.nLat, offset);
    uint64_t oc = 0, pc = 0;
    for (uint64_t f = 0; f < o.nFinest; f++) {
        uint64_t R = o.nCoarse + f;
        double lonMin = b.lonMin[R], lonMax = b.lonMax[R], latMin = b.latMin[R], latMax = b.latMax[R];
        obs_offsets[f] = oc;
        for (uint64_t i = 0; i < n; i++) {
            if (lon[i] >= lonMin && lon[i] < lonMax && lat[i] >= latMin && lat[i] < latMax) {
                bool skip = false;
                uint64_t anc = R;
                for (int k = 0; k < M - 1 && !skip; k++) {
                    anc = (anc - 1) / J;
                    for (int q = 0; q < o.nKnots; q++)
                        if (knot_lon[anc * o.nKnots + q] == lon[i] && knot_lat[anc * o.nKnots + q] == lat[i]) {
                            skip = true;
                            break;
                        }
                }
                if (!skip) {
                    obs_lon[oc] = lon[i];
                    obs_lat[oc] = lat[i];
                    obs_val[oc] = val[i];
                    obs_index[oc] = i;
                    oc++;
                }
            }
        }
        if (pred_offsets) {
            pred_offsets[f] = pc;
            for (uint64_t i = 0; i < np; i++)
                if (plon[i] >= lonMin && plon[i] < lonMax && plat[i] >= latMin && plat[i] < latMax) {
                    pred_lon[pc] = plon[i];
                    pred_lat[pc] = plat[i];
                    pred_index[pc] = i;
                    pc++;
                }
        }
    }
    obs_offsets[o.nFinest] = oc;
    if (pred_offsets) pred_offsets[o.nFinest] = pc;
    return 0;
}

}  // extern "C"
