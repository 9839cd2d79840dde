/*
 * CPU FP64 oracle for the nonstationary MRA likelihood / prediction path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library restates, region by region and in
 * the same operation order, what the reference program computes.  It is the
 * checker that the HIP product path is compared against; nothing in the
 * product links, loads or calls it.  Only tests/, __graft_entry__.smoke() and
 * the cpu_baseline leg of bench.py may use it.
 *
 * Each function cites the reference file:line it follows.  How each part is
 * pinned (closed forms, the dense-covariance brute force in oracle/dense.py,
 * the reference's example data set) is listed in oracle/README.md.
 */
#ifndef MRA_ORACLE_H
#define MRA_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Flat tree: the same layout as mra_tree_desc in include/mra.h, restated here
 * so that the oracle does not depend on the product's headers. */
typedef struct orc_tree {
    int32_t num_levels;      /* M */
    int32_t num_partitions;  /* J */
    int32_t num_knots;       /* r (knots per coarse region) */
    int32_t various_j;       /* 1: finest parents come from a parent file */
    uint64_t n_coarse;       /* regions at levels 0..M-2 */
    uint64_t n_finest;       /* regions at level M-1 */
    uint64_t n_obs, n_pred;
    const double *knot_lon, *knot_lat;   /* [n_coarse*r], region-major */
    const uint64_t *finest_parent;       /* [n_finest], global coarse id */
    const uint64_t *obs_offsets;         /* [n_finest+1] */
    const double *obs_lon, *obs_lat, *obs_val;
    const uint64_t *pred_offsets;        /* [n_finest+1] or NULL */
    const double *pred_lon, *pred_lat;
} orc_tree;

typedef struct orc_field {
    int32_t kind;            /* 0 constant, 1 Wendland basis, 2 linear trend */
    int32_t pad;
    uint64_t n_basis;
    double intercept;
    const double *basis_lon, *basis_lat, *basis_weight;
    double slope_x, slope_y;
} orc_field;

typedef struct orc_cov {
    int32_t nonstationary;   /* 1 Paciorek exponential, 0 stationary exponential */
    int32_t metric;          /* 0 chordal (x6371), 1 great circle (x6371), 2 planar */
    double wendland_len;
    orc_field sigmasq, beta, tausq;
} orc_cov;

typedef struct orc_params {
    double sill, range, nugget;                 /* stationary model */
    double sill_scale, range_scale, nugget_scale; /* nonstationary model */
} orc_params;

/* Pairwise covariances.  out[i1*n2+i2] as in the reference. */
int orc_cross_covariance(const orc_cov* cov, const orc_params* p,
                         const double* lon1, const double* lat1, uint64_t n1,
                         const double* lon2, const double* lat2, uint64_t n2, double* out);
int orc_variance_covariance(const orc_cov* cov, const orc_params* p,
                            const double* lon, const double* lat, uint64_t n, double* out);
/* Scaled field values sigma^2, beta, tau^2 at points (without the scales when
 * scales are 1). */
int orc_fields(const orc_cov* cov, const double* lon, const double* lat, uint64_t n,
               double* sig2, double* beta, double* tau2);

/* Full evaluation: loglikelihood D (= -2 loglik - n log(2 pi)); if tree has
 * prediction points and mean/var are non-NULL, posterior predictive mean and
 * variance in finest-region order.  Returns 0, or a negative code. */
int orc_evaluate(const orc_tree* t, const orc_cov* cov, const orc_params* p,
                 double* D, double* mean, double* var, int nthreads);

/* Per-level partial sums of D (finest level first, then levels M-2..0);
 * out has M entries.  Same order of summation as orc_evaluate. */
/* D of the likelihood evaluation (no prediction) and the wall seconds of
 * its phases (measurement only): phase_s[l] coarse prior level l (l < L),
 * phase_s[L] the finest level, phase_s[L + 1 + l] posterior level l; L = M - 1. */
int orc_evaluate_timed(const orc_tree* t, const orc_cov* cov, const orc_params* p, double* D, double* phase_s,
                       int nthreads);
int orc_evaluate_levels(const orc_tree* t, const orc_cov* cov, const orc_params* p,
                        double* level_terms, int nthreads);

/* Partition restatement (reference: build_partition_by_program).
 * Inputs: raw locations (NaN values mark prediction-only locations, which
 * the caller removes before this call), prediction locations.  Outputs are
 * written into caller buffers sized by orc_partition_program_sizes. */
int orc_partition_program_sizes(const double* lon, const double* lat, uint64_t n,
                                const double* plon, const double* plat, uint64_t np,
                                int M, int J, int r, double offset,
                                const double* bounds, uint64_t* sizes /*[4]: n_coarse, n_finest, n_obs_kept, n_pred_kept*/);
int orc_partition_program(const double* lon, const double* lat, const double* val, uint64_t n,
                          const double* plon, const double* plat, uint64_t np,
                          int M, int J, int r, double offset, const double* bounds,
                          double* knot_lon, double* knot_lat,
                          uint64_t* obs_offsets, double* obs_lon, double* obs_lat, double* obs_val, uint64_t* obs_index,
                          uint64_t* pred_offsets, double* pred_lon, double* pred_lat, uint64_t* pred_index);

/* Select the dense kernels: 0 = the oracle's own loops, 1 = MKL (libmkl_rt,
 * loaded at run time).  Returns the backend actually in use. */
int orc_set_blas(int use_mkl);

#ifdef __cplusplus
}
#endif
#endif
