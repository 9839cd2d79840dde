/*
 * mra.h — C ABI of the MI355X-native nonstationary MRA library (libmra_hip.so).
 *
 * Plain pointers and sizes only.  Host arrays passed in are copied at the
 * call; outputs go to caller-allocated host buffers.  A handle owns all of its
 * device memory and one HIP stream.  A handle is not thread-safe; distinct
 * handles are independent.
 *
 * Each entry point names the reference interface it replaces
 * (hhuang90/nonstationary_mra, file:line).
 */
#ifndef MRA_H
#define MRA_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ----------------------------------------------------------------- status */
enum {
    MRA_OK = 0,
    MRA_E_ARG = -1,         /* invalid argument / descriptor */
    MRA_E_HIP = -2,         /* HIP runtime error */
    MRA_E_NOT_PD = -3,      /* a Cholesky pivot was not positive; see mra_last_error */
    MRA_E_STATE = -4,       /* call order (e.g. predict without prediction points) */
    MRA_E_NOMEM = -5,       /* device allocation failed */
    MRA_E_UNSUPPORTED = -6  /* configuration outside what this build supports */
};

enum { MRA_METRIC_CHORDAL = 0, MRA_METRIC_GREAT_CIRCLE = 1, MRA_METRIC_PLANAR = 2 };
enum { MRA_FIELD_CONSTANT = 0, MRA_FIELD_WENDLAND = 1, MRA_FIELD_LINEAR = 2 };
enum { MRA_PRECISION_FP64 = 0, MRA_PRECISION_MIXED = 1 };

typedef struct mra_handle mra_handle;

/* Flattened MRA structure (the arrays of Partition,
 * include/class_partition.hpp:14-61).  Region ids follow the reference's heap
 * order: coarse regions 0..n_coarse-1 (root 0, children of i are J*i+1..J*i+J),
 * finest regions n_coarse..n_coarse+n_finest-1.  Observations are the knots of
 * the finest level (OBS_AS_KNOTS_FLAG) and are grouped by finest region. */
typedef struct mra_tree_desc {
    int32_t num_levels;      /* M (NUM_LEVELS_M) */
    int32_t num_partitions;  /* J (NUM_PARTITIONS_J), 2 or 4 */
    int32_t num_knots;       /* r: knots per coarse region (Partition::nKnots) */
    int32_t various_j;       /* 1: finest parents come from a parent file */
    uint64_t n_coarse;       /* regions at levels 0..M-2 */
    uint64_t n_finest;       /* regions at level M-1 */
    uint64_t n_obs;          /* observations (sum of finest knot counts) */
    uint64_t n_pred;         /* prediction locations (0: none) */
    const double* knot_lon;  /* [n_coarse*r], knot q of region R at R*r+q */
    const double* knot_lat;
    const uint64_t* finest_parent; /* [n_finest] coarse id of each finest region's parent */
    const uint64_t* obs_offsets;   /* [n_finest+1] */
    const double* obs_lon;         /* [n_obs] grouped by finest region */
    const double* obs_lat;
    const double* obs_val;         /* residual values (Partition::knotsResidual) */
    const uint64_t* pred_offsets;  /* [n_finest+1] or NULL */
    const double* pred_lon;        /* [n_pred] grouped by finest region */
    const double* pred_lat;
} mra_tree_desc;

/* One log-parameter field (class_nonstationary_covariance.hpp:8-49):
 *   WENDLAND: log f(s) = intercept + sum_i w_i psi(d(s,c_i)/wendland_len),
 *             psi(d) = (1-d)^6 (35 d^2 + 18 d + 3)/3 for d < 1;
 *   LINEAR:   log f(s) = intercept + slope_x*lon + slope_y*lat  (planar fields);
 *   CONSTANT: log f(s) = intercept. */
typedef struct mra_field_desc {
    int32_t kind;
    int32_t reserved;
    uint64_t n_basis;
    double intercept;
    const double* basis_lon;
    const double* basis_lat;
    const double* basis_weight;
    double slope_x, slope_y;
} mra_field_desc;

/* Covariance model.  nonstationary=1: Paciorek exponential
 * C = sqrt(s1 s2) (b1 b2/bm^2)^{3/2} exp(-d/bm), bm^2 = (b1^2+b2^2)/2
 * (src/evaluate_covariance.cpp:155-313); nonstationary=0: sill*exp(-d/range)
 * (src/evaluate_covariance.cpp:56-132).  Metric: chordal or great-circle on
 * the unit sphere times 6371 (lon/lat in degrees), or planar (unit 1). */
typedef struct mra_cov_desc {
    int32_t nonstationary;
    int32_t metric;
    double wendland_len;
    mra_field_desc sigmasq, beta, tausq;
} mra_cov_desc;

/* ----------------------------------------------------------- lifecycle */
/* Upload the structure (replaces the Partition arrays consumed by
 * Approximation, src/class_approximation.cpp:16-34,410-463). */
int mra_create(const mra_tree_desc* tree, mra_handle** out);
void mra_destroy(mra_handle* h);
/* Message of the last error on this handle (or a global one if h is NULL). */
const char* mra_last_error(const mra_handle* h);

/* Covariance model (Nonstationary_covariance::load_data,
 * src/class_nonstationary_covariance.cpp:46-429, plus the global flags
 * NONSTATIONARY_FLAG / CHORDAL_DISTANCE_FLAG / WENDLAND_LEN). */
int mra_set_covariance(mra_handle* h, const mra_cov_desc* cov);

/* Approximation(sill, range, nugget, sill_scale, range_scale, nugget_scale)
 * (src/class_approximation.cpp:396-408). */
int mra_set_params(mra_handle* h, double sill, double range, double nugget,
                   double sill_scale, double range_scale, double nugget_scale);

/* MRA_PRECISION_FP64 (default) or MRA_PRECISION_MIXED: covariance evaluation
 * (finest observation blocks and the coarse knot covariances) in FP32
 * arithmetic and the conditioning products of the finest level on FP32
 * matrix cores; knot Cholesky factors, conditioning of knots and
 * log-determinants in FP64.  Finest regions above MRA_OPT_MIXED_FP64_ABOVE observations
 * (default 128) are conditioned in FP64 in either mode.  In mra_likelihood
 * (not mra_predict) with 64-knot blocks the row Gram products of the
 * posterior (X_j^T X_k, X_l^T X_j) also run on FP32 matrix cores, so the two
 * entry points' D agree to the mixed tolerance, not bit for bit.  The
 * rounding error of the mixed mode is relative to D: about 1e-8 of |D| per
 * observation on data drawn from the model, and larger where D is dominated
 * by a large quadratic form (data far from the model's parameters). */
int mra_set_precision(mra_handle* h, int precision);

/* Approximation::likelihood() (src/class_approximation.cpp:466-470):
 * writes D = Approximation::loglikelihood (= -2 loglik - n log(2*3.14159265359)). */
int mra_likelihood(mra_handle* h, double* D);

/* likelihood() + predict() (src/class_approximation-predict.cpp:15-200) in
 * one pass: D, and the posterior predictive mean / variance at the
 * prediction points in the handle's (finest-region) order. */
int mra_predict(mra_handle* h, double* D, double* mean, double* var);

/* likelihood() and the log-likelihood the reference prints
 * (src/class_approximation-posterior-inference.cpp:244):
 * loglik = -0.5 (D + n log(2*3.14159265359)) with n = n_observations
 * (NUM_OBSERVATIONS: every observation loaded, including any the program
 * partition dropped on ancestor knots) or, when 0, the handle's observation
 * count.  D may be NULL. */
int mra_loglikelihood(mra_handle* h, uint64_t n_observations, double* D, double* loglik);

/* Per-level partial sums of D: [0] finest level, [1..M-1] levels M-2..0. */
int mra_likelihood_levels(mra_handle* h, double* terms);

/* Draw y ~ N(0, Sigma_MRA) at the observation locations with the current
 * parameters (no counterpart in the reference: the synthetic fields of the
 * benchmark configurations are sampled from the MRA prior).  Writes n_obs
 * values in the handle's observation order and also replaces the handle's
 * observation values with them when replace != 0. */
int mra_sample_prior(mra_handle* h, uint64_t seed, double* y, int replace);

/* Replace the observation values (same order as obs_val at create). */
int mra_set_observations(mra_handle* h, const double* values);

/* Capture one complete evaluation (likelihood, or likelihood + prediction)
 * as a HIP graph and replay it on later calls (1), or launch kernels
 * directly (0).  A graph is captured per (likelihood | predict, precision)
 * on first use; parameter changes (mra_set_params) are picked up without
 * re-capture; mra_set_covariance and mra_set_option discard the graphs.
 * With graphs, mra_last_timings reports the whole replay as MRA_T_GRAPH. */
int mra_use_graph(mra_handle* h, int enable);

/* Engine options (no reference counterpart).
 * MRA_OPT_FACTORED_POSTERIOR: 1 (default) propagates the finest-but-one
 * level's posterior messages in factored form (A~ = X'^T X', an exact
 * reformulation by the Woodbury identity); 0 uses the reference's dense
 * update at every level.  mra_get_option reports whether the factored form is in effect.
 * MRA_OPT_GRAM_KERNEL: 1 (default) computes the posterior message blocks
 * A~_jk = X_j^T X_k - P_j^T P_k of 64-knot regions whose messages come from
 * stacked rows in a separate kernel, one wave per block; 0 computes them
 * inside the posterior kernel (same values up to
 * rounding).
 * MRA_OPT_WAVE_LOCAL16: 1 (default) runs the finest-level local step
 * (Sigma_F, its Cholesky, u, H) of regions with at most 16 observations one
 * wave per region; 0 one work group per region (same values up to
 * rounding).
 * MRA_OPT_CROSS16: 1 (default) conditions the last observations of each
 * finest-level segment (the observations below one finest-but-one region)
 * in the narrowest tile of 16, 32 or 64 columns that holds them; 0 uses the
 * segment's tile width throughout (same values).
 * MRA_OPT_CONCURRENT: 1 (default) issues the independent launches of one
 * step (the finest-level tile widths and length buckets, the dense and the
 * factored posterior regions of the finest-but-one level) on two more
 * streams of the handle, forked and joined by events; 0 issues everything
 * on the handle's stream (same values).  Handles evaluated through
 * mra_batch_likelihood use their one stream.
 * MRA_OPT_SUM_MESSAGES: 1 (default) has the message-block kernel
 * (MRA_OPT_GRAM_KERNEL) store the sum of the messages of each parent's
 * children when the parent level sums dense child messages, so that the
 * parent reads one block set instead of one per child; 0 stores every
 * region's message (same values up to rounding).  mra_get_option reports
 * whether the summed form is in effect.
 * MRA_OPT_XCD_MAP: 1 (default) has the coarse prior and finest conditioning
 * launches hand out their regions in 8 contiguous ranges, one per XCD
 * (work groups are dispatched to the XCDs round robin), so that neighbouring
 * regions, which read the same ancestor blocks, run under one L2; 0 takes
 * them in launch order (same values).
 * MRA_OPT_MIXED_FP64_ABOVE: in mixed precision, finest regions with more
 * than this many observations (default 128, in [64, 256]) are conditioned
 * and factored in FP64: in large, strongly correlated regions the FP32
 * rounding amplified through Sigma_F exceeds |dloglik|/n = 1e-6.
 * MRA_OPT_PERSISTENT: 1 launches the finest-level kernels (conditioning tiles
 * and every local length bucket) on a resident grid (as many work groups as
 * the device holds at once for that kernel) whose work groups walk the
 * length-bucketed work lists, longest first, by grid stride; 0 (default) one
 * work group (one wave for regions of at most 16 observations) per item.
 * Same values bit for bit.
 * MRA_OPT_PRIOR_WAVE: 0 (default, and the only accepted value) computes the
 * coarse prior one work group per region.  1 would select the one-wave-per-
 * region form for 64-knot regions (k_prior_wave + k_prior_chol); it is
 * refused with MRA_E_UNSUPPORTED: a GPU run with it stopped making progress
 * and the cause is not established.
 * MRA_OPT_FINEST_WAVE: 1 (default) runs the mixed-precision finest
 * conditioning of 32- and 64-knot regions one wave per tile (k_finest_cross_w: the
 * tile in the registers of the FP32 matrix instructions, no work-group
 * barrier); 0 one work group per tile (same values up to FP32 rounding).
 * Persistent launches (MRA_OPT_PERSISTENT) use the work-group form.
 * MRA_OPT_LOCAL_LL: the finest local step of regions of more than 64
 * observations as a left-looking Cholesky on 32-wide block columns
 * (k_finest_local_ll: panel tiles formed from G_F and the final rows of L_F,
 * no Sigma_F workspace, H and u by forward substitution in the same pass).
 * 1 (default): for every such region except those of 65-128 observations
 * when more than 256 ancestor knot columns enter their Gram product (those
 * stay on the workspace kernel, faster there); 2: for every such region;
 * 0: the earlier forms (a padded Sigma_F workspace with 16 x 16 blocks up to
 * 256 observations, an unblocked factorisation above).  Same values up to
 * rounding. */
enum {
    MRA_OPT_FACTORED_POSTERIOR = 1,
    MRA_OPT_GRAM_KERNEL = 2,
    MRA_OPT_WAVE_LOCAL16 = 3,
    MRA_OPT_CROSS16 = 4,
    MRA_OPT_CONCURRENT = 5,
    MRA_OPT_SUM_MESSAGES = 6,
    MRA_OPT_XCD_MAP = 7,
    MRA_OPT_MIXED_FP64_ABOVE = 8,
    MRA_OPT_PERSISTENT = 9,
    MRA_OPT_PRIOR_WAVE = 10,
    MRA_OPT_FINEST_WAVE = 11,
    MRA_OPT_LOCAL_LL = 12
};
int mra_set_option(mra_handle* h, int option, int value);
int mra_get_option(mra_handle* h, int option, int* value);

/* Timing of the last evaluation: milliseconds per launch group, measured
 * with HIP events on the handle's stream, indexed by the MRA_T_* ids. */
enum {
    MRA_T_SETUP = 0,          /* parameter scaling, clears */
    MRA_T_FINEST_CROSS = 1,   /* k_finest_cross */
    MRA_T_FINEST_LOCAL = 2,   /* k_finest_local */
    MRA_T_PRED_PRE = 3,
    MRA_T_PRED_POST = 4,
    MRA_T_REDUCE = 5,
    MRA_T_GRAPH = 6,          /* one replay of the captured evaluation */
    MRA_T_PRIOR0 = 8,         /* + level l: coarse prior of level l (l < 24) */
    MRA_T_POST0 = 32,         /* + level l: k_posterior of level l */
    MRA_T_COUNT = 56
};
int mra_last_timings(mra_handle* h, double* ms, int n);

/* ------------------------------------------------ single-kernel entries */
/* evaluate_nonstationary_cross_covariance / evaluate_cross_covariance
 * (src/evaluate_covariance.cpp:56-84,155-234): out[i1*n2+i2] = C(s1_i1, s2_i2).
 * Host pointers; fields and scales from the covariance model. */
int mra_cross_covariance(const mra_cov_desc* cov, double sill, double range, double nugget,
                         double sill_scale, double range_scale, double nugget_scale,
                         const double* lon1, const double* lat1, uint64_t n1,
                         const double* lon2, const double* lat2, uint64_t n2,
                         double* out);
/* evaluate_nonstationary_variance_covariance / evaluate_variance_covariance
 * (src/evaluate_covariance.cpp:100-132,248-313): nugget on the diagonal. */
int mra_variance_covariance(const mra_cov_desc* cov, double sill, double range, double nugget,
                            double sill_scale, double range_scale, double nugget_scale,
                            const double* lon, const double* lat, uint64_t n, double* out);
/* The same with the caller's unit-sphere coordinates x, y, z (the
 * reference's evaluate_* take them, src/evaluate_covariance.cpp:155-313):
 * used as given for the chordal metric; NULL (or another metric) derives
 * them from lon/lat. */
int mra_cross_covariance_xyz(const mra_cov_desc* cov, double sill, double range, double nugget,
                             double sill_scale, double range_scale, double nugget_scale,
                             const double* lon1, const double* lat1, const double* x1, const double* y1,
                             const double* z1, uint64_t n1, const double* lon2, const double* lat2,
                             const double* x2, const double* y2, const double* z2, uint64_t n2, double* out);
int mra_variance_covariance_xyz(const mra_cov_desc* cov, double sill, double range, double nugget,
                                double sill_scale, double range_scale, double nugget_scale,
                                const double* lon, const double* lat, const double* x, const double* y,
                                const double* z, uint64_t n, double* out);
/* Allocation-free form of mra_cross_covariance (no counterpart in the
 * reference; for callers that cannot allocate device memory: code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments and the same values bit for bit; every device buffer it needs is
 * carved from one static device array of
 * mra_cross_covariance_noalloc_limit() bytes (32 MiB) per GPU, held until
 * the process ends, so it calls no hipMalloc / hipFree.  Inputs beyond that
 * array (8 n1 n2 + 88 (n1 + n2) + 48 x basis functions bytes, each piece
 * rounded up to 256) are refused with MRA_E_ARG.  Calls that overlap on one
 * GPU share the array and race; run them one at a time per GPU. */
int mra_cross_covariance_noalloc(const mra_cov_desc* cov, double sill, double range, double nugget,
                                 double sill_scale, double range_scale, double nugget_scale, const double* lon1,
                                 const double* lat1, uint64_t n1, const double* lon2, const double* lat2,
                                 uint64_t n2, double* out);
uint64_t mra_cross_covariance_noalloc_limit(void);
/* Nonstationary_covariance::get_sigmasq / get_beta / get_tausq, both
 * overloads (src/class_nonstationary_covariance.cpp:479-573): the fields
 * exp(intercept + sum_i w_i psi(d_i / wendland_len)) (or the linear /
 * constant forms) at n points, unscaled.  lon/lat are required except for
 * the chordal metric with x, y, z given (used as given); any output may be
 * NULL. */
int mra_fields(const mra_cov_desc* cov, const double* lon, const double* lat, const double* x, const double* y,
               const double* z, uint64_t n, double* sigmasq, double* beta, double* tausq);
/* Device-resident form of evaluate_nonstationary_cross_covariance
 * (src/evaluate_covariance.cpp:155-234), the pairwise covariance kernel on
 * its own: d_out[i*n2+j] (row-major n1 x n2, as the reference writes it) for
 * points whose coordinates (d_xyz = [x | y | z], n each) and scaled sqrt-sill
 * and range (d_sb = [s | b]) are already on the device.  precision selects
 * FP64 (d_out is double*) or FP32 (MRA_PRECISION_MIXED, float*) output;
 * stream may be NULL.  n1 is limited to 65535*64 rows per call. */
int mra_cov_tile_device(int precision, int metric, const double* d_xyz1, const double* d_sb1, uint64_t n1,
                        const double* d_xyz2, const double* d_sb2, uint64_t n2, void* d_out, void* stream);
/* Throughput measurement of mra_cov_tile_device (no counterpart in the
 * reference, which prints wall-clock seconds per phase,
 * src/class_approximation.cpp:161-164): one untimed launch, then `launches`
 * launches on a private stream between two HIP events; *ms_per_launch is
 * the average.  Synchronous. */
int mra_cov_tile_timed(int precision, int metric, const double* d_xyz1, const double* d_sb1, uint64_t n1,
                       const double* d_xyz2, const double* d_sb2, uint64_t n2, void* d_out, int launches,
                       double* ms_per_launch);
/* Device-buffer utilities for callers without their own HIP allocator
 * (no counterpart in the reference, whose arrays are host memory): plain
 * hipMalloc / hipFree and a synchronous copy (to_device = 1: host to
 * device, 0: device to host). */
int mra_device_alloc(uint64_t bytes, void** d_ptr);
int mra_device_free(void* d_ptr);
int mra_device_copy(void* dst, const void* src, uint64_t bytes, int to_device);

/* ------------------------------------------------------------ optimizer */
/* Positive status: the minimiser stopped at its evaluation limit (the
 * reference's dlib::find_min_bobyqa throws bobyqa_failure there). */
enum { MRA_W_MAX_EVALS = 1 };
typedef double (*mra_objective_fn)(const double* x, void* ctx);
/* Bound-constrained derivative-free minimisation of fn over lo <= x <= hi
 * (src/optimization.cpp:80,169: dlib::find_min_bobyqa with npt = 2n+1):
 * quadratic interpolation models with least-Frobenius-norm updates, trust
 * region with bounds, resolution rho from rho_begin down to rho_end.  x is
 * the start on entry and the best point on return.  npt: 0 (= 2n+1, as the
 * reference) or any value from 2n+1 to (n+1)(n+2)/2 (a full quadratic model;
 * the points beyond 2n+1 step along pairs of coordinates).
 * Requires hi[i] - lo[i] >= 2 rho_begin and max_evals >= 2n+2.  Returns
 * MRA_OK, MRA_W_MAX_EVALS or a negative error. */
int mra_bobyqa(int n, int npt, mra_objective_fn fn, void* ctx, double* x, const double* lo, const double* hi,
               double rho_begin, double rho_end, int max_evals, double* f_best, int* n_evals);

/* The same with a starting curvature: h0 (n x n row-major, symmetrised; NULL
 * = none) is the Hessian the first quadratic model starts from, the first
 * model being the interpolant of the initial points whose Hessian is closest
 * to h0 in the Frobenius norm (no counterpart in the reference; the learned
 * warm start supplies it). */
int mra_bobyqa_h0(int n, int npt, mra_objective_fn fn, void* ctx, double* x, const double* lo, const double* hi,
                  double rho_begin, double rho_end, int max_evals, const double* h0, double* f_best, int* n_evals);

/* Maximum-likelihood fit of the six parameters of the planar linear-trend
 * nonstationary model, theta = (a_s, bx_s, by_s, a_b, bx_b, log_tau2):
 *   log sigma^2 = a_s + bx_s x + by_s y,  log beta = a_b + bx_b x,
 *   tau^2 = exp(log_tau2),
 * minimising D on the handle's observations (objective_nonstationary,
 * src/optimization.cpp:107-156, extended from three scales to six field
 * parameters).  theta: start on entry, best on return; the handle is left
 * at the best parameters.  metric: MRA_METRIC_*. */
enum { MRA_TREND_NPARAM = 6 };
int mra_fit_linear_trend(mra_handle* h, int metric, double* theta, const double* lo, const double* hi,
                         double rho_begin, double rho_end, int max_evals, double* D_best, int* n_evals);
/* The same with the minimiser's interpolation set size npt (0 = 2n+1 = 13,
 * up to 28 = a full quadratic model in the six parameters) and an optional
 * starting curvature h0 (6 x 6, as mra_bobyqa_h0; NULL = none). */
int mra_fit_linear_trend_npt(mra_handle* h, int metric, double* theta, const double* lo, const double* hi,
                             double rho_begin, double rho_end, int max_evals, int npt, const double* h0,
                             double* D_best, int* n_evals);

/* --------------------------------------------------------------- autotuner */
/* Work-list shape of a handle (no reference counterpart):
 *   x32max: segments (observations below one finest-but-one region) with at
 *           most x32max observations use 32-column tiles in the finest
 *           conditioning kernel, larger ones 64-column tiles (0, 32 or 64;
 *           -1: every segment uses 32-column tiles);
 *   lrmax:  finest-but-one regions with at most lrmax rows use the factored
 *           posterior, larger ones the dense update (64 or 128);
 *   fb16:   1 = finest regions of at most 16 observations have their own
 *           length bucket, 0 = they join the 32 bucket.
 * Results are the same up to rounding for every shape. */
int mra_set_tuning(mra_handle* h, int x32max, int lrmax, int fb16);
int mra_get_tuning(mra_handle* h, int* x32max, int* lrmax, int* fb16);
/* Times a likelihood evaluation for every candidate shape (covariance and
 * parameters must be set) and keeps the fastest; returns it and its time. */
int mra_autotune(mra_handle* h, int* x32max, int* lrmax, int* fb16, double* best_ms);

/* ------------------------------------------------------------------ batch */
/* Many independent fields (one handle each, any sizes) evaluated as one
 * launch sequence (no reference counterpart; the reference runs one data set
 * per MPI job): one kernel scatters every field's parameters, the fields'
 * evaluations run on `streams` (1..16) batch streams (largest first, each to
 * the least loaded stream), one kernel gathers every D and status flag.
 * Streams run side by side only on as many hardware queues: HIP's
 * GPU_MAX_HW_QUEUES (4 by default) must be set before the process first
 * uses the GPU to let more than four run at once.  With
 * graph = 1 the whole sequence is captured once as one HIP graph and
 * replayed; it is captured again when a handle's work lists, options,
 * covariance model or precision change (or, for a stationary model, its
 * sill or range).  The host waits once per call.  The handles must be
 * distinct, must stay alive while the batch exists and must not be used
 * concurrently from elsewhere; the batch does not use their own streams. */
typedef struct mra_batch mra_batch;
int mra_batch_create(mra_handle* const* handles, int n, mra_batch** out);
void mra_batch_destroy(mra_batch* b);
/* Message of the last error of this batch (or of mra_batch_create if b is NULL). */
const char* mra_batch_last_error(const mra_batch* b);
int mra_batch_set(mra_batch* b, int streams, int graph);
int mra_batch_get(mra_batch* b, int* streams, int* graph);
/* D[i] of handle i; status[i] = MRA_OK or MRA_E_NOT_PD (status may be NULL). */
int mra_batch_likelihood(mra_batch* b, double* D, int* status);
/* Times one batch evaluation for streams in {1,2,4,8,16} x graph in {0,1} and
 * keeps the fastest setting (returned; best_ms = its wall time). */
int mra_batch_autotune(mra_batch* b, int* streams, int* graph, double* best_ms);

/* ---------------------------------------------------- learned warm start */
/* (No reference counterpart: the reference starts its optimizer from a
 * fixed guess, src/optimization.cpp:160-169.)
 * Features of the handle's current observation values (planar metric):
 * log empirical semivariogram in 4 quadrants of the observations' bounding
 * box x 16 lag bins (edges hmax (k/16)^2, hmax = half the box's larger side),
 * from up to 1024 points per quadrant; features[q*16 + b]. */
int mra_variogram_features(mra_handle* h, float* features);

/* MLP 64 -> 256 -> 256 -> 6 with GELU (exact, erf form), FP32 parameters
 * trained by Adam with the library's own forward/backward kernels.  Flat
 * parameter order: W1 (256x64), b1, W2 (256x256), b2, W3 (6x256), b3; W row-major
 * (out x in).  Batches of 1 to 1024 rows; X row-major batch x 64, targets /
 * outputs batch x 6.  Loss: mean squared error over batch and outputs. */
typedef struct mra_mlp mra_mlp;
int mra_mlp_create(uint64_t seed, mra_mlp** out);
void mra_mlp_destroy(mra_mlp* m);
const char* mra_mlp_last_error(const mra_mlp* m);
int mra_mlp_param_count(void);
int mra_mlp_set_params(mra_mlp* m, const float* params);  /* also resets Adam state */
int mra_mlp_get_params(mra_mlp* m, float* params);
/* bf16 != 0: weights and activations rounded to bfloat16, FP32 accumulation */
int mra_mlp_forward(mra_mlp* m, const float* X, int batch, float* Y, int bf16);
/* loss and its gradient with respect to the flat parameters (grad may be NULL) */
int mra_mlp_loss_grad(mra_mlp* m, const float* X, const float* T, int batch, float* loss, float* grad);
/* one Adam step (beta1 0.9, beta2 0.999, eps 1e-8); loss before the step */
int mra_mlp_train_step(mra_mlp* m, const float* X, const float* T, int batch, float lr, float* loss);

/* ------------------------------------------------------------ partition */
/* Sort-based replacement of Partition::build_partition_by_program
 * (src/class_partition.cpp:459-654; same knot grid, half-open boxes and
 * ancestor-knot skip; O(n log n) instead of the reference's O(n*F) scan).
 * bounds = {lonMin, lonMax, latMin, latMax} as Data::load_data pads them
 * (src/class_data.cpp:149-163).  Two-step: sizes first, then fill. */
int mra_partition_program_sizes(const double* lon, const double* lat, uint64_t n,
                                const double* plon, const double* plat, uint64_t np,
                                int M, int J, int r, double offset, const double* bounds,
                                uint64_t* sizes /* [4]: n_coarse*rk, n_finest, n_obs_kept, n_pred_kept */);
int mra_partition_program(const double* lon, const double* lat, const double* val, uint64_t n,
                          const double* plon, const double* plat, uint64_t np,
                          int M, int J, int r, double offset, const double* bounds,
                          double* knot_lon, double* knot_lat, uint64_t* finest_parent,
                          uint64_t* obs_offsets, double* obs_lon, double* obs_lat, double* obs_val,
                          uint64_t* obs_index, uint64_t* pred_offsets, double* pred_lon,
                          double* pred_lat, uint64_t* pred_index);

/* ------------------------------------------------ inputs and outputs (host) */
/* The reference's run configuration (the globals of include/constants.hpp set
 * from the `user_parameters` file by src/read_user_parameters.cpp:10-341).
 * Flags are 0/1; num_levels_m = MRA_LEVELS_DEFAULT means "default" (chosen
 * by mra_default_levels at partition time). */
enum { MRA_LEVELS_DEFAULT = -99 };
typedef struct mra_user_params {
    char data_file_name[1024];
    char prediction_location_file[1024];
    char prediction_results_file_name[1024];
    char sigmasq_file_name[1024], beta_file_name[1024], tausq_file_name[1024];
    char knots_file_name[1024];
    char parent_region_file_name[1024];
    char tmp_directory[1024];
    char calculation_mode[64];   /* prediction | optimization | likelihood | build_structure_only */
    char prediction_location_mode;  /* 'D', 'N' or 'A' */
    char pad_[3];
    int32_t elimination_duplicates_flag, regression_flag, nonstationary_flag, print_detail_flag;
    int32_t dump_prediction_results_flag, save_to_disk_flag, dynamic_schedule_flag, chordal_distance_flag;
    int32_t provided_knots_flag, obs_as_knots_flag, various_j_finest_level_flag;
    int32_t num_partitions_j, num_knots_r, num_levels_m, num_knots_finest, max_iterations;
    double offset, wendland_len;
    double sigmasq, beta, tausq, sigmasq_scale, beta_scale, tausq_scale;
    double sigmasq_initial_guess, sigmasq_lower_bound, sigmasq_upper_bound;
    double beta_initial_guess, beta_lower_bound, beta_upper_bound;
    double tausq_initial_guess, tausq_lower_bound, tausq_upper_bound;
} mra_user_params;
/* Message of the last failed call of this group (the calling thread's). */
const char* mra_io_last_error(void);
/* Zero every field, as the reference's globals start (src/constants.cpp:8-120). */
void mra_user_params_init(mra_user_params* p);
/* read_from_file() + check_errors() (src/read_user_parameters.cpp:10-341,
 * 428-637): `name=value` lines, '#' comments, spaces, tabs and quotes
 * removed, flags "true"/"false", OFFSET=default -> e/100, NUM_LEVELS_M=default
 * -> MRA_LEVELS_DEFAULT; unknown names ignored.  Fields not named keep their
 * value in *p (call mra_user_params_init first for the reference's start). */
int mra_read_user_parameters(const char* path, mra_user_params* p);
/* check_errors (src/read_user_parameters.cpp:428-637) for one process. */
int mra_check_user_parameters(const mra_user_params* p);

/* Data::load_data (src/class_data.cpp:15-358): reads DATA_FILE_NAME
 * (u64 n; f64 lon[n], lat[n], val[n]; u64 region[n] iff
 * provided_knots_flag), takes the non-NaN values as observations, the domain
 * bounds over every location padded by 1e-6 of the extent, and in prediction
 * mode the prediction locations of PREDICTION_LOCATION_MODE: 'D' every data
 * location, 'N' the NaN ones, 'A' PREDICTION_LOCATION_FILE (u64 n; f64
 * lon[n], lat[n]; u64 region[n] iff provided_knots_flag).  Then duplicate
 * elimination (eliminate_duplicate :400-506: locations sorted by (lon, lat),
 * the first of equal locations kept; region indices travel with their
 * locations) and the regression detrend (:333-345: least squares on
 * [1, lon, lat], residuals replace the values). */
typedef struct mra_data mra_data;
typedef struct mra_data_view {
    uint64_t n_locations, n_obs_raw, n_pred_raw, n_obs, n_pred;
    const double *obs_lon, *obs_lat, *obs_val;  /* obs_val: residuals when regression = 1 */
    const uint64_t* obs_region;                 /* NULL without region indices */
    const double *pred_lon, *pred_lat;
    const uint64_t* pred_region;
    double bounds[4];                           /* lonMin, lonMax, latMin, latMax (padded) */
    double coefficients[3];                     /* intercept, lon, lat (regression = 1) */
    int32_t regression, predicting;
} mra_data_view;
int mra_data_load(const mra_user_params* p, mra_data** out);
void mra_data_destroy(mra_data* d);
/* Pointers stay valid while d lives. */
int mra_data_get(const mra_data* d, mra_data_view* v);

/* Partition (src/class_partition.cpp): the number of levels (the default
 * rule :782-797 when num_levels_m = MRA_LEVELS_DEFAULT), the finest level
 * (heap children, or the parent file of VARIOUS_J_FINEST_LEVEL_FLAG :659-770:
 * u64 start; u64 end; u64 parent[end-start+1]), the knots (program grid
 * :459-654 through mra_partition_program, or the knots file of
 * PROVIDED_KNOTS_FLAG :93-263: u64 nr; f64 lon[nr], lat[nr]; u64 region[nr]
 * in runs of r per coarse region) and the assignment of observations and
 * prediction locations to finest regions (program: point in half-open box,
 * observations on an ancestor's knot dropped; user files: the region index
 * of each location, :263-454).  Locations keep their order within a region. */
typedef struct mra_partition mra_partition;
int mra_partition_build(const mra_user_params* p, const mra_data* d, mra_partition** out);
void mra_partition_destroy(mra_partition* P);
/* The structure as the descriptor of mra_create; pointers into P. */
int mra_partition_tree(const mra_partition* P, mra_tree_desc* t);
/* Index into the loaded data of each observation / prediction location of
 * the tree (finest-region order); *pred_index is NULL without predictions. */
int mra_partition_indices(const mra_partition* P, const uint64_t** obs_index, const uint64_t** pred_index);
/* Partition::nRegionsAtEachLevel: counts[l] for l < num_levels. */
int mra_partition_level_counts(const mra_partition* P, uint64_t* counts);
/* Partition::dump_structure_information (src/class_partition.cpp:1433-1498):
 * the same text; path NULL or empty = "structure_information.txt". */
int mra_partition_dump_structure(const mra_partition* P, const char* path);
/* The default number of levels: ceil(log(ceil(n/r)) / log J) + 1
 * (src/class_partition.cpp:782-797); MRA_E_ARG for n = 0. */
int mra_default_levels(uint64_t n_obs, int J, int r);

/* Nonstationary_covariance::load_data (src/class_nonstationary_covariance.cpp:
 * 46-170): the three basis files (u64 nb; f64 intercept; f64 lon[nb], lat[nb],
 * w[nb]) as Wendland fields, metric chordal or great circle from
 * chordal_distance_flag, wendland_len; constant fields when
 * nonstationary_flag = 0.  The descriptor's arrays point into the model. */
typedef struct mra_cov_model mra_cov_model;
int mra_cov_model_load(const mra_user_params* p, mra_cov_model** out);
void mra_cov_model_destroy(mra_cov_model* c);
int mra_cov_model_desc(const mra_cov_model* c, mra_cov_desc* desc);

/* Approximation::dump_prediction_result's file (src/class_approximation.cpp:
 * 473-550): u64 n; f64 lon[n], lat[n], mean[n], var[n]; coef non-NULL adds
 * the regression trend coef[0] + coef[1] lon + coef[2] lat to the mean. */
int mra_write_prediction_result(const char* path, uint64_t n, const double* lon, const double* lat,
                                const double* mean, const double* var, const double* coef);

/* The same partition built on the GPU (one thread per location descends the
 * region boxes; a stable radix sort groups the locations by finest region):
 * identical knots, assignment and order.  Unlike the host form it needs no
 * sizes call: obs_lon / obs_lat / obs_val / obs_index must hold n entries and
 * pred_* np entries; sizes[4] receives what mra_partition_program_sizes
 * reports (knot count, finest regions, kept observations, kept prediction
 * locations).  Up to 2^31 - 1 locations of each kind. */
int mra_partition_program_gpu(const double* lon, const double* lat, const double* val, uint64_t n,
                              const double* plon, const double* plat, uint64_t np, int M, int J, int r,
                              double offset, const double* bounds, double* knot_lon, double* knot_lat,
                              uint64_t* finest_parent, uint64_t* obs_offsets, double* obs_lon, double* obs_lat,
                              double* obs_val, uint64_t* obs_index, uint64_t* pred_offsets, double* pred_lon,
                              double* pred_lat, uint64_t* pred_index, uint64_t* sizes);
const char* mra_partition_gpu_last_error(void);

/* Self test of the device building blocks (no reference counterpart; test
 * hook): one work group computes Cout = C0 - A^T B with the FP64 matrix-core
 * GEMM (A, B: K x 64 row-major; only the Mv x Nv leading part of Cout is
 * updated), Lout = lower Cholesky of the SPD 64 x 64 matrix S (row-major;
 * entries above the diagonal keep S), Xout = L^{-1} Bt; *bad = 1 on a
 * non-positive pivot. */
int mra_selftest(int K, int Mv, int Nv, const double* A, const double* B, const double* C0,
                 const double* S, const double* Bt, double* Cout, double* Lout, double* Xout, int* bad);

/* FP32 variant of the GEMM self test: Cout = C0 - A^T B with the FP32
 * matrix-core GEMM and FP32 accumulation (no reference counterpart). */
int mra_selftest_f32(int K, int Mv, int Nv, const float* A, const float* B, const float* C0, float* Cout);
/* Blocked LDS Cholesky / triangular inverse on one n x n SPD matrix S
 * (n = 16, 32, 48, 64 or 128; row-major): Lout = lower Cholesky factor,
 * Xout = L^{-1} (lower, zero strict upper), *bad = 1 on a non-positive pivot. */
int mra_selftest_blk(int n, const double* S, double* Lout, double* Xout, int* bad);

/* Library version string. */
const char* mra_version(void);

#ifdef __cplusplus
}
#endif
#endif /* MRA_H */
