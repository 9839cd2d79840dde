/*
 * mra_optimization.hpp — the reference's optimization mode
 * (include/optimization.hpp, src/optimization.cpp:13-192) over the library:
 * the same objectives (D of one Approximation per evaluation), start points,
 * bounds and resolution schedule, minimised by mra_bobyqa (the library's
 * bound-constrained quadratic-model minimiser with npt = 7 interpolation
 * points, as the reference's dlib::find_min_bobyqa call) instead of dlib.
 *
 * Results are left in mra::last_optimization (parameters, D, the printed
 * log-likelihood, evaluation count, status) and in OPTIMIZATION_ITERATION.
 */
#ifndef MRA_OPTIMIZATION_HPP
#define MRA_OPTIMIZATION_HPP

#include <cstdio>

#include "mra_approximation.hpp"

inline int OPTIMIZATION_ITERATION = 0;

namespace mra {

struct OptimizationResult {
    double parameter[3] = {};  // sill, range, nugget or the three scales
    double D = 0.0;            // Approximation::loglikelihood at the optimum
    double loglikelihood = 0.0;
    int evaluations = 0;
    int status = MRA_OK;  // MRA_OK or MRA_W_MAX_EVALS (dlib throws bobyqa_failure there)
};
inline OptimizationResult last_optimization;

inline double objective_stationary(const double* x, void*) {
    OPTIMIZATION_ITERATION++;
    Approximation approximation(x[0], x[1], x[2], 0.0, 0.0, 0.0);  // src/optimization.cpp:52-53
    approximation.likelihood();
    return approximation.loglikelihood;
}

inline double objective_nonstationary(const double* x, void*) {
    OPTIMIZATION_ITERATION++;
    Approximation approximation(0.0, 0.0, 0.0, x[0], x[1], x[2]);  // src/optimization.cpp:144-145
    approximation.likelihood();
    return approximation.loglikelihood;
}

inline void minimise(mra_objective_fn fn, double* x, const double* lo, const double* hi, double rho_begin,
                     double rho_end) {
    OPTIMIZATION_ITERATION = 0;
    double f = 0.0;
    int ne = 0;
    const int rc = mra_bobyqa(3, 7, fn, nullptr, x, lo, hi, rho_begin, rho_end, MAX_ITERATIONS, &f, &ne);
    if (rc < 0) throw Error(rc, "mra_bobyqa failed");
    last_optimization = OptimizationResult{};
    for (int k = 0; k < 3; k++) last_optimization.parameter[k] = x[k];
    last_optimization.D = f;
    last_optimization.loglikelihood = -0.5 * (f + NUM_OBSERVATIONS * std::log(2 * 3.14159265359));
    last_optimization.evaluations = ne;
    last_optimization.status = rc;
    if (rc == MRA_W_MAX_EVALS)
        std::printf("The optimization stopped at MAX_ITERATIONS = %d evaluations.\n", MAX_ITERATIONS);
    std::printf("The optimal parameters are %.10g %.10g %.10g with loglikelihood %.10g (%d evaluations)\n", x[0],
                x[1], x[2], last_optimization.loglikelihood, ne);
}

}  // namespace mra

// src/optimization.cpp:60-105: sill, range, nugget of the stationary model
// from the *_INITIAL_GUESS globals within the *_BOUND globals, resolution
// 1e-4 down to 1e-8.
inline void get_optimal_parameters() {
    double x[3] = {SIGMASQ_INITIAL_GUESS, BETA_INITIAL_GUESS, TAUSQ_INITIAL_GUESS};
    const double lo[3] = {SIGMASQ_LOWER_BOUND, BETA_LOWER_BOUND, TAUSQ_LOWER_BOUND};
    const double hi[3] = {SIGMASQ_UPPER_BOUND, BETA_UPPER_BOUND, TAUSQ_UPPER_BOUND};
    mra::minimise(mra::objective_stationary, x, lo, hi, 1e-4, 1e-8);
}

// src/optimization.cpp:158-192: the three scales of the nonstationary model
// from (1, 1, 1) within [1e-2, 1e2], resolution 0.1 down to 1e-6.
inline void get_optimal_nonstationary_parameters() {
    double x[3] = {1.0, 1.0, 1.0};
    const double lo[3] = {1e-2, 1e-2, 1e-2};
    const double hi[3] = {1e2, 1e2, 1e2};
    mra::minimise(mra::objective_nonstationary, x, lo, hi, 0.1, 1e-6);
}

#endif
