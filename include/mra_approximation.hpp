/*
 * mra_approximation.hpp — header-only C++ layer over the C ABI (mra.h) with
 * the reference's class and function names, so that a main.cpp /
 * optimization.cpp written against the reference compiles against this
 * library with its includes swapped.
 *
 * The reference reads its structure and covariance model from globals
 * (`partition`, `nonstat_cov`, NONSTATIONARY_FLAG, ...; include/constants.hpp).
 * Here those live in one mra::Context made current with mra::Context::use();
 * Approximation and the evaluate_* functions read the current context.
 *
 * Error behaviour: the reference prints and exits (or Armadillo throws on a
 * non positive definite matrix); this layer throws mra::Error carrying the
 * library's status code and message.
 */
#ifndef MRA_APPROXIMATION_HPP
#define MRA_APPROXIMATION_HPP

#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "mra.h"

namespace mra {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void check(int rc, const mra_handle* h) {
    if (rc != MRA_OK) throw Error(rc, mra_last_error(h));
}

// The structure (Partition arrays) and covariance model of one data set,
// uploaded once; the stand-in for the reference's globals.
class Context {
  public:
    Context(const mra_tree_desc& tree, const mra_cov_desc& cov) {
        check(mra_create(&tree, &h_), nullptr);
        n_obs_ = tree.n_obs;
        n_pred_ = tree.pred_offsets ? tree.n_pred : 0;
        n_finest_ = tree.n_finest;
        if (tree.pred_offsets) {
            pred_offsets_.assign(tree.pred_offsets, tree.pred_offsets + tree.n_finest + 1);
            pred_lon_.assign(tree.pred_lon, tree.pred_lon + n_pred_);
            pred_lat_.assign(tree.pred_lat, tree.pred_lat + n_pred_);
        }
        cov_ = cov;
        try {
            check(mra_set_covariance(h_, &cov), h_);
        } catch (...) {
            mra_destroy(h_);
            throw;
        }
    }
    ~Context() {
        if (current() == this) current() = nullptr;
        mra_destroy(h_);
    }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;

    void use() { current() = this; }
    static Context*& current() {
        static Context* c = nullptr;
        return c;
    }
    static Context& get() {
        if (!current()) throw Error(MRA_E_STATE, "no mra::Context in use");
        return *current();
    }

    mra_handle* handle() const { return h_; }
    const mra_cov_desc& cov() const { return cov_; }
    uint64_t n_obs() const { return n_obs_; }
    uint64_t n_pred() const { return n_pred_; }
    uint64_t n_finest() const { return n_finest_; }
    const std::vector<uint64_t>& pred_offsets() const { return pred_offsets_; }
    const std::vector<double>& pred_lon() const { return pred_lon_; }
    const std::vector<double>& pred_lat() const { return pred_lat_; }
    // MRA_PRECISION_FP64 (default) or MRA_PRECISION_MIXED for later evaluations
    void set_precision(int p) { check(mra_set_precision(h_, p), h_); }
    void use_graph(bool on) { check(mra_use_graph(h_, on ? 1 : 0), h_); }

  private:
    mra_handle* h_ = nullptr;
    mra_cov_desc cov_{};
    uint64_t n_obs_ = 0, n_pred_ = 0, n_finest_ = 0;
    std::vector<uint64_t> pred_offsets_;
    std::vector<double> pred_lon_, pred_lat_;
};

}  // namespace mra

// include/class_approximation.hpp:7-60.  One evaluation object per parameter
// set, as in the reference (src/main.cpp:124-133, src/optimization.cpp:144);
// the structure stays resident in the context between objects.
class Approximation {
  public:
    double sill, range, nugget;
    double sill_scale, range_scale, nugget_scale;
    double loglikelihood = 0.0;  // D = -2 loglik - n log(2*3.14159265359)
    // posterior predictive mean / variance, flattened in finest-region order;
    // region F occupies [posteriorPredictionOffsets[F], posteriorPredictionOffsets[F+1])
    std::vector<double> posteriorPredictionMean, posteriorPredictionVariance;
    std::vector<uint64_t> posteriorPredictionOffsets;

    Approximation(const double& sill_, const double& range_, const double& nugget_, const double& sill_scale_,
                  const double& range_scale_, const double& nugget_scale_)
        : sill(sill_), range(range_), nugget(nugget_), sill_scale(sill_scale_), range_scale(range_scale_),
          nugget_scale(nugget_scale_) {}

    // src/class_approximation.cpp:466-470
    void likelihood() {
        mra::Context& c = mra::Context::get();
        set_params(c);
        mra::check(mra_likelihood(c.handle(), &loglikelihood), c.handle());
    }

    // src/class_approximation-predict.cpp:15-200 (re-evaluates the prior
    // and posterior with the prediction pre-pass, then predicts)
    void predict() {
        mra::Context& c = mra::Context::get();
        set_params(c);
        posteriorPredictionMean.assign(c.n_pred(), 0.0);
        posteriorPredictionVariance.assign(c.n_pred(), 0.0);
        posteriorPredictionOffsets = c.pred_offsets();
        mra::check(mra_predict(c.handle(), &loglikelihood, posteriorPredictionMean.data(),
                               posteriorPredictionVariance.data()),
                   c.handle());
    }

    // src/class_approximation.cpp:473-550: u64 n; f64 lon[n], lat[n],
    // mean[n], var[n] in finest-region order (no regression trend added)
    void dump_prediction_result(const std::string& path = "prediction_result.bin") const {
        mra::Context& c = mra::Context::get();
        const uint64_t n = posteriorPredictionMean.size();
        FILE* f = std::fopen(path.c_str(), "wb");
        if (!f) throw mra::Error(MRA_E_ARG, "cannot open " + path);
        bool ok = std::fwrite(&n, sizeof(n), 1, f) == 1;
        ok = ok && std::fwrite(c.pred_lon().data(), sizeof(double), n, f) == n;
        ok = ok && std::fwrite(c.pred_lat().data(), sizeof(double), n, f) == n;
        ok = ok && std::fwrite(posteriorPredictionMean.data(), sizeof(double), n, f) == n;
        ok = ok && std::fwrite(posteriorPredictionVariance.data(), sizeof(double), n, f) == n;
        std::fclose(f);
        if (!ok) throw mra::Error(MRA_E_ARG, "short write to " + path);
    }

  private:
    void set_params(mra::Context& c) const {
        mra::check(mra_set_params(c.handle(), sill, range, nugget, sill_scale, range_scale, nugget_scale), c.handle());
    }
};

// include/evaluate_covariance.hpp:25-88.  The x, y, z arguments are accepted
// for signature compatibility; the library derives its coordinates from
// lon/lat and the context's metric.  Output layout as in the reference:
// covarianceMatrix[i1 * n2 + i2].
inline void evaluate_cross_covariance(double* covarianceMatrix, double* lon1, double* lat1, double*, double*, double*,
                                      double* lon2, double* lat2, double*, double*, double*,
                                      const unsigned long& nLocations1, const unsigned long& nLocations2,
                                      const double& sill, const double& range, const double& nugget) {
    mra_cov_desc cov = mra::Context::get().cov();
    cov.nonstationary = 0;
    mra::check(mra_cross_covariance(&cov, sill, range, nugget, 1.0, 1.0, 1.0, lon1, lat1, nLocations1, lon2, lat2,
                                    nLocations2, covarianceMatrix),
               nullptr);
}

inline void evaluate_variance_covariance(double* covarianceMatrix, double* lon, double* lat, double*, double*, double*,
                                         const unsigned long& nLocations, const double& sill, const double& range,
                                         const double& nugget) {
    mra_cov_desc cov = mra::Context::get().cov();
    cov.nonstationary = 0;
    mra::check(mra_variance_covariance(&cov, sill, range, nugget, 1.0, 1.0, 1.0, lon, lat, nLocations,
                                       covarianceMatrix),
               nullptr);
}

inline void evaluate_nonstationary_cross_covariance(double* covarianceMatrix, double* lon1, double* lat1, double*,
                                                    double*, double*, double* lon2, double* lat2, double*, double*,
                                                    double*, const unsigned long& nLocations1,
                                                    const unsigned long& nLocations2, const double& sill_scale,
                                                    const double& range_scale, const double& nugget_scale) {
    mra_cov_desc cov = mra::Context::get().cov();
    cov.nonstationary = 1;
    mra::check(mra_cross_covariance(&cov, 0.0, 1.0, 0.0, sill_scale, range_scale, nugget_scale, lon1, lat1,
                                    nLocations1, lon2, lat2, nLocations2, covarianceMatrix),
               nullptr);
}

inline void evaluate_nonstationary_variance_covariance(double* covarianceMatrix, double* lon, double* lat, double*,
                                                       double*, double*, const unsigned long& nLocations,
                                                       const double& sill_scale, const double& range_scale,
                                                       const double& nugget_scale) {
    mra_cov_desc cov = mra::Context::get().cov();
    cov.nonstationary = 1;
    mra::check(mra_variance_covariance(&cov, 0.0, 1.0, 0.0, sill_scale, range_scale, nugget_scale, lon, lat,
                                       nLocations, covarianceMatrix),
               nullptr);
}

#endif
