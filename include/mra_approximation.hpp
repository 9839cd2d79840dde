/*
 * mra_approximation.hpp — header-only C++ layer over the C ABI (mra.h) with
 * the reference's class and function names, so that a main.cpp /
 * optimization.cpp written against the reference compiles against this
 * library with its includes swapped.
 *
 * The reference reads its structure and covariance model from globals
 * (`partition`, `nonstat_cov`, NONSTATIONARY_FLAG, ...; include/constants.hpp).
 * Those globals and classes are in mra_classes.hpp; when no mra::Context has
 * been made current with mra::Context::use(), Approximation and the
 * evaluate_* functions build one from `partition`, `nonstat_cov` and the
 * flags (and rebuild it when those objects change).
 *
 * Error behaviour: the reference prints and exits (or Armadillo throws on a
 * non positive definite matrix); this layer throws mra::Error carrying the
 * library's status code and message.
 */
#ifndef MRA_APPROXIMATION_HPP
#define MRA_APPROXIMATION_HPP

#include <cstdio>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "mra.h"
#include "mra_classes.hpp"

namespace mra {

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
    // The current context, or one built from the reference's globals
    // (`partition`, `nonstat_cov`, NONSTATIONARY_FLAG, CHORDAL_DISTANCE_FLAG,
    // WENDLAND_LEN), rebuilt when those objects or flags change.
    static Context& get() {
        if (current()) return *current();
        return from_globals();
    }
    // The context built from the globals (destroy it with release_globals()
    // before the program ends; it is not destroyed at exit).
    static Context*& global_slot() {
        static Context* c = nullptr;
        return c;
    }
    static void release_globals() {
        delete global_slot();
        global_slot() = nullptr;
    }
    static Context& from_globals() {
        struct Key {
            const void *p = nullptr, *c = nullptr;
            bool ns = false, chordal = false;
            double wlen = 0;
        };
        static Key key;
        Context*& ctx = global_slot();
        if (!::partition) throw Error(MRA_E_STATE, "no mra::Context in use and no global partition");
        const Key k{::partition, ::nonstat_cov, NONSTATIONARY_FLAG, CHORDAL_DISTANCE_FLAG, WENDLAND_LEN};
        if (!ctx || k.p != key.p || k.c != key.c || k.ns != key.ns || k.chordal != key.chordal || k.wlen != key.wlen) {
            release_globals();
            mra_cov_desc cov{};
            if (NONSTATIONARY_FLAG) {
                if (!::nonstat_cov) throw Error(MRA_E_STATE, "NONSTATIONARY_FLAG is true but nonstat_cov is not loaded");
                cov = ::nonstat_cov->desc_for(CHORDAL_DISTANCE_FLAG);
            } else {
                cov.nonstationary = 0;
                cov.metric = CHORDAL_DISTANCE_FLAG ? MRA_METRIC_CHORDAL : MRA_METRIC_GREAT_CIRCLE;
                cov.wendland_len = 1.0;
            }
            ctx = new Context(::partition->tree(), cov);
            key = k;
        }
        return *ctx;
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
    // posterior predictive mean / variance per finest region (index = finest
    // region ordinal, as the reference's vec arrays), filled by predict()
    mra::vec* posteriorPredictionMean = nullptr;
    mra::vec* posteriorPredictionVariance = nullptr;

    Approximation(const double& sill_, const double& range_, const double& nugget_, const double& sill_scale_,
                  const double& range_scale_, const double& nugget_scale_)
        : sill(sill_), range(range_), nugget(nugget_), sill_scale(sill_scale_), range_scale(range_scale_),
          nugget_scale(nugget_scale_) {}
    Approximation(const Approximation&) = delete;
    Approximation& operator=(const Approximation&) = delete;

    // src/class_approximation.cpp:466-470.  Also sets printedLoglikelihood,
    // the value the reference prints
    // (src/class_approximation-posterior-inference.cpp:244), with n =
    // NUM_OBSERVATIONS when a Data set was loaded.
    void likelihood() {
        mra::Context& c = mra::Context::get();
        set_params(c);
        mra::check(mra_loglikelihood(c.handle(), NUM_OBSERVATIONS, &loglikelihood, &printedLoglikelihood),
                   c.handle());
    }
    double printedLoglikelihood = 0.0;

    // src/class_approximation-predict.cpp:15-200 (re-evaluates the prior
    // and posterior with the prediction pre-pass, then predicts)
    void predict() {
        mra::Context& c = mra::Context::get();
        set_params(c);
        mean_.assign(c.n_pred(), 0.0);
        var_.assign(c.n_pred(), 0.0);
        mra::check(mra_predict(c.handle(), &loglikelihood, mean_.data(), var_.data()), c.handle());
        const std::vector<uint64_t>& off = c.pred_offsets();
        const size_t nF = off.empty() ? 0 : off.size() - 1;
        mv_.assign(nF, mra::vec{});
        vv_.assign(nF, mra::vec{});
        for (size_t f = 0; f < nF; f++) {
            mv_[f] = mra::vec{mean_.data() + off[f], (unsigned long)(off[f + 1] - off[f])};
            vv_[f] = mra::vec{var_.data() + off[f], (unsigned long)(off[f + 1] - off[f])};
        }
        posteriorPredictionMean = mv_.data();
        posteriorPredictionVariance = vv_.data();
    }

    // The predictions in finest-region order, all regions concatenated.
    const std::vector<double>& prediction_mean() const { return mean_; }
    const std::vector<double>& prediction_variance() const { return var_; }

    // src/class_approximation.cpp:473-550: u64 n; f64 lon[n], lat[n],
    // mean[n], var[n] in finest-region order into PREDICTION_RESULTS_FILE_NAME
    // (or `path`); with REGRESSION_FLAG and a loaded Data set the mean adds
    // the regression trend.
    void dump_prediction_result(const std::string& path = "") const {
        mra::Context& c = mra::Context::get();
        const std::string fn = !path.empty() ? path : PREDICTION_RESULTS_FILE_NAME;
        if (fn.empty()) throw mra::Error(MRA_E_ARG, "no prediction results file name");
        const bool trend = REGRESSION_FLAG && ::data && ::data->coefficients.n_elem == 3;
        mra::check_io(mra_write_prediction_result(fn.c_str(), mean_.size(), c.pred_lon().data(), c.pred_lat().data(),
                                                  mean_.data(), var_.data(),
                                                  trend ? ::data->coefficients.memptr() : nullptr));
    }

  private:
    void set_params(mra::Context& c) const {
        mra::check(mra_set_params(c.handle(), sill, range, nugget, sill_scale, range_scale, nugget_scale), c.handle());
    }
    std::vector<double> mean_, var_;
    std::vector<mra::vec> mv_, vv_;
};

// include/evaluate_covariance.hpp:25-88.  Output layout as in the
// reference: covarianceMatrix[i1 * n2 + i2].  With the chordal metric the
// x, y, z arguments are the coordinates used (as in the reference); the
// covariance model's metric and fields come from the current context.
inline mra_cov_desc context_cov(int nonstationary) {
    mra_cov_desc cov = mra::Context::get().cov();
    cov.nonstationary = nonstationary;
    return cov;
}

inline void evaluate_cross_covariance(double* covarianceMatrix, double* lon1, double* lat1, double* x1, double* y1,
                                      double* z1, double* lon2, double* lat2, double* x2, double* y2, double* z2,
                                      const unsigned long& nLocations1, const unsigned long& nLocations2,
                                      const double& sill, const double& range, const double& nugget) {
    const mra_cov_desc cov = context_cov(0);
    mra::check(mra_cross_covariance_xyz(&cov, sill, range, nugget, 1.0, 1.0, 1.0, lon1, lat1, x1, y1, z1, nLocations1,
                                        lon2, lat2, x2, y2, z2, nLocations2, covarianceMatrix),
               nullptr);
}

inline void evaluate_variance_covariance(double* covarianceMatrix, double* lon, double* lat, double* x, double* y,
                                         double* z, const unsigned long& nLocations, const double& sill,
                                         const double& range, const double& nugget) {
    const mra_cov_desc cov = context_cov(0);
    mra::check(mra_variance_covariance_xyz(&cov, sill, range, nugget, 1.0, 1.0, 1.0, lon, lat, x, y, z, nLocations,
                                           covarianceMatrix),
               nullptr);
}

inline void evaluate_nonstationary_cross_covariance(double* covarianceMatrix, double* lon1, double* lat1, double* x1,
                                                    double* y1, double* z1, double* lon2, double* lat2, double* x2,
                                                    double* y2, double* z2, const unsigned long& nLocations1,
                                                    const unsigned long& nLocations2, const double& sill_scale,
                                                    const double& range_scale, const double& nugget_scale) {
    const mra_cov_desc cov = context_cov(1);
    mra::check(mra_cross_covariance_xyz(&cov, 0.0, 1.0, 0.0, sill_scale, range_scale, nugget_scale, lon1, lat1, x1,
                                        y1, z1, nLocations1, lon2, lat2, x2, y2, z2, nLocations2, covarianceMatrix),
               nullptr);
}

inline void evaluate_nonstationary_variance_covariance(double* covarianceMatrix, double* lon, double* lat, double* x,
                                                       double* y, double* z, const unsigned long& nLocations,
                                                       const double& sill_scale, const double& range_scale,
                                                       const double& nugget_scale) {
    const mra_cov_desc cov = context_cov(1);
    mra::check(mra_variance_covariance_xyz(&cov, 0.0, 1.0, 0.0, sill_scale, range_scale, nugget_scale, lon, lat, x,
                                           y, z, nLocations, covarianceMatrix),
               nullptr);
}

#endif
