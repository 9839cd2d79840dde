/*
 * mra_classes.hpp — header-only C++ layer with the reference's run
 * configuration globals (include/constants.hpp:10-205), read_user_parameters()
 * (include/read_user_parameters.hpp), and the classes Data
 * (include/class_data.hpp:8-53), Nonstationary_covariance
 * (include/class_nonstationary_covariance.hpp:8-68) and Partition
 * (include/class_partition.hpp:9-112), implemented over the C ABI of mra.h.
 *
 * A main.cpp written against the reference compiles against these headers
 * (with mra_approximation.hpp for Approximation) once its MPI start-up and
 * shut-down are removed: the globals carry the same names and types, the
 * classes the same public members, and `data`, `partition` and `nonstat_cov`
 * are the same global pointers.  The classes read the globals when their
 * load / build functions run, as the reference does.
 *
 * Differences from the reference: errors throw mra::Error instead of
 * printing and exiting; Data::coefficients is a 3-element mra::vec instead of
 * an Armadillo vec; per-region arrays point into contiguous storage owned by
 * the object (do not delete[] them); there is one process, so the MPI members
 * (origninalWorker, dynamic_schedule) are absent.
 */
#ifndef MRA_CLASSES_HPP
#define MRA_CLASSES_HPP

#include <cmath>
#include <cstdio>
#include <cstring>
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

inline void check_io(int rc) {
    if (rc != MRA_OK) throw Error(rc, mra_io_last_error());
}

// A view of n_elem doubles with the element access of an Armadillo vec.
struct vec {
    double* mem = nullptr;
    unsigned long n_elem = 0;
    double& operator[](unsigned long i) { return mem[i]; }
    const double& operator[](unsigned long i) const { return mem[i]; }
    double& operator()(unsigned long i) { return mem[i]; }
    const double& operator()(unsigned long i) const { return mem[i]; }
    double* memptr() { return mem; }
    const double* memptr() const { return mem; }
};

}  // namespace mra

class Data;
class Partition;
class Nonstationary_covariance;

// ------------------------------------------------------ the run configuration
// include/constants.hpp: the globals the reference's classes read.
inline std::string DATA_FILE_NAME;
inline bool ELIMINATION_DUPLICATES_FLAG = false;
inline bool REGRESSION_FLAG = false;
inline bool NONSTATIONARY_FLAG = false;
inline double OFFSET = 0.0;
inline int NUM_PARTITIONS_J = 0;
inline int NUM_KNOTS_r = 0;
inline int NUM_LEVELS_M = 0;
inline std::string CALCULATION_MODE;
inline bool PRINT_DETAIL_FLAG = false;
inline char PREDICTION_LOCATION_MODE = 0;
inline std::string PREDICTION_LOCATION_FILE;
inline bool DUMP_PREDICTION_RESULTS_FLAG = false;
inline std::string PREDICTION_RESULTS_FILE_NAME;
inline bool SAVE_TO_DISK_FLAG = false;
inline bool DYNAMIC_SCHEDULE_FLAG = false;
inline std::string TMP_DIRECTORY;
inline double SIGMASQ = 0, BETA = 0, TAUSQ = 0;
inline double SIGMASQ_SCALE = 0, BETA_SCALE = 0, TAUSQ_SCALE = 0;
inline int MAX_ITERATIONS = 0;
inline double SIGMASQ_INITIAL_GUESS = 0, SIGMASQ_LOWER_BOUND = 0, SIGMASQ_UPPER_BOUND = 0;
inline double BETA_INITIAL_GUESS = 0, BETA_LOWER_BOUND = 0, BETA_UPPER_BOUND = 0;
inline double TAUSQ_INITIAL_GUESS = 0, TAUSQ_LOWER_BOUND = 0, TAUSQ_UPPER_BOUND = 0;
inline std::string SIGMASQ_FILE_NAME, BETA_FILE_NAME, TAUSQ_FILE_NAME;
inline bool CHORDAL_DISTANCE_FLAG = false;
inline bool PROVIDED_KNOTS_FLAG = false;
inline std::string KNOTS_FILE_NAME;
inline bool OBS_AS_KNOTS_FLAG = false;
inline int NUM_KNOTS_FINEST = 0;
inline bool VARIOUS_J_FINEST_LEVEL_FLAG = false;
inline std::string PARENT_REGION_FILE_NAME;
inline double WENDLAND_LEN = 0;
inline unsigned long NUM_OBSERVATIONS = 0;
inline unsigned long NUM_PREDICTIONS = 0;
inline Data* data = nullptr;
inline Partition* partition = nullptr;
inline Nonstationary_covariance* nonstat_cov = nullptr;

namespace mra {

// The globals as the C ABI's parameter struct.
inline mra_user_params params_from_globals() {
    mra_user_params p;
    mra_user_params_init(&p);
    auto s = [](char* dst, size_t cap, const std::string& v) { std::snprintf(dst, cap, "%s", v.c_str()); };
    s(p.data_file_name, sizeof(p.data_file_name), DATA_FILE_NAME);
    s(p.prediction_location_file, sizeof(p.prediction_location_file), PREDICTION_LOCATION_FILE);
    s(p.prediction_results_file_name, sizeof(p.prediction_results_file_name), PREDICTION_RESULTS_FILE_NAME);
    s(p.sigmasq_file_name, sizeof(p.sigmasq_file_name), SIGMASQ_FILE_NAME);
    s(p.beta_file_name, sizeof(p.beta_file_name), BETA_FILE_NAME);
    s(p.tausq_file_name, sizeof(p.tausq_file_name), TAUSQ_FILE_NAME);
    s(p.knots_file_name, sizeof(p.knots_file_name), KNOTS_FILE_NAME);
    s(p.parent_region_file_name, sizeof(p.parent_region_file_name), PARENT_REGION_FILE_NAME);
    s(p.tmp_directory, sizeof(p.tmp_directory), TMP_DIRECTORY);
    s(p.calculation_mode, sizeof(p.calculation_mode), CALCULATION_MODE);
    p.prediction_location_mode = PREDICTION_LOCATION_MODE;
    p.elimination_duplicates_flag = ELIMINATION_DUPLICATES_FLAG;
    p.regression_flag = REGRESSION_FLAG;
    p.nonstationary_flag = NONSTATIONARY_FLAG;
    p.print_detail_flag = PRINT_DETAIL_FLAG;
    p.dump_prediction_results_flag = DUMP_PREDICTION_RESULTS_FLAG;
    p.save_to_disk_flag = SAVE_TO_DISK_FLAG;
    p.dynamic_schedule_flag = DYNAMIC_SCHEDULE_FLAG;
    p.chordal_distance_flag = CHORDAL_DISTANCE_FLAG;
    p.provided_knots_flag = PROVIDED_KNOTS_FLAG;
    p.obs_as_knots_flag = OBS_AS_KNOTS_FLAG;
    p.various_j_finest_level_flag = VARIOUS_J_FINEST_LEVEL_FLAG;
    p.num_partitions_j = NUM_PARTITIONS_J;
    p.num_knots_r = NUM_KNOTS_r;
    p.num_levels_m = NUM_LEVELS_M;
    p.num_knots_finest = NUM_KNOTS_FINEST;
    p.max_iterations = MAX_ITERATIONS;
    p.offset = OFFSET;
    p.wendland_len = WENDLAND_LEN;
    p.sigmasq = SIGMASQ;
    p.beta = BETA;
    p.tausq = TAUSQ;
    p.sigmasq_scale = SIGMASQ_SCALE;
    p.beta_scale = BETA_SCALE;
    p.tausq_scale = TAUSQ_SCALE;
    p.sigmasq_initial_guess = SIGMASQ_INITIAL_GUESS;
    p.sigmasq_lower_bound = SIGMASQ_LOWER_BOUND;
    p.sigmasq_upper_bound = SIGMASQ_UPPER_BOUND;
    p.beta_initial_guess = BETA_INITIAL_GUESS;
    p.beta_lower_bound = BETA_LOWER_BOUND;
    p.beta_upper_bound = BETA_UPPER_BOUND;
    p.tausq_initial_guess = TAUSQ_INITIAL_GUESS;
    p.tausq_lower_bound = TAUSQ_LOWER_BOUND;
    p.tausq_upper_bound = TAUSQ_UPPER_BOUND;
    return p;
}

inline void globals_from_params(const mra_user_params& p) {
    DATA_FILE_NAME = p.data_file_name;
    PREDICTION_LOCATION_FILE = p.prediction_location_file;
    PREDICTION_RESULTS_FILE_NAME = p.prediction_results_file_name;
    SIGMASQ_FILE_NAME = p.sigmasq_file_name;
    BETA_FILE_NAME = p.beta_file_name;
    TAUSQ_FILE_NAME = p.tausq_file_name;
    KNOTS_FILE_NAME = p.knots_file_name;
    PARENT_REGION_FILE_NAME = p.parent_region_file_name;
    TMP_DIRECTORY = p.tmp_directory;
    CALCULATION_MODE = p.calculation_mode;
    PREDICTION_LOCATION_MODE = p.prediction_location_mode;
    ELIMINATION_DUPLICATES_FLAG = p.elimination_duplicates_flag;
    REGRESSION_FLAG = p.regression_flag;
    NONSTATIONARY_FLAG = p.nonstationary_flag;
    PRINT_DETAIL_FLAG = p.print_detail_flag;
    DUMP_PREDICTION_RESULTS_FLAG = p.dump_prediction_results_flag;
    SAVE_TO_DISK_FLAG = p.save_to_disk_flag;
    DYNAMIC_SCHEDULE_FLAG = p.dynamic_schedule_flag;
    CHORDAL_DISTANCE_FLAG = p.chordal_distance_flag;
    PROVIDED_KNOTS_FLAG = p.provided_knots_flag;
    OBS_AS_KNOTS_FLAG = p.obs_as_knots_flag;
    VARIOUS_J_FINEST_LEVEL_FLAG = p.various_j_finest_level_flag;
    NUM_PARTITIONS_J = p.num_partitions_j;
    NUM_KNOTS_r = p.num_knots_r;
    NUM_LEVELS_M = p.num_levels_m;
    NUM_KNOTS_FINEST = p.num_knots_finest;
    MAX_ITERATIONS = p.max_iterations;
    OFFSET = p.offset;
    WENDLAND_LEN = p.wendland_len;
    SIGMASQ = p.sigmasq;
    BETA = p.beta;
    TAUSQ = p.tausq;
    SIGMASQ_SCALE = p.sigmasq_scale;
    BETA_SCALE = p.beta_scale;
    TAUSQ_SCALE = p.tausq_scale;
    SIGMASQ_INITIAL_GUESS = p.sigmasq_initial_guess;
    SIGMASQ_LOWER_BOUND = p.sigmasq_lower_bound;
    SIGMASQ_UPPER_BOUND = p.sigmasq_upper_bound;
    BETA_INITIAL_GUESS = p.beta_initial_guess;
    BETA_LOWER_BOUND = p.beta_lower_bound;
    BETA_UPPER_BOUND = p.beta_upper_bound;
    TAUSQ_INITIAL_GUESS = p.tausq_initial_guess;
    TAUSQ_LOWER_BOUND = p.tausq_lower_bound;
    TAUSQ_UPPER_BOUND = p.tausq_upper_bound;
}

// Unit-sphere coordinates of lon/lat in degrees (src/class_partition.cpp:507-517).
inline void to_xyz(double lon, double lat, double& x, double& y, double& z) {
    const double cosLat = std::cos(lat * M_PI / 180), cosLon = std::cos(lon * M_PI / 180);
    const double sinLat = std::sin(lat * M_PI / 180), sinLon = std::sin(lon * M_PI / 180);
    x = cosLat * cosLon;
    y = cosLat * sinLon;
    z = sinLat;
}

}  // namespace mra

// read_user_parameters() (src/read_user_parameters.cpp:639-658): the file
// "user_parameters" of the working directory (or `path`) into the globals,
// then check_errors().
inline void read_user_parameters(const char* path = "user_parameters") {
    mra_user_params p = mra::params_from_globals();
    mra::check_io(mra_read_user_parameters(path, &p));
    mra::globals_from_params(p);
}

// include/class_data.hpp:8-53
class Data {
  public:
    unsigned long nLocations = 0, numObservationsInRawData = 0, numPredictionsInRawData = 0;
    double* observationResiduals = nullptr;
    double *observationLon = nullptr, *observationLat = nullptr;
    unsigned long* observationRegion = nullptr;
    double *predictionLon = nullptr, *predictionLat = nullptr;
    unsigned long* predictionRegion = nullptr;
    mra::vec coefficients;
    double domainBoundaries[4] = {};

    Data() = default;
    Data(const Data&) = delete;
    Data& operator=(const Data&) = delete;
    ~Data() { mra_data_destroy(d_); }

    // src/class_data.cpp:15-358; sets NUM_OBSERVATIONS and NUM_PREDICTIONS
    void load_data() {
        mra_data_destroy(d_);
        d_ = nullptr;
        const mra_user_params p = mra::params_from_globals();
        mra::check_io(mra_data_load(&p, &d_));
        mra_data_view v;
        mra::check_io(mra_data_get(d_, &v));
        nLocations = v.n_locations;
        numObservationsInRawData = v.n_obs_raw;
        numPredictionsInRawData = v.n_pred_raw;
        observationResiduals = const_cast<double*>(v.obs_val);
        observationLon = const_cast<double*>(v.obs_lon);
        observationLat = const_cast<double*>(v.obs_lat);
        observationRegion = reinterpret_cast<unsigned long*>(const_cast<uint64_t*>(v.obs_region));
        predictionLon = const_cast<double*>(v.pred_lon);
        predictionLat = const_cast<double*>(v.pred_lat);
        predictionRegion = reinterpret_cast<unsigned long*>(const_cast<uint64_t*>(v.pred_region));
        for (int k = 0; k < 4; k++) domainBoundaries[k] = v.bounds[k];
        coef_[0] = v.coefficients[0];
        coef_[1] = v.coefficients[1];
        coef_[2] = v.coefficients[2];
        coefficients.mem = coef_;
        coefficients.n_elem = v.regression ? 3 : 0;
        NUM_OBSERVATIONS = v.n_obs;
        NUM_PREDICTIONS = v.n_pred;
    }

    void print_data_summary() const {
        std::printf(">>The number of locations: %lu\n", nLocations);
        std::printf(">>The number of valid observations (non-NaN values) in the raw data: %lu\n\n",
                    numObservationsInRawData);
        std::printf(">>The boundry of the spatial domain\n  longitude: minimal %g, maximal %g\n"
                    "  latitude : minimal %g, maximal %g\n\n",
                    domainBoundaries[0], domainBoundaries[1], domainBoundaries[2], domainBoundaries[3]);
    }

    const mra_data* handle() const { return d_; }

  private:
    mra_data* d_ = nullptr;
    double coef_[3] = {};
};

// include/class_nonstationary_covariance.hpp:8-68
class Nonstationary_covariance {
  public:
    unsigned long numSigmasqBasis = 0, numBetaBasis = 0, numTausqBasis = 0;
    double interceptSigmasq = 0, interceptBeta = 0, interceptTausq = 0;
    double *sigmasqLon = nullptr, *sigmasqLat = nullptr, *sigmasqWeight = nullptr;
    double *sigmasqX = nullptr, *sigmasqY = nullptr, *sigmasqZ = nullptr;
    double *betaLon = nullptr, *betaLat = nullptr, *betaWeight = nullptr;
    double *betaX = nullptr, *betaY = nullptr, *betaZ = nullptr;
    double *tausqLon = nullptr, *tausqLat = nullptr, *tausqWeight = nullptr;
    double *tausqX = nullptr, *tausqY = nullptr, *tausqZ = nullptr;

    Nonstationary_covariance() = default;
    Nonstationary_covariance(const Nonstationary_covariance&) = delete;
    Nonstationary_covariance& operator=(const Nonstationary_covariance&) = delete;
    ~Nonstationary_covariance() { mra_cov_model_destroy(c_); }

    // src/class_nonstationary_covariance.cpp:46-429: the three basis files
    void load_data() {
        mra_cov_model_destroy(c_);
        c_ = nullptr;
        mra_user_params p = mra::params_from_globals();
        p.nonstationary_flag = 1;
        mra::check_io(mra_cov_model_load(&p, &c_));
        mra::check_io(mra_cov_model_desc(c_, &desc_));
        const mra_field_desc* f[3] = {&desc_.sigmasq, &desc_.beta, &desc_.tausq};
        unsigned long* nb[3] = {&numSigmasqBasis, &numBetaBasis, &numTausqBasis};
        double* ic[3] = {&interceptSigmasq, &interceptBeta, &interceptTausq};
        double** ll[3][3] = {{&sigmasqLon, &sigmasqLat, &sigmasqWeight},
                             {&betaLon, &betaLat, &betaWeight},
                             {&tausqLon, &tausqLat, &tausqWeight}};
        double** xyz[3][3] = {{&sigmasqX, &sigmasqY, &sigmasqZ}, {&betaX, &betaY, &betaZ}, {&tausqX, &tausqY, &tausqZ}};
        for (int k = 0; k < 3; k++) {
            *nb[k] = f[k]->n_basis;
            *ic[k] = f[k]->intercept;
            *ll[k][0] = const_cast<double*>(f[k]->basis_lon);
            *ll[k][1] = const_cast<double*>(f[k]->basis_lat);
            *ll[k][2] = const_cast<double*>(f[k]->basis_weight);
            xyz_[k].assign(3 * f[k]->n_basis, 0.0);
            for (uint64_t i = 0; i < f[k]->n_basis; i++)
                mra::to_xyz(f[k]->basis_lon[i], f[k]->basis_lat[i], xyz_[k][i], xyz_[k][f[k]->n_basis + i],
                            xyz_[k][2 * f[k]->n_basis + i]);
            for (int c = 0; c < 3; c++) *xyz[k][c] = xyz_[k].data() + c * f[k]->n_basis;
        }
    }

    void print_summary() const {
        std::printf(">>Nonstationary model: %lu / %lu / %lu basis functions for sill / range / nugget\n\n",
                    numSigmasqBasis, numBetaBasis, numTausqBasis);
    }

    // src/class_nonstationary_covariance.cpp:479-573 (one point per call;
    // fields() evaluates many at once)
    double get_sigmasq(double lon, double lat) const { return one(0, lon, lat); }
    double get_sigmasq(double x, double y, double z) const { return one_xyz(0, x, y, z); }
    double get_beta(double lon, double lat) const { return one(1, lon, lat); }
    double get_beta(double x, double y, double z) const { return one_xyz(1, x, y, z); }
    double get_tausq(double lon, double lat) const { return one(2, lon, lat); }
    double get_tausq(double x, double y, double z) const { return one_xyz(2, x, y, z); }

    // The three fields at n points (lon/lat in degrees) in one call.
    void fields(unsigned long n, const double* lon, const double* lat, double* sigmasq, double* beta,
                double* tausq) const {
        mra_cov_desc d = desc_for(false);
        mra::check(mra_fields(&d, lon, lat, nullptr, nullptr, nullptr, n, sigmasq, beta, tausq), nullptr);
    }

    // The model as the C ABI's descriptor: the lon/lat overloads measure the
    // great-circle distance, the x/y/z overloads the chordal one.
    mra_cov_desc desc_for(bool chordal) const {
        if (!c_) throw mra::Error(MRA_E_STATE, "Nonstationary_covariance::load_data has not run");
        mra_cov_desc d = desc_;
        d.metric = chordal ? MRA_METRIC_CHORDAL : MRA_METRIC_GREAT_CIRCLE;
        d.wendland_len = WENDLAND_LEN;
        return d;
    }

  private:
    double pick(int k, const double* s, const double* b, const double* t) const { return k == 0 ? *s : k == 1 ? *b : *t; }
    double one(int k, double lon, double lat) const {
        double s, b, t;
        mra_cov_desc d = desc_for(false);
        mra::check(mra_fields(&d, &lon, &lat, nullptr, nullptr, nullptr, 1, &s, &b, &t), nullptr);
        return pick(k, &s, &b, &t);
    }
    double one_xyz(int k, double x, double y, double z) const {
        double s, b, t;
        mra_cov_desc d = desc_for(true);
        mra::check(mra_fields(&d, nullptr, nullptr, &x, &y, &z, 1, &s, &b, &t), nullptr);
        return pick(k, &s, &b, &t);
    }
    mra_cov_model* c_ = nullptr;
    mra_cov_desc desc_{};
    std::vector<double> xyz_[3];
};

// include/class_partition.hpp:9-112.  Region R's arrays: knotsLon[R] ...
// (nKnots values for coarse regions; for a finest region the observations,
// nKnotsAtFinestLevel[R - indexStartFinestLevel] of them); per finest
// region f: knotsResidual[f], predictionLon[f], predictionLat[f] (and the
// unit-sphere X/Y/Z companions when CHORDAL_DISTANCE_FLAG).
class Partition {
  public:
    unsigned long nRegionsInTotal = 0;
    int nKnots = 0;
    unsigned long* nRegionsAtEachLevel = nullptr;
    double **knotsX = nullptr, **knotsY = nullptr, **knotsZ = nullptr, **knotsLon = nullptr, **knotsLat = nullptr;
    double** knotsResidual = nullptr;
    double **predictionX = nullptr, **predictionY = nullptr, **predictionZ = nullptr, **predictionLon = nullptr,
           **predictionLat = nullptr;
    unsigned long *nKnotsAtFinestLevel = nullptr, *nPredictionsAtFinestLevel = nullptr;
    unsigned long indexStartFinestLevel = 0, indexStartSecondFinestLevel = 0;
    unsigned long nRegionsAtFinestLevel = 0;
    unsigned long* parent = nullptr;
    unsigned long *childrenStart = nullptr, *childrenEnd = nullptr;

    Partition() = default;
    Partition(const Partition&) = delete;
    Partition& operator=(const Partition&) = delete;
    ~Partition() { mra_partition_destroy(p_); }

    // Partition() + build_partition() (src/class_partition.cpp:778-1040):
    // reads the globals and the global `data`; sets NUM_LEVELS_M when it was
    // "default".
    void build_partition() {
        if (!::data || !::data->handle()) throw mra::Error(MRA_E_STATE, "Data::load_data has not run");
        mra_partition_destroy(p_);
        p_ = nullptr;
        const mra_user_params prm = mra::params_from_globals();
        mra::check_io(mra_partition_build(&prm, ::data->handle(), &p_));
        mra::check_io(mra_partition_tree(p_, &tree_));
        NUM_LEVELS_M = tree_.num_levels;
        const int M = tree_.num_levels;
        const unsigned long nC = tree_.n_coarse, nF = tree_.n_finest;
        nKnots = tree_.num_knots;
        levels_.assign(M, 0);
        mra::check_io(mra_partition_level_counts(p_, reinterpret_cast<uint64_t*>(levels_.data())));
        nRegionsAtEachLevel = levels_.data();
        nRegionsInTotal = nC + nF;
        nRegionsAtFinestLevel = nF;
        indexStartFinestLevel = nC;
        indexStartSecondFinestLevel = nC - levels_[M - 2];
        parent = const_cast<unsigned long*>(reinterpret_cast<const unsigned long*>(tree_.finest_parent));
        // children of the second finest level (read_parent_region_index :737-769)
        const unsigned long n2 = levels_[M - 2];
        chS_.assign(n2, 0);
        chE_.assign(n2, 0);
        for (unsigned long f = 0; f < nF; f++) {
            const unsigned long k = parent[f] - indexStartSecondFinestLevel, R = f + nC;
            if (chS_[k] == 0 || chS_[k] > R) chS_[k] = R;
            if (chE_[k] == 0 || chE_[k] < R) chE_[k] = R;
        }
        childrenStart = chS_.data();
        childrenEnd = chE_.data();
        // per-region pointers into the tree's contiguous arrays
        const bool chordal = CHORDAL_DISTANCE_FLAG;
        rLon_.assign(nRegionsInTotal, nullptr);
        rLat_.assign(nRegionsInTotal, nullptr);
        rX_.assign(chordal ? nRegionsInTotal : 0, nullptr);
        rY_ = rZ_ = rX_;
        nKf_.assign(nF, 0);
        nPf_.assign(nF, 0);
        resid_.assign(nF, nullptr);
        pLon_.assign(nF, nullptr);
        pLat_.assign(nF, nullptr);
        pX_.assign(chordal ? nF : 0, nullptr);
        pY_ = pZ_ = pX_;
        auto mut = [](const double* p) { return const_cast<double*>(p); };
        xyzK_.assign(chordal ? 3 * nC * nKnots : 0, 0.0);
        xyzO_.assign(chordal ? 3 * tree_.n_obs : 0, 0.0);
        xyzP_.assign(chordal ? 3 * tree_.n_pred : 0, 0.0);
        for (unsigned long R = 0; R < nC; R++) {
            rLon_[R] = mut(tree_.knot_lon + R * nKnots);
            rLat_[R] = mut(tree_.knot_lat + R * nKnots);
            if (chordal) {
                for (int q = 0; q < nKnots; q++) {
                    const unsigned long i = R * nKnots + q;
                    mra::to_xyz(tree_.knot_lon[i], tree_.knot_lat[i], xyzK_[i], xyzK_[nC * nKnots + i],
                                xyzK_[2 * nC * nKnots + i]);
                }
                rX_[R] = xyzK_.data() + R * nKnots;
                rY_[R] = xyzK_.data() + nC * nKnots + R * nKnots;
                rZ_[R] = xyzK_.data() + 2 * nC * nKnots + R * nKnots;
            }
        }
        const unsigned long no = tree_.n_obs, np = tree_.n_pred;
        if (chordal) {
            for (unsigned long i = 0; i < no; i++)
                mra::to_xyz(tree_.obs_lon[i], tree_.obs_lat[i], xyzO_[i], xyzO_[no + i], xyzO_[2 * no + i]);
            for (unsigned long i = 0; i < np; i++)
                mra::to_xyz(tree_.pred_lon[i], tree_.pred_lat[i], xyzP_[i], xyzP_[np + i], xyzP_[2 * np + i]);
        }
        for (unsigned long f = 0; f < nF; f++) {
            const unsigned long o = tree_.obs_offsets[f], R = nC + f;
            nKf_[f] = tree_.obs_offsets[f + 1] - o;
            rLon_[R] = mut(tree_.obs_lon + o);
            rLat_[R] = mut(tree_.obs_lat + o);
            resid_[f] = mut(tree_.obs_val + o);
            if (chordal) {
                rX_[R] = xyzO_.data() + o;
                rY_[R] = xyzO_.data() + no + o;
                rZ_[R] = xyzO_.data() + 2 * no + o;
            }
            if (tree_.pred_offsets) {
                const unsigned long q = tree_.pred_offsets[f];
                nPf_[f] = tree_.pred_offsets[f + 1] - q;
                pLon_[f] = mut(tree_.pred_lon + q);
                pLat_[f] = mut(tree_.pred_lat + q);
                if (chordal) {
                    pX_[f] = xyzP_.data() + q;
                    pY_[f] = xyzP_.data() + np + q;
                    pZ_[f] = xyzP_.data() + 2 * np + q;
                }
            }
        }
        knotsLon = rLon_.data();
        knotsLat = rLat_.data();
        knotsX = chordal ? rX_.data() : nullptr;
        knotsY = chordal ? rY_.data() : nullptr;
        knotsZ = chordal ? rZ_.data() : nullptr;
        knotsResidual = resid_.data();
        nKnotsAtFinestLevel = nKf_.data();
        nPredictionsAtFinestLevel = nPf_.data();
        predictionLon = pLon_.data();
        predictionLat = pLat_.data();
        predictionX = chordal ? pX_.data() : nullptr;
        predictionY = chordal ? pY_.data() : nullptr;
        predictionZ = chordal ? pZ_.data() : nullptr;
    }

    void print_partition_summary() const {
        std::printf(">>The number of levels: %d\n\n>>The total number of regions: %lu\n\n", NUM_LEVELS_M,
                    nRegionsInTotal);
    }

    // src/class_partition.cpp:1433-1498: "structure_information.txt"
    void dump_structure_information(const char* path = "structure_information.txt") const {
        if (!p_) throw mra::Error(MRA_E_STATE, "Partition::build_partition has not run");
        mra::check_io(mra_partition_dump_structure(p_, path));
    }

    // The structure as the descriptor of mra_create (pointers into this object).
    const mra_tree_desc& tree() const {
        if (!p_) throw mra::Error(MRA_E_STATE, "Partition::build_partition has not run");
        return tree_;
    }

  private:
    mra_partition* p_ = nullptr;
    mra_tree_desc tree_{};
    std::vector<unsigned long> levels_, chS_, chE_, nKf_, nPf_;
    std::vector<double*> rLon_, rLat_, rX_, rY_, rZ_, resid_, pLon_, pLat_, pX_, pY_, pZ_;
    std::vector<double> xyzK_, xyzO_, xyzP_;
};

#endif
