// Inputs and outputs either side of the likelihood path, on the host:
//   * the `user_parameters` key file (src/read_user_parameters.cpp:10-341,
//     checks :428-637);
//   * Data::load_data (src/class_data.cpp:15-358): the data file, NaN values
//     as prediction-only locations, PREDICTION_LOCATION_MODE D / N / A, the
//     padded domain bounds, duplicate elimination (:400-506) and the linear
//     regression detrend (:333-345);
//   * Partition (src/class_partition.cpp): the default number of levels
//     (:782-797), the program partition (:459-654, through the sort-based
//     builder of mra_partition.cpp), the user-files partition with its knots
//     and parent files (:93-454, :659-770) and dump_structure_information
//     (:1433-1498);
//   * Nonstationary_covariance::load_data's basis files
//     (src/class_nonstationary_covariance.cpp:46-170);
//   * Approximation::dump_prediction_result's output file
//     (src/class_approximation.cpp:473-550).
// Errors are status codes with a message (mra_io_last_error), never exit().
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <numeric>
#include <sstream>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "../../include/mra.h"

namespace {

thread_local std::string g_io_err;

int fail(int code, const std::string& msg) {
    g_io_err = msg;
    return code;
}

void copy_str(char* dst, size_t cap, const std::string& s) {
    std::snprintf(dst, cap, "%s", s.c_str());
}

bool is_file(const char* path) {
    struct stat st;
    return path && path[0] && stat(path, &st) == 0 && S_ISREG(st.st_mode);
}

// A little-endian binary file read in order, every read checked.
struct Reader {
    FILE* f = nullptr;
    std::string path;
    explicit Reader(const char* p) : path(p ? p : "") { f = p ? std::fopen(p, "rb") : nullptr; }
    ~Reader() {
        if (f) std::fclose(f);
    }
    template <typename T>
    bool read(T* dst, uint64_t n) {
        return n == 0 || (f && std::fread(dst, sizeof(T), n, f) == n);
    }
    template <typename T>
    bool read_vec(std::vector<T>& v, uint64_t n) {
        v.resize(n);
        return read(v.data(), n);
    }
};

}  // namespace

extern "C" const char* mra_io_last_error(void) { return g_io_err.c_str(); }

// ------------------------------------------------------------ user_parameters
extern "C" void mra_user_params_init(mra_user_params* p) {
    // the reference's parameters are zero-initialised globals
    // (src/constants.cpp:8-120): false, 0, empty strings
    if (p) std::memset(p, 0, sizeof(*p));
}

extern "C" int mra_read_user_parameters(const char* path, mra_user_params* p) {
    if (!path || !p) return fail(MRA_E_ARG, "null argument");
    std::ifstream file(path);
    if (!file.is_open()) return fail(MRA_E_ARG, std::string("cannot open the user parameters file ") + path);
    std::string line;
    while (std::getline(file, line)) {
        size_t k = 0;
        while (k < line.size() && (line[k] == ' ' || line[k] == '\t')) k++;
        if (k >= line.size() || line[k] == '#') continue;
        // name=value with spaces, tabs and quotes removed (read_user_parameters.cpp:26-46)
        std::string name, value;
        bool inName = true;
        for (size_t i = k; i < line.size(); i++) {
            const char c = line[i];
            if (c == ' ' || c == '\t' || c == '\'' || c == '"' || c == '\r') continue;
            if (c == '=') {
                inName = false;
                continue;
            }
            (inName ? name : value) += c;
        }
        if (name.empty()) continue;
        auto num = [&](double& d) {
            std::istringstream t(value);
            t >> d;
        };
        auto inum = [&](int32_t& i) {
            std::istringstream t(value);
            long v = 0;
            t >> v;
            i = (int32_t)v;
        };
        auto flag = [&](int32_t& b) {
            // operator>> with std::boolalpha: "true" / "false"; anything else
            // fails and leaves false
            b = value == "true" ? 1 : 0;
        };
        struct S {
            const char* key;
            char* dst;
            size_t cap;
        } strs[] = {{"DATA_FILE_NAME", p->data_file_name, sizeof(p->data_file_name)},
                    {"PREDICTION_RESULTS_FILE_NAME", p->prediction_results_file_name,
                     sizeof(p->prediction_results_file_name)},
                    {"SIGMASQ_FILE_NAME", p->sigmasq_file_name, sizeof(p->sigmasq_file_name)},
                    {"BETA_FILE_NAME", p->beta_file_name, sizeof(p->beta_file_name)},
                    {"TAUSQ_FILE_NAME", p->tausq_file_name, sizeof(p->tausq_file_name)},
                    {"CALCULATION_MODE", p->calculation_mode, sizeof(p->calculation_mode)},
                    {"PREDICTION_LOCATION_FILE", p->prediction_location_file, sizeof(p->prediction_location_file)},
                    {"TMP_DIRECTORY", p->tmp_directory, sizeof(p->tmp_directory)},
                    {"KNOTS_FILE_NAME", p->knots_file_name, sizeof(p->knots_file_name)},
                    {"PARENT_REGION_FILE_NAME", p->parent_region_file_name, sizeof(p->parent_region_file_name)}};
        bool done = false;
        for (auto& s : strs)
            if (name == s.key) {
                copy_str(s.dst, s.cap, value);
                done = true;
            }
        if (done) continue;
        struct B {
            const char* key;
            int32_t* dst;
        } flags[] = {{"NONSTATIONARY_FLAG", &p->nonstationary_flag},
                     {"PRINT_DETAIL_FLAG", &p->print_detail_flag},
                     {"DUMP_PREDICTION_RESULTS_FLAG", &p->dump_prediction_results_flag},
                     {"SAVE_TO_DISK_FLAG", &p->save_to_disk_flag},
                     {"DYNAMIC_SCHEDULE_FLAG", &p->dynamic_schedule_flag},
                     {"REGRESSION_FLAG", &p->regression_flag},
                     {"ELIMINATION_DUPLICATES_FLAG", &p->elimination_duplicates_flag},
                     {"CHORDAL_DISTANCE_FLAG", &p->chordal_distance_flag},
                     {"PROVIDED_KNOTS_FLAG", &p->provided_knots_flag},
                     {"OBS_AS_KNOTS_FLAG", &p->obs_as_knots_flag},
                     {"VARIOUS_J_FINEST_LEVEL_FLAG", &p->various_j_finest_level_flag}};
        for (auto& b : flags)
            if (name == b.key) {
                flag(*b.dst);
                done = true;
            }
        if (done) continue;
        struct D {
            const char* key;
            double* dst;
        } dbls[] = {{"SIGMASQ", &p->sigmasq},
                    {"BETA", &p->beta},
                    {"TAUSQ", &p->tausq},
                    {"SIGMASQ_SCALE", &p->sigmasq_scale},
                    {"BETA_SCALE", &p->beta_scale},
                    {"TAUSQ_SCALE", &p->tausq_scale},
                    {"SIGMASQ_LOWER_BOUND", &p->sigmasq_lower_bound},
                    {"BETA_LOWER_BOUND", &p->beta_lower_bound},
                    {"TAUSQ_LOWER_BOUND", &p->tausq_lower_bound},
                    {"SIGMASQ_UPPER_BOUND", &p->sigmasq_upper_bound},
                    {"BETA_UPPER_BOUND", &p->beta_upper_bound},
                    {"TAUSQ_UPPER_BOUND", &p->tausq_upper_bound},
                    {"SIGMASQ_INITIAL_GUESS", &p->sigmasq_initial_guess},
                    {"BETA_INITIAL_GUESS", &p->beta_initial_guess},
                    {"TAUSQ_INITIAL_GUESS", &p->tausq_initial_guess},
                    {"WENDLAND_LEN", &p->wendland_len}};
        for (auto& d : dbls)
            if (name == d.key) {
                num(*d.dst);
                done = true;
            }
        if (done) continue;
        if (name == "PREDICTION_LOCATION_MODE") {
            p->prediction_location_mode = value.empty() ? 0 : value[0];
        } else if (name == "OFFSET") {
            if (value == "default") {
                p->offset = std::exp(1.0) / 100;
            } else {
                num(p->offset);
                if (p->offset <= 0 || p->offset >= 1)
                    return fail(MRA_E_ARG, "OFFSET must be in (0, 1) or \"default\"");
            }
        } else if (name == "NUM_PARTITIONS_J") {
            inum(p->num_partitions_j);
        } else if (name == "NUM_KNOTS_r") {
            inum(p->num_knots_r);
        } else if (name == "NUM_LEVELS_M") {
            if (value == "default")
                p->num_levels_m = MRA_LEVELS_DEFAULT;
            else
                inum(p->num_levels_m);
        } else if (name == "MAX_ITERATIONS") {
            inum(p->max_iterations);
        } else if (name == "NUM_KNOTS_FINEST") {
            inum(p->num_knots_finest);
        }
        // unknown names are ignored, as in the reference
    }
    return mra_check_user_parameters(p);
}

// check_errors (src/read_user_parameters.cpp:428-637), single process
extern "C" int mra_check_user_parameters(const mra_user_params* p) {
    if (!p) return fail(MRA_E_ARG, "null argument");
    const std::string mode = p->calculation_mode;
    if (p->nonstationary_flag && p->wendland_len < 0)
        return fail(MRA_E_ARG, "WENDLAND_LEN must be positive when NONSTATIONARY_FLAG = true");
    if (!p->obs_as_knots_flag) return fail(MRA_E_UNSUPPORTED, "OBS_AS_KNOTS_FLAG = false is not implemented");
    if (p->various_j_finest_level_flag && !p->provided_knots_flag)
        return fail(MRA_E_ARG, "VARIOUS_J_FINEST_LEVEL_FLAG = true requires PROVIDED_KNOTS_FLAG = true");
    if (p->num_partitions_j != 2 && p->num_partitions_j != 4)
        return fail(MRA_E_ARG, "NUM_PARTITIONS_J must be 2 or 4");
    if (p->num_levels_m < 2 && p->num_levels_m != MRA_LEVELS_DEFAULT)
        return fail(MRA_E_ARG, "NUM_LEVELS_M must be larger than 1");
    if (mode != "prediction" && mode != "optimization" && mode != "likelihood" && mode != "build_structure_only")
        return fail(MRA_E_ARG, "CALCULATION_MODE must be prediction, optimization, likelihood or build_structure_only");
    const char m = p->prediction_location_mode;
    if (m != 'D' && m != 'N' && m != 'A') return fail(MRA_E_ARG, "PREDICTION_LOCATION_MODE must be 'D', 'N' or 'A'");
    if (p->save_to_disk_flag) {
        struct stat st;
        if (stat(p->tmp_directory, &st) != 0 || !S_ISDIR(st.st_mode))
            return fail(MRA_E_ARG, std::string("TMP_DIRECTORY ") + p->tmp_directory + " is not a directory");
    }
    if (mode == "optimization") {
        if (p->sigmasq_upper_bound < p->sigmasq_lower_bound || p->beta_upper_bound < p->beta_lower_bound ||
            p->tausq_upper_bound < p->tausq_lower_bound)
            return fail(MRA_E_ARG, "an upper bound is smaller than its lower bound");
        if (p->sigmasq_initial_guess < p->sigmasq_lower_bound || p->beta_initial_guess < p->beta_lower_bound ||
            p->tausq_initial_guess < p->tausq_lower_bound || p->sigmasq_initial_guess > p->sigmasq_upper_bound ||
            p->beta_initial_guess > p->beta_upper_bound || p->tausq_initial_guess > p->tausq_upper_bound)
            return fail(MRA_E_ARG, "an initial guess lies outside its bounds");
    } else if (p->nonstationary_flag) {
        if (p->sigmasq_scale <= 0 || p->beta_scale <= 0 || p->tausq_scale <= 0)
            return fail(MRA_E_ARG, "SIGMASQ_SCALE, BETA_SCALE and TAUSQ_SCALE must be positive");
    } else if (p->sigmasq <= 0 || p->beta <= 0 || p->tausq <= 0) {
        return fail(MRA_E_ARG, "SIGMASQ, BETA and TAUSQ must be positive");
    }
    return MRA_OK;
}

// ------------------------------------------------------------------- data
struct mra_data {
    uint64_t nLoc = 0, nObsRaw = 0, nPredRaw = 0;
    std::vector<double> olon, olat, oval;
    std::vector<uint64_t> oreg;
    std::vector<double> plon, plat;
    std::vector<uint64_t> preg;
    bool hasReg = false, predicting = false, regression = false;
    double bounds[4] = {};
    double coef[3] = {};
};

namespace {

// Indices of the distinct (lon, lat) pairs in increasing (lon, lat) order, the
// first occurrence kept (Data::eliminate_duplicate, src/class_data.cpp:400-506).
std::vector<uint64_t> unique_order(const std::vector<double>& lon, const std::vector<double>& lat) {
    std::vector<uint64_t> idx(lon.size());
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](uint64_t i, uint64_t j) {
        return lon[i] < lon[j] || (lon[i] == lon[j] && lat[i] < lat[j]);
    });
    std::vector<uint64_t> out;
    out.reserve(idx.size());
    for (size_t k = 0; k < idx.size(); k++)
        if (k == 0 || lon[idx[k]] != lon[idx[k - 1]] || lat[idx[k]] != lat[idx[k - 1]]) out.push_back(idx[k]);
    return out;
}

template <typename T>
void permute(std::vector<T>& v, const std::vector<uint64_t>& order) {
    if (v.empty()) return;
    std::vector<T> o(order.size());
    for (size_t k = 0; k < order.size(); k++) o[k] = v[order[k]];
    v.swap(o);
}

// Least squares fit of val on [1, lon, lat] by Householder QR (the solve()
// of src/class_data.cpp:333-345); returns false for a rank-deficient design.
bool regress(const std::vector<double>& lon, const std::vector<double>& lat, const std::vector<double>& val,
             double coef[3]) {
    const size_t n = val.size();
    if (n < 3) return false;
    std::vector<double> A(3 * n), b(val);
    for (size_t i = 0; i < n; i++) {
        A[i] = 1.0;
        A[n + i] = lon[i];
        A[2 * n + i] = lat[i];
    }
    double R[3][3] = {};
    for (int k = 0; k < 3; k++) {
        double* a = &A[k * n];
        double nrm = 0;
        for (size_t i = k; i < n; i++) nrm += a[i] * a[i];
        nrm = std::sqrt(nrm);
        if (nrm == 0) return false;
        const double alpha = a[k] > 0 ? -nrm : nrm;
        a[k] -= alpha;  // v = a - alpha e_k, stored in place
        double vv = 0;
        for (size_t i = k; i < n; i++) vv += a[i] * a[i];
        auto reflect = [&](double* x) {
            double d = 0;
            for (size_t i = k; i < n; i++) d += a[i] * x[i];
            const double s = 2 * d / vv;
            for (size_t i = k; i < n; i++) x[i] -= s * a[i];
        };
        for (int c = k + 1; c < 3; c++) reflect(&A[c * n]);
        reflect(b.data());
        R[k][k] = alpha;
        for (int c = k + 1; c < 3; c++) R[k][c] = A[c * n + k];
    }
    for (int k = 2; k >= 0; k--) {
        double s = b[k];
        for (int c = k + 1; c < 3; c++) s -= R[k][c] * coef[c];
        if (std::fabs(R[k][k]) < 1e-300) return false;
        coef[k] = s / R[k][k];
    }
    return true;
}

}  // namespace

extern "C" int mra_data_load(const mra_user_params* p, mra_data** out) {
    if (!p || !out) return fail(MRA_E_ARG, "null argument");
    *out = nullptr;
    if (!is_file(p->data_file_name))
        return fail(MRA_E_ARG, std::string("the data file ") + p->data_file_name + " does not exist");
    const bool predicting = std::string(p->calculation_mode) == "prediction";
    const char mode = p->prediction_location_mode;
    const bool regions = p->provided_knots_flag != 0;
    Reader rd(p->data_file_name);
    uint64_t n = 0;
    std::vector<double> lon, lat, val;
    std::vector<uint64_t> reg;
    const std::string bad = std::string("fail reading data file ") + p->data_file_name + " for ";
    if (!rd.read(&n, 1)) return fail(MRA_E_ARG, bad + "the number of locations");
    if (!rd.read_vec(lon, n)) return fail(MRA_E_ARG, bad + "longitude");
    if (!rd.read_vec(lat, n)) return fail(MRA_E_ARG, bad + "latitude");
    if (!rd.read_vec(val, n)) return fail(MRA_E_ARG, bad + "observation values");
    if (regions && !rd.read_vec(reg, n)) return fail(MRA_E_ARG, bad + "observation region indices");
    mra_data* d = new mra_data();
    d->nLoc = n;
    d->hasReg = regions;
    d->predicting = predicting;
    // bounds over every location, padded by 1e-6 of the extent (class_data.cpp:149-163)
    double mnl = 1e9, mna = 1e9, mxl = -1e9, mxa = -1e9;
    for (uint64_t i = 0; i < n; i++) {
        mnl = std::min(mnl, lon[i]);
        mna = std::min(mna, lat[i]);
        mxl = std::max(mxl, lon[i]);
        mxa = std::max(mxa, lat[i]);
    }
    d->bounds[0] = mnl;
    d->bounds[1] = mxl + 1e-6 * (mxl - mnl);
    d->bounds[2] = mna;
    d->bounds[3] = mxa + 1e-6 * (mxa - mna);
    // observations are the non-NaN values; mode N: the NaN ones are the prediction locations
    for (uint64_t i = 0; i < n; i++) {
        if (!std::isnan(val[i])) {
            d->olon.push_back(lon[i]);
            d->olat.push_back(lat[i]);
            d->oval.push_back(val[i]);
            if (regions) d->oreg.push_back(reg[i]);
        } else if (predicting && mode == 'N') {
            d->plon.push_back(lon[i]);
            d->plat.push_back(lat[i]);
            if (regions) d->preg.push_back(reg[i]);
        }
    }
    d->nObsRaw = d->olon.size();
    if (predicting && mode == 'D') {
        d->plon = lon;
        d->plat = lat;
        if (regions) d->preg = reg;
    }
    if (predicting && mode == 'A') {
        if (!is_file(p->prediction_location_file)) {
            delete d;
            return fail(MRA_E_ARG, std::string("the prediction location file ") + p->prediction_location_file +
                                       " does not exist");
        }
        Reader pr(p->prediction_location_file);
        uint64_t np = 0;
        const std::string pbad = std::string("fail reading prediction location file ") + p->prediction_location_file;
        bool ok = pr.read(&np, 1) && pr.read_vec(d->plon, np) && pr.read_vec(d->plat, np) &&
                  (!regions || pr.read_vec(d->preg, np));
        if (!ok) {
            delete d;
            return fail(MRA_E_ARG, pbad);
        }
    }
    d->nPredRaw = d->plon.size();
    if (p->elimination_duplicates_flag) {
        // the region indices travel with their locations (the reference leaves
        // them unpermuted, src/class_data.cpp:442-462)
        const std::vector<uint64_t> o = unique_order(d->olon, d->olat);
        permute(d->olon, o);
        permute(d->olat, o);
        permute(d->oval, o);
        permute(d->oreg, o);
        if (predicting) {
            const std::vector<uint64_t> q = unique_order(d->plon, d->plat);
            permute(d->plon, q);
            permute(d->plat, q);
            permute(d->preg, q);
        }
    }
    if (p->regression_flag) {
        d->regression = true;
        if (!regress(d->olon, d->olat, d->oval, d->coef)) {
            delete d;
            return fail(MRA_E_ARG, "the regression design [1, lon, lat] is rank deficient");
        }
        for (size_t i = 0; i < d->oval.size(); i++)
            d->oval[i] -= d->coef[0] + d->coef[1] * d->olon[i] + d->coef[2] * d->olat[i];
    }
    *out = d;
    return MRA_OK;
}

extern "C" void mra_data_destroy(mra_data* d) { delete d; }

extern "C" int mra_data_get(const mra_data* d, mra_data_view* v) {
    if (!d || !v) return fail(MRA_E_ARG, "null argument");
    std::memset(v, 0, sizeof(*v));
    v->n_locations = d->nLoc;
    v->n_obs_raw = d->nObsRaw;
    v->n_pred_raw = d->nPredRaw;
    v->n_obs = d->olon.size();
    v->n_pred = d->plon.size();
    v->obs_lon = d->olon.data();
    v->obs_lat = d->olat.data();
    v->obs_val = d->oval.data();
    v->obs_region = d->hasReg ? d->oreg.data() : nullptr;
    v->pred_lon = d->plon.data();
    v->pred_lat = d->plat.data();
    v->pred_region = d->hasReg && !d->preg.empty() ? d->preg.data() : nullptr;
    for (int k = 0; k < 4; k++) v->bounds[k] = d->bounds[k];
    v->regression = d->regression ? 1 : 0;
    for (int k = 0; k < 3; k++) v->coefficients[k] = d->coef[k];
    v->predicting = d->predicting ? 1 : 0;
    return MRA_OK;
}

// -------------------------------------------------------------- partition
struct mra_partition {
    int M = 0, J = 0, r = 0;
    bool various = false, userKnots = false;
    uint64_t nC = 0, nF = 0;
    std::vector<uint64_t> nPerLevel;
    std::vector<double> klon, klat;
    std::vector<uint64_t> parent, ooff, poff, oidx, pidx;
    std::vector<double> olon, olat, oval, plon, plat;
    bool hasPred = false;
};

// The default number of levels (Partition::Partition, src/class_partition.cpp:782-797):
// the smallest M with J^(M-1) >= ceil(n / r).
extern "C" int mra_default_levels(uint64_t n_obs, int J, int r) {
    if (n_obs == 0 || r < 1 || (J != 2 && J != 4)) return MRA_E_ARG;
    const double nF = std::ceil((double)n_obs / (double)r);
    const int M = (int)std::ceil(std::log(nF) / std::log((double)J)) + 1;
    return M;
}

namespace {

// Locations grouped by finest region; reg[i] = local finest index or -1.
// Stable: a region keeps its locations in input order.
void group_by(const std::vector<int64_t>& reg, uint64_t nF, std::vector<uint64_t>& offs,
              std::vector<uint64_t>& order) {
    offs.assign(nF + 1, 0);
    for (int64_t f : reg)
        if (f >= 0) offs[(uint64_t)f + 1]++;
    for (uint64_t f = 0; f < nF; f++) offs[f + 1] += offs[f];
    order.assign(offs[nF], 0);
    std::vector<uint64_t> pos(offs.begin(), offs.end() - 1);
    for (uint64_t i = 0; i < reg.size(); i++)
        if (reg[i] >= 0) order[pos[(uint64_t)reg[i]]++] = i;
}

int build_user_files(const mra_user_params* p, const mra_data* d, mra_partition* P) {
    // knots file: u64 nr; f64 lon[nr]; f64 lat[nr]; u64 region[nr], r knots per
    // coarse region in consecutive runs (src/class_partition.cpp:122-263)
    if (!is_file(p->knots_file_name))
        return fail(MRA_E_ARG, std::string("the knots file ") + p->knots_file_name + " does not exist");
    Reader rd(p->knots_file_name);
    uint64_t nk = 0;
    std::vector<double> lon, lat;
    std::vector<uint64_t> reg;
    if (!rd.read(&nk, 1) || !rd.read_vec(lon, nk) || !rd.read_vec(lat, nk) || !rd.read_vec(reg, nk))
        return fail(MRA_E_ARG, std::string("fail reading knots file ") + p->knots_file_name);
    const int r = P->r;
    P->klon.assign(P->nC * r, 0.0);
    P->klat.assign(P->nC * r, 0.0);
    for (uint64_t i = 0; i < nk;) {
        if (reg[i] >= P->nC) break;  // finest-level knots follow only when OBS_AS_KNOTS_FLAG = false
        if (i + r > nk) return fail(MRA_E_ARG, "the knots file ends inside a region's run of knots");
        for (int q = 0; q < r; q++) {
            const uint64_t R = reg[i + q];
            if (R >= P->nC) return fail(MRA_E_ARG, "a run of knots crosses into the finest level");
            P->klon[R * r + q] = lon[i + q];
            P->klat[R * r + q] = lat[i + q];
        }
        i += r;
    }
    if (!d->hasReg) return fail(MRA_E_ARG, "PROVIDED_KNOTS_FLAG = true needs region indices in the data file");
    const uint64_t F0 = P->nC;
    auto local = [&](const std::vector<uint64_t>& rg) {
        std::vector<int64_t> f(rg.size(), -1);
        for (size_t i = 0; i < rg.size(); i++)
            if (rg[i] >= F0 && rg[i] < F0 + P->nF) f[i] = (int64_t)(rg[i] - F0);
        return f;
    };
    std::vector<uint64_t> order;
    group_by(local(d->oreg), P->nF, P->ooff, order);
    P->oidx = order;
    for (uint64_t i : order) {
        P->olon.push_back(d->olon[i]);
        P->olat.push_back(d->olat[i]);
        P->oval.push_back(d->oval[i]);
    }
    if (d->predicting) {
        P->hasPred = true;
        if (d->preg.size() != d->plon.size()) return fail(MRA_E_ARG, "prediction locations lack region indices");
        group_by(local(d->preg), P->nF, P->poff, order);
        P->pidx = order;
        for (uint64_t i : order) {
            P->plon.push_back(d->plon[i]);
            P->plat.push_back(d->plat[i]);
        }
    }
    return MRA_OK;
}

int read_parent_file(const mra_user_params* p, mra_partition* P) {
    // u64 regionStart; u64 regionEnd; u64 parent[end - start + 1]
    // (Partition::read_parent_region_index, src/class_partition.cpp:659-770)
    if (!is_file(p->parent_region_file_name))
        return fail(MRA_E_ARG, std::string("the parent region file ") + p->parent_region_file_name + " does not exist");
    Reader rd(p->parent_region_file_name);
    uint64_t s = 0, e = 0;
    if (!rd.read(&s, 1) || !rd.read(&e, 1) || e < s)
        return fail(MRA_E_ARG, std::string("fail reading parent region file ") + p->parent_region_file_name);
    if (s != P->nC)
        return fail(MRA_E_ARG, "the finest region indices must start right after the coarse regions (" +
                                   std::to_string(P->nC) + "), the parent file starts at " + std::to_string(s));
    P->nF = e - s + 1;
    if (!rd.read_vec(P->parent, P->nF))
        return fail(MRA_E_ARG, std::string("fail reading parent region file ") + p->parent_region_file_name);
    const uint64_t second = P->nC - P->nPerLevel[P->M - 2];
    for (uint64_t par : P->parent)
        if (par < second || par >= P->nC)
            return fail(MRA_E_ARG, "a parent region index is not a region of the second finest level");
    return MRA_OK;
}

}  // namespace

extern "C" int mra_partition_build(const mra_user_params* p, const mra_data* d, mra_partition** out) {
    if (!p || !d || !out) return fail(MRA_E_ARG, "null argument");
    *out = nullptr;
    const int J = p->num_partitions_j, r = p->num_knots_r;
    if ((J != 2 && J != 4) || r < 1) return fail(MRA_E_ARG, "NUM_PARTITIONS_J must be 2 or 4 and NUM_KNOTS_r positive");
    int M = p->num_levels_m;
    if (M == MRA_LEVELS_DEFAULT) {
        M = mra_default_levels(d->olon.size(), J, r);
        if (M < 2)
            return fail(MRA_E_ARG, "the default number of levels is smaller than 2: use fewer knots or more data");
    }
    if (M < 2 || M > 24) return fail(MRA_E_ARG, "NUM_LEVELS_M must be in 2..24");
    mra_partition* P = new mra_partition();
    P->M = M;
    P->J = J;
    P->nPerLevel.assign(M, 0);
    uint64_t cnt = 1;
    P->nC = 0;
    for (int l = 0; l < M - 1; l++) {
        P->nPerLevel[l] = cnt;
        P->nC += cnt;
        cnt *= (uint64_t)J;
    }
    P->various = p->various_j_finest_level_flag != 0;
    P->userKnots = p->provided_knots_flag != 0;
    int rc = MRA_OK;
    if (P->various) {
        rc = read_parent_file(p, P);
    } else {
        P->nF = cnt;
        P->parent.resize(P->nF);
        for (uint64_t f = 0; f < P->nF; f++) P->parent[f] = (P->nC + f - 1) / (uint64_t)J;
    }
    if (!rc) P->nPerLevel[M - 1] = P->nF;
    if (!rc && P->userKnots) {
        P->r = r;
        rc = build_user_files(p, d, P);
    } else if (!rc) {
        // program-placed knots (build_partition_by_program) through the sort-based builder
        const double offset = p->offset > 0 ? p->offset : std::exp(1.0) / 100;
        const uint64_t n = d->olon.size(), np = d->predicting ? d->plon.size() : 0;
        uint64_t sz[4];
        rc = mra_partition_program_sizes(d->olon.data(), d->olat.data(), n, d->plon.data(), d->plat.data(), np, M, J,
                                         r, offset, d->bounds, sz);
        if (!rc) {
            P->r = (int)(sz[0] / P->nC);
            P->klon.resize(sz[0]);
            P->klat.resize(sz[0]);
            P->ooff.resize(P->nF + 1);
            P->olon.resize(sz[2]);
            P->olat.resize(sz[2]);
            P->oval.resize(sz[2]);
            P->oidx.resize(sz[2]);
            P->hasPred = d->predicting;
            P->poff.resize(P->nF + 1);
            P->plon.resize(sz[3]);
            P->plat.resize(sz[3]);
            P->pidx.resize(sz[3]);
            rc = mra_partition_program(d->olon.data(), d->olat.data(), d->oval.data(), n, d->plon.data(),
                                       d->plat.data(), np, M, J, r, offset, d->bounds, P->klon.data(), P->klat.data(),
                                       nullptr, P->ooff.data(), P->olon.data(), P->olat.data(), P->oval.data(),
                                       P->oidx.data(), P->hasPred ? P->poff.data() : nullptr, P->plon.data(),
                                       P->plat.data(), P->pidx.data());
            if (rc) fail(rc, "program partition failed");
        }
    }
    if (rc) {
        delete P;
        return rc;
    }
    *out = P;
    return MRA_OK;
}

extern "C" void mra_partition_destroy(mra_partition* P) { delete P; }

extern "C" int mra_partition_tree(const mra_partition* P, mra_tree_desc* t) {
    if (!P || !t) return fail(MRA_E_ARG, "null argument");
    std::memset(t, 0, sizeof(*t));
    t->num_levels = P->M;
    t->num_partitions = P->J;
    t->num_knots = P->r;
    t->various_j = P->various ? 1 : 0;
    t->n_coarse = P->nC;
    t->n_finest = P->nF;
    t->n_obs = P->olon.size();
    t->knot_lon = P->klon.data();
    t->knot_lat = P->klat.data();
    t->finest_parent = P->parent.data();
    t->obs_offsets = P->ooff.data();
    t->obs_lon = P->olon.data();
    t->obs_lat = P->olat.data();
    t->obs_val = P->oval.data();
    if (P->hasPred) {
        t->n_pred = P->plon.size();
        t->pred_offsets = P->poff.data();
        t->pred_lon = P->plon.data();
        t->pred_lat = P->plat.data();
    }
    return MRA_OK;
}

extern "C" int mra_partition_indices(const mra_partition* P, const uint64_t** obs_index, const uint64_t** pred_index) {
    if (!P) return fail(MRA_E_ARG, "null argument");
    if (obs_index) *obs_index = P->oidx.data();
    if (pred_index) *pred_index = P->hasPred ? P->pidx.data() : nullptr;
    return MRA_OK;
}

extern "C" int mra_partition_level_counts(const mra_partition* P, uint64_t* counts) {
    if (!P || !counts) return fail(MRA_E_ARG, "null argument");
    for (int l = 0; l < P->M; l++) counts[l] = P->nPerLevel[l];
    return MRA_OK;
}

// Partition::dump_structure_information (src/class_partition.cpp:1433-1498),
// the same text line for line.
extern "C" int mra_partition_dump_structure(const mra_partition* P, const char* path) {
    if (!P) return fail(MRA_E_ARG, "null argument");
    const char* fn = path && path[0] ? path : "structure_information.txt";
    FILE* f = std::fopen(fn, "w");
    if (!f) return fail(MRA_E_ARG, std::string("cannot open ") + fn);
    const uint64_t nTot = P->nC + P->nF;
    std::vector<uint64_t> cnt(P->nF);
    for (uint64_t k = 0; k < P->nF; k++) cnt[k] = P->ooff[k + 1] - P->ooff[k];
    std::fprintf(f, "The number of levels: %d\n\n", P->M);
    std::fprintf(f, "The total number of regions: %llu\n\n", (unsigned long long)nTot);
    std::fprintf(f, "The actual number of knots in regions before the finest level: %d\n\n", P->r);
    uint64_t mn = cnt.empty() ? 0 : cnt[0], mx = mn, zero = 0;
    for (uint64_t c : cnt) {
        mn = std::min(mn, c);
        mx = std::max(mx, c);
        zero += c == 0;
    }
    std::fprintf(f, "The minimal number of knots in regions at the finest level: %llu\n\n", (unsigned long long)mn);
    std::fprintf(f, "The maximal number of knots in regions at the finest level: %llu\n\n", (unsigned long long)mx);
    std::fprintf(f, "The number of in regions at the finest level that have zero knots: %llu\n\n",
                 (unsigned long long)zero);
    // ten intervals of width (max - min) / 10 in integer arithmetic
    const uint64_t width = (mx - mn) / 10;
    uint64_t left = mn;
    std::fprintf(f, "The number of regions at the finest level with number of knots in the following intervals:\n");
    for (int i = 0; i < 10; i++) {
        uint64_t right = i == 9 ? mx : left + width;
        uint64_t k = 0;
        for (uint64_t c : cnt) k += (left <= c && c <= right);
        std::fprintf(f, "[ %llu , %llu ]: %llu\n", (unsigned long long)left, (unsigned long long)right,
                     (unsigned long long)k);
        left = right + 1;
    }
    std::fprintf(f, "\n");
    std::fprintf(f, "The number of regions in each level: \n");
    for (int l = 0; l < P->M; l++) std::fprintf(f, "Level %d: %llu\n", l + 1, (unsigned long long)P->nPerLevel[l]);
    std::fprintf(f, "\n");
    std::fprintf(f, "The number of knots in each region at the finest level: \n");
    for (uint64_t k = 0; k < P->nF; k++)
        std::fprintf(f, "Region %llu: %llu\n", (unsigned long long)(nTot - P->nF + k), (unsigned long long)cnt[k]);
    const bool ok = std::fclose(f) == 0;
    return ok ? MRA_OK : fail(MRA_E_ARG, std::string("short write to ") + fn);
}

// ------------------------------------------------------- covariance model
struct mra_cov_model {
    std::vector<double> lon[3], lat[3], w[3];
    double intercept[3] = {};
    mra_cov_desc desc{};
};

namespace {

int read_basis(const char* path, std::vector<double>& lon, std::vector<double>& lat, std::vector<double>& w,
               double& intercept) {
    // u64 nb; f64 intercept; f64 lon[nb]; f64 lat[nb]; f64 w[nb]
    // (src/class_nonstationary_covariance.cpp:84-170)
    if (!is_file(path)) return fail(MRA_E_ARG, std::string("the basis file ") + (path ? path : "") + " does not exist");
    Reader rd(path);
    uint64_t nb = 0;
    if (!rd.read(&nb, 1) || !rd.read(&intercept, 1) || !rd.read_vec(lon, nb) || !rd.read_vec(lat, nb) ||
        !rd.read_vec(w, nb))
        return fail(MRA_E_ARG, std::string("fail reading basis file ") + path);
    return MRA_OK;
}

}  // namespace

extern "C" int mra_cov_model_load(const mra_user_params* p, mra_cov_model** out) {
    if (!p || !out) return fail(MRA_E_ARG, "null argument");
    *out = nullptr;
    mra_cov_model* c = new mra_cov_model();
    c->desc.nonstationary = p->nonstationary_flag ? 1 : 0;
    c->desc.metric = p->chordal_distance_flag ? MRA_METRIC_CHORDAL : MRA_METRIC_GREAT_CIRCLE;
    c->desc.wendland_len = p->wendland_len;
    mra_field_desc* f[3] = {&c->desc.sigmasq, &c->desc.beta, &c->desc.tausq};
    const char* files[3] = {p->sigmasq_file_name, p->beta_file_name, p->tausq_file_name};
    for (int k = 0; k < 3; k++) {
        f[k]->kind = MRA_FIELD_CONSTANT;
        if (!p->nonstationary_flag) continue;
        const int rc = read_basis(files[k], c->lon[k], c->lat[k], c->w[k], c->intercept[k]);
        if (rc) {
            delete c;
            return rc;
        }
        f[k]->kind = MRA_FIELD_WENDLAND;
        f[k]->n_basis = c->lon[k].size();
        f[k]->intercept = c->intercept[k];
        f[k]->basis_lon = c->lon[k].data();
        f[k]->basis_lat = c->lat[k].data();
        f[k]->basis_weight = c->w[k].data();
    }
    *out = c;
    return MRA_OK;
}

extern "C" void mra_cov_model_destroy(mra_cov_model* c) { delete c; }

extern "C" int mra_cov_model_desc(const mra_cov_model* c, mra_cov_desc* desc) {
    if (!c || !desc) return fail(MRA_E_ARG, "null argument");
    *desc = c->desc;
    return MRA_OK;
}

// ------------------------------------------------------- prediction output
// u64 n; f64 lon[n]; f64 lat[n]; f64 mean[n]; f64 var[n]; with coef non-NULL
// the mean adds the regression trend coef[0] + coef[1] lon + coef[2] lat
// (src/class_approximation.cpp:473-550, one process).
extern "C" int mra_write_prediction_result(const char* path, uint64_t n, const double* lon, const double* lat,
                                           const double* mean, const double* var, const double* coef) {
    if (!path || (n && (!lon || !lat || !mean || !var))) return fail(MRA_E_ARG, "null argument");
    FILE* f = std::fopen(path, "wb");
    if (!f) return fail(MRA_E_ARG, std::string("fail to open the output file ") + path);
    bool ok = std::fwrite(&n, sizeof(n), 1, f) == 1;
    ok = ok && std::fwrite(lon, sizeof(double), n, f) == n;
    ok = ok && std::fwrite(lat, sizeof(double), n, f) == n;
    if (coef) {
        for (uint64_t i = 0; ok && i < n; i++) {
            // dot([1, lon, lat], coef) in the order of the reference's dot()
            const double m = mean[i] + (1.0 * coef[0] + lon[i] * coef[1] + lat[i] * coef[2]);
            ok = std::fwrite(&m, sizeof(double), 1, f) == 1;
        }
    } else {
        ok = ok && std::fwrite(mean, sizeof(double), n, f) == n;
    }
    ok = ok && std::fwrite(var, sizeof(double), n, f) == n;
    ok = (std::fclose(f) == 0) && ok;
    return ok ? MRA_OK : fail(MRA_E_ARG, std::string("short write to ") + path);
}
