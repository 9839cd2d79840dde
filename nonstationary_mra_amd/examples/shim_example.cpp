// Reference-style driver over the C++ layer (include/mra_approximation.hpp):
// the "prediction" mode of src/main.cpp:127-133 with the globals replaced by
// an mra::Context.
//
//   shim_example DATA.bin PRED.bin OUT.bin M J r [mixed]
//
// DATA.bin / PRED.bin use the reference's formats (u64 n; f64 lon[n],
// lat[n][, val[n]]).  The covariance model is the planar nonstationary
// exponential with linear-trend log fields
//   log sigma^2 = 0.5 x - 0.3 y,  log beta = log 0.05 + 0.8 x,  tau^2 = 0.05.
// Prints "D <value>" and writes the prediction file with
// Approximation::dump_prediction_result.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "mra_approximation.hpp"

static std::vector<double> read_f64(FILE* f, uint64_t n) {
    std::vector<double> v(n);
    if (n && std::fread(v.data(), sizeof(double), n, f) != n) throw mra::Error(MRA_E_ARG, "short read");
    return v;
}

int main(int argc, char** argv) {
    if (argc < 7) {
        std::fprintf(stderr, "usage: %s DATA.bin PRED.bin OUT.bin M J r [mixed]\n", argv[0]);
        return 2;
    }
    try {
        FILE* fd = std::fopen(argv[1], "rb");
        FILE* fp = std::fopen(argv[2], "rb");
        if (!fd || !fp) throw mra::Error(MRA_E_ARG, "cannot open input files");
        uint64_t n = 0, np = 0;
        if (std::fread(&n, 8, 1, fd) != 1 || std::fread(&np, 8, 1, fp) != 1) throw mra::Error(MRA_E_ARG, "bad header");
        std::vector<double> lon = read_f64(fd, n), lat = read_f64(fd, n), val = read_f64(fd, n);
        std::vector<double> plon = read_f64(fp, np), plat = read_f64(fp, np);
        std::fclose(fd);
        std::fclose(fp);
        const int M = std::atoi(argv[4]), J = std::atoi(argv[5]), r = std::atoi(argv[6]);
        const double bounds[4] = {0.0, 1.0, 0.0, 1.0};
        const double offset = std::exp(1.0) / 100.0;
        // Partition::build_partition (program-placed knots)
        uint64_t sizes[4];
        mra::check(mra_partition_program_sizes(lon.data(), lat.data(), n, plon.data(), plat.data(), np, M, J, r, offset,
                                               bounds, sizes),
                   nullptr);
        uint64_t nC = 0, pw = 1;
        for (int l = 0; l < M - 1; l++, pw *= J) nC += pw;
        const uint64_t nk = sizes[0], nF = sizes[1], no = sizes[2], nq = sizes[3];
        std::vector<double> kl(nk), kt(nk), ol(no), ot(no), ov(no), pl(nq), pt(nq);
        std::vector<uint64_t> parent(nF), oo(nF + 1), po(nF + 1), oi(no), pi(nq);
        mra::check(mra_partition_program(lon.data(), lat.data(), val.data(), n, plon.data(), plat.data(), np, M, J, r,
                                         offset, bounds, kl.data(), kt.data(), parent.data(), oo.data(), ol.data(),
                                         ot.data(), ov.data(), oi.data(), po.data(), pl.data(), pt.data(), pi.data()),
                   nullptr);
        mra_tree_desc tree{};
        tree.num_levels = M;
        tree.num_partitions = J;
        tree.num_knots = (int)(nk / nC);
        tree.n_coarse = nC;
        tree.n_finest = nF;
        tree.n_obs = no;
        tree.n_pred = nq;
        tree.knot_lon = kl.data();
        tree.knot_lat = kt.data();
        tree.finest_parent = parent.data();
        tree.obs_offsets = oo.data();
        tree.obs_lon = ol.data();
        tree.obs_lat = ot.data();
        tree.obs_val = ov.data();
        tree.pred_offsets = po.data();
        tree.pred_lon = pl.data();
        tree.pred_lat = pt.data();
        // Nonstationary_covariance::load_data
        mra_cov_desc cov{};
        cov.nonstationary = 1;
        cov.metric = MRA_METRIC_PLANAR;
        cov.wendland_len = 1.0;
        cov.sigmasq.kind = MRA_FIELD_LINEAR;
        cov.sigmasq.intercept = 0.0;
        cov.sigmasq.slope_x = 0.5;
        cov.sigmasq.slope_y = -0.3;
        cov.beta.kind = MRA_FIELD_LINEAR;
        cov.beta.intercept = std::log(0.05);
        cov.beta.slope_x = 0.8;
        cov.tausq.kind = MRA_FIELD_CONSTANT;
        cov.tausq.intercept = std::log(0.05);
        mra::Context ctx(tree, cov);
        ctx.use();
        if (argc > 7) ctx.set_precision(MRA_PRECISION_MIXED);
        // src/main.cpp:127-133
        Approximation approximation(0.0, 1.0, 0.0, 1.0, 1.0, 1.0);
        approximation.likelihood();
        std::printf("D %.17g\n", approximation.loglikelihood);
        approximation.predict();
        approximation.dump_prediction_result(argv[3]);
        // evaluate_nonstationary_variance_covariance on the first few observations
        const unsigned long k = n < 4 ? n : 4;
        std::vector<double> C(k * k);
        evaluate_nonstationary_variance_covariance(C.data(), lon.data(), lat.data(), nullptr, nullptr, nullptr, k, 1.0,
                                                   1.0, 1.0);
        std::printf("C00 %.17g\n", C[0]);
    } catch (const mra::Error& e) {
        std::fprintf(stderr, "mra error %d: %s\n", e.code, e.what());
        return 1;
    }
    return 0;
}
