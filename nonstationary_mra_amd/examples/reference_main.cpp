// The reference's driver (src/main.cpp:104-148) compiled against this
// library's C++ layer: the same sequence of calls, without the MPI start-up
// and shut-down.  Reads "user_parameters" from the working directory (or the
// file named by the first argument) and runs its CALCULATION_MODE:
// build_structure_only, likelihood, prediction or optimization.
//
// Prints "D <value>" and "loglik <value>" for the likelihood and prediction
// modes (the reference prints the latter as "The obtained loglikelihood is").
#include <cstdio>

#include "mra_optimization.hpp"

int main(int argc, char** argv) {
    try {
        read_user_parameters(argc > 1 ? argv[1] : "user_parameters");
        data = new Data;
        data->load_data();
        if (NONSTATIONARY_FLAG) {
            nonstat_cov = new Nonstationary_covariance;
            nonstat_cov->load_data();
        }
        partition = new Partition;
        partition->build_partition();
        if (CALCULATION_MODE == "build_structure_only") {
            partition->dump_structure_information();
            std::printf("structure_information.txt written\n");
        }
        if (CALCULATION_MODE == "likelihood") {
            Approximation approximation(SIGMASQ, BETA, TAUSQ, SIGMASQ_SCALE, BETA_SCALE, TAUSQ_SCALE);
            approximation.likelihood();
            std::printf("D %.17g\nloglik %.17g\n", approximation.loglikelihood, approximation.printedLoglikelihood);
        }
        if (CALCULATION_MODE == "prediction") {
            Approximation approximation(SIGMASQ, BETA, TAUSQ, SIGMASQ_SCALE, BETA_SCALE, TAUSQ_SCALE);
            approximation.likelihood();
            std::printf("D %.17g\nloglik %.17g\n", approximation.loglikelihood, approximation.printedLoglikelihood);
            approximation.predict();
            if (DUMP_PREDICTION_RESULTS_FLAG) approximation.dump_prediction_result();
            // the per-region views of the reference's vec arrays
            double s = 0.0;
            for (unsigned long f = 0; f < partition->nRegionsAtFinestLevel; f++)
                for (unsigned long i = 0; i < approximation.posteriorPredictionMean[f].n_elem; i++)
                    s += approximation.posteriorPredictionMean[f][i];
            std::printf("mean_sum %.17g\n", s);
        }
        if (CALCULATION_MODE == "optimization") {
            if (NONSTATIONARY_FLAG)
                get_optimal_nonstationary_parameters();
            else
                get_optimal_parameters();
            const mra::OptimizationResult& r = mra::last_optimization;
            std::printf("opt %.17g %.17g %.17g D %.17g evals %d\n", r.parameter[0], r.parameter[1], r.parameter[2],
                        r.D, r.evaluations);
        }
        mra::Context::release_globals();
        delete partition;
        delete nonstat_cov;
        delete data;
    } catch (const mra::Error& e) {
        std::fprintf(stderr, "mra error %d: %s\n", e.code, e.what());
        return 1;
    }
    return 0;
}
