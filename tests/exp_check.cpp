// Host check of mra::exp_neg (nonstationary_mra_amd/csrc/mra_exp.hpp)
// against std::exp: prints the largest error in units in the last place over
// a dense sweep of [-708, 0] and at special arguments.
#include <cmath>
#include <cstdio>
#include <cstdint>
#include <cstring>

#include "mra_exp.hpp"

static double ulp_err(double got, double ref) {
    if (got == ref) return 0.0;
    const double u = std::nextafter(ref, INFINITY) - ref;
    return std::fabs(got - ref) / u;
}

int main() {
    double worst = 0.0, wx = 0.0;
    const int N = 4000000;
    for (int i = 0; i <= N; i++) {
        const double x = -708.0 * i / N;
        const double e = ulp_err(mra::exp_neg(x), std::exp(x));
        if (e > worst) {
            worst = e;
            wx = x;
        }
    }
    // small arguments, table boundaries, the underflow threshold
    const double sp[] = {0.0, -0.0, -1e-300, -1e-17, -1e-8, -0.0054, -0.0055, -0.693147, -1.0, -10.0, -707.99};
    for (double x : sp) {
        const double e = ulp_err(mra::exp_neg(x), std::exp(x));
        if (e > worst) {
            worst = e;
            wx = x;
        }
    }
    std::printf("%.3f %.17g %d\n", worst, wx, mra::exp_neg(-709.0) == 0.0 ? 1 : 0);
    return 0;
}
