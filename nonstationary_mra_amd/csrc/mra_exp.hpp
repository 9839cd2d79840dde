// e^x for x <= 0 in FP64, shared by the device covariance code and a host
// test.  Table-driven method of P. T. P. Tang, "Table-driven implementation
// of the exponential function in IEEE floating-point arithmetic", ACM TOMS
// 15(2), 1989, with 64 table entries:
//   x = k ln2/64 + r, k = rint(64 x / ln2), |r| <= ln2/128 (ln2/64 in two
//   parts, the first with 21 trailing zero bits so that k times it is exact),
//   e^r = degree-5 Taylor polynomial in Horner form (truncation below
//   3.5e-17 relative on that interval),
//   e^x = 2^(k >> 6) * T[k & 63] * e^r,  T[m] = 2^(m/64) rounded to double.
// x < -708 (below the normal range) gives 0.
#pragma once

#if defined(__HIPCC__)
#define MRA_HD __host__ __device__
#else
#define MRA_HD
#endif

namespace mra {

struct Exp2Frac {
    // 2^(m/64), m = 0..63, correctly rounded
    static constexpr double T[64] = {
    1, 1.0108892860517005, 1.0218971486541166, 1.0330248790212284,
    1.0442737824274138, 1.0556451783605572, 1.0671404006768237, 1.0787607977571199,
    1.0905077326652577, 1.1023825833078409, 1.1143867425958924, 1.1265216186082418,
    1.1387886347566916, 1.1511892299529827, 1.1637248587775775, 1.1763969916502812,
    1.189207115002721, 1.2021567314527031, 1.215247359980469, 1.22848053610687,
    1.241857812073484, 1.2553807570246911, 1.2690509571917332, 1.2828700160787783,
    1.2968395546510096, 1.3109612115247644, 1.3252366431597413, 1.3396675240533029,
    1.3542555469368927, 1.3690024229745905, 1.383909881963832, 1.3989796725383112,
    1.4142135623730951, 1.42961333839197, 1.4451808069770467, 1.460917794180647,
    1.4768261459394993, 1.4929077282912648, 1.5091644275934228, 1.5255981507445384,
    1.5422108254079407, 1.5590044002378369, 1.5759808451078865, 1.593142151342267,
    1.6104903319492543, 1.6280274218573478, 1.6457554781539649, 1.6636765803267364,
    1.681792830507429, 1.7001063537185235, 1.7186192981224779, 1.7373338352737062,
    1.7562521603732995, 1.7753764925265212, 1.7947090750031072, 1.8142521755003989,
    1.8340080864093424, 1.8539791250833855, 1.8741676341103, 1.8945759815869656,
    1.9152065613971474, 1.9360617934922943, 1.9571441241754002, 1.9784560263879509,
    };
};

// The two halves of exp_neg, so that a caller can issue several table loads
// before it needs the first: the reduction and polynomial (p = e^r, k), and
// the scaling by the table entry T[k & 63] and 2^(k >> 6).
struct ExpParts {
    double p;
    int k;
};

MRA_HD inline ExpParts exp_neg_parts(double x) {
    const double kd = __builtin_rint(x * 92.332482616893658);
    double r = __builtin_fma(kd, -0.01083042469326756, x);
    r = __builtin_fma(kd, -2.9815858269852933e-12, r);
    double p = __builtin_fma(r, 1.0 / 120.0, 1.0 / 24.0);
    p = __builtin_fma(p, r, 1.0 / 6.0);
    p = __builtin_fma(p, r, 0.5);
    p = __builtin_fma(p, r, 1.0);
    p = __builtin_fma(p, r, 1.0);
    return ExpParts{p, (int)kd};
}

MRA_HD inline double exp_neg_finish(double tk, const ExpParts& e, double x) {
    const double v = __builtin_ldexp(tk * e.p, e.k >> 6);
    return x < -708.0 ? 0.0 : v;
}

MRA_HD inline double exp_neg(double x) {
    const ExpParts e = exp_neg_parts(x);
    return exp_neg_finish(Exp2Frac::T[e.k & 63], e, x);
}

}  // namespace mra
