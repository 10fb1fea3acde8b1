// Branch-free FP64 exp / expm1 / log / pow / tanh for the hot path.
//
// The device math library's FP64 tanh and pow are long, branchy sequences
// (≈165 and ≈215 VALU instructions on gfx950); the hot path calls them about 32
// times per cell-iteration.  These replacements use textbook methods:
//   * exp:   x = k ln2 + r with a two-part ln2 (Cody & Waite), |r| <= ln2/2,
//            e^r - 1 by its Taylor series to r^14/14! (truncation < 1e-17),
//            2^k applied with ldexp;
//   * expm1: the same reduction, result (2^k - 1) + 2^k (e^r - 1), exact scaling;
//   * tanh:  expm1(2|x|) / (expm1(2|x|) + 2), saturating at |x| >= 20 where
//            tanh rounds to 1;
//   * log:   x = m 2^e with m in [sqrt(1/2), sqrt(2)), log m = 2 atanh(s),
//            s = (m-1)/(m+1), atanh by its odd series to s^21/21;
//   * pow:   exp(y log x) with the product's rounding error carried by FMA.
// Accuracy against glibc is checked on the CPU (tests/test_tmath.py): a few
// ulps at most.  Written for __host__ __device__ so the same code is tested.
#pragma once

#include <cmath>

#if defined(__HIPCC__)
#define TDT_HD __host__ __device__
#else
#define TDT_HD
#endif

namespace tdt {
namespace tm {

// ln 2 rounded to 32 significant bits (k * kLn2Hi exact for |k| < 2^21) and the rest
constexpr double kLn2Hi = 0x1.62e42ff000000p-1;
constexpr double kLn2Lo = -0x1.718432a1b0e26p-35;
constexpr double kInvLn2 = 1.4426950408889634;
constexpr double kSqrtHalf = 0.7071067811865476;

// a / b.  On the device: the refined-reciprocal sequence of the compiler's
// FP64 division without its scaling and special-case steps (not needed for the
// normal, finite operands below); same bits as a / b there.
TDT_HD inline double quot(double a, double b) {
#if defined(__HIP_DEVICE_COMPILE__)
    const double r0 = __builtin_amdgcn_rcp(b);
    const double r1 = __builtin_fma(r0, __builtin_fma(-b, r0, 1.0), r0);
    const double r = __builtin_fma(r1, __builtin_fma(-b, r1, 1.0), r1);
    const double q = a * r;
    return __builtin_fma(__builtin_fma(-b, q, a), r, q);
#else
    return a / b;
#endif
}

// e^r - 1 for |r| <= ln2/2: r (1 + r/2! + r^2/3! + ... + r^13/14!)
TDT_HD inline double expm1_reduced(double r) {
    double p = 1.1470745597729725e-11;          // 1/14!
    p = std::fma(p, r, 1.6059043836821613e-10);  // 1/13!
    p = std::fma(p, r, 2.08767569878681e-09);    // 1/12!
    p = std::fma(p, r, 2.505210838544172e-08);   // 1/11!
    p = std::fma(p, r, 2.755731922398589e-07);   // 1/10!
    p = std::fma(p, r, 2.7557319223985893e-06);  // 1/9!
    p = std::fma(p, r, 2.48015873015873e-05);    // 1/8!
    p = std::fma(p, r, 0.0001984126984126984);   // 1/7!
    p = std::fma(p, r, 0.001388888888888889);    // 1/6!
    p = std::fma(p, r, 0.008333333333333333);    // 1/5!
    p = std::fma(p, r, 0.041666666666666664);    // 1/4!
    p = std::fma(p, r, 0.16666666666666666);     // 1/3!
    p = std::fma(p, r, 0.5);                     // 1/2!
    p = std::fma(p, r, 1.0);
    return r * p;
}

// x = k ln2 + r
TDT_HD inline void reduce_ln2(double x, double &k, double &r) {
    k = std::rint(x * kInvLn2);
    r = std::fma(-k, kLn2Hi, x);
    r = std::fma(-k, kLn2Lo, r);
}

TDT_HD inline double exp(double x) {
    // outside [-1100, 1100] the result is 0 or inf anyway; keeps k an int
    const double xc = x > 1100.0 ? 1100.0 : (x < -1100.0 ? -1100.0 : x);
    double k, r;
    reduce_ln2(xc, k, r);
    const double e = expm1_reduced(r) + 1.0;
    return std::ldexp(e, (int)k);
}

// e^y - 1 for y in [0, 40]
TDT_HD inline double expm1_pos(double y) {
    double k, r;
    reduce_ln2(y, k, r);
    const double q = expm1_reduced(r);
    const double s = std::ldexp(1.0, (int)k);
    return (s - 1.0) + s * q;
}

TDT_HD inline double tanh(double x) {
    double y = 2.0 * std::fabs(x);
    y = y > 40.0 ? 40.0 : y;        // tanh(20) rounds to 1; NaN passes through
    const double e = expm1_pos(y);
    return std::copysign(quot(e, e + 2.0), x);
}

// natural log of a positive, finite, normal x
TDT_HD inline double log_pos(double x) {
    int ex;
    double m = std::frexp(x, &ex);   // x = m 2^ex, m in [0.5, 1)
    const bool low = m < kSqrtHalf;
    m = low ? m + m : m;
    const double e = (double)(low ? ex - 1 : ex);
    const double f = m - 1.0;        // exact
    const double s = quot(f, m + 1.0);
    const double z = s * s;
    double p = 0.047619047619047616;             // 1/21
    p = std::fma(p, z, 0.05263157894736842);     // 1/19
    p = std::fma(p, z, 0.058823529411764705);    // 1/17
    p = std::fma(p, z, 0.06666666666666667);     // 1/15
    p = std::fma(p, z, 0.07692307692307693);     // 1/13
    p = std::fma(p, z, 0.09090909090909091);     // 1/11
    p = std::fma(p, z, 0.1111111111111111);      // 1/9
    p = std::fma(p, z, 0.14285714285714285);     // 1/7
    p = std::fma(p, z, 0.2);                     // 1/5
    p = std::fma(p, z, 0.3333333333333333);      // 1/3
    const double s2 = s + s;
    const double lm = std::fma(s2 * z, p, s2);   // 2s + 2s z p
    return std::fma(e, kLn2Hi, std::fma(e, kLn2Lo, lm));
}

// x^y for the hot path's arguments (x > 0 finite, y finite); pow's special
// cases for x <= 0, inf and NaN are kept
TDT_HD inline double pow(double x, double y) {
    const bool ok = x > 0.0 && x < INFINITY;
    const double l = log_pos(ok ? x : 1.0);
    const double p = y * l;
    const double pe = std::fma(y, l, -p);        // rounding error of y * l
    const double r = tm::exp(p);
    double res = std::fma(r, pe, r);             // e^(p + pe)
    if (!ok) {
        if (x == 0.0) res = y > 0.0 ? 0.0 : (y < 0.0 ? INFINITY : 1.0);
        else if (x == INFINITY) res = y > 0.0 ? INFINITY : (y < 0.0 ? 0.0 : 1.0);
        else res = y == 0.0 ? 1.0 : NAN;         // x < 0 (non-integer y) or NaN
    }
    return res;
}

}  // namespace tm
}  // namespace tdt
