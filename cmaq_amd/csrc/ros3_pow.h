// Correctly rounded pow for positive normal arguments, in double-double
// arithmetic, for the step/order selection of SMVGEAR.
//
// SMVGEAR's discrete control (order changes, the 1.3/1.2/1.4 step-ratio
// choices, the RDELMAX / FRACDEC tests) turns a last-bit difference of one
// pow into another step sequence in a few stiff cells.  A pow whose result is
// the double nearest the exact x^y is the one every correct library agrees
// with up to its own misroundings, so the device uses this one rather than a
// 1-ulp library pow.
//
// Method: log x = e ln2 + 2 atanh(s), s = (m-1)/(m+1), m in [sqrt(1/2),
// sqrt(2)), summed as a double-double series; w = y log x; exp w = 2^k exp r,
// r = w - k ln2, exp r = (1 + expm1(r/256))^256 with the squarings done on
// expm1 (2e + e^2) so that no precision is lost near 1.  The relative error of
// the double-double result is below 2^-95 for |y log x| < 700; the final
// rounding of hi + lo then misrounds only when x^y lies within that distance
// of a midpoint between two doubles.  Zero, negative, infinite and NaN
// arguments and results outside the normal range go to the library pow,
// whose results there are exact or irrelevant (0, inf, NaN).
#pragma once

#include <math.h>

#ifndef ROS3_HD
#if defined(__HIPCC__)
#define ROS3_HD __host__ __device__
#else
#define ROS3_HD
#endif
#endif

namespace ros3 {
namespace crpow {

struct dd {
    double hi, lo;
};

ROS3_HD inline dd two_sum(double a, double b) {
    const double s = a + b;
    const double bb = s - a;
    return {s, (a - (s - bb)) + (b - bb)};
}

ROS3_HD inline dd fast_two_sum(double a, double b) {   // |a| >= |b|
    const double s = a + b;
    return {s, b - (s - a)};
}

ROS3_HD inline dd two_prod(double a, double b) {
    const double p = a * b;
    return {p, fma(a, b, -p)};
}

ROS3_HD inline dd add(dd a, dd b) {
    dd s = two_sum(a.hi, b.hi);
    const dd t = two_sum(a.lo, b.lo);
    s = fast_two_sum(s.hi, s.lo + t.hi);
    return fast_two_sum(s.hi, s.lo + t.lo);
}

ROS3_HD inline dd mul(dd a, dd b) {
    const dd p = two_prod(a.hi, b.hi);
    return fast_two_sum(p.hi, p.lo + (a.hi * b.lo + a.lo * b.hi));
}

ROS3_HD inline dd mul_d(dd a, double b) {
    const dd p = two_prod(a.hi, b);
    return fast_two_sum(p.hi, p.lo + a.lo * b);
}

// ln 2 = kLn2Hi + kLn2Lo to 2^-107
constexpr double kLn2Hi = 0x1.62e42fefa39efp-1;
constexpr double kLn2Lo = 0x1.abc9e3b39803fp-56;

// 1/(2k+1), k = 1..9, as hi + lo (each to 2^-106)
constexpr double kInvOddHi[9] = {0x1.5555555555555p-2, 0x1.999999999999ap-3, 0x1.2492492492492p-3,
                                 0x1.c71c71c71c71cp-4, 0x1.745d1745d1746p-4, 0x1.3b13b13b13b14p-4,
                                 0x1.1111111111111p-4, 0x1.e1e1e1e1e1e1ep-5, 0x1.af286bca1af28p-5};
constexpr double kInvOddLo[9] = {0x1.5555555555555p-56, -0x1.999999999999ap-57, 0x1.2492492492492p-57,
                                 0x1.c71c71c71c71cp-58, -0x1.745d1745d1746p-59, -0x1.3b13b13b13b14p-58,
                                 0x1.1111111111111p-60, 0x1.e1e1e1e1e1e1ep-59, 0x1.af286bca1af28p-59};

// 1/6 and 1/24 as hi + lo
constexpr double kSixthHi = 0x1.5555555555555p-3, kSixthLo = 0x1.5555555555555p-57;
constexpr double kTwentyFourthHi = 0x1.5555555555555p-5, kTwentyFourthLo = 0x1.5555555555555p-59;

// log x, x positive and normal, to about 2^-102 relative
ROS3_HD inline dd log_dd(double x) {
    int e;
    double m = frexp(x, &e);                 // x = m 2^e, m in [1/2, 1)
    if (m < 0.70710678118654752440) {
        m *= 2.0;
        e -= 1;
    }
    const double num = m - 1.0;              // exact (Sterbenz)
    const dd den = two_sum(m, 1.0);
    const double sh = num / den.hi;
    const double rem = fma(-sh, den.hi, num) - sh * den.lo;
    const dd s = fast_two_sum(sh, rem / den.hi);
    const dd t = mul(s, s);                  // s^2 <= 0.0295
    // 1/21 + t/23 + ... + t^10/41 in double: its terms weigh t^10 < 2^-50
    double q = 1.0 / 41.0;
    for (int k = 19; k >= 10; --k) q = q * t.hi + 1.0 / (double)(2 * k + 1);
    dd acc = {q, 0.0};
    for (int k = 8; k >= 0; --k) acc = add(dd{kInvOddHi[k], kInvOddLo[k]}, mul(t, acc));
    acc = add(dd{1.0, 0.0}, mul(t, acc));    // atanh(s)/s
    dd lm = mul(s, acc);
    lm.hi *= 2.0;
    lm.lo *= 2.0;
    const double fe = (double)e;
    dd le = two_prod(fe, kLn2Hi);
    le = fast_two_sum(le.hi, le.lo + fe * kLn2Lo);
    return add(le, lm);
}

// exp of a double-double argument with |w| < 700, as a double-double scaled
// by 2^k (result = (hi + lo) 2^k)
ROS3_HD inline dd exp_dd(dd w, int *k_out) {
    const double k = rint(w.hi * 1.44269504088896340736);
    const dd kh = two_prod(k, kLn2Hi);
    dd r = add(w, dd{-kh.hi, -kh.lo});
    r = add(r, dd{-k * kLn2Lo, 0.0});        // |r| <= 0.35
    r.hi *= 0x1p-8;
    r.lo *= 0x1p-8;                          // |r| <= 0.0014
    // expm1(r) = r (1 + r/2 + r^2/6 + r^3/24 + r^4/120 + ... + r^8/9!)
    double p = 1.0 / 362880.0;
    p = p * r.hi + 1.0 / 40320.0;
    p = p * r.hi + 1.0 / 5040.0;
    p = p * r.hi + 1.0 / 720.0;
    p = p * r.hi + 1.0 / 120.0;
    dd acc = {p, 0.0};
    acc = add(dd{kTwentyFourthHi, kTwentyFourthLo}, mul(r, acc));
    acc = add(dd{kSixthHi, kSixthLo}, mul(r, acc));
    acc = add(dd{0.5, 0.0}, mul(r, acc));
    acc = add(dd{1.0, 0.0}, mul(r, acc));
    dd em1 = mul(r, acc);
    for (int i = 0; i < 8; ++i) {            // (1+e)^2 - 1 = 2e + e^2
        const dd e2 = mul(em1, em1);
        em1 = add(dd{2.0 * em1.hi, 2.0 * em1.lo}, e2);
    }
    *k_out = (int)k;
    return add(dd{1.0, 0.0}, em1);
}

// x^y rounded to nearest
ROS3_HD inline double pow_cr(double x, double y) {
    if (!(x > 0.0) || !(x < INFINITY) || !(fabs(y) < INFINITY) || x < 2.2250738585072014e-308) return pow(x, y);
    if (x == 1.0 || y == 0.0) return 1.0;
    const dd l = log_dd(x);
    const dd w = mul_d(l, y);
    if (!(fabs(w.hi) < 700.0)) return pow(x, y);
    int k;
    const dd e = exp_dd(w, &k);
    return ldexp(e.hi, k);                   // e.hi = round(hi + lo); 2^k keeps it exact
}

}  // namespace crpow
}  // namespace ros3
