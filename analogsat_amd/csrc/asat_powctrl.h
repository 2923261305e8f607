// Correctly rounded pow for the RK45 step controller's two exponents.
//
// The reference controller calls std::pow(errmax, PGROW = -0.2) after an
// accepted step and std::pow(errmax, PSHRNK = -0.25) after a rejected one
// (AnalogSAT/cpu_sat/CpuRungeKutta.cpp:236-252).  Here the result is x^y
// rounded to nearest, y = -1/q for q = 4 (exact) and q = 5 (-0.2 in double is
// -(1/5 + d) with d = 1.1102230246251565e-17).
//
// The same source builds for the device (hipcc, ASAT_HD = __host__ __device__)
// and for the host (any C++17 compiler, ASAT_HD empty), with the same
// operations in the same order, so a host build computes what the device
// computes and can be checked against a multiprecision pow.
//
//   1. seed r0 = x^(-1/q) from the bit pattern of x (the exponent field plus
//      the mantissa is log2 x to within 0.09 between powers of two): within
//      8 % for every normal x.
//   2. five Newton steps for x r^q = 1 in double, r += (r / q) (1 - x r^q):
//      8e-2 -> 2e-2 -> 8e-4 -> 2e-6 -> 1e-11 -> rounding level.
//   3. one more Newton step with x r^q formed in double-double (FMA two
//      products), so 1 - x r^q is exact to ~1e-32, and for q = 5 the factor
//      x^-d = 1 - d ln x (ln x to 1e-9 from the exponent and an atanh series);
//      r + corr is x^y to ~1e-31 relative, rounded once.
// Outside (1e-300, 1e300), and for NaN, the library pow is used.
#pragma once

#include <cmath>

#ifndef ASAT_HD
#define ASAT_HD
#endif

namespace asat {

ASAT_HD inline double pc_from_bits(long long b) { return __builtin_bit_cast(double, b); }
ASAT_HD inline long long pc_to_bits(double x) { return __builtin_bit_cast(long long, x); }

// (hi, lo) *= b in double-double
ASAT_HD inline void dd_mul_d(double& hi, double& lo, double b)
{
    double p = hi * b;
    double e = fma(hi, b, -p);
    e = fma(lo, b, e);
    double s = p + e;
    lo = e - (s - p);
    hi = s;
}

// ln x for normal x to ~1e-9 absolute: x = 2^e m with m in [sqrt(1/2), sqrt(2)),
// ln m = 2 atanh(t), t = (m - 1) / (m + 1), |t| < 0.172, series to t^9
ASAT_HD inline double pc_ln(double x)
{
    const long long b = pc_to_bits(x);
    long long e = ((b >> 52) & 0x7FF) - 1023;
    double m = pc_from_bits((b & 0x000FFFFFFFFFFFFFLL) | 0x3FF0000000000000LL);
    if (m > 1.4142135623730951) {
        m *= 0.5;
        e += 1;
    }
    const double t = (m - 1.0) / (m + 1.0), t2 = t * t;
    const double s = 1.0 + t2 * (1.0 / 3.0 + t2 * (0.2 + t2 * (1.0 / 7.0 + t2 * (1.0 / 9.0))));
    return (double)e * 0.69314718055994531 + 2.0 * t * s;
}

// x^-0.25 (q = 4) or x^-0.2 (q = 5), rounded to nearest
ASAT_HD inline double pow_ctrl(double x, int q)
{
    const double y = q == 4 ? -0.25 : -0.2;
    if (!(x > 1.0e-300) || !(x < 1.0e300)) return pow(x, y);
    const long long one = 0x3FF0000000000000LL;
    double r = pc_from_bits(one - (pc_to_bits(x) - one) / q);
    const double invq = q == 4 ? 0.25 : 0.2;
    for (int i = 0; i < 5; i++) {
        // x r^q as (x r^(q-4)) (r^2)^2 and r / q beside it: six dependent
        // operations per step
        const double r2 = r * r, xr = q == 5 ? x * r : x, rq = r * invq;
        const double t = xr * (r2 * r2);
        r = r + rq * (1.0 - t);
    }
    double hi = r, lo = 0.0;
    for (int i = 1; i < q; i++) dd_mul_d(hi, lo, r);
    dd_mul_d(hi, lo, x);
    double e = (1.0 - hi) - lo;              // 1 - x r^q, exact leading part (Sterbenz)
    double corr = r * e * invq;              // (1/5 rounded: 1e-16 of a 1e-16 correction)
    if (q == 5) corr -= r * (1.1102230246251565404e-17 * pc_ln(x));
    return r + corr;
}

}  // namespace asat
