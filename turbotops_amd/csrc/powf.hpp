// SPDX-License-Identifier: MIT
// powf.hpp — x^p for the Minkowski Lp kernels: the reference calls powf per
// term and for the root (include/turbotops/metric.hpp:204, :206).
//
// Method: table-driven logarithm and exponential after P. T. P. Tang
// ("Table-driven implementation of the exponential function in IEEE
// floating-point arithmetic", ACM TOMS 15(2), 1989; "... of the logarithm
// function", ACM TOMS 16(4), 1990), carried in double precision and rounded
// once to float:
//   log2(x) = e + log2(m),  x = 2^e m, m in [1, 2)
//           = e + T_j + log2(1 + r),   r = m * inv_c_j - 1, |r| <= 2^-8,
//     inv_c_j the double nearest 1 / (1 + (j + 1/2) / 128), j = the top 7
//     mantissa bits, T_j = -log2(inv_c_j); log2(1 + r) by its Taylor series
//     to r^6 (truncation below 2^-58);
//   2^y   = 2^k * 2^(i/64) * exp(g ln 2),  64 y = 64 k + i + 64 g, g in
//     [0, 1/64); exp by its Taylor series to u^6, u = g ln 2 <= 0.0109
//     (truncation below 2^-57).
// Error bound: every step is a double-precision operation on values whose
// rounding errors are at most 2^-53 relative; the largest is the absolute
// error of log2(x) (|log2 x| <= 150 for float x) times p, so the double
// result is within 2^-43 relative of x^p wherever x^p is at least the
// smallest normal float. After the final rounding the float is within
// 0.5 + 2^-19 ulp of x^p: the correctly rounded value except when x^p lies
// within 2^-43 relative of a midpoint between floats. glibc's powf, which
// the reference calls, is within 0.82 ulp of x^p (its own documented bound),
// so the two differ by at most one ulp; an exhaustive sweep over every float
// x in (0, 1] (terms) and [1, 16] (roots) checks that bound.
//
// The tables are computed at compile time (constexpr series in double) and
// live in the code object's read-only data on the device.
#pragma once

#include <cstdint>

#ifdef __HIPCC__
#define TTG_HD __host__ __device__
#else
#define TTG_HD
#endif

namespace ttg {
namespace powf_detail {

constexpr double kLn2 = 0x1.62e42fefa39efp-1;
constexpr double kInvLn2 = 0x1.71547652b82fep0;

// ln(v) for v in [1/2, 2]: 2 atanh((v - 1) / (v + 1)), |t| <= 1/3
constexpr double ln_series(double v) {
    const double t = (v - 1.0) / (v + 1.0);
    const double t2 = t * t;
    double term = t, sum = 0.0;
    for (int k = 0; k < 40; ++k) {
        sum += term / double(2 * k + 1);
        term *= t2;
    }
    return 2.0 * sum;
}

// e^u for |u| <= 1 by its Taylor series
constexpr double exp_series(double u) {
    double term = 1.0, sum = 1.0;
    for (int k = 1; k < 30; ++k) {
        term *= u / double(k);
        sum += term;
    }
    return sum;
}

struct Tables {
    double inv_c[128];   // ~ 1 / (1 + (j + 1/2) / 128)
    double t[128];       // -log2(inv_c[j]), for the stored inv_c[j]
    double e[64];        // 2^(i / 64)
};

constexpr Tables make_tables() {
    Tables tb{};
    for (int j = 0; j < 128; ++j) {
        tb.inv_c[j] = 1.0 / (1.0 + (double(j) + 0.5) / 128.0);
        tb.t[j] = -ln_series(tb.inv_c[j]) * kInvLn2;
    }
    for (int i = 0; i < 64; ++i) tb.e[i] = exp_series(double(i) * kLn2 / 64.0);
    return tb;
}

inline constexpr Tables kTables = make_tables();

}  // namespace powf_detail

// x^p as a float, for finite x >= 0 and p > 0 with x^p below 2^127 (the Lp
// kernels: terms x in [0, 1], p >= 1; roots s in [1, dim], 1/p in (0, 1]).
TTG_HD inline float pow_f32(float x, float p) {
    using namespace powf_detail;
    // x = 0 reads as 2^-1023: y far below -160, and the result 0
    const double xd = double(x);
    const uint64_t b = __builtin_bit_cast(uint64_t, xd);
    const int e = int(b >> 52) - 1023;   // float x is a normal double
    const unsigned j = unsigned(b >> 45) & 127u;
    const double m = __builtin_bit_cast(double, (b & 0x000fffffffffffffull) | 0x3ff0000000000000ull);
    const double r = __builtin_fma(m, kTables.inv_c[j], -1.0);
    double q = -1.0 / 6.0;
    q = __builtin_fma(q, r, 1.0 / 5.0);
    q = __builtin_fma(q, r, -1.0 / 4.0);
    q = __builtin_fma(q, r, 1.0 / 3.0);
    q = __builtin_fma(q, r, -1.0 / 2.0);
    q = __builtin_fma(q, r, 1.0);
    const double lg = (double(e) + kTables.t[j]) + (q * r) * kInvLn2;
    // below -160 every result rounds to 0; the clamp keeps 64 y in int range
    const double y = __builtin_fmax(double(p) * lg, -2000.0);
    const double ys = y * 64.0;
    const double nf = __builtin_floor(ys);
    const int ni = int(nf);
    const double u = ((ys - nf) * (1.0 / 64.0)) * kLn2;
    double pe = 1.0 / 720.0;
    pe = __builtin_fma(pe, u, 1.0 / 120.0);
    pe = __builtin_fma(pe, u, 1.0 / 24.0);
    pe = __builtin_fma(pe, u, 1.0 / 6.0);
    pe = __builtin_fma(pe, u, 1.0 / 2.0);
    pe = __builtin_fma(pe, u, 1.0);
    pe = __builtin_fma(pe, u, 1.0);
    return float(__builtin_ldexp(kTables.e[ni & 63] * pe, ni >> 6));
}

// x^(n + 1/2) as a float, for finite x >= 0: sqrt(x) x^n in double, rounded
// once. The square root is one Newton step from the reciprocal square root
// y0 = (1 + e0) / sqrt(x): s1 = s0 + (x - s0^2) y0 / 2 with s0 = x y0 is
// sqrt(x) (1 - 3/2 e0^2 - e0^3 / 2) up to the steps' own roundings (each at
// most 2^-53 relative). On the device y0 is the hardware's v_rsq_f64, whose
// error on every float x in (0, 1] is below 2^-24.1 (measured over all of
// them), so s1 is within 2^-47 relative of sqrt(x) and the product within
// 2^-46 of x^(n + 1/2): the float result is the correctly rounded one except
// within 2^-22 ulp of a midpoint, and within one ulp of glibc's powf. On the
// host (the test build of this source) y0 is 1 / sqrt(x).
TTG_HD inline float pow_half_f32(float x, unsigned n) {
    const double xd = double(x);
#if defined(__HIP_DEVICE_COMPILE__)
    const double y0 = __builtin_amdgcn_rsq(xd);
#else
    const double y0 = 1.0 / __builtin_sqrt(xd);
#endif
    const double s0 = xd * y0;
    const double r = __builtin_fma(-s0, s0, xd);
    double v = __builtin_fma(r, 0.5 * y0, s0);
    for (unsigned t = 0; t < n; ++t) v *= xd;
    return x == 0.0f ? 0.0f : float(v);
}

}  // namespace ttg
