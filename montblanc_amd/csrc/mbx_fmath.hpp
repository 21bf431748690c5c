// Float64 shape-factor arithmetic for the matrix-core passes: exp(-x) for
// x >= 0 and d^(-3/2) for d >= 1, in fewer vector instructions than the
// general library routines (which handle every range and special value).
#pragma once

#include <hip/hip_runtime.h>

namespace mbx {

// Entries of the exp_neg table: 2^(j / 256), j = 0..255.
constexpr int kExpTab = 256;

// Fill the table (LDS) with threads [0, nth) of a workgroup; the caller
// synchronises before use.
__device__ __forceinline__ void exp_table_fill(double *tab, int tid, int nth) {
    for (int j = tid; j < kExpTab; j += nth) tab[j] = exp2(double(j) / kExpTab);
}

// exp(-x), x >= 0, by Tang's table-driven method (P. T. P. Tang, "Table-
// driven implementation of the exponential function in IEEE floating-point
// arithmetic", ACM TOMS 15(2), 1989): -x = (256 m + j) ln2 / 256 + r with
// |r| <= ln2 / 512, the reduction in two steps (Cody and Waite) with ln2
// split into a 32-bit head, whose products with k are exact, and a tail;
// exp(r) by its Taylor polynomial to degree 4 (truncation below 4e-17 at
// |r| = ln2 / 512); 2^(j / 256) from the table; 2^m by ldexp. Arguments
// above 745.5 give 0 (as exp does, past the smallest subnormal).
__device__ __forceinline__ double exp_neg(double x, const double *tab) {
    constexpr double kInvStep = 369.32993046757462751;      // 256 / ln 2
    constexpr double kStepHi = 0x1.62e42fee00000p-1 / 256; // ln 2 / 256, head
    constexpr double kStepLo = 0x1.a39ef35793c76p-33 / 256; // ln 2 / 256, tail
    // x = min(x, 745.5) on the high word (signed: monotonic for x >= +0;
    // +inf and NaN clamp too); the low word is kept, so a clamped x lies in
    // [745.5, 745.5005), where the result is 0 as well
    const long long xb = __builtin_bit_cast(long long, x);
    const int xhi = min(int(xb >> 32), 0x40874C00);   // high word of 745.5
    x = __builtin_bit_cast(double, (static_cast<long long>(xhi) << 32) | static_cast<unsigned int>(xb));
    // k = round(-x 256 / ln 2) by the 1.5 2^52 shift: the sum's low word
    // holds k (|k| < 2^19) as a two's complement integer
    constexpr double kShift = 0x1.8p52;
    const double kd = fma(x, -kInvStep, kShift);
    const int ki = int(static_cast<unsigned int>(__builtin_bit_cast(unsigned long long, kd)));
    const double k = kd - kShift;
    double r = fma(-k, kStepHi, -x);
    r = fma(-k, kStepLo, r);
    double p = fma(r, 1.0 / 24, 1.0 / 6);
    p = fma(p, r, 0.5);
    p = fma(p, r, 1.0);
    p = fma(p, r, 1.0);
    return ldexp(p * tab[ki & (kExpTab - 1)], ki >> 8);   // ki = 256 m + j, m = floor(ki / 256)
}

// 1 / sqrt(d) for d >= 1: the hardware reciprocal square root refined by
// three Newton steps y <- y (3 - d y^2) / 2, each of which squares the
// relative error. d is capped at 1e300 so that d y^2 stays finite (the
// sersic factors built on it underflow to 0 there, as the exact ones do).
__device__ __forceinline__ double rsq_refined(double d) {
    d = fmin(d, 1e300);
    double y = __builtin_amdgcn_rsq(d);
#pragma unroll
    for (int i = 0; i < 3; ++i) y *= fma(-0.5 * d * y, y, 1.5);
    return y;
}

// d^(-3/2) for d >= 1 (the sersic factor).
__device__ __forceinline__ double rpow32(double d) {
    const double y = rsq_refined(d);
    return y * y * y;
}

}  // namespace mbx
