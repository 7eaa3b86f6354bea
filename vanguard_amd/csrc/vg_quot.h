// Correctly rounded quotients a / b for the EXACT conversions, cheaper than the
// compiler's division where the operands are in range.
//
// The compiler's correctly rounded FP64 division on gfx950 is the sequence
//
//     s = div_scale(b), t = div_scale(a)       (scale out-of-range operands)
//     y = rcp(s); two Newton steps y += y (1 - s y)
//     q = t y;  r = t - s q                     (exact residual, one FMA)
//     div_fixup(div_fmas(r, y, q), b, a)        (q + r y rounded once, then the
//                                                special cases)
//
// For operands in range -- a zero or 2^-600 <= |a| <= 2^600, 2^-100 <= |b| <=
// 2^100, or more generally: b and 1/b normal, a / b normal, exponent(a) -
// exponent(b) < 768 and a not within 2^53 of the smallest normal -- the two
// scaling steps return their operands unchanged (and clear VCC), div_fmas is a
// plain FMA and div_fixup returns its first operand, so the division is
//
//     y = rcp(b); two Newton steps; q = a y; r = b q - a; q - r y
//
// and the reciprocal part depends on b only: several quotients by one density
// (u = rho u / rho, v, E / rho, ...) share it and cost one multiply and two
// FMAs each.  The residual is formed as b q - a and its product with y
// subtracted, so a zero numerator keeps its sign (-0 / b = -0) as div_fixup
// would give it.  tests/test_quotient_gpu.py compares these sequences on the
// GPU with the compiler's division over random operands across the range,
// quotients next to rounding midpoints and signed zeros.
//
// Outside that range the shared form falls back to the IEEE division (a / b)
// for divisors beyond 2^-100 <= b <= 2^100 (NaN, negative and other
// densities).  Numerators are not checked: the tests show the short form
// equal to the IEEE quotient for every numerator from 2^-1000 to 2^600 in
// magnitude (and zero); below 2^-1000 (subnormal territory) it may be one
// unit in the last place off, and above 2^600 it is not covered -- no
// momentum or energy of a state the solvers reach.
//
// The host build (used by tests/quotient_check.cpp) keeps y = RN(1 / b) and two
// residual corrections, q = RN(a y), q += (a - b q) y twice, which Markstein's
// theorem (P. W. Markstein, IBM J. Res. Dev. 34(1), 1990; J.-M. Muller et al.,
// Handbook of Floating-Point Arithmetic) makes equal to RN(a / b) in the same
// range.
#pragma once

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define VG_QUOT_FN __host__ __device__ __forceinline__
#else
#include <cmath>
#define VG_QUOT_FN inline
#endif

namespace vg {

struct SharedDivisor {
    double b;    // the divisor
    double y;    // device: the division sequence's reciprocal estimate; host: RN(1 / b)
    bool fast;   // 2^-100 <= b <= 2^100: the shared form applies
};

#ifdef __HIPCC__
// The reciprocal estimate of the division sequence: rcp and two Newton steps.
__device__ __forceinline__ double div_reciprocal(double b)
{
#if defined(__HIP_DEVICE_COMPILE__)
    double y = __builtin_amdgcn_rcp(b);
#else
    double y = 1.0 / b;   // host pass of the HIP build: declared, never called
#endif
    double e = __builtin_fma(-b, y, 1.0);
    y = __builtin_fma(y, e, y);
    e = __builtin_fma(-b, y, 1.0);
    return __builtin_fma(y, e, y);
}

// The rest of the division sequence: RN(a / b) from the estimate y of 1 / b.
__device__ __forceinline__ double div_finish(double a, double b, double y)
{
    const double q = a * y;
    const double r = __builtin_fma(b, q, -a);
    return __builtin_fma(-r, y, q);
}

// RN(a / b) for operands in range (see the header comment); the caller
// guarantees the range.
__device__ __forceinline__ double div_in_range(double a, double b)
{
    return div_finish(a, b, div_reciprocal(b));
}
#endif

VG_QUOT_FN SharedDivisor shared_divisor(double b)
{
    SharedDivisor d;
    d.b = b;
    d.fast = (b >= 0x1p-100) && (b <= 0x1p+100);
#if defined(__HIP_DEVICE_COMPILE__)
    d.y = div_reciprocal(b);   // unused when !fast
#else
    d.y = 1.0 / b;
#endif
    return d;
}

// RN(a / d.b); valid when d.fast and a is zero or 2^-600 <= |a| <= 2^600.
VG_QUOT_FN double quot_refined(double a, const SharedDivisor &d)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return div_finish(a, d.b, d.y);
#elif defined(__HIPCC__)
    double q = a * d.y;
    double r = __builtin_fma(d.b, q, -a);
    q = __builtin_fma(r, -d.y, q);
    r = __builtin_fma(d.b, q, -a);
    return __builtin_fma(r, -d.y, q);
#else
    double q = a * d.y;
    double r = std::fma(d.b, q, -a);
    q = std::fma(r, -d.y, q);
    r = std::fma(d.b, q, -a);
    return std::fma(r, -d.y, q);
#endif
}

// RN(a / d.b) for any divisor.
VG_QUOT_FN double quot(double a, const SharedDivisor &d)
{
    return d.fast ? quot_refined(a, d) : a / d.b;
}

// q[k] = RN(a[k] / d.b) for a group of numerators sharing the divisor: one
// branch for the group, taken the same way by every lane of physical states.
template <int N>
VG_QUOT_FN void quot_group(const double (&a)[N], const SharedDivisor &d, double (&q)[N])
{
    if (d.fast) {
#pragma unroll
        for (int k = 0; k < N; k++) q[k] = quot_refined(a[k], d);
    } else {
#pragma unroll
        for (int k = 0; k < N; k++) q[k] = a[k] / d.b;
    }
}

}  // namespace vg
