// Correctly rounded quotients a / b that share one reciprocal of b.
//
// Several quotients of a state divide by the same density (u = rho u / rho,
// v = rho v / rho, e_t = E / rho, ...).  A correctly rounded IEEE division costs
// a scale, a reciprocal estimate, Newton steps and a fix-up on the AMDGPU
// (about 11 dependent instructions, two of them tied to VCC); here the
// reciprocal y = RN(1 / b) is taken once with that division, and each quotient
// is then refined from q0 = RN(a y) by two residual corrections with fused
// multiply-adds:
//
//     r = a - b q          (exact: the FMA's product is not rounded, and the
//                           residual of a quotient within 1.5 ulp fits a double)
//     q <- RN(q + r y)
//
// q0 is within 1.5 ulp of a / b (y is within half an ulp of 1 / b).  After the
// first correction q is within 1 ulp: its exact value is a/b + (a/b - q0)(b y - 1)
// and |b y - 1| <= 2^-53.  The second correction then returns RN(a / b): this is
// Markstein's theorem (P. W. Markstein, "Computation of elementary functions on
// the IBM RISC System/6000 processor", IBM J. Res. Dev. 34(1), 1990; also J.-M.
// Muller et al., Handbook of Floating-Point Arithmetic, the section on
// Newton-Raphson-based division): if y is within half an ulp of 1/b and q within
// one ulp of a/b, RN(q + (a - bq) y) = RN(a/b).
//
// The residual is formed as b q - a and its product with y subtracted (the
// negation on the operand y, never on a rounded result), so a zero numerator
// keeps its sign (-0 / b = -0), which the RN(q + r y) form would lose.  The theorem
// needs every intermediate normal: the shared reciprocal is used only for
// 2^-100 <= b <= 2^100 (densities; NaN, negative and other divisors take the
// IEEE division), and the result equals IEEE a / b for every numerator a that
// is zero or of magnitude in [2^-800, 2^800] -- the momenta and energies of any
// state the solvers can reach.  A numerator beyond 2^800 may give NaN where
// IEEE gives an infinity (a state already lost), and one below 2^-800 (under
// 1e-240, not a physical momentum) may differ in the last bit.
// tests/test_quotient.py compares this code with IEEE division over random,
// near-midpoint and edge operands.
#pragma once

#ifdef __HIPCC__
#define VG_QUOT_FN __host__ __device__ __forceinline__
#else
#include <cmath>
#define VG_QUOT_FN inline
#endif

namespace vg {

struct SharedDivisor {
    double b;    // the divisor
    double y;    // RN(1 / b)
    bool fast;   // 2^-100 <= b <= 2^100: the FMA refinement applies
};

VG_QUOT_FN SharedDivisor shared_divisor(double b)
{
    SharedDivisor d;
    d.b = b;
    d.y = 1.0 / b;
    d.fast = (b >= 0x1p-100) && (b <= 0x1p+100);
    return d;
}

// RN(a / d.b) by the two corrections; valid when d.fast.
VG_QUOT_FN double quot_refined(double a, const SharedDivisor &d)
{
#ifdef __HIPCC__
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
