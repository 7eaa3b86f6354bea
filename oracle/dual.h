// Forward-mode dual numbers and the branch-decision hash shared by the
// oracle's dual-number restatements (euler_dual.cpp, hpr_dual.cpp) -- TEST
// INFRASTRUCTURE ONLY (oracle).
//
// A Dual carries a value and one tangent; every operation propagates the
// tangent by the chain rule, so a run over Duals returns the exact
// Jacobian-vector product of the discrete map.  Branches read the value part.
#pragma once

#include <cmath>
#include <cstdint>

namespace orc_dual {

struct Dual {
    double v, d;
    Dual(double a = 0.0, double b = 0.0) : v(a), d(b) {}
};
inline Dual operator+(Dual a, Dual b) { return Dual(a.v + b.v, a.d + b.d); }
inline Dual operator-(Dual a, Dual b) { return Dual(a.v - b.v, a.d - b.d); }
inline Dual operator-(Dual a) { return Dual(-a.v, -a.d); }
inline Dual operator*(Dual a, Dual b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
inline Dual operator/(Dual a, Dual b) { return Dual(a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)); }
inline Dual operator+(Dual a, double b) { return Dual(a.v + b, a.d); }
inline Dual operator+(double a, Dual b) { return Dual(a + b.v, b.d); }
inline Dual operator-(Dual a, double b) { return Dual(a.v - b, a.d); }
inline Dual operator-(double a, Dual b) { return Dual(a - b.v, -b.d); }
inline Dual operator*(Dual a, double b) { return Dual(a.v * b, a.d * b); }
inline Dual operator*(double a, Dual b) { return Dual(a * b.v, a * b.d); }
inline Dual operator/(Dual a, double b) { return Dual(a.v / b, a.d / b); }
inline Dual operator/(double a, Dual b) { return Dual(a / b.v, -a * b.d / (b.v * b.v)); }
inline double val(double x) { return x; }
inline double val(Dual x) { return x.v; }
inline double fabs_v(double x) { return std::fabs(x); }
inline double fabs_v(Dual x) { return std::fabs(x.v); }
inline double sqrt_t(double x) { return std::sqrt(x); }
inline Dual sqrt_t(Dual x)
{
    const double r = std::sqrt(x.v);
    return Dual(r, x.d / (2.0 * r));
}
inline double pow_t(double x, double e) { return std::pow(x, e); }
inline Dual pow_t(Dual x, double e) { return Dual(std::pow(x.v, e), e * std::pow(x.v, e - 1.0) * x.d); }
// A value of the scalar type with a given tangent (dropped for double).
template <class T>
inline T make(double v, double d);
template <>
inline double make<double>(double v, double) { return v; }
template <>
inline Dual make<Dual>(double v, double d) { return Dual(v, d); }

// Hash of every branch decision a run takes (slope clamps, limiter cases, min
// selections, interface branches): two runs with equal hashes evaluated the
// same smooth pieces, so a finite difference between them crosses no kink.
inline uint64_t &branch_state()
{
    static uint64_t h = 0;
    return h;
}
inline void reset_branches() { branch_state() = 0xcbf29ce484222325ull; }
inline bool note(bool b)
{
    branch_state() = (branch_state() ^ (uint64_t)(b ? 0x9e37u : 0x51u)) * 0x100000001b3ull;
    return b;
}

}  // namespace orc_dual
