// The SLIC / FORCE building blocks, generic over the physics P (Euler, MHD,
// HPR) and the sweep direction, on register arrays.  EXACT mode follows the
// reference's operation order (see the citations) with every product and sum
// rounded separately; FAST mode is the same scheme with the primitive/conserved
// round trips removed and multiply-adds fused (fp contract(fast) in its
// branches only).
#pragma once

#include "vg_device.h"
#include "vg_quot.h"

namespace vg {

enum { MODE_EXACT = 0, MODE_FAST = 1 };

// Physics whose EXACT conversions can hand the shared divisor of rho from a
// primitive conversion to the flux of the same density (cons_to_prim_d, flux_d).
template <class P, class = void>
struct HasSharedDivisor {
    static constexpr bool value = false;
};
template <class P>
struct HasSharedDivisor<P, decltype((void)&P::template flux_d<0>)> {
    static constexpr bool value = true;
};

// Correctly rounded division: the compiler's IEEE division, or the division
// sequence without its scaling steps where the caller has shown the operands
// in range (vg_quot.h).
struct IeeeDiv {
    __device__ __forceinline__ double operator()(double a, double b) const { return a / b; }
};
struct InRangeDiv {
    __device__ __forceinline__ double operator()(double a, double b) const { return div_in_range(a, b); }
};

// SlopeLimiters::computeRCoefficient and computeSlopeLimiter
// (Solvers/slopelimiters.cpp:7-74): 0 superbee, 1 van Leer, anything else minbee.
template <class Div = IeeeDiv>
__device__ __forceinline__ double r_coefficient(double r, double bias, Div div = Div())
{
    return div(2.0, (1.0 - bias) + ((1.0 + bias) * r));
}

template <class Div = IeeeDiv>
__device__ __forceinline__ double limiter_value(double r, double bias, int limiter, Div div = Div())
{
    if (r < 0.0) return 0.0;
    if (limiter == 0) {
        if (r < 0.5) return 2.0 * r;
        if (r < 1.0) return 1.0;
        return smin(smin(r, r_coefficient(r, bias, div)), 2.0);
    }
    if (limiter == 1) return smin(div(2.0 * r, 1.0 + r), r_coefficient(r, bias, div));
    if (r < 1.0) return r;
    return smin(1.0, r_coefficient(r, bias, div));
}

// The limiter's divisions are in the range of the short division sequence when
// every slope difference is at most 2^600 in magnitude (they are at least 1e-5
// after the clamp) and -1 <= bias <= 1: then 2^-617 <= |r| <= 2^617, and the
// denominators 1 + r and (1 - bias) + (1 + bias) r of the positive-r branches
// lie in [2^-617, 2^619] with numerators 2 and 2r.
__device__ __forceinline__ bool limiter_in_range(double num, double den)
{
    return smax(fabs(num), fabs(den)) <= 0x1p600;
}

// FAST-mode limiter: the same piecewise functions with Newton-refined reciprocals.
__device__ __forceinline__ double limiter_value_fast(double r, double bias, int limiter)
{
#pragma clang fp contract(fast)
    if (r < 0.0) return 0.0;
    if (limiter == 0) {
        if (r < 0.5) return 2.0 * r;
        if (r < 1.0) return 1.0;
        return smin(smin(r, 2.0 * rcp_nr((1.0 - bias) + (1.0 + bias) * r)), 2.0);
    }
    if (limiter == 1) {
        const double R = 2.0 * rcp_nr((1.0 - bias) + (1.0 + bias) * r);
        return smin(2.0 * r * rcp_nr(1.0 + r), R);
    }
    if (r < 1.0) return r;
    return smin(1.0, 2.0 * rcp_nr((1.0 - bias) + (1.0 + bias) * r));
}

// Conserved state plus its physical flux in the sweep direction.
template <int NF>
struct ConsFlux {
    double U[NF];
    double F[NF];
};

template <class P, int DIR, int MODE>
struct Scheme {
    static constexpr int NF = P::NF;
    using Params = typename P::Params;

    __device__ static __forceinline__ void cons(const Params &m, const double (&w)[NF], double (&U)[NF])
    {
        if (MODE == MODE_EXACT) P::prim_to_cons(m, w, U);
        else P::prim_to_cons_fast(m, w, U);
    }

    // Flux of a conserved vector.
    __device__ static __forceinline__ void flux_of_cons(const Params &m, const double (&U)[NF], double (&F)[NF])
    {
        if (MODE == MODE_EXACT) {
            P::template flux<DIR>(m, U, F);
        } else {
            double w[NF];
            P::cons_to_prim_fast(m, U, w);
            P::template flux_prim_fast<DIR>(m, w, U, F);
        }
    }

    // Conserved vector and flux of a cell given as primitives.  EXACT: the
    // reference's state-overload flux, computeXFluxVector(computeConserved...)
    // (Euler/eulerstatevector.cpp:122-125).
    __device__ static __forceinline__ void cons_flux_of_prim(const Params &m, const double (&w)[NF], ConsFlux<NF> &cf)
    {
        cons(m, w, cf.U);
        if (MODE == MODE_EXACT) P::template flux<DIR>(m, cf.U, cf.F);
        else P::template flux_prim_fast<DIR>(m, w, cf.U, cf.F);
    }

    // MUSCL-Hancock half-step evolution of the middle cell's two boundary
    // extrapolations: SlopeLimiters::computeSlopeVector (slopelimiters.cpp:217-246)
    // then Solvers::evolveStateByHalfX/YTimeStep (Solvers/solvers.cpp:454-466,
    // 502-514) with computeEvolutionVector (:601-609) and evolveStateByHalfTimeStep
    // (:550-565).  Produces, for the SLIC face fluxes, the conserved vector and
    // flux of both evolved states: side 0 (left) and side 1 (right).
    // half_c = 0.5 * (ts / dx).
    __device__ static __forceinline__ void evolve(const Params &m, const double (&UL)[NF], const double (&UM)[NF],
                                                  const double (&UR)[NF], double bias, int limiter, double half_c,
                                                  ConsFlux<NF> &left, ConsFlux<NF> &right)
    {
        double xl[NF], xr[NF];
        const double cl = 0.5 * (1.0 + bias), cr = 0.5 * (1.0 - bias);
        if constexpr (MODE == MODE_EXACT) {
            double num[NF], den[NF], d[NF];
            bool in_range = (bias >= -1.0) && (bias <= 1.0);
#pragma unroll
            for (int k = 0; k < NF; k++) {
                const double dl = UM[k] - UL[k];
                const double dr = UR[k] - UM[k];
                d[k] = (cl * dl) + (cr * dr);
                // |difference| < 1e-5 is replaced by +1e-5 (sign dropped) before the ratio.
                num[k] = (fabs(dl) < 1e-5) ? 1e-5 : dl;
                den[k] = (fabs(dr) < 1e-5) ? 1e-5 : dr;
                in_range = in_range && limiter_in_range(num[k], den[k]);
            }
            if (in_range) {
                // Two cases need no division: slope differences of opposite
                // signs (r < 0, the limiters' zero piece) and equal ones (r = 1
                // exactly, e.g. both clamped to 1e-5 in a flat region), whose
                // limiter value is the same for every cell.
                const double f1 = limiter_value(1.0, bias, limiter, InRangeDiv());
#pragma unroll
                for (int k = 0; k < NF; k++) {
                    double lim;
                    if (num[k] * den[k] < 0.0) lim = 0.0;
                    else if (num[k] == den[k]) lim = f1;
                    else lim = limiter_value(div_in_range(num[k], den[k]), bias, limiter, InRangeDiv());
                    d[k] *= lim;
                }
            } else {
#pragma unroll
                for (int k = 0; k < NF; k++) d[k] *= limiter_value(num[k] / den[k], bias, limiter);
            }
#pragma unroll
            for (int k = 0; k < NF; k++) {
                xl[k] = UM[k] - (0.5 * d[k]);
                xr[k] = UM[k] + (0.5 * d[k]);
            }
            double Fl[NF], Fr[NF];
            flux_of_cons(m, xl, Fl);
            flux_of_cons(m, xr, Fr);
#pragma unroll
            for (int k = 0; k < NF; k++) {
                const double evo = half_c * (Fl[k] - Fr[k]);
                xl[k] = xl[k] + evo;
                xr[k] = xr[k] + evo;
            }
        } else {
#pragma clang fp contract(fast)
#pragma unroll
            for (int k = 0; k < NF; k++) {
                const double dl = UM[k] - UL[k];
                const double dr = UR[k] - UM[k];
                double d = (cl * dl) + (cr * dr);
                const double num = (fabs(dl) < 1e-5) ? 1e-5 : dl;
                const double den = (fabs(dr) < 1e-5) ? 1e-5 : dr;
                d *= limiter_value_fast(num * rcp_nr(den), bias, limiter);
                xl[k] = UM[k] - (0.5 * d);
                xr[k] = UM[k] + (0.5 * d);
            }
            double Fl[NF], Fr[NF];
            flux_of_cons(m, xl, Fl);
            flux_of_cons(m, xr, Fr);
#pragma unroll
            for (int k = 0; k < NF; k++) {
                const double evo = half_c * (Fl[k] - Fr[k]);
                xl[k] = xl[k] + evo;
                xr[k] = xr[k] + evo;
            }
        }
        if (MODE == MODE_EXACT) {
            // setConservedVariableVector, then the FORCE flux converts the state
            // back: computeConservedVariableVector + state-overload flux.  The
            // round trip returns rho unchanged, so the flux divides by the
            // density whose reciprocal the conversion already holds.
            double w[NF];
            if constexpr (HasSharedDivisor<P>::value) {
                const SharedDivisor dl = P::cons_to_prim_d(m, xl, w);
                cons(m, w, left.U);
                P::template flux_d<DIR>(m, left.U, dl, left.F);
                const SharedDivisor dr = P::cons_to_prim_d(m, xr, w);
                cons(m, w, right.U);
                P::template flux_d<DIR>(m, right.U, dr, right.F);
            } else {
                P::cons_to_prim(m, xl, w);
                cons_flux_of_prim(m, w, left);
                P::cons_to_prim(m, xr, w);
                cons_flux_of_prim(m, w, right);
            }
        } else {
            double w[NF];
#pragma unroll
            for (int k = 0; k < NF; k++) { left.U[k] = xl[k]; right.U[k] = xr[k]; }
            P::cons_to_prim_fast(m, xl, w);
            P::template flux_prim_fast<DIR>(m, w, left.U, left.F);
            P::cons_to_prim_fast(m, xr, w);
            P::template flux_prim_fast<DIR>(m, w, right.U, right.F);
        }
    }

    // FORCE flux between two states: FirstOrderSolver::computeLaxFriedrichsFlux
    // (firstordersolver.cpp:79-85), computeRichtmyerFlux (:165-171) followed by the
    // flux of the intermediate state (:96-97), and computeFORCEFlux (:227-230).
    // lf_c = dx / ts, ri_c = ts / dx.
    __device__ static __forceinline__ void force_face(const Params &m, const ConsFlux<NF> &a, const ConsFlux<NF> &b,
                                                      double lf_c, double ri_c, double (&F)[NF])
    {
        double us[NF], lf[NF], ri[NF];
        if constexpr (MODE == MODE_EXACT) {
#pragma unroll
            for (int k = 0; k < NF; k++) {
                lf[k] = 0.5 * ((a.F[k] + b.F[k]) + (lf_c * (a.U[k] - b.U[k])));
                us[k] = 0.5 * ((a.U[k] + b.U[k]) + (ri_c * (a.F[k] - b.F[k])));
            }
            flux_of_cons(m, us, ri);
#pragma unroll
            for (int k = 0; k < NF; k++) F[k] = 0.5 * (lf[k] + ri[k]);
        } else {
#pragma clang fp contract(fast)
#pragma unroll
            for (int k = 0; k < NF; k++) {
                lf[k] = 0.5 * ((a.F[k] + b.F[k]) + (lf_c * (a.U[k] - b.U[k])));
                us[k] = 0.5 * ((a.U[k] + b.U[k]) + (ri_c * (a.F[k] - b.F[k])));
            }
            flux_of_cons(m, us, ri);
#pragma unroll
            for (int k = 0; k < NF; k++) F[k] = 0.5 * (lf[k] + ri[k]);
        }
    }

    // FirstOrderSolver::computeFORCEUpdate (firstordersolver.cpp:428-433) followed
    // by setConservedVariableVector.  up_c = ts / dx.
    __device__ static __forceinline__ void update(const Params &m, const double (&U)[NF], const double (&Fm)[NF],
                                                  const double (&Fp)[NF], double up_c, double (&w)[NF])
    {
        double N[NF];
        if constexpr (MODE == MODE_EXACT) {
#pragma unroll
            for (int k = 0; k < NF; k++) N[k] = U[k] + (up_c * (Fm[k] - Fp[k]));
            P::cons_to_prim(m, N, w);
        } else {
#pragma clang fp contract(fast)
#pragma unroll
            for (int k = 0; k < NF; k++) N[k] = U[k] + (up_c * (Fm[k] - Fp[k]));
            P::cons_to_prim_fast(m, N, w);
        }
    }
};

}  // namespace vg
