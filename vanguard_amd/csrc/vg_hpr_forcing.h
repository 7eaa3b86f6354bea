// HPR stiff-relaxation source step: HPRForcingSolver::computeRungeKuttaTimeStep[2D]
// (Solvers/hprforcingsolver.cpp:341-356, 394-417), one thread per cell.
//
// Each cell advances its conserved vector by classical RK4 on the source
// S(U) (hprstatevector.cpp:312-366).  Every stage first predicts the cell's
// state with fractional MUSCL steps from its FIXED neighbours
// (evolveConservedVariableVector[2D], hprforcingsolver.cpp:7-44, 134-192:
// X(1/2) in 1D; X(1/4) Y(1/2) X(1/4) in 2D for stages 1-2 and X(1/2) Y(1) X(1/2)
// for stage 3), so a launch is a read-only 3- or 5-point stencil over the
// input planes and a write of the output planes.
//
// Only the A slots (and the J slots with heat conduction) of S are non-zero;
// the other slots' RK4 increments are exactly +0, so they are not stored.
#pragma once

#include "vg_phys_hpr.h"
#include "vg_scheme.h"
#include "vg_sweep.h"

namespace vg {

struct ForcingArgs {
    const double *in;
    double *out;
    long long plane;
    int ny, nx, dim, bc, limiter, subcycles;
    double dx, bias;
    int fuse_max;
    HPRPhys::Params m;
    Clock *clk;
    unsigned long long *partials;
};

constexpr int kForceBS = 128;

// evolveStateByFractionalX/YTimeStep (hprsolvers.cpp:606-620, 673-687) on the
// conserved vectors of the three cells: slope, half-slope extrapolations, flux
// difference times frac * (ts / dx) added to the middle state, converted back to
// primitives (setConservedVariableVector).
template <int DIR>
__device__ __forceinline__ void hpr_fractional(const HPRPhys::Params &m, const double (&UL)[17], const double (&UM)[17],
                                               const double (&UR)[17], double bias, int limiter, double c,
                                               double (&w)[17])
{
    double xl[17], xr[17];
    const double cl = 0.5 * (1.0 + bias), cr = 0.5 * (1.0 - bias);
#pragma unroll
    for (int k = 0; k < 17; k++) {
        const double dl = UM[k] - UL[k];
        const double dr = UR[k] - UM[k];
        double d = (cl * dl) + (cr * dr);
        const double num = (fabs(dl) < 1e-5) ? 1e-5 : dl;
        const double den = (fabs(dr) < 1e-5) ? 1e-5 : dr;
        d *= limiter_value(num / den, bias, limiter);
        xl[k] = UM[k] - (0.5 * d);
        xr[k] = UM[k] + (0.5 * d);
    }
    double F[17];
    HPRPhys::flux<DIR>(m, xl, F);
#pragma unroll
    for (int k = 0; k < 17; k++) xl[k] = F[k];   // xl now holds F(left extrapolation)
    HPRPhys::flux<DIR>(m, xr, F);
    double N[17];
#pragma unroll
    for (int k = 0; k < 17; k++) N[k] = UM[k] + (c * (xl[k] - F[k]));
    HPRPhys::cons_to_prim(m, N, w);
}

// Slots of S that can be non-zero: A (2..10) and J (14..16).
__device__ __forceinline__ constexpr int src_slot(int q) { return q < 9 ? 2 + q : 14 + (q - 9); }

template <int DIM>
__global__ __launch_bounds__(kForceBS) void hpr_forcing_kernel(ForcingArgs a)
{
    using P = HPRPhys;
    if (a.clk->done) return;
    const P::Params &m = a.m;
    const double ts = 0.5 * (a.clk->dt / a.subcycles);
    const long long cells = (long long)a.ny * a.nx;
    const long long c = (long long)blockIdx.x * kForceBS + threadIdx.x;
    unsigned long long key = 0;
    if (c < cells) {
        const int i = (int)(c / a.nx), j = (int)(c - (long long)i * a.nx);
        auto cons_rt = [&](int ii, int jj, double (&U)[17]) {
            // computeRungeKuttaTimeStep converts every cell to conserved form;
            // the neighbours re-enter the predictors as primitive states, so a
            // neighbour's conserved vector is p2c(c2p(p2c(w))).
            double w[17], t[17];
            load_prim<17>(a.in, a.plane, (long long)ii * a.nx + jj, w);
            P::prim_to_cons(m, w, t);
            P::cons_to_prim(m, t, w);
            P::prim_to_cons(m, w, U);
        };
        const int jl = ghost_index(j - 1, a.nx, a.bc), jr = ghost_index(j + 1, a.nx, a.bc);
        double UL[17], UR[17], UT[17], UB[17];
        cons_rt(i, jl, UL);
        cons_rt(i, jr, UR);
        if (DIM == 2) {
            cons_rt(ghost_index(i - 1, a.ny, a.bc), j, UT);
            cons_rt(ghost_index(i + 1, a.ny, a.bc), j, UB);
        }
        double w0[17], Um[17];
        load_prim<17>(a.in, a.plane, c, w0);
        P::prim_to_cons(m, w0, Um);

        const int nsrc = m.is_thermal ? 12 : 9;
        double K[3][12];
        double S[17];
        P::source(m, Um, S);
#pragma unroll
        for (int q = 0; q < 12; q++) K[0][q] = ts * S[src_slot(q)];
        double k4[12];
        const double cx = ts / a.dx;
#pragma unroll
        for (int st = 0; st < 3; st++) {
            double Us[17];
#pragma unroll
            for (int k = 0; k < 17; k++) Us[k] = Um[k] + 0.0;
#pragma unroll
            for (int q = 0; q < 12; q++) {
                if (q < nsrc) {
                    const double kp = K[st][q];
                    Us[src_slot(q)] = Um[src_slot(q)] + (st < 2 ? (0.5 * kp) : kp);
                }
            }
            // The stage state enters the predictor as a primitive state
            // (evolveConservedVariableVector takes states): c2p then p2c.
            double wm[17], UM[17], e[17];
            P::cons_to_prim(m, Us, wm);
            P::prim_to_cons(m, wm, UM);
            if (DIM == 1) {
                hpr_fractional<0>(m, UL, UM, UR, a.bias, a.limiter, (st < 2 ? 0.5 : 1.0) * cx, e);
            } else {
                const double fx = (st < 2) ? 0.25 : 0.5, fy = (st < 2) ? 0.5 : 1.0;
                hpr_fractional<0>(m, UL, UM, UR, a.bias, a.limiter, fx * cx, e);
                P::prim_to_cons(m, e, UM);
                hpr_fractional<1>(m, UT, UM, UB, a.bias, a.limiter, fy * cx, e);
                P::prim_to_cons(m, e, UM);
                hpr_fractional<0>(m, UL, UM, UR, a.bias, a.limiter, fx * cx, e);
            }
            double Ue[17];
            P::prim_to_cons(m, e, Ue);
            P::source(m, Ue, S);
            if (st < 2) {
#pragma unroll
                for (int q = 0; q < 12; q++) K[st + 1][q] = ts * S[src_slot(q)];
            } else {
#pragma unroll
                for (int q = 0; q < 12; q++) k4[q] = ts * S[src_slot(q)];
            }
        }
        double N[17];
#pragma unroll
        for (int k = 0; k < 17; k++) N[k] = Um[k] + 0.0;
#pragma unroll
        for (int q = 0; q < 12; q++) {
            if (q < nsrc) {
                const int k = src_slot(q);
                N[k] = Um[k] + ((1.0 / 6.0) * (K[0][q] + ((2.0 * K[1][q]) + ((2.0 * K[2][q]) + k4[q]))));
            }
        }
        double wn[17];
        P::cons_to_prim(m, N, wn);
        if (!m.is_thermal) {   // the cell's own J is left untouched
            wn[14] = w0[14];
            wn[15] = w0[15];
            wn[16] = w0[16];
        }
        store_prim<17>(a.out, a.plane, c, wn);
        if (a.fuse_max) key = max_key(P::wave_speed(m, wn, a.dim));
    }
    if (a.fuse_max) block_max_to<kForceBS>(a.partials, key);
}

}  // namespace vg
