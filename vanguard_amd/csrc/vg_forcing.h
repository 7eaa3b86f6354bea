// Source sub-cycles of the second-order solvers, one thread per cell:
// HPRForcingSolver::computeRungeKuttaTimeStep[2D] (Solvers/hprforcingsolver.cpp:
// 341-356, 394-417) and MHDForcingSolver::computeRungeKuttaTimeStep[2D]
// (Solvers/mhdforcingsolver.cpp:341-356, 394-417).
//
// Each cell advances its conserved vector by classical RK4 on the source S(U)
// (HPR: stiff relaxation of A and J, hprstatevector.cpp:312-366; MHD: GLM
// damping of psi, mhdstatevector.cpp:222-234).  Every stage first predicts the
// cell's state with fractional MUSCL steps from its FIXED neighbours
// (evolveConservedVariableVector[2D], hprforcingsolver.cpp:7-44, 134-192,
// mhdforcingsolver.cpp:7-44, 134-194: X(1/2) in 1D; X(1/4) Y(1/2) X(1/4) in 2D
// for stages 1-2 and X(1/2) Y(1) X(1/2) for stage 3), so a launch is a
// read-only 3- or 5-point stencil over the input planes and a write of the
// output planes.
//
// Only a few slots of S can be non-zero (ForcingTraits<P>::slot); the other
// slots' RK4 increments are exactly +0, so they are not stored.
#pragma once

#include "vg_scheme.h"
#include "vg_sweep.h"

namespace vg {

// Per physics: kMaxSlots / slot(q) / count(m) (the non-zero slots of S),
// finish(m, w_old, w_new) (fix-ups of the new primitive state) and the timer name.
template <class P>
struct ForcingTraits;

template <class P>
struct ForcingArgs {
    const double *in;
    double *out;
    long long plane;
    int ny, nx, dim, bc, limiter, subcycles;
    double dx, bias;
    int fuse_max;
    typename P::Params m;
    Clock *clk;
    unsigned long long *partials;
};

constexpr int kForceBS = 128;

// Solvers::evolveStateByFractionalX/YTimeStep (hprsolvers.cpp:606-620, 673-687;
// mhdsolvers.cpp:612-626) on the conserved vectors of the three cells: slope,
// half-slope extrapolations, flux difference times frac * (ts / dx) added to the
// middle state, converted back to primitives (setConservedVariableVector).
template <class P, int DIR>
__device__ __forceinline__ void fractional_step(const typename P::Params &m, const double (&UL)[P::NF],
                                                const double (&UM)[P::NF], const double (&UR)[P::NF], double bias,
                                                int limiter, double c, double (&w)[P::NF])
{
    constexpr int NF = P::NF;
    double xl[NF], xr[NF];
    const double cl = 0.5 * (1.0 + bias), cr = 0.5 * (1.0 - bias);
#pragma unroll
    for (int k = 0; k < NF; k++) {
        const double dl = UM[k] - UL[k];
        const double dr = UR[k] - UM[k];
        double d = (cl * dl) + (cr * dr);
        const double num = (fabs(dl) < 1e-5) ? 1e-5 : dl;
        const double den = (fabs(dr) < 1e-5) ? 1e-5 : dr;
        d *= limiter_value(num / den, bias, limiter);
        xl[k] = UM[k] - (0.5 * d);
        xr[k] = UM[k] + (0.5 * d);
    }
    double F[NF];
    P::template flux<DIR>(m, xl, F);
#pragma unroll
    for (int k = 0; k < NF; k++) xl[k] = F[k];   // xl now holds F(left extrapolation)
    P::template flux<DIR>(m, xr, F);
    double N[NF];
#pragma unroll
    for (int k = 0; k < NF; k++) N[k] = UM[k] + (c * (xl[k] - F[k]));
    P::cons_to_prim(m, N, w);
}

template <class P, int DIM>
__global__ __launch_bounds__(kForceBS) void forcing_kernel(ForcingArgs<P> a)
{
    constexpr int NF = P::NF;
    using T = ForcingTraits<P>;
    constexpr int NS = T::kMaxSlots;
    if (a.clk->done) return;
    typename P::Params m = a.m;
    P::begin_step(m, a.clk->S);   // sub-cycles run in second-order steps only
    const double ts = 0.5 * (a.clk->dt / a.subcycles);
    const long long cells = (long long)a.ny * a.nx;
    const long long c = (long long)blockIdx.x * kForceBS + threadIdx.x;
    unsigned long long key = 0;
    if (c < cells) {
        const int i = (int)(c / a.nx), j = (int)(c - (long long)i * a.nx);
        auto cons_rt = [&](int ii, int jj, double (&U)[NF]) {
            // computeRungeKuttaTimeStep converts every cell to conserved form;
            // the neighbours re-enter the predictors as primitive states, so a
            // neighbour's conserved vector is p2c(c2p(p2c(w))).
            double w[NF], t[NF];
            load_prim<NF>(a.in, a.plane, (long long)ii * a.nx + jj, w);
            P::prim_to_cons(m, w, t);
            P::cons_to_prim(m, t, w);
            P::prim_to_cons(m, w, U);
        };
        const int jl = ghost_index(j - 1, a.nx, a.bc), jr = ghost_index(j + 1, a.nx, a.bc);
        double UL[NF], UR[NF], UT[NF], UB[NF];
        cons_rt(i, jl, UL);
        cons_rt(i, jr, UR);
        if (DIM == 2) {
            cons_rt(ghost_index(i - 1, a.ny, a.bc), j, UT);
            cons_rt(ghost_index(i + 1, a.ny, a.bc), j, UB);
        }
        double w0[NF], Um[NF];
        load_prim<NF>(a.in, a.plane, c, w0);
        P::prim_to_cons(m, w0, Um);

        const int nsrc = T::count(m);
        double K[3][NS];
        double S[NF];
        P::source(m, Um, S);
#pragma unroll
        for (int q = 0; q < NS; q++) K[0][q] = ts * S[T::slot(q)];
        double k4[NS];
        const double cx = ts / a.dx;
#pragma unroll
        for (int st = 0; st < 3; st++) {
            double Us[NF];
#pragma unroll
            for (int k = 0; k < NF; k++) Us[k] = Um[k] + 0.0;
#pragma unroll
            for (int q = 0; q < NS; q++) {
                if (q < nsrc) {
                    const double kp = K[st][q];
                    Us[T::slot(q)] = Um[T::slot(q)] + (st < 2 ? (0.5 * kp) : kp);
                }
            }
            // The stage state enters the predictor as a primitive state
            // (evolveConservedVariableVector takes states): c2p then p2c.
            double wm[NF], UM[NF], e[NF];
            P::cons_to_prim(m, Us, wm);
            P::prim_to_cons(m, wm, UM);
            if (DIM == 1) {
                fractional_step<P, 0>(m, UL, UM, UR, a.bias, a.limiter, (st < 2 ? 0.5 : 1.0) * cx, e);
            } else {
                const double fx = (st < 2) ? 0.25 : 0.5, fy = (st < 2) ? 0.5 : 1.0;
                fractional_step<P, 0>(m, UL, UM, UR, a.bias, a.limiter, fx * cx, e);
                P::prim_to_cons(m, e, UM);
                fractional_step<P, 1>(m, UT, UM, UB, a.bias, a.limiter, fy * cx, e);
                P::prim_to_cons(m, e, UM);
                fractional_step<P, 0>(m, UL, UM, UR, a.bias, a.limiter, fx * cx, e);
            }
            double Ue[NF];
            P::prim_to_cons(m, e, Ue);
            P::source(m, Ue, S);
            if (st < 2) {
#pragma unroll
                for (int q = 0; q < NS; q++) K[st + 1][q] = ts * S[T::slot(q)];
            } else {
#pragma unroll
                for (int q = 0; q < NS; q++) k4[q] = ts * S[T::slot(q)];
            }
        }
        // computeEvolvedConservedVariableVector (hprforcingsolver.cpp:332-339)
        double N[NF];
#pragma unroll
        for (int k = 0; k < NF; k++) N[k] = Um[k] + 0.0;
#pragma unroll
        for (int q = 0; q < NS; q++) {
            if (q < nsrc) {
                const int k = T::slot(q);
                N[k] = Um[k] + ((1.0 / 6.0) * (K[0][q] + ((2.0 * K[1][q]) + ((2.0 * K[2][q]) + k4[q]))));
            }
        }
        double wn[NF];
        P::cons_to_prim(m, N, wn);
        T::finish(m, w0, wn);
        store_prim<NF>(a.out, a.plane, c, wn);
        if (a.fuse_max) key = max_key(P::wave_speed(m, wn, a.dim));
    }
    if (a.fuse_max) block_max_to<kForceBS>(a.partials, key);
}

// The grid's source stage for a physics with ForcingTraits: one launch per
// sub-cycle (vg_grid.h calls it around the SLIC step of second-order runs).
template <class P>
struct ForcingStage {
    static constexpr bool kActive = true;
    template <class GS>
    static long long blocks(const GS &g)
    {
        return (g.cells + kForceBS - 1) / kForceBS;
    }
    template <class GS>
    static int launch(GS &g, const double *in, double *out, int fuse, int bc, double dx, double bias, int limiter,
                      const typename P::Params &m)
    {
        ForcingArgs<P> a;
        a.in = in;
        a.out = out;
        a.plane = g.plane;
        a.ny = g.ny;
        a.nx = g.nx;
        a.dim = g.dim;
        a.bc = bc;
        a.limiter = limiter;
        a.subcycles = g.subcycles;
        a.dx = dx;
        a.bias = bias;
        a.fuse_max = fuse;
        a.m = m;
        a.clk = g.clk;
        a.partials = g.partials;
        const long long nb = blocks(g);
        if (fuse) g.n_partials = (int)nb;
        return g.ctx->timed(ForcingTraits<P>::kTimer, [&] {
            if (g.dim == 1)
                forcing_kernel<P, 1><<<(unsigned)nb, kForceBS, 0, g.ctx->stream>>>(a);
            else
                forcing_kernel<P, 2><<<(unsigned)nb, kForceBS, 0, g.ctx->stream>>>(a);
        });
    }
};

}  // namespace vg
