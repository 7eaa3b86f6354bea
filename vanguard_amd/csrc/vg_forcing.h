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
    int fixed_params = 0;   // keep m as given: no per-step parameters from the clock
};

constexpr int kForceBS = 128;
#ifndef VG_FORCING_WAVES
#define VG_FORCING_WAVES 2
#endif

// A neighbour's conserved vector in the workgroup's LDS tile (field f at
// p[f * stride]), read where the fractional steps use it instead of being held
// in registers for the whole sub-cycle.
struct LdsNeighbour {
    const double *p;
    int stride;
    __device__ __forceinline__ double operator[](int f) const { return p[f * stride]; }
};
// Before each fractional step: values read from the LDS tile are read again
// (a compiler barrier, no instruction), so none stays live across steps.
template <int NF>
__device__ __forceinline__ void reload_point(const double (&)[NF])
{
}
__device__ __forceinline__ void reload_point(const LdsNeighbour &) { asm volatile("" ::: "memory"); }

// Solvers::evolveStateByFractionalX/YTimeStep (hprsolvers.cpp:606-620, 673-687;
// mhdsolvers.cpp:612-626) on the conserved vectors of the three cells: slope,
// half-slope extrapolations, flux difference times frac * (ts / dx) added to the
// middle state, converted back to primitives (setConservedVariableVector).
template <class P, int DIR, class NB>
__device__ __forceinline__ void fractional_step(const typename P::Params &m, const NB &UL, const double (&UM)[P::NF],
                                                const NB &UR, double bias, int limiter, double c, double (&w)[P::NF])
{
    constexpr int NF = P::NF;
    static_assert(StoredFields<P>::n == NF, "this kernel reads and writes all NF primitive planes");
    reload_point(UL);
    double xl[NF], xr[NF];
    const double cl = 0.5 * (1.0 + bias), cr = 0.5 * (1.0 - bias);
    // finite slope differences of opposite signs (r < 0) or equal ones (r = 1
    // exactly) need no division: the limiter's zero piece, or its value at 1
    const double f1 = limiter_value(1.0, bias, limiter);
#pragma unroll
    for (int k = 0; k < NF; k++) {
        const double dl = UM[k] - UL[k];
        const double dr = UR[k] - UM[k];
        double d = (cl * dl) + (cr * dr);
        const double num = (fabs(dl) < 1e-5) ? 1e-5 : dl;
        const double den = (fabs(dr) < 1e-5) ? 1e-5 : dr;
        double lim;
        if (limiter_in_range(num, den) && num * den < 0.0) lim = 0.0;
        else if (limiter_in_range(num, den) && num == den) lim = f1;
        else lim = limiter_value(num / den, bias, limiter);
        d *= lim;
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

// Recorder of the stage states of forcing_cell (the forward records nothing;
// the adjoint keeps the states it linearises at).
struct NoStages {
    template <int NF>
    __device__ __forceinline__ void mid(int, int, const double (&)[NF])
    {
    }
    template <int NF>
    __device__ __forceinline__ void end(int, const double (&)[NF])
    {
    }
};

// One cell's RK4 sub-cycle: N (conserved) from its own conserved vector Um and
// its neighbours' (UT, UB unused in 1D).  K[st][q] are the stage increments of
// the non-zero slots.  rec.mid(st, phase, UM) sees the middle state entering
// each fractional step of stage st, rec.end(st, Ue) the state whose source gives
// the next increment.
template <class P, int DIM, class R, class NB>
__device__ __forceinline__ void forcing_cell(const typename P::Params &m, const NB &UL, const NB &UR, const NB &UT,
                                             const NB &UB, const double (&Um)[P::NF], double ts, double dx,
                                             double bias, int limiter, double (&N)[P::NF],
                                             double (&K)[4][ForcingTraits<P>::kMaxSlots], R &rec)
{
    constexpr int NF = P::NF;
    static_assert(StoredFields<P>::n == NF, "this kernel reads and writes all NF primitive planes");
    using T = ForcingTraits<P>;
    constexpr int NS = T::kMaxSlots;
    const int nsrc = T::count(m);
    double S[NF];
    P::source(m, Um, S);
#pragma unroll
    for (int q = 0; q < NS; q++) K[0][q] = ts * S[T::slot(q)];
    const double cx = ts / dx;
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
        rec.mid(st, 0, UM);
        if (DIM == 1) {
            fractional_step<P, 0>(m, UL, UM, UR, bias, limiter, (st < 2 ? 0.5 : 1.0) * cx, e);
        } else {
            const double fx = (st < 2) ? 0.25 : 0.5, fy = (st < 2) ? 0.5 : 1.0;
            fractional_step<P, 0>(m, UL, UM, UR, bias, limiter, fx * cx, e);
            P::prim_to_cons(m, e, UM);
            rec.mid(st, 1, UM);
            fractional_step<P, 1>(m, UT, UM, UB, bias, limiter, fy * cx, e);
            P::prim_to_cons(m, e, UM);
            rec.mid(st, 2, UM);
            fractional_step<P, 0>(m, UL, UM, UR, bias, limiter, fx * cx, e);
        }
        double Ue[NF];
        P::prim_to_cons(m, e, Ue);
        rec.end(st, Ue);
        P::source(m, Ue, S);
#pragma unroll
        for (int q = 0; q < NS; q++) K[st + 1][q] = ts * S[T::slot(q)];
    }
    // computeEvolvedConservedVariableVector (hprforcingsolver.cpp:332-339)
#pragma unroll
    for (int k = 0; k < NF; k++) N[k] = Um[k] + 0.0;
#pragma unroll
    for (int q = 0; q < NS; q++) {
        if (q < nsrc) {
            const int k = T::slot(q);
            N[k] = Um[k] + ((1.0 / 6.0) * (K[0][q] + ((2.0 * K[1][q]) + ((2.0 * K[2][q]) + K[3][q]))));
        }
    }
}

// The neighbour states of forcing_cell: computeRungeKuttaTimeStep converts
// every cell to conserved form; the neighbours re-enter the predictors as
// primitive states, so a neighbour's conserved vector is p2c(c2p(p2c(w))).
template <class P>
__device__ __forceinline__ void forcing_neighbour(const typename P::Params &m, const double *in, long long plane,
                                                  long long idx, double (&U)[P::NF])
{
    double w[P::NF], t[P::NF];
    load_prim<P::NF>(in, plane, idx, w);
    P::prim_to_cons(m, w, t);
    P::cons_to_prim(m, t, w);
    P::prim_to_cons(m, w, U);
}

// Workgroup tile of the source stage: FTX x FTY cells (2D: rows of 32 cells,
// 1D: 128 cells of the line).  The neighbours' conserved vectors of the tile
// and its one-cell rim (corners excluded) are computed once into LDS, so a
// cell's sub-cycle keeps only its own state in registers.
template <int DIM>
struct ForcingTile {
    static constexpr int tx = (DIM == 2) ? 32 : kForceBS, ty = kForceBS / tx;
    static constexpr int lx = tx + 2, ly = (DIM == 2) ? ty + 2 : 1, slots = lx * ly;
};

// This workgroup's tile: its first row and column (i0, j0) and the thread's
// cell (i, j).
template <int DIM>
__device__ __forceinline__ void forcing_tile_origin(int nx, int &i0, int &j0, int &i, int &j)
{
    using Tile = ForcingTile<DIM>;
    const int tid = threadIdx.x;
    if (DIM == 2) {
        const int bx = (nx + Tile::tx - 1) / Tile::tx;
        i0 = (blockIdx.x / bx) * Tile::ty;
        j0 = (blockIdx.x % bx) * Tile::tx;
        i = i0 + tid / Tile::tx;
        j = j0 + tid % Tile::tx;
    } else {
        const long long c0 = (long long)blockIdx.x * kForceBS;
        i0 = (int)(c0 / nx);
        j0 = (int)(c0 - (long long)i0 * nx);
        i = i0;
        j = j0 + tid;   // 1D grids have one row (ny = 1)
    }
}

// Fills the tile's slots: slot (ly, lx) holds the conserved vector that
// cell_vector(idx, U) gives for the cell at (i0 - 1 + ly, j0 - 1 + lx) (1D:
// (i0, j0 - 1 + lx)), positions outside the grid taking the boundary's ghost
// cell; slots no cell of the tile reads are skipped.  Ends with a barrier.
template <int NF, int DIM, class F>
__device__ __forceinline__ void fill_neighbour_tile(double *s_nb, int i0, int j0, int ny, int nx, int bc,
                                                    const F &cell_vector)
{
    using Tile = ForcingTile<DIM>;
    constexpr int SL = Tile::slots;
    for (int sidx = threadIdx.x; sidx < SL; sidx += kForceBS) {
        const int sy = sidx / Tile::lx, sx = sidx - sy * Tile::lx;
        const int gi = (DIM == 2) ? i0 - 1 + sy : i0, gj = j0 - 1 + sx;
        const bool rim_x = (sx == 0 || sx == Tile::lx - 1), rim_y = (DIM == 2) && (sy == 0 || sy == Tile::ly - 1);
        const bool need = !(rim_x && rim_y) && gi >= -1 && gi <= ny && gj >= -1 && gj <= nx;
        if (!need) continue;
        const long long idx = (long long)ghost_index(gi, ny, bc) * nx + ghost_index(gj, nx, bc);
        double U[NF];
        cell_vector(idx, U);
#pragma unroll
        for (int f = 0; f < NF; f++) s_nb[f * SL + sidx] = U[f];
    }
    __syncthreads();
}

// The four neighbours of cell (i, j) in the tile.
template <int DIM>
struct TileNeighbours {
    LdsNeighbour l, r, t, b;
    __device__ __forceinline__ TileNeighbours(const double *s_nb, int i, int j, int i0, int j0)
    {
        using Tile = ForcingTile<DIM>;
        const int sy = (DIM == 2) ? i - i0 + 1 : 0, sx = j - j0 + 1;
        const int me = sy * Tile::lx + sx;
        l = {s_nb + me - 1, Tile::slots};
        r = {s_nb + me + 1, Tile::slots};
        t = {s_nb + me - Tile::lx * (DIM == 2), Tile::slots};
        b = {s_nb + me + Tile::lx * (DIM == 2), Tile::slots};
    }
};

inline long long forcing_blocks(int dim, int ny, int nx)
{
    if (dim == 2) {
        using T = ForcingTile<2>;
        return (long long)((nx + T::tx - 1) / T::tx) * ((ny + T::ty - 1) / T::ty);
    }
    return ((long long)ny * nx + kForceBS - 1) / kForceBS;
}

template <class P, int DIM>
__global__ __launch_bounds__(kForceBS, VG_FORCING_WAVES) void forcing_kernel(ForcingArgs<P> a)
{
    constexpr int NF = P::NF;
    static_assert(StoredFields<P>::n == NF, "this kernel reads and writes all NF primitive planes");
    using T = ForcingTraits<P>;
    using Tile = ForcingTile<DIM>;
    constexpr int NS = T::kMaxSlots, SL = Tile::slots;
    __shared__ double s_nb[NF * SL];   // neighbour conserved vectors: field f of slot s at f * SL + s
    if (a.clk->done) return;
    typename P::Params m = a.m;
    if (!a.fixed_params) P::begin_step(m, a.clk->S);   // sub-cycles run in second-order steps only
    const double ts = 0.5 * (a.clk->dt / a.subcycles);
    int i, j, i0, j0;
    forcing_tile_origin<DIM>(a.nx, i0, j0, i, j);
    fill_neighbour_tile<NF, DIM>(s_nb, i0, j0, a.ny, a.nx, a.bc, [&](long long idx, double (&U)[NF]) {
        forcing_neighbour<P>(m, a.in, a.plane, idx, U);
    });
    unsigned long long key = 0;
    if (i < a.ny && j < a.nx) {
        const TileNeighbours<DIM> nb(s_nb, i, j, i0, j0);
        const LdsNeighbour &UL = nb.l, &UR = nb.r, &UT = nb.t, &UB = nb.b;
        const long long c = (long long)i * a.nx + j;
        double w0[NF], Um[NF];
        load_prim<NF>(a.in, a.plane, c, w0);
        P::prim_to_cons(m, w0, Um);
        double N[NF], K[4][NS];
        NoStages rec;
        forcing_cell<P, DIM>(m, UL, UR, UT, UB, Um, ts, a.dx, a.bias, a.limiter, N, K, rec);
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
        return forcing_blocks(g.dim, g.ny, g.nx);
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
