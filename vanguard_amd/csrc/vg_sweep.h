// Dimensionally split sweep kernels, generic over the physics P.
//
// Data layout: structure-of-arrays primitive planes in HBM, plane f of cell
// (i, j) at in[f * plane + i * nx + j], x (j) contiguous.  No ghost cells are
// stored: the transmissive/periodic ghosts of the reference's
// insertBoundaryCells[2D] (Solvers/solvers.cpp:7-143) are produced by index
// mapping at load time (ghost_index).
//
// X sweep: one thread per cell along a row, tiles of BS cells overlapping by the
// stencil halo; neighbour states and face fluxes are exchanged through LDS so
// that each face flux is computed once (the reference computes it once per
// adjacent cell, secondordersolver.cpp:90-93; both evaluations are identical).
// Y sweep: one thread per column, sliding down H rows with the stencil window in
// registers -- every load is a coalesced row segment, no LDS and no barriers.
// The last sweep of a step can fold the next step's CFL wave-speed maximum into
// its stores (fuse_max).
#pragma once

#include "vg_device.h"
#include "vg_scheme.h"

namespace vg {

template <class P>
struct SweepArgs {
    const double *in;
    double *out;
    long long plane;
    int ny, nx;
    int bc, limiter;
    double dx, bias;
    double dt_scale;   // 0.5 for the half-step X sweeps of the 2D split, else 1
    int dim;           // 1 or 2: selects the wave-speed formula of the fused max
    int fuse_max;
    typename P::Params m;
    Clock *clk;
};

// Per-step scalars, computed identically in every thread.
struct StepScalars {
    double half_c, lf_c, ri_c, up_c;
    __device__ __forceinline__ explicit StepScalars(double ts, double dx)
    {
        half_c = 0.5 * (ts / dx);  // computeEvolutionVector (solvers.cpp:601-609)
        lf_c = dx / ts;            // computeLaxFriedrichsFlux (firstordersolver.cpp:83)
        ri_c = ts / dx;            // computeRichtmyerFlux (:169), computeFORCEUpdate (:431)
        up_c = ts / dx;
    }
};

template <int NF>
__device__ __forceinline__ void load_prim(const double *__restrict__ in, long long plane, long long idx, double (&w)[NF])
{
#pragma unroll
    for (int f = 0; f < NF; f++) w[f] = __builtin_nontemporal_load(in + f * plane + idx);
}

template <int NF>
__device__ __forceinline__ void store_prim(double *__restrict__ out, long long plane, long long idx, const double (&w)[NF])
{
#pragma unroll
    for (int f = 0; f < NF; f++) out[f * plane + idx] = w[f];
}

__device__ __forceinline__ void flush_max(Clock *clk, unsigned long long key)
{
    key = wave_max_u64(key);
    if ((threadIdx.x & (kWave - 1)) == 0 && key) atomicMax(&clk->smax, key);
}

template <class P, int ORDER, int MODE, int BS>
__global__ __launch_bounds__(BS) void sweep_x_kernel(SweepArgs<P> a)
{
    constexpr int NF = P::NF;
    constexpr int G = (ORDER == 2) ? 2 : 1;
    constexpr int OUT = BS - 2 * G;
    using S = Scheme<P, 0, MODE>;
    __shared__ double s_a[NF * BS];       // cell conserved vectors, later face fluxes
    __shared__ double s_e[2 * NF * BS];   // left evolved state (SLIC) / cell state (FORCE): U and F

    if (a.clk->done) return;
    const double ts = a.clk->dt * a.dt_scale;
    const StepScalars sc(ts, a.dx);
    const typename P::Params m = a.m;

    const int tid = threadIdx.x;
    const int i = blockIdx.y;
    const int j = blockIdx.x * OUT - G + tid;
    const int nx = a.nx;
    const int jj = (j >= -G && j < nx + G) ? ghost_index(j, nx, a.bc) : (j < 0 ? 0 : nx - 1);
    const long long row = (long long)i * nx;

    double w[NF], U[NF];
    load_prim<NF>(a.in, a.plane, row + jj, w);
    double Fp[NF];
    unsigned long long key = 0;

    if (ORDER == 2) {
        S::cons(m, w, U);
#pragma unroll
        for (int f = 0; f < NF; f++) s_a[f * BS + tid] = U[f];
        __syncthreads();
        ConsFlux<NF> left, right;
        if (tid >= 1 && tid <= BS - 2) {
            double UL[NF], UR[NF];
#pragma unroll
            for (int f = 0; f < NF; f++) {
                UL[f] = s_a[f * BS + tid - 1];
                UR[f] = s_a[f * BS + tid + 1];
            }
            S::evolve(m, UL, U, UR, a.bias, a.limiter, sc.half_c, left, right);
#pragma unroll
            for (int f = 0; f < NF; f++) {
                s_e[f * BS + tid] = left.U[f];
                s_e[(NF + f) * BS + tid] = left.F[f];
            }
        }
        __syncthreads();
        if (tid >= 1 && tid <= BS - 3) {
            ConsFlux<NF> nb;
#pragma unroll
            for (int f = 0; f < NF; f++) {
                nb.U[f] = s_e[f * BS + tid + 1];
                nb.F[f] = s_e[(NF + f) * BS + tid + 1];
            }
            S::force_face(m, right, nb, sc.lf_c, sc.ri_c, Fp);   // face j + 1/2
#pragma unroll
            for (int f = 0; f < NF; f++) s_a[f * BS + tid] = Fp[f];
        }
        __syncthreads();
    } else {
        ConsFlux<NF> cf;
        S::cons_flux_of_prim(m, w, cf);
#pragma unroll
        for (int f = 0; f < NF; f++) {
            U[f] = cf.U[f];
            s_e[f * BS + tid] = cf.U[f];
            s_e[(NF + f) * BS + tid] = cf.F[f];
        }
        __syncthreads();
        if (tid <= BS - 2) {
            ConsFlux<NF> nb;
#pragma unroll
            for (int f = 0; f < NF; f++) {
                nb.U[f] = s_e[f * BS + tid + 1];
                nb.F[f] = s_e[(NF + f) * BS + tid + 1];
            }
            S::force_face(m, cf, nb, sc.lf_c, sc.ri_c, Fp);
#pragma unroll
            for (int f = 0; f < NF; f++) s_a[f * BS + tid] = Fp[f];
        }
        __syncthreads();
    }

    if (tid >= G && tid < BS - G && j < nx) {
        double Fm[NF], wn[NF];
#pragma unroll
        for (int f = 0; f < NF; f++) Fm[f] = s_a[f * BS + tid - 1];
        S::update(m, U, Fm, Fp, sc.up_c, wn);
        store_prim<NF>(a.out, a.plane, row + j, wn);
        if (a.fuse_max) key = max_key(P::wave_speed(m, wn, a.dim));
    }
    if (a.fuse_max) flush_max(a.clk, key);
}

template <class P, int ORDER, int MODE, int BS, int H>
__global__ __launch_bounds__(BS) void sweep_y_kernel(SweepArgs<P> a)
{
    constexpr int NF = P::NF;
    using S = Scheme<P, 1, MODE>;

    if (a.clk->done) return;
    const double ts = a.clk->dt * a.dt_scale;
    const StepScalars sc(ts, a.dx);
    const typename P::Params m = a.m;

    const int j = blockIdx.x * BS + threadIdx.x;
    const int ny = a.ny, nx = a.nx;
    const int i0 = blockIdx.y * H;
    const int i1 = (i0 + H < ny) ? i0 + H : ny;
    unsigned long long key = 0;

    if (j < nx) {
        auto row_of = [&](int r) -> long long { return (long long)ghost_index(r, ny, a.bc) * nx + j; };
        if (ORDER == 2) {
            double A[NF], B[NF], C[NF], w[NF];
            load_prim<NF>(a.in, a.plane, row_of(i0 - 2), w);
            S::cons(m, w, A);
            load_prim<NF>(a.in, a.plane, row_of(i0 - 1), w);
            S::cons(m, w, B);
            ConsFlux<NF> er_prev;
            double Fprev[NF];
            for (int e = i0 - 1; e <= i1; e++) {
                load_prim<NF>(a.in, a.plane, row_of(e + 1), w);
                S::cons(m, w, C);
                ConsFlux<NF> L, R;
                S::evolve(m, A, B, C, a.bias, a.limiter, sc.half_c, L, R);
                double Fnew[NF];
                if (e >= i0) S::force_face(m, er_prev, L, sc.lf_c, sc.ri_c, Fnew);   // face e - 1/2
                if (e >= i0 + 1) {
                    double wn[NF];
                    S::update(m, A, Fprev, Fnew, sc.up_c, wn);                       // row e - 1
                    store_prim<NF>(a.out, a.plane, (long long)(e - 1) * nx + j, wn);
                    if (a.fuse_max) {
                        const unsigned long long k2 = max_key(P::wave_speed(m, wn, a.dim));
                        key = (k2 > key) ? k2 : key;
                    }
                }
#pragma unroll
                for (int f = 0; f < NF; f++) {
                    Fprev[f] = Fnew[f];
                    er_prev.U[f] = R.U[f];
                    er_prev.F[f] = R.F[f];
                    A[f] = B[f];
                    B[f] = C[f];
                }
            }
        } else {
            double w[NF];
            ConsFlux<NF> prev;
            load_prim<NF>(a.in, a.plane, row_of(i0 - 1), w);
            S::cons_flux_of_prim(m, w, prev);
            double Fprev[NF];
            for (int e = i0; e <= i1; e++) {
                ConsFlux<NF> cur;
                load_prim<NF>(a.in, a.plane, row_of(e), w);
                S::cons_flux_of_prim(m, w, cur);
                double Fnew[NF];
                S::force_face(m, prev, cur, sc.lf_c, sc.ri_c, Fnew);                 // face e - 1/2
                if (e >= i0 + 1) {
                    double wn[NF];
                    S::update(m, prev.U, Fprev, Fnew, sc.up_c, wn);                  // row e - 1
                    store_prim<NF>(a.out, a.plane, (long long)(e - 1) * nx + j, wn);
                    if (a.fuse_max) {
                        const unsigned long long k2 = max_key(P::wave_speed(m, wn, a.dim));
                        key = (k2 > key) ? k2 : key;
                    }
                }
#pragma unroll
                for (int f = 0; f < NF; f++) {
                    Fprev[f] = Fnew[f];
                    prev.U[f] = cur.U[f];
                    prev.F[f] = cur.F[f];
                }
            }
        }
    }
    if (a.fuse_max) flush_max(a.clk, key);
}

// Wave-speed maximum of a whole grid (the first step, or after the caller
// changed the state).
template <class P>
__global__ __launch_bounds__(256) void wave_max_kernel(const double *__restrict__ in, long long plane, long long cells,
                                                       int dim, typename P::Params m, Clock *clk)
{
    constexpr int NF = P::NF;
    unsigned long long key = 0;
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < cells; c += (long long)gridDim.x * blockDim.x) {
        double w[NF];
        load_prim<NF>(in, plane, c, w);
        const unsigned long long k2 = max_key(P::wave_speed(m, w, dim));
        key = (k2 > key) ? k2 : key;
    }
    flush_max(clk, key);
}

// Default-state wave speed (2D transmissive corner ghosts) into clk->corner.
template <class P>
__global__ void corner_speed_kernel(typename P::Params m, int dim, Clock *clk)
{
    double w[P::NF];
    P::default_state(w);
    clk->corner = P::wave_speed(m, w, dim);
}

// Start of a step: finish the previous one (t += dt; it += 1 as in
// secondordersolver.cpp:307-308), evaluate the loop condition, then
// dt = CFL * (dx / S), x0.2 for the first six steps, clipped to the final time
// (Solvers/solvers.cpp:392, 437-452).
__global__ void clock_begin_kernel(Clock *c, double *dt_log, int dt_cap)
{
    if (c->pending) {
        c->t += c->dt;
        c->it += 1;
        c->pending = 0;
    }
    if (c->done) return;
    if (!(c->t < c->t_final) || (c->max_steps > 0 && c->it >= c->max_steps)) {
        c->done = 1;
        return;
    }
    double S = __longlong_as_double((long long)c->smax);
    if (c->corner > S) S = c->corner;
    c->smax = 0ull;
    double dt = c->cfl * (c->dx / S);
    if (c->it <= 5) dt *= 0.2;
    if ((c->t + dt) > c->t_final) dt = c->t_final - c->t;
    c->dt = dt;
    if (dt_log && c->steps_this_run < dt_cap) dt_log[c->steps_this_run] = dt;
    if (c->steps_this_run == 0 || dt < c->dt_min) c->dt_min = dt;
    if (c->steps_this_run == 0 || dt > c->dt_max) c->dt_max = dt;
    c->steps_this_run += 1;
    c->pending = 1;
}

__global__ void clock_finish_kernel(Clock *c)
{
    if (c->pending) {
        c->t += c->dt;
        c->it += 1;
        c->pending = 0;
    }
}

// Array-of-structures (the reference's layout) <-> structure-of-arrays planes.
template <int NF>
__global__ __launch_bounds__(256) void aos_to_soa_kernel(const double *__restrict__ aos, double *__restrict__ soa,
                                                         long long plane, long long cells)
{
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < cells; c += (long long)gridDim.x * blockDim.x) {
#pragma unroll
        for (int f = 0; f < NF; f++) soa[f * plane + c] = aos[c * NF + f];
    }
}

template <int NF>
__global__ __launch_bounds__(256) void soa_to_aos_kernel(const double *__restrict__ soa, double *__restrict__ aos,
                                                         long long plane, long long cells)
{
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < cells; c += (long long)gridDim.x * blockDim.x) {
#pragma unroll
        for (int f = 0; f < NF; f++) aos[c * NF + f] = soa[f * plane + c];
    }
}

// Counts cells whose state is not finite or whose density is not positive, and
// records the smallest such flat index.
template <int NF>
__global__ __launch_bounds__(256) void nonfinite_kernel(const double *__restrict__ soa, long long plane, long long cells,
                                                        unsigned long long *count, unsigned long long *first)
{
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < cells; c += (long long)gridDim.x * blockDim.x) {
        bool bad = !(soa[c] > 0.0);
#pragma unroll
        for (int f = 0; f < NF; f++) bad = bad || !isfinite(soa[f * plane + c]);
        if (bad) {
            atomicAdd(count, 1ull);
            atomicMin(first, (unsigned long long)c);
        }
    }
}

}  // namespace vg
