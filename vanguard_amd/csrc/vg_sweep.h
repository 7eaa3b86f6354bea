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
    unsigned long long *partials;   // one wave-speed key per block when fuse_max
    int fixed_params = 0;           // keep m as given (a per-sweep call): no per-step parameters from the clock
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
    for (int f = 0; f < NF; f++) w[f] = in[f * plane + idx];
}

template <int NF>
__device__ __forceinline__ void store_prim(double *__restrict__ out, long long plane, long long idx, const double (&w)[NF])
{
#pragma unroll
    for (int f = 0; f < NF; f++) out[f * plane + idx] = w[f];
}

// Block-level maximum of the wave-speed keys, written to this block's slot of
// the partials array (no atomics: every block owns one slot, and the 1-block
// clock kernel reduces the slots, so the result does not depend on ordering).
template <int BS>
__device__ __forceinline__ void block_max_to(unsigned long long *partials, unsigned long long key)
{
    __shared__ unsigned long long s_w[BS / kWave];
    key = wave_max_u64(key);
    if ((threadIdx.x & (kWave - 1)) == 0) s_w[threadIdx.x / kWave] = key;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long mx = 0;
#pragma unroll
        for (int k = 0; k < BS / kWave; k++) mx = (s_w[k] > mx) ? s_w[k] : mx;
        partials[(long long)blockIdx.y * gridDim.x + blockIdx.x] = mx;
    }
}

// One tile of an x sweep: thread tid holds the primitive state w of cell
// j = j0 - G + tid (boundary-mapped by the caller); threads G .. BS-G-1 with
// `write` set compute and store the updated cell at out[f * plane + out_idx].
// Returns the wave-speed key of the stored cell when want_key (else 0).  Ends
// with a barrier, so the LDS arrays can be reused by the next tile.
template <class P, int ORDER, int MODE, int BS>
__device__ __forceinline__ unsigned long long row_tile(const typename P::Params &m, const StepScalars &sc, double bias,
                                                       int limiter, const double (&w)[P::NF], bool write,
                                                       double *__restrict__ out, long long plane, long long out_idx,
                                                       int dim, bool want_key, double *s_a, double *s_e)
{
    constexpr int NF = P::NF;
    constexpr int G = (ORDER == 2) ? 2 : 1;
    using S = Scheme<P, 0, MODE>;
    const int tid = threadIdx.x;
    double U[NF], Fp[NF];
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
            S::evolve(m, UL, U, UR, bias, limiter, sc.half_c, left, right);
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
    unsigned long long key = 0;
    if (tid >= G && tid < BS - G && write) {
        double Fm[NF], wn[NF];
#pragma unroll
        for (int f = 0; f < NF; f++) Fm[f] = s_a[f * BS + tid - 1];
        S::update(m, U, Fm, Fp, sc.up_c, wn);
        store_state<P>(out, plane, out_idx, wn);
        if (want_key) key = max_key(P::wave_speed(m, wn, dim));
    }
    __syncthreads();
    return key;
}

template <class P, int ORDER, int MODE, int BS, int RX>
__global__ __launch_bounds__(BS) void sweep_x_kernel(SweepArgs<P> a)
{
    constexpr int NF = P::NF;
    constexpr int G = (ORDER == 2) ? 2 : 1;
    constexpr int OUT = BS - 2 * G;
    __shared__ double s_a[NF * BS];       // cell conserved vectors, later face fluxes
    __shared__ double s_e[2 * NF * BS];   // left evolved state (SLIC) / cell state (FORCE): U and F

    if (a.clk->done) return;
    const double ts = a.clk->dt * a.dt_scale;
    const StepScalars sc(ts, a.dx);
    typename P::Params m = a.m;
    if (ORDER == 2 && !a.fixed_params) P::begin_step(m, a.clk->S);   // first-order solvers never set c_h

    const int tid = threadIdx.x;
    const int j = blockIdx.x * OUT - G + tid;
    const int nx = a.nx;
    const int jj = (j >= -G && j < nx + G) ? ghost_index(j, nx, a.bc) : (j < 0 ? 0 : nx - 1);
    const int i_begin = blockIdx.y * RX;
    const int i_end = (i_begin + RX < a.ny) ? i_begin + RX : a.ny;
    unsigned long long key = 0;

    // Each block walks RX rows; the next row's cells are loaded while the
    // current row is computed.
    double wnext[NF];
    load_state<P>(a.in, a.plane, (long long)i_begin * nx + jj, wnext);
    for (int i = i_begin; i < i_end; i++) {
        const long long row = (long long)i * nx;
        double w[NF];
#pragma unroll
        for (int f = 0; f < NF; f++) w[f] = wnext[f];
        if (i + 1 < i_end) load_state<P>(a.in, a.plane, row + nx + jj, wnext);
        const unsigned long long k2 = row_tile<P, ORDER, MODE, BS>(m, sc, a.bias, a.limiter, w, j < nx, a.out, a.plane,
                                                                   row + j, a.dim, a.fuse_max != 0, s_a, s_e);
        key = (k2 > key) ? k2 : key;
    }
    if (a.fuse_max) block_max_to<BS>(a.partials, key);
}

// CPT columns per thread (j, j + BS, ...): their computations are independent,
// so the instruction stream of a wave interleaves CPT dependency chains (the
// divide sequences of EXACT mode are long chains); the arithmetic of every
// cell is the same.  Columns past the grid's edge compute on the last column
// and store nothing.
template <class P, int ORDER, int MODE, int BS, int H, int CPT = 1>
__global__ __launch_bounds__(BS) void sweep_y_kernel(SweepArgs<P> a)
{
    constexpr int NF = P::NF;
    using S = Scheme<P, 1, MODE>;

    if (a.clk->done) return;
    const double ts = a.clk->dt * a.dt_scale;
    const StepScalars sc(ts, a.dx);
    typename P::Params m = a.m;
    if (ORDER == 2 && !a.fixed_params) P::begin_step(m, a.clk->S);   // first-order solvers never set c_h

    const int j0 = blockIdx.x * BS * CPT + threadIdx.x;
    const int ny = a.ny, nx = a.nx;
    const int i0 = blockIdx.y * H;
    const int i1 = (i0 + H < ny) ? i0 + H : ny;
    unsigned long long key = 0;

    if (j0 < nx) {
        int jc[CPT];
        bool st[CPT];
#pragma unroll
        for (int c = 0; c < CPT; c++) {
            const int j = j0 + c * BS;
            st[c] = j < nx;
            jc[c] = st[c] ? j : nx - 1;
        }
        auto row_of = [&](int r, int c) -> long long { return (long long)ghost_index(r, ny, a.bc) * nx + jc[c]; };
        if (ORDER == 2) {
            double A[CPT][NF], B[CPT][NF], C[CPT][NF], w[CPT][NF];
#pragma unroll
            for (int c = 0; c < CPT; c++) {
                load_state<P>(a.in, a.plane, row_of(i0 - 2, c), w[c]);
                S::cons(m, w[c], A[c]);
                load_state<P>(a.in, a.plane, row_of(i0 - 1, c), w[c]);
                S::cons(m, w[c], B[c]);
                load_state<P>(a.in, a.plane, row_of(i0, c), w[c]);   // row e + 1 of the first iteration
            }
            ConsFlux<NF> er_prev[CPT];
            double Fprev[CPT][NF];
            for (int e = i0 - 1; e <= i1; e++) {
#pragma unroll
                for (int c = 0; c < CPT; c++) {
                    S::cons(m, w[c], C[c]);
                    if (e + 2 <= i1 + 1) load_state<P>(a.in, a.plane, row_of(e + 2, c), w[c]);   // prefetch
                    ConsFlux<NF> L, R;
                    S::evolve(m, A[c], B[c], C[c], a.bias, a.limiter, sc.half_c, L, R);
                    double Fnew[NF];
                    if (e >= i0) S::force_face(m, er_prev[c], L, sc.lf_c, sc.ri_c, Fnew);   // face e - 1/2
                    if (e >= i0 + 1) {
                        double wn[NF];
                        S::update(m, A[c], Fprev[c], Fnew, sc.up_c, wn);                    // row e - 1
                        if (st[c]) {
                            store_state<P>(a.out, a.plane, (long long)(e - 1) * nx + jc[c], wn);
                            if (a.fuse_max) {
                                const unsigned long long k2 = max_key(P::wave_speed(m, wn, a.dim));
                                key = (k2 > key) ? k2 : key;
                            }
                        }
                    }
#pragma unroll
                    for (int f = 0; f < NF; f++) {
                        Fprev[c][f] = Fnew[f];
                        er_prev[c].U[f] = R.U[f];
                        er_prev[c].F[f] = R.F[f];
                        A[c][f] = B[c][f];
                        B[c][f] = C[c][f];
                    }
                }
            }
        } else {
            double w[CPT][NF];
            ConsFlux<NF> prev[CPT];
            double Fprev[CPT][NF];
#pragma unroll
            for (int c = 0; c < CPT; c++) {
                load_state<P>(a.in, a.plane, row_of(i0 - 1, c), w[c]);
                S::cons_flux_of_prim(m, w[c], prev[c]);
            }
            for (int e = i0; e <= i1; e++) {
#pragma unroll
                for (int c = 0; c < CPT; c++) {
                    ConsFlux<NF> cur;
                    load_state<P>(a.in, a.plane, row_of(e, c), w[c]);
                    S::cons_flux_of_prim(m, w[c], cur);
                    double Fnew[NF];
                    S::force_face(m, prev[c], cur, sc.lf_c, sc.ri_c, Fnew);                 // face e - 1/2
                    if (e >= i0 + 1) {
                        double wn[NF];
                        S::update(m, prev[c].U, Fprev[c], Fnew, sc.up_c, wn);               // row e - 1
                        if (st[c]) {
                            store_state<P>(a.out, a.plane, (long long)(e - 1) * nx + jc[c], wn);
                            if (a.fuse_max) {
                                const unsigned long long k2 = max_key(P::wave_speed(m, wn, a.dim));
                                key = (k2 > key) ? k2 : key;
                            }
                        }
                    }
#pragma unroll
                    for (int f = 0; f < NF; f++) {
                        Fprev[c][f] = Fnew[f];
                        prev[c].U[f] = cur.U[f];
                        prev[c].F[f] = cur.F[f];
                    }
                }
            }
        }
    }
    if (a.fuse_max) block_max_to<BS>(a.partials, key);
}

// Wave-speed maximum of a whole grid (the first step, or after the caller
// changed the state).
template <class P>
__global__ __launch_bounds__(256) void wave_max_kernel(const double *__restrict__ in, long long plane, long long cells,
                                                       int dim, typename P::Params m, unsigned long long *partials)
{
    constexpr int NF = P::NF;
    unsigned long long key = 0;
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < cells; c += (long long)gridDim.x * blockDim.x) {
        double w[NF];
        load_prim<NF>(in, plane, c, w);
        const unsigned long long k2 = max_key(P::wave_speed(m, w, dim));
        key = (k2 > key) ? k2 : key;
    }
    block_max_to<256>(partials, key);
}

// Reduces n partial keys with one block of kReduceThreads threads; returns the
// maximum in thread 0.
constexpr int kReduceThreads = 1024;

__device__ __forceinline__ unsigned long long reduce_partials(const unsigned long long *partials, int n)
{
    __shared__ unsigned long long s_r[kReduceThreads / kWave];
    unsigned long long key = 0;
    for (int k = threadIdx.x; k < n; k += kReduceThreads) key = (partials[k] > key) ? partials[k] : key;
    key = wave_max_u64(key);
    if ((threadIdx.x & (kWave - 1)) == 0) s_r[threadIdx.x / kWave] = key;
    __syncthreads();
    unsigned long long mx = 0;
    if (threadIdx.x == 0) {
        for (int k = 0; k < kReduceThreads / kWave; k++) mx = (s_r[k] > mx) ? s_r[k] : mx;
    }
    return mx;
}

static __global__ __launch_bounds__(kReduceThreads) void reduce_max_kernel(const unsigned long long *partials, int n, Clock *clk)
{
    const unsigned long long mx = reduce_partials(partials, n);
    if (threadIdx.x == 0) clk->smax = mx;
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
static __global__ __launch_bounds__(kReduceThreads) void clock_begin_kernel(Clock *c, const unsigned long long *partials,
                                                                    int n_partials, double *dt_log, int dt_cap)
{
    if (c->done) return;   // uniform: nothing pending once done
    const unsigned long long mx = reduce_partials(partials, n_partials);
    if (threadIdx.x != 0) return;
    c->smax = mx;
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
    double S = __longlong_as_double((long long)mx);
    if (c->corner > S) S = c->corner;
    double dt = c->cfl * (c->dx / S);
    if (c->it <= 5) dt *= 0.2;
    if ((c->t + dt) > c->t_final) dt = c->t_final - c->t;
    if (!(dt > 0.0) || !(c->t + dt > c->t)) {
        // The reference would repeat this step forever (t never reaches t_final);
        // stop and report instead.
        c->done = 1;
        c->stalled = 1;
        return;
    }
    c->dt = dt;
    c->S = S;
    if (dt_log && c->steps_this_run < dt_cap) dt_log[c->steps_this_run] = dt;
    if (c->steps_this_run == 0 || dt < c->dt_min) c->dt_min = dt;
    if (c->steps_this_run == 0 || dt > c->dt_max) c->dt_max = dt;
    c->steps_this_run += 1;
    c->pending = 1;
}

static __global__ void clock_finish_kernel(Clock *c)
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

// Counts cells whose state is not finite or not admissible for the physics
// (P::admissible: e.g. a real sound speed), and records the smallest such flat index.
template <class P>
__global__ __launch_bounds__(256) void nonfinite_kernel(const double *__restrict__ soa, long long plane, long long cells,
                                                        typename P::Params m, unsigned long long *count,
                                                        unsigned long long *first)
{
    constexpr int NF = P::NF;
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < cells; c += (long long)gridDim.x * blockDim.x) {
        double w[NF];
        bool bad = false;
#pragma unroll
        for (int f = 0; f < NF; f++) {
            w[f] = (f < StoredFields<P>::n) ? soa[f * plane + c] : 0.0;
            bad = bad || !isfinite(w[f]);
        }
        bad = bad || !P::admissible(m, w);
        if (bad) {
            atomicAdd(count, 1ull);
            atomicMin(first, (unsigned long long)c);
        }
    }
}

}  // namespace vg
