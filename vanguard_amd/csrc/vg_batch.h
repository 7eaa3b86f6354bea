// Batched 1D solves in one launch.
//
// Thousands of independent 1D problems of uneven size (each the reference's
// whole solve(), e.g. MHDSecondOrderSolver::solve, mhdsecondordersolver.cpp:318-360)
// run in one kernel: workgroup b takes instance order[b] (instances sorted by
// decreasing estimated work, so the hardware dispatcher, which hands out
// workgroups in index order as CUs free up, schedules longest-first) and runs
// that instance's complete time loop -- CFL reduction, dt, SLIC/FORCE sweep in
// tiles of BS cells.  Per-instance time steps and parameters (MHD: c_h) are kept
// in registers; the grid lives in HBM/L2 as per-instance structure-of-arrays
// planes, ping-ponged between two buffers.
//
// There is deliberately no software work queue: a loop of the form
// `for (;;) { if (tid == 0) s_q = atomicAdd(..); __syncthreads(); ... }` is
// restructured by the compiler so that lane 0 leaves the loop to fetch while
// the other lanes of its wave re-enter the barrier, and the block never
// reconverges.  Every barrier below sits in wave-uniform control flow.
#pragma once

#include "vg_sweep.h"

namespace vg {

// Upper bound on the steps of one instance in one launch; every workgroup's
// loop is bounded, so the grid always drains.
constexpr int kStepGuard = 50000000;

template <class P>
struct BatchArgs {
    int n_inst;
    const int *order;                 // processing order (instance ids)
    const long long *offsets;         // n_inst + 1 cell offsets
    const double *dx;
    const typename P::Params *params;
    const double *t_final;
    double cfl, bias;
    int limiter, max_steps, bc;
    double *buf0, *buf1;              // instance q: planes of n_q doubles at NF * offsets[q]
    int *steps_out;
    double *t_out;
    int *where;                       // 0: result in buf0, 1: in buf1
    int *stalled;
    double *dt_log;                   // single-instance runs: the dt sequence
    int dt_cap;
    double work_budget;               // > 0: an instance of n cells stops after max(1024, ceil(budget / n)) steps
};

template <int BS>
__device__ __forceinline__ unsigned long long block_allreduce_max(unsigned long long key, unsigned long long *s_red)
{
    key = wave_max_u64(key);
    if ((threadIdx.x & (kWave - 1)) == 0) s_red[threadIdx.x / kWave] = key;
    __syncthreads();
    unsigned long long mx = 0;
#pragma unroll
    for (int k = 0; k < BS / kWave; k++) mx = (s_red[k] > mx) ? s_red[k] : mx;
    __syncthreads();
    return mx;
}

template <class P, int ORDER, int MODE, int BS>
__global__ __launch_bounds__(BS) void batch_solve_kernel(BatchArgs<P> a)
{
    constexpr int NF = P::NF;
    static_assert(StoredFields<P>::n == NF, "this kernel reads and writes all NF primitive planes");
    constexpr int G = (ORDER == 2) ? 2 : 1;
    constexpr int OUT = BS - 2 * G;
    __shared__ double s_a[NF * BS];
    __shared__ double s_e[2 * NF * BS];
    __shared__ unsigned long long s_red[BS / kWave];
    const int tid = threadIdx.x;
    const int q = blockIdx.x;
    if (q >= a.n_inst) return;
    const int inst = a.order ? a.order[q] : q;
    const long long off = a.offsets[inst];
    const int n = (int)(a.offsets[inst + 1] - off);
    typename P::Params m = a.params[inst];
    const double dx = a.dx[inst];
    const double T = a.t_final[inst];
    double *A = a.buf0 + (long long)NF * off;
    double *B = a.buf1 + (long long)NF * off;

    unsigned long long key = 0;
    for (int j = tid; j < n; j += BS) {
        double w[NF];
        load_prim<NF>(A, n, j, w);
        const unsigned long long k2 = max_key(P::wave_speed(m, w, 1));
        key = (k2 > key) ? k2 : key;
    }
    key = block_allreduce_max<BS>(key, s_red);

    // The solve loop (secondordersolver.cpp:295-311): dt from the maximum
    // wave speed, x0.2 for the first six steps, clipped to the final time.
    // This instance's step cap: the launch's, lowered for a large instance by
    // the work budget (cells x steps), so no single workgroup holds the launch
    // much past the others; the host continues such an instance elsewhere.
    int cap = a.max_steps;
    if (a.work_budget > 0.0) {
        const double c = ceil(a.work_budget / (double)n);
        const int cq = (c < 1024.0) ? 1024 : (c > 2.0e9 ? 2000000000 : (int)c);
        if (cap == 0 || cq < cap) cap = cq;
    }
    double t = 0.0;
    int it = 0, stalled = 0;
    for (;;) {
        if (!(t < T) || (cap > 0 && it >= cap)) break;
        const double S = __longlong_as_double((long long)key);
        double dt = a.cfl * (dx / S);
        if (it <= 5) dt *= 0.2;
        if ((t + dt) > T) dt = T - t;
        // A step that cannot advance t would make the reference loop forever;
        // stop the instance and report it instead (also past kStepGuard steps).
        if (!(dt > 0.0) || !(t + dt > t) || it >= kStepGuard) {
            stalled = 1;
            break;
        }
        if (a.dt_log && it < a.dt_cap && tid == 0) a.dt_log[it] = dt;
        if (ORDER == 2) P::begin_step(m, S);   // first-order solvers never set c_h
        const StepScalars sc(dt, dx);
        key = 0;
        for (int j0 = 0; j0 < n; j0 += OUT) {
            const int j = j0 - G + tid;
            const int jj = (j >= -G && j < n + G) ? ghost_index(j, n, a.bc) : (j < 0 ? 0 : n - 1);
            double w[NF];
            load_prim<NF>(A, n, jj, w);
            const unsigned long long k2 =
                row_tile<P, ORDER, MODE, BS>(m, sc, a.bias, a.limiter, w, j < n, B, n, j, 1, true, s_a, s_e);
            key = (k2 > key) ? k2 : key;
        }
        key = block_allreduce_max<BS>(key, s_red);   // its barriers also publish B for the next step
        double *tmp = A;
        A = B;
        B = tmp;
        t += dt;
        it += 1;
    }
    if (tid == 0) {
        a.steps_out[inst] = it;
        a.t_out[inst] = t;
        a.where[inst] = it & 1;
        a.stalled[inst] = stalled;
    }
}

// Initial maximum wave speed of each instance (one block per instance), used to
// order the work queue.
template <class P>
__global__ __launch_bounds__(256) void batch_speed_kernel(const long long *offsets, const typename P::Params *params,
                                                          const double *buf, double *speed)
{
    constexpr int NF = P::NF;
    static_assert(StoredFields<P>::n == NF, "this kernel reads and writes all NF primitive planes");
    __shared__ unsigned long long s_red[256 / kWave];
    const int inst = blockIdx.x;
    const long long off = offsets[inst];
    const int n = (int)(offsets[inst + 1] - off);
    const double *A = buf + (long long)NF * off;
    unsigned long long key = 0;
    for (int j = threadIdx.x; j < n; j += 256) {
        double w[NF];
        load_prim<NF>(A, n, j, w);
        const unsigned long long k2 = max_key(P::wave_speed(params[inst], w, 1));
        key = (k2 > key) ? k2 : key;
    }
    key = block_allreduce_max<256>(key, s_red);
    if (threadIdx.x == 0) speed[inst] = __longlong_as_double((long long)key);
}

// Array-of-structures cells of all instances <-> per-instance planes.
__device__ __forceinline__ int find_instance(const long long *offsets, int n_inst, long long c)
{
    int lo = 0, hi = n_inst - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (offsets[mid] <= c) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

template <int NF>
__global__ __launch_bounds__(256) void batch_aos_to_planes(const double *__restrict__ aos, double *__restrict__ buf,
                                                           const long long *offsets, int n_inst, long long cells)
{
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < cells; c += (long long)gridDim.x * blockDim.x) {
        const int q = find_instance(offsets, n_inst, c);
        const long long off = offsets[q], n = offsets[q + 1] - off;
#pragma unroll
        for (int f = 0; f < NF; f++) buf[NF * off + f * n + (c - off)] = aos[c * NF + f];
    }
}

template <int NF>
__global__ __launch_bounds__(256) void batch_planes_to_aos(const double *__restrict__ buf0, const double *__restrict__ buf1,
                                                           const int *where, double *__restrict__ aos,
                                                           const long long *offsets, int n_inst, long long cells)
{
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < cells; c += (long long)gridDim.x * blockDim.x) {
        const int q = find_instance(offsets, n_inst, c);
        const long long off = offsets[q], n = offsets[q + 1] - off;
        const double *src = where[q] ? buf1 : buf0;
#pragma unroll
        for (int f = 0; f < NF; f++) aos[c * NF + f] = src[NF * off + f * n + (c - off)];
    }
}

}  // namespace vg
