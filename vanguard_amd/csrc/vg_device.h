// Device-side building blocks shared by every sweep kernel (gfx950, wave64).
//
// All arithmetic here is IEEE binary64 with no contraction: the library is
// compiled with -ffp-contract=off, so a*b+c stays a multiply and an add, and
// '/' and sqrt are the correctly rounded AMDGPU expansions.  That is what makes
// EXACT mode bit-identical to the reference's x86-64 SSE2 arithmetic.
#pragma once

#include <type_traits>

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vg {

constexpr int kWave = 64;

// std::min / std::max with libstdc++'s argument order (they differ from fmin/fmax
// for NaN and signed zeros).
__device__ __forceinline__ double smin(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ double smax(double a, double b) { return (a < b) ? b : a; }

// Reciprocal for FAST mode: the hardware estimate refined by two Newton steps
// (explicit FMAs), within an ulp of 1/x for normal x > 0; not correctly rounded,
// so never used in EXACT mode.
__device__ __forceinline__ double rcp_nr(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    double e = __builtin_fma(-x, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-x, r, 1.0);
    return __builtin_fma(r, e, r);
}

// Ghost-cell index for a cell index k in [-2, n+1].  Transmissive ghosts are
// mirror-ordered copies of the interior (Solvers/solvers.cpp:17-23, 112-131:
// g[-1] = c[0], g[-2] = c[1], g[n] = c[n-1], g[n+1] = c[n-2]); the one-layer
// copy (:12-16, 94-109) is the same map.  Periodic wraps.
__device__ __forceinline__ int ghost_index(int k, int n, int bc)
{
    if (k >= 0 && k < n) return k;
    if (bc == 1) return k < 0 ? k + n : k - n;
    return k < 0 ? -1 - k : 2 * n - 1 - k;
}

// Key for an order-independent maximum of non-negative doubles: for x >= 0 the
// IEEE bit pattern is monotone in x, so an unsigned 64-bit max of the bits is
// the max of the values.  NaN (and anything not > 0) maps to 0, which is how the
// reference's `waveSpeed > maximumWaveSpeed` test skips it
// (Solvers/solvers.cpp:269-272).
__device__ __forceinline__ unsigned long long max_key(double s)
{
    return (s > 0.0) ? (unsigned long long)__double_as_longlong(s) : 0ull;
}

__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long v)
{
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) {
        unsigned long long o = __shfl_xor(v, off, kWave);
        v = (o > v) ? o : v;
    }
    return v;
}

// Tile shapes of the grid sweeps for a physics P: x-sweep block size and rows
// per block, y-sweep block size and rows per column strip.  The x tile keeps
// 3 * NF * bsx doubles in LDS, so wide states use narrower tiles.  `tune`: the
// y-sweep shape is chosen per grid by the autotuner among fixed candidates.
template <class P>
struct Tiles {
    static constexpr int bsx = 256, rx = 4, bsy = 256, hy = 64;
    static constexpr bool tune = true;
};

// Plane of field f of a physics' state in the grid's structure of arrays: the
// identity, except for a physics that carries a subset of a wider state.
template <class P>
struct PlaneMap {
    __device__ static constexpr int of(int f) { return f; }
};

// Fields a physics stores per cell: P::NF, or P::NS when it keeps fewer
// primitive values than it has conserved variables (the two-material Euler
// models: 8 or 7 primitives for 9 or 8 conserved); the remaining slots of its
// register state are zero.
template <class P, class = void>
struct StoredFields {
    static constexpr int n = P::NF;
};
template <class P>
struct StoredFields<P, std::void_t<decltype(P::NS)>> {
    static constexpr int n = P::NS;
};

template <class P>
__device__ __forceinline__ void load_state(const double *__restrict__ in, long long plane, long long idx,
                                           double (&w)[P::NF])
{
    constexpr int NS = StoredFields<P>::n;
#pragma unroll
    for (int f = 0; f < P::NF; f++) w[f] = (f < NS) ? in[PlaneMap<P>::of(f) * plane + idx] : 0.0;
}

template <class P>
__device__ __forceinline__ void store_state(double *__restrict__ out, long long plane, long long idx,
                                            const double (&w)[P::NF])
{
    constexpr int NS = StoredFields<P>::n;
#pragma unroll
    for (int f = 0; f < NS; f++) out[PlaneMap<P>::of(f) * plane + idx] = w[f];
}

// The narrower physics a grid of P may run with when some of its fields are
// constant (void: none).  Chosen per run by an exact check of those fields.
template <class P>
struct ReducedPhys {
    using type = void;
};

// Per-grid time-step clock kept in device memory so that a whole run can be
// enqueued (or captured in a graph) without host round trips.  The step logic
// is Solvers::computeStableTimeStep (Solvers/solvers.cpp:389-395, 413-419,
// 437-452) and the solve loop condition `while (currentTime < finalTime)`
// (Solvers/secondordersolver.cpp:295, 399).
struct Clock {
    double t;            // current simulated time
    double t_final;
    double dt;           // time step of the step in progress
    double cfl, dx;
    double corner;       // wave speed of the default-state corner ghosts (2D transmissive), else 0
    unsigned long long smax;  // max_key of the wave speeds of the current grid
    int it;              // steps completed
    int max_steps;       // 0 = no cap; else stop when it reaches this (absolute)
    int done;            // set once the loop condition fails
    int pending;         // a step was started and t/it not yet advanced
    double dt_min, dt_max;
    int steps_this_run;
    int stalled;         // set when a step would not advance t (dt <= 0 or NaN while t < t_final)
    double S;            // maximum wave speed of the step in progress (MHD: its c_h)
};

}  // namespace vg
