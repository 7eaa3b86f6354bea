// Generic device-resident grid and time loop for one physics P.
//
// The loop mirrors the reference's solve/solve2D drivers
// (Solvers/secondordersolver.cpp:288-314, 392-424; firstordersolver.cpp:435-461,
// 521-553): each step computes dt from the wave-speed maximum of the current
// grid, then runs one X sweep (1D) or the Strang split X(dt/2) Y(dt) X(dt/2)
// (2D).  Everything stays on the device: dt lives in a Clock in device memory,
// the maximum is folded into the last sweep of the previous step, and the host
// only enqueues launches, reading the clock back once per chunk of steps.
#pragma once

#include <chrono>
#include <cmath>
#include <algorithm>
#include <cstring>
#include <type_traits>
#include <vector>

#include "vg_runtime.h"
#include "vg_sweep.h"

namespace vg {

// Clock step with a prescribed time step: dts[it] instead of the CFL formula
// (the forward of a frozen-dt derivative check, and recomputation in the
// adjoint).  Same loop test and t / it bookkeeping as clock_begin_kernel.
static __global__ void clock_fixed_kernel(Clock *c, const double *dts, int n_dts)
{
    if (c->done) return;
    if (c->pending) {
        c->t += c->dt;
        c->it += 1;
        c->pending = 0;
    }
    if (c->it >= n_dts || (c->max_steps > 0 && c->it >= c->max_steps)) {
        c->done = 1;
        return;
    }
    const double dt = dts[c->it];
    c->dt = dt;
    if (c->steps_this_run == 0 || dt < c->dt_min) c->dt_min = dt;
    if (c->steps_this_run == 0 || dt > c->dt_max) c->dt_max = dt;
    c->steps_this_run += 1;
    c->pending = 1;
}

// Optional per-step source stage (the HPR relaxation and MHD GLM damping
// sub-cycles): launched `subcycles` times before and after the sweeps of every
// second-order step, each launch reading one buffer and writing the other.  The
// first-order solvers never sub-cycle.  The default has none.
template <class P>
struct SourceStage {
    static constexpr bool kActive = false;
    template <class GS>
    static int launch(GS &, const double *, double *, int, int, double, double, int, const typename P::Params &)
    {
        return VG_OK;
    }
    // Blocks of one launch (the CFL partials it writes when it is the last of a step).
    template <class GS>
    static long long blocks(const GS &)
    {
        return 0;
    }
};

// Optional per-step stage after the sweeps of a second-order step, in place on
// the step's result (the two-material models' volume-fraction
// reinitialisation).  When active the step's last sweep does not fold in the
// CFL maximum; the grid recomputes it after this stage.  `prepare` runs before a
// run (outside any graph capture) and may allocate what `launch` needs in
// GridState::post_buf.  The default has none.
template <class P>
struct PostStep {
    static bool active(const typename P::Params &) { return false; }
    template <class GS>
    static int prepare(GS &, const typename P::Params &)
    {
        return VG_OK;
    }
    template <class GS>
    static int launch(GS &, double *, const typename P::Params &)
    {
        return VG_OK;
    }
};

// Whether the planes a narrower physics drops hold their constant bit
// patterns in every cell: any mismatch clears *ok (every writer stores 0).
template <class R>
__global__ __launch_bounds__(256) void constant_planes_kernel(const double *__restrict__ buf, long long plane,
                                                              long long cells, int *ok)
{
    for (long long c = (long long)blockIdx.x * 256 + threadIdx.x; c < cells; c += (long long)gridDim.x * 256) {
        bool same = true;
#pragma unroll
        for (int i = 0; i < R::kConst; i++)
            same = same && (__double_as_longlong(buf[R::plane(i) * plane + c]) == (long long)R::bits(i));
        if (!same) *ok = 0;
    }
}

// ---- y-sweep shapes and their autotuner --------------------------------------

// Candidate (threads per block, rows per strip, columns per thread) of the y
// sweep: variant 0 is Tiles<P>'s shape; 1-3 trade strip length for more
// workgroups on small grids; 4-5 give each thread two columns (two independent
// dependency chains per wave); 6-7 are longer strips, whose wave counts on
// 4096-row grids (3072 and 6144 per 4096 columns) fill a device running three
// waves per SIMD in whole rounds.
constexpr int kYVariants = 8;

template <class P, int ORDER, int MODE, int BS, int H, int CPT = 1>
int launch_y_shape(vg_ctx *ctx, const SweepArgs<P> &a)
{
    dim3 grid((a.nx + BS * CPT - 1) / (BS * CPT), (a.ny + H - 1) / H);
    return ctx->timed(ORDER == 2 ? "sweep_y_slic" : "sweep_y_force",
                      [&] { sweep_y_kernel<P, ORDER, MODE, BS, H, CPT><<<grid, BS, 0, ctx->stream>>>(a); });
}

template <class P, int ORDER, int MODE>
int launch_y(vg_ctx *ctx, int variant, const SweepArgs<P> &a)
{
    if constexpr (Tiles<P>::tune) {
        switch (variant) {
        case 1: return launch_y_shape<P, ORDER, MODE, 256, 16>(ctx, a);
        case 2: return launch_y_shape<P, ORDER, MODE, 128, 32>(ctx, a);
        case 3: return launch_y_shape<P, ORDER, MODE, 64, 64>(ctx, a);
        case 4: return launch_y_shape<P, ORDER, MODE, 64, 32, 2>(ctx, a);
        case 5: return launch_y_shape<P, ORDER, MODE, 64, 64, 2>(ctx, a);
        case 6: return launch_y_shape<P, ORDER, MODE, 256, 86>(ctx, a);
        case 7: return launch_y_shape<P, ORDER, MODE, 256, 43>(ctx, a);
        default: break;
        }
    }
    return launch_y_shape<P, ORDER, MODE, Tiles<P>::bsy, Tiles<P>::hy>(ctx, a);
}

inline std::string tune_key(const char *phys, int order, int mode, int ny, int nx)
{
    return std::string(phys) + "/" + std::to_string(order) + "/" + std::to_string(mode) + "/" + std::to_string(ny) +
           "/" + std::to_string(nx);
}

// Picks the y-sweep shape for a grid: the context's table when the shape was
// tuned before; otherwise each candidate runs once to warm up and three times
// timed on the caller's grid (a.in -> a.out, with a.clk a scratch clock holding
// a small positive dt, no CFL fusion), and the fastest is recorded.
template <class P, int ORDER, int MODE>
int tune_y(vg_ctx *ctx, const std::string &key, SweepArgs<P> a, int &variant)
{
    auto it = ctx->tuned.find(key);
    if (it != ctx->tuned.end()) {
        variant = it->second;
        return VG_OK;
    }
    variant = 0;
    if (!Tiles<P>::tune || !ctx->autotune || ctx->timing) return VG_OK;
    a.fuse_max = 0;
    hipEvent_t e0, e1;
    VG_HIP(hipEventCreate(&e0));
    VG_HIP(hipEventCreate(&e1));
    float best = 0.0f;
    for (int v = 0; v < kYVariants; v++) {
        int rc = launch_y<P, ORDER, MODE>(ctx, v, a);
        VG_HIP(hipEventRecord(e0, ctx->stream));
        for (int r = 0; r < 3 && !rc; r++) rc = launch_y<P, ORDER, MODE>(ctx, v, a);
        VG_HIP(hipEventRecord(e1, ctx->stream));
        VG_HIP(hipEventSynchronize(e1));
        if (rc) return launch_status("autotune");
        float ms = 0.0f;
        VG_HIP(hipEventElapsedTime(&ms, e0, e1));
        if (v == 0 || ms < best) {
            best = ms;
            variant = v;
        }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    ctx->tuned[key] = variant;
    return VG_OK;
}

template <class P>
struct GridState {
    static constexpr int NF = P::NF;
    static constexpr int NS = StoredFields<P>::n;   // planes per buffer (stored fields per cell)
    static constexpr int kBSX = Tiles<P>::bsx, kRX = Tiles<P>::rx, kBSY = Tiles<P>::bsy, kHY = Tiles<P>::hy;
    vg_ctx *ctx = nullptr;
    int dim = 2, ny = 1, nx = 1;
    long long plane = 0, cells = 0;
    double *buf[2] = {nullptr, nullptr};
    int cur = 0;
    Clock *clk = nullptr;          // device
    double *staging = nullptr;     // AoS staging for host transfers (lazy)
    double *dt_dev = nullptr;      // device dt log (lazy)
    int dt_dev_cap = 0;
    unsigned long long *scan = nullptr;  // 2 counters for the non-finite scan
    unsigned long long *partials = nullptr;  // per-block wave-speed keys
    int n_partials = 0;                      // slots written by the last producer
    bool max_valid = false;
    int subcycles = 0;             // source sub-cycles per half step (SourceStage)
    typename P::Params last_params{};
    int last_bc = -1;
    int y_variant = 0;             // y-sweep shape (autotuned per grid shape)
    bool reduced = false;          // this run sweeps with ReducedPhys<P> (constant planes checked)
    bool planar_known = false;     // both buffers hold the constant planes (checked since the last load)
    int *flag = nullptr;           // device flag of the constant-plane check (lazy)
    Clock *tune_clk = nullptr;     // scratch clock for tuning launches (lazy)
    Clock *sweep_clk = nullptr;    // scratch clock of the per-sweep entry points (lazy)
    void *post_buf = nullptr;      // scratch of the PostStep stage (lazy)
    hipGraphExec_t graph = nullptr;  // captured chunk of steps
    std::string graph_key;           // arguments the graph was captured with
    static constexpr int kGraphSteps = 8;   // steps per captured chunk (even: buffers return to their start)

    int create(vg_ctx *c, int d, int y, int x)
    {
        ctx = c;
        dim = d;
        ny = y;
        nx = x;
        cells = (long long)ny * nx;
        plane = (cells + 63) / 64 * 64;
        VG_HIP(hipMalloc(&buf[0], sizeof(double) * NS * plane));
        VG_HIP(hipMalloc(&buf[1], sizeof(double) * NS * plane));
        VG_HIP(hipMalloc(&clk, sizeof(Clock)));
        VG_HIP(hipMalloc(&scan, 2 * sizeof(unsigned long long)));
        const long long gx = (nx + kBSX - 4 - 1) / (kBSX - 4) + 1;
        // one per block of the largest producer: x sweep tiles, 64-cell wave-speed
        // blocks, or the source stage's tiles of 32 x 4 cells
        const long long tiles = (long long)((nx + 31) / 32) * ((ny + 3) / 4);
        const long long slots = std::max<long long>(std::max<long long>(gx * ny, cells / 64 + 1), std::max(tiles, 4096LL));
        VG_HIP(hipMalloc(&partials, sizeof(unsigned long long) * slots));
        // stream-ordered copies only: a synchronous one would join the legacy
        // stream, which HIP refuses while another thread captures a graph
        Clock h{};
        VG_HIP(hipMemcpyAsync(clk, &h, sizeof(Clock), hipMemcpyHostToDevice, ctx->stream));
        VG_HIP(hipStreamSynchronize(ctx->stream));
        return VG_OK;
    }

    void destroy()
    {
        hipFree(buf[0]);
        hipFree(buf[1]);
        hipFree(clk);
        hipFree(staging);
        hipFree(dt_dev);
        hipFree(scan);
        hipFree(partials);
        hipFree(tune_clk);
        hipFree(sweep_clk);
        hipFree(flag);
        hipFree(post_buf);
        if (graph) hipGraphExecDestroy(graph);
    }

    // Takes over the device storage of an empty-or-not grid of another physics
    // with the same fields (a variant of P's equations): buffers, clock and
    // scratch move here and `o` is left empty; what depends on the physics (the
    // captured graph, the tuned y shape, the cached wave-speed maximum) starts
    // afresh.  *this must hold no storage.
    template <class Q>
    void adopt(GridState<Q> &o)
    {
        static_assert(Q::NF == NF, "adopt needs the same state layout");
        ctx = o.ctx;
        dim = o.dim;
        ny = o.ny;
        nx = o.nx;
        plane = o.plane;
        cells = o.cells;
        buf[0] = o.buf[0];
        buf[1] = o.buf[1];
        cur = o.cur;
        clk = o.clk;
        staging = o.staging;
        dt_dev = o.dt_dev;
        dt_dev_cap = o.dt_dev_cap;
        scan = o.scan;
        partials = o.partials;
        n_partials = 0;
        max_valid = false;
        subcycles = o.subcycles;
        last_params = typename P::Params{};
        last_bc = -1;
        y_variant = 0;
        reduced = false;
        planar_known = false;
        flag = o.flag;
        tune_clk = o.tune_clk;
        sweep_clk = o.sweep_clk;
        post_buf = o.post_buf;
        graph = nullptr;
        graph_key.clear();
        if (o.graph) hipGraphExecDestroy(o.graph);
        o = GridState<Q>{};
    }

    int ensure_staging()
    {
        if (!staging) VG_HIP(hipMalloc(&staging, sizeof(double) * NS * cells));
        return VG_OK;
    }

    static int blocks_for(long long n) { return (int)((n + 255) / 256 > 4096 ? 4096 : (n + 255) / 256); }

    int load(const double *prim, int on_device)
    {
        const double *src = prim;
        if (!on_device) {
            int rc = ensure_staging();
            if (rc) return rc;
            VG_HIP(hipMemcpyAsync(staging, prim, sizeof(double) * NS * cells, hipMemcpyHostToDevice, ctx->stream));
            src = staging;
        }
        aos_to_soa_kernel<NS><<<blocks_for(cells), 256, 0, ctx->stream>>>(src, buf[0], plane, cells);
        int rc = launch_status("aos_to_soa");
        if (rc) return rc;
        cur = 0;
        planar_known = false;
        Clock h{};
        VG_HIP(hipMemcpyAsync(clk, &h, sizeof(Clock), hipMemcpyHostToDevice, ctx->stream));
        VG_HIP(hipStreamSynchronize(ctx->stream));
        max_valid = false;
        return VG_OK;
    }

    int store(double *prim, int on_device)
    {
        double *dst = prim;
        if (!on_device) {
            int rc = ensure_staging();
            if (rc) return rc;
            dst = staging;
        }
        soa_to_aos_kernel<NS><<<blocks_for(cells), 256, 0, ctx->stream>>>(buf[cur], dst, plane, cells);
        int rc = launch_status("soa_to_aos");
        if (rc) return rc;
        if (!on_device)
            VG_HIP(hipMemcpyAsync(prim, staging, sizeof(double) * NS * cells, hipMemcpyDeviceToHost, ctx->stream));
        VG_HIP(hipStreamSynchronize(ctx->stream));
        return VG_OK;
    }

    // Blocks of the launch that ends a step and folds in the next step's CFL
    // maximum: the last source sub-cycle, else the last x sweep.
    int fused_count(int order) const
    {
        if (SourceStage<P>::kActive && subcycles > 0 && order == 2) return (int)SourceStage<P>::blocks(*this);
        const int G = (order == 2) ? 2 : 1;
        return (int)(((nx + (kBSX - 2 * G) - 1) / (kBSX - 2 * G)) * (long long)((ny + kRX - 1) / kRX));
    }

    // Wave-speed maximum of the current grid into clk->smax (and the corner
    // term), written as `blocks` per-block partials: the same count the step's
    // fused producer writes, so that every clock_begin of a run -- and of a
    // replayed graph -- reduces the same number of slots.
    int refresh_max(int bc, const typename P::Params &m, int blocks)
    {
        if (blocks < 1) blocks = 1;
        wave_max_kernel<P><<<blocks, 256, 0, ctx->stream>>>(buf[cur], plane, cells, dim, m, partials);
        int rc = launch_status("wave_max");
        if (rc) return rc;
        n_partials = blocks;
        reduce_max_kernel<<<1, kReduceThreads, 0, ctx->stream>>>(partials, n_partials, clk);
        rc = launch_status("reduce_max");
        if (rc) return rc;
        if (dim == 2 && bc == VG_BC_TRANSMISSIVE) {
            corner_speed_kernel<P><<<1, 1, 0, ctx->stream>>>(m, dim, clk);
            rc = launch_status("corner_speed");
            if (rc) return rc;
        } else {
            VG_HIP(hipMemsetAsync(&clk->corner, 0, sizeof(double), ctx->stream));
        }
        max_valid = true;
        last_params = m;
        last_bc = bc;
        return VG_OK;
    }

    // One sweep with physics Q (P, or its reduced physics on the same planes).
    template <class Q, int ORDER, int MODE>
    int launch_sweep_as(int dir, const double *in, double *out, double dt_scale, int fuse, int bc, double dx,
                        double bias, int limiter, const typename P::Params &m, Clock *use_clk = nullptr,
                        int fixed_params = 0)
    {
        SweepArgs<Q> a;
        a.in = in;
        a.out = out;
        a.plane = plane;
        a.ny = ny;
        a.nx = nx;
        a.bc = bc;
        a.limiter = limiter;
        a.dx = dx;
        a.bias = bias;
        a.dt_scale = dt_scale;
        a.dim = dim;
        a.fuse_max = fuse;
        a.m = m;
        a.clk = use_clk ? use_clk : clk;
        a.partials = partials;
        a.fixed_params = fixed_params;
        if (dir == 0) {
            constexpr int G = (ORDER == 2) ? 2 : 1;
            dim3 grid((nx + (kBSX - 2 * G) - 1) / (kBSX - 2 * G), (ny + kRX - 1) / kRX);
            if (fuse) n_partials = (int)(grid.x * grid.y);
            return ctx->timed(ORDER == 2 ? "sweep_x_slic" : "sweep_x_force", [&] {
                sweep_x_kernel<Q, ORDER, MODE, kBSX, kRX><<<grid, kBSX, 0, ctx->stream>>>(a);
            });
        }
        // The y sweep never ends a step, so it never writes the CFL partials.
        a.fuse_max = 0;
        return launch_y<Q, ORDER, MODE>(ctx, y_variant, a);
    }

    template <int ORDER, int MODE>
    int launch_sweep(int dir, const double *in, double *out, double dt_scale, int fuse, int bc, double dx,
                     double bias, int limiter, const typename P::Params &m, Clock *use_clk = nullptr,
                     int fixed_params = 0)
    {
        using R = typename ReducedPhys<P>::type;
        if constexpr (!std::is_void_v<R>) {
            if (reduced)
                return launch_sweep_as<R, ORDER, MODE>(dir, in, out, dt_scale, fuse, bc, dx, bias, limiter, m, use_clk,
                                                       fixed_params);
        }
        return launch_sweep_as<P, ORDER, MODE>(dir, in, out, dt_scale, fuse, bc, dx, bias, limiter, m, use_clk,
                                               fixed_params);
    }

    // One sweep of the current grid in direction dir with the caller's time step
    // dt and parameters as given (the per-sweep entry points: the reference's
    // computeSLICTimeStep / computeX|YSLICTimeStep2D and FORCE twins).  The
    // grid's own clock is left alone: dt goes into a scratch clock.
    template <int ORDER, int MODE>
    int sweep_once(int dir, double dt, int bc, double dx, double bias, int limiter, const typename P::Params &m)
    {
        if (!sweep_clk) VG_HIP(hipMalloc(&sweep_clk, sizeof(Clock)));
        Clock h{};
        h.dt = dt;
        VG_HIP(hipMemcpyAsync(sweep_clk, &h, sizeof(Clock), hipMemcpyHostToDevice, ctx->stream));
        int rc = choose_physics();
        if (rc) return rc;
        rc = launch_sweep<ORDER, MODE>(dir, buf[cur], buf[1 - cur], 1.0, 0, bc, dx, bias, limiter, m, sweep_clk, 1);
        if (rc) return launch_status("sweep");
        cur = 1 - cur;
        max_valid = false;
        VG_HIP(hipStreamSynchronize(ctx->stream));
        return VG_OK;
    }
    int sweep_once(int order, int mode, int dir, double dt, int bc, double dx, double bias, int limiter,
                   const typename P::Params &m)
    {
        if (order == 2)
            return mode == VG_MODE_EXACT ? sweep_once<2, 0>(dir, dt, bc, dx, bias, limiter, m)
                                         : sweep_once<2, 1>(dir, dt, bc, dx, bias, limiter, m);
        return mode == VG_MODE_EXACT ? sweep_once<1, 0>(dir, dt, bc, dx, bias, limiter, m)
                                     : sweep_once<1, 1>(dir, dt, bc, dx, bias, limiter, m);
    }

    // Decides this run's physics: the reduced one when the context allows it and
    // every cell of the current buffer holds the constant planes; the other
    // buffer then receives those planes, so the sweeps may leave them alone.
    int choose_physics()
    {
        reduced = false;
        using R = typename ReducedPhys<P>::type;
        if constexpr (!std::is_void_v<R>) {
            if (!ctx->reduce_fields) return VG_OK;
            // Once checked, the planes stay constant until the grid is loaded
            // again or its planes are handed out: the reduced sweeps leave them
            // alone and the full sweeps keep them at the same bits.
            if (planar_known) {
                reduced = true;
                return VG_OK;
            }
            if (!flag) VG_HIP(hipMalloc(&flag, sizeof(int)));
            const int one = 1;
            VG_HIP(hipMemcpyAsync(flag, &one, sizeof(int), hipMemcpyHostToDevice, ctx->stream));
            constant_planes_kernel<ReducedPhys<P>><<<blocks_for(cells), 256, 0, ctx->stream>>>(buf[cur], plane, cells,
                                                                                            flag);
            int rc = launch_status("constant_planes");
            if (rc) return rc;
            int ok = 0;
            VG_HIP(hipMemcpyAsync(&ok, flag, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
            VG_HIP(hipStreamSynchronize(ctx->stream));
            if (!ok) return VG_OK;
            for (int i = 0; i < ReducedPhys<P>::kConst; i++) {
                const long long off = (long long)ReducedPhys<P>::plane(i) * plane;
                VG_HIP(hipMemcpyAsync(buf[1 - cur] + off, buf[cur] + off, sizeof(double) * cells,
                                      hipMemcpyDeviceToDevice, ctx->stream));
            }
            reduced = true;
            planar_known = true;
        }
        return VG_OK;
    }

    template <int ORDER, int MODE>
    int enqueue_step(int bc, double dx, double bias, int limiter, const typename P::Params &m, double *dt_log, int dt_cap)
    {
        clock_begin_kernel<<<1, kReduceThreads, 0, ctx->stream>>>(clk, partials, n_partials, dt_log, dt_cap);
        int rc = launch_status("clock_begin");
        if (rc) return rc;
        double *a = buf[cur], *b = buf[1 - cur];
        // Source half step before the sweeps (hprsecondordersolver.cpp:326-330, 514-520).
        const int nsub = (SourceStage<P>::kActive && ORDER == 2) ? subcycles : 0;
        for (int s = 0; s < nsub; s++) {
            rc = SourceStage<P>::launch(*this, a, b, 0, bc, dx, bias, limiter, m);
            if (rc) return rc;
            std::swap(a, b);
        }
        const bool post = ORDER == 2 && PostStep<P>::active(m);
        const int fuse = (nsub == 0 && !post) ? 1 : 0;
        if (dim == 1) {
            rc = launch_sweep<ORDER, MODE>(0, a, b, 1.0, fuse, bc, dx, bias, limiter, m);
        } else {
            rc = launch_sweep<ORDER, MODE>(0, a, b, 0.5, 0, bc, dx, bias, limiter, m);
            if (!rc) rc = launch_sweep<ORDER, MODE>(1, b, a, 1.0, 0, bc, dx, bias, limiter, m);
            if (!rc) rc = launch_sweep<ORDER, MODE>(0, a, b, 0.5, fuse, bc, dx, bias, limiter, m);
        }
        if (rc) return launch_status("sweep");
        std::swap(a, b);
        // ... and after them (:337-341, 531-537); the last launch folds in the CFL maximum.
        for (int s = 0; s < nsub; s++) {
            rc = SourceStage<P>::launch(*this, a, b, s == nsub - 1, bc, dx, bias, limiter, m);
            if (rc) return rc;
            std::swap(a, b);
        }
        if (post) {
            // the stage, then the next step's CFL maximum of its result, written as
            // the same number of partials a fused sweep would write
            rc = PostStep<P>::launch(*this, a, m);
            if (rc) return rc;
            wave_max_kernel<P><<<fused_count(ORDER), 256, 0, ctx->stream>>>(a, plane, cells, dim, m, partials);
            rc = launch_status("wave_max");
            if (rc) return rc;
        }
        // Every step moves the grid an odd number of times (1 + 2 nsub).
        cur = (a == buf[0]) ? 0 : 1;
        return VG_OK;
    }

    // The y-sweep shape for physics Q on this grid (tune_y on the current buffer).
    template <class Q, int ORDER, int MODE>
    int tune_as(int bc, double dx, double bias, int limiter, const typename P::Params &m)
    {
        SweepArgs<Q> ta;
        ta.in = buf[cur];
        ta.out = buf[1 - cur];
        ta.plane = plane;
        ta.ny = ny;
        ta.nx = nx;
        ta.bc = bc;
        ta.limiter = limiter;
        ta.dx = dx;
        ta.bias = bias;
        ta.dt_scale = 1.0;
        ta.dim = dim;
        ta.fuse_max = 0;
        ta.m = m;
        ta.clk = tune_clk;
        ta.partials = nullptr;
        return tune_y<Q, ORDER, MODE>(ctx, tune_key(Q::name, ORDER, MODE, ny, nx), ta, y_variant);
    }

    template <int ORDER, int MODE>
    int run_impl(double dx, double cfl, double t_final, int max_steps, double bias, int limiter, int bc,
                 const typename P::Params &m, double *dt_log, int dt_cap, vg_run_stats *stats)
    {
        auto w0 = std::chrono::steady_clock::now();
        {
            int rc = choose_physics();
            if (rc) return rc;
        }
        const int fused = fused_count(ORDER);
        if (!max_valid || last_bc != bc || std::memcmp(&last_params, &m, sizeof(m)) != 0 || n_partials != fused) {
            int rc = refresh_max(bc, m, fused);
            if (rc) return rc;
        }
        if (ORDER == 2 && PostStep<P>::active(m)) {
            int rc = PostStep<P>::prepare(*this, m);
            if (rc) return rc;
        }
        if (dim == 2) {
            if (!tune_clk) VG_HIP(hipMalloc(&tune_clk, sizeof(Clock)));
            Clock tc{};
            tc.dt = 1e-3 * dx;
            VG_HIP(hipMemcpyAsync(tune_clk, &tc, sizeof(Clock), hipMemcpyHostToDevice, ctx->stream));
            using R = typename ReducedPhys<P>::type;
            int rc = VG_OK;
            if constexpr (!std::is_void_v<R>) {
                if (reduced) rc = tune_as<R, ORDER, MODE>(bc, dx, bias, limiter, m);
                else rc = tune_as<P, ORDER, MODE>(bc, dx, bias, limiter, m);
            } else {
                rc = tune_as<P, ORDER, MODE>(bc, dx, bias, limiter, m);
            }
            if (rc) return rc;
        }
        Clock h;
        VG_HIP(hipMemcpyAsync(&h, clk, sizeof(Clock), hipMemcpyDeviceToHost, ctx->stream));
        VG_HIP(hipStreamSynchronize(ctx->stream));
        const int start_it = h.it;
        h.t_final = t_final;
        h.cfl = cfl;
        h.dx = dx;
        h.max_steps = (max_steps > 0) ? start_it + max_steps : 0;
        h.done = 0;
        h.pending = 0;
        h.steps_this_run = 0;
        h.dt_min = h.dt_max = 0.0;
        VG_HIP(hipMemcpyAsync(clk, &h, sizeof(Clock), hipMemcpyHostToDevice, ctx->stream));
        if (dt_log && dt_cap > 0 && dt_dev_cap < dt_cap) {
            hipFree(dt_dev);
            dt_dev = nullptr;
            VG_HIP(hipMalloc(&dt_dev, sizeof(double) * dt_cap));
            dt_dev_cap = dt_cap;
        }
        double *dlog = (dt_log && dt_cap > 0) ? dt_dev : nullptr;

        hipEvent_t e0, e1;
        VG_HIP(hipEventCreate(&e0));
        VG_HIP(hipEventCreate(&e1));
        VG_HIP(hipEventRecord(e0, ctx->stream));
        const int start_cur = cur;
        const int chunk = (max_steps > 0 && max_steps <= 4096) ? max_steps : 256;
        if (ctx->graphs && !ctx->timing) {
            // Capture kGraphSteps steps once (all kernels; no memory operations),
            // replay until the clock stops.  kGraphSteps is even, so every replay
            // starts from the same buffer; steps past the stop do nothing.
            std::string key(reinterpret_cast<const char *>(&m), sizeof(m));
            const double num[2] = {dx, bias};
            key.append(reinterpret_cast<const char *>(num), sizeof(num));
            key += "|" + std::to_string(ORDER) + "|" + std::to_string(MODE) + "|" + std::to_string(bc) + "|" +
                   std::to_string(limiter) + "|" + std::to_string((unsigned long long)(uintptr_t)dlog) + "|" +
                   std::to_string(dt_cap) + "|" + std::to_string(subcycles) + "|" + std::to_string(y_variant) + "|" +
                   std::to_string(cur) + "|" + std::to_string(n_partials) + "|" + std::to_string((int)reduced);
            if (!graph || key != graph_key) {
                if (graph) {
                    hipGraphExecDestroy(graph);
                    graph = nullptr;
                }
                hipGraph_t g = nullptr;
                VG_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
                int rc = VG_OK;
                for (int s = 0; s < kGraphSteps && !rc; s++) rc = enqueue_step<ORDER, MODE>(bc, dx, bias, limiter, m, dlog, dt_cap);
                hipError_t ec = hipStreamEndCapture(ctx->stream, &g);
                if (rc) {
                    if (g) hipGraphDestroy(g);
                    return rc;
                }
                if (ec != hipSuccess) {
                    set_error(std::string("hipStreamEndCapture: ") + hipGetErrorString(ec));
                    return VG_ERR_HIP;
                }
                ec = hipGraphInstantiate(&graph, g, nullptr, nullptr, 0);
                hipGraphDestroy(g);
                if (ec != hipSuccess) {
                    graph = nullptr;
                    set_error(std::string("hipGraphInstantiate: ") + hipGetErrorString(ec));
                    return VG_ERR_HIP;
                }
                graph_key = key;
            }
            const int reps = std::max(1, chunk / kGraphSteps);
            for (;;) {
                for (int r = 0; r < reps; r++) VG_HIP(hipGraphLaunch(graph, ctx->stream));
                VG_HIP(hipMemcpyAsync(&h, clk, sizeof(Clock), hipMemcpyDeviceToHost, ctx->stream));
                VG_HIP(hipStreamSynchronize(ctx->stream));
                if (h.done) break;
                if (max_steps > 0 && h.steps_this_run >= max_steps) break;
            }
        } else {
            for (;;) {
                for (int s = 0; s < chunk; s++) {
                    int rc = enqueue_step<ORDER, MODE>(bc, dx, bias, limiter, m, dlog, dt_cap);
                    if (rc) return rc;
                }
                VG_HIP(hipMemcpyAsync(&h, clk, sizeof(Clock), hipMemcpyDeviceToHost, ctx->stream));
                VG_HIP(hipStreamSynchronize(ctx->stream));
                if (h.done) break;
                if (max_steps > 0 && h.steps_this_run >= max_steps) break;
            }
        }
        clock_finish_kernel<<<1, 1, 0, ctx->stream>>>(clk);
        int rc = launch_status("clock_finish");
        if (rc) return rc;
        VG_HIP(hipEventRecord(e1, ctx->stream));
        VG_HIP(hipMemcpyAsync(&h, clk, sizeof(Clock), hipMemcpyDeviceToHost, ctx->stream));
        VG_HIP(hipStreamSynchronize(ctx->stream));
        float dev_ms = 0.0f;
        VG_HIP(hipEventElapsedTime(&dev_ms, e0, e1));
        hipEventDestroy(e0);
        hipEventDestroy(e1);
        // Each executed step moved the grid to the other buffer; steps enqueued
        // after the loop condition failed did nothing.
        cur = start_cur ^ (h.steps_this_run & 1);
        if (dlog) {
            const int n = h.steps_this_run < dt_cap ? h.steps_this_run : dt_cap;
            if (n > 0) {
                VG_HIP(hipMemcpyAsync(dt_log, dt_dev, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream));
                VG_HIP(hipStreamSynchronize(ctx->stream));
            }
        }
        if (ctx->timing) {
            rc = ctx->resolve_timing();
            if (rc) return rc;
        }
        if (stats) {
            stats->steps = h.steps_this_run;
            stats->total_steps = h.it;
            stats->t = h.t;
            stats->dt_min = h.dt_min;
            stats->dt_max = h.dt_max;
            stats->device_ms = dev_ms;
            stats->wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - w0).count();
            stats->nonfinite_cells = 0;
            stats->first_nonfinite = -1;
        }
        return VG_OK;
    }

    int run(int order, int mode, double dx, double cfl, double t_final, int max_steps, double bias, int limiter,
            int bc, const typename P::Params &m, double *dt_log, int dt_cap, vg_run_stats *stats)
    {
        if (order == 2 && mode == 0) return run_impl<2, 0>(dx, cfl, t_final, max_steps, bias, limiter, bc, m, dt_log, dt_cap, stats);
        if (order == 2 && mode == 1) return run_impl<2, 1>(dx, cfl, t_final, max_steps, bias, limiter, bc, m, dt_log, dt_cap, stats);
        if (order == 1 && mode == 0) return run_impl<1, 0>(dx, cfl, t_final, max_steps, bias, limiter, bc, m, dt_log, dt_cap, stats);
        if (order == 1 && mode == 1) return run_impl<1, 1>(dx, cfl, t_final, max_steps, bias, limiter, bc, m, dt_log, dt_cap, stats);
        set_error("run: order must be 1 or 2 and mode 0 or 1");
        return VG_ERR_ARG;
    }

    // Scan of the current grid for non-finite or inadmissible cells.
    int scan_nonfinite(const typename P::Params &m, int64_t *count, int64_t *first)
    {
        unsigned long long init[2] = {0ull, ~0ull};
        VG_HIP(hipMemcpyAsync(scan, init, sizeof(init), hipMemcpyHostToDevice, ctx->stream));
        nonfinite_kernel<P><<<blocks_for(cells), 256, 0, ctx->stream>>>(buf[cur], plane, cells, m, scan, scan + 1);
        int rc = launch_status("nonfinite");
        if (rc) return rc;
        unsigned long long res[2];
        VG_HIP(hipMemcpyAsync(res, scan, sizeof(res), hipMemcpyDeviceToHost, ctx->stream));
        VG_HIP(hipStreamSynchronize(ctx->stream));
        *count = (int64_t)res[0];
        *first = res[0] ? (int64_t)res[1] : -1;
        return VG_OK;
    }
};


// The reference's per-sweep calls on a grid the caller holds on the host with
// its ghost cells (computeSLICTimeStep(currentCells, currentCellsWithBoundary,
// ...), secondordersolver.cpp:80-98, 140-162, 214-236; computeFORCETimeStep and
// computeX|YFORCETimeStep2D, firstordersolver.cpp:232-378): `with_boundary` is
// the ghosted snapshot as given ((n + 2G) cells in 1D, (ny + 2G) x (nx + 2G) in
// 2D, G = 2 for SLIC, 1 for FORCE), `cells` receives the ny x nx updated
// interior.  The whole ghosted array is swept on the device (its ghost rows and
// columns with index-mapped neighbours, then discarded); every interior cell
// reads exactly the ghosted neighbours the reference's loop reads.
template <class P>
int ghosted_time_step(vg_ctx *ctx, int dim, double *cells, const double *with_boundary, int ny, int nx, int dir,
                      double dx, double dt, double bias, int limiter, int order, int mode,
                      const typename P::Params &m)
{
    constexpr int NF = StoredFields<P>::n;   // fields per cell of the caller's arrays
    const int G = (order == 2) ? 2 : 1;
    const int gy = (dim == 1) ? 1 : ny + 2 * G, gx = nx + 2 * G;
    GridState<P> g;
    int rc = g.create(ctx, dim, gy, gx);
    if (!rc) rc = g.load(with_boundary, 0);
    if (!rc) rc = g.sweep_once(order, mode, dir, dt, VG_BC_TRANSMISSIVE, dx, bias, limiter, m);
    std::vector<double> out;
    if (!rc) {
        out.resize((size_t)gy * gx * NF);
        rc = g.store(out.data(), 0);
    }
    g.destroy();
    if (rc) return rc;
    const int oy = (dim == 1) ? 0 : G;
    for (int i = 0; i < ny; i++)
        std::memcpy(cells + (size_t)i * nx * NF, out.data() + ((size_t)(i + oy) * gx + G) * NF, sizeof(double) * nx * NF);
    return VG_OK;
}

// Argument checks shared by the per-sweep entry points.
inline int check_time_step_args(int dim, int ny, int nx, int dir, double dx, double dt, int order, int mode,
                                const void *cells, const void *with_boundary, const void *mp)
{
    VG_CHECK(cells != nullptr && with_boundary != nullptr && mp != nullptr, "null argument");
    VG_CHECK(order == 1 || order == 2, "order must be 1 (FORCE) or 2 (SLIC)");
    VG_CHECK(mode == VG_MODE_EXACT || mode == VG_MODE_FAST, "mode must be VG_MODE_EXACT or VG_MODE_FAST");
    VG_CHECK(dim == 1 ? (ny == 1 && dir == 0) : (ny >= 1 && (dir == 0 || dir == 1)),
             "dir must be 0 (x) or 1 (y; 2D only)");
    VG_CHECK(nx >= 1, "grid needs at least 1 cell");
    VG_CHECK((long long)(ny + 4) * (nx + 4) <= (1ll << 31), "grid too large");
    VG_CHECK(dx > 0.0 && std::isfinite(dx), "cell spacing must be positive");
    VG_CHECK(std::isfinite(dt), "time step must be finite");
    return VG_OK;
}

}  // namespace vg
