// Adjoint C ABI for the Euler SLIC solve, and the fixed-time-step forward it
// differentiates.
//
// vg_euler_adjoint_{1d,2d}[_device]: when every sweep input of a capped run
// fits in device memory, the forward writes each sweep into its own buffer and
// the adjoint sweeps read them (adjoint_recorded).  Otherwise: forward solve
// (the same kernels and arithmetic as vg_euler_solve_*, EXACT mode) that
// records every dt and keeps a checkpoint of the grid every K steps (K = 1 when a grid per step fits in free
// device memory, else ceil(sqrt(steps)), unless given); then, segment by
// segment from the end, the segment is recomputed from its checkpoint with the
// recorded dts, storing the input of every sweep (the last sweep of a segment
// is not needed), and the adjoint sweeps run backwards through it: 2D steps
// X(dt/2) Y(dt) X(dt/2) reverse as X*(dt/2) Y*(dt) X*(dt/2).  Memory: (steps / K)
// checkpoints + 3K - 1 sweep inputs (2D) of one grid each, kept in the
// context's workspace between calls.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "vg_adjoint.h"
#include "vg_adjoint_host.h"
#include "vg_grid.h"
#include "vg_phys_euler.h"

namespace {

using vg::EulerPhys;
using EGrid = vg::GridState<EulerPhys>;
using vg::adj_sweep;
using vg::adjoint_args;
using vg::kAdjDefault;
using vg::tune_adjoint;

EulerPhys::Params to_params(const vg_euler_params *p)
{
    EulerPhys::Params m;
    m.gamma = p->gamma;
    m.p_inf = p->p_inf;
    m.cv = p->cv;
    m.e_form = p->e_form;
    m.t_ign = p->t_ign;
    m.k_rate = p->k_rate;
    m.gm1 = p->gamma - 1;
    m.gpinf = p->gamma * p->p_inf;
    m.inv_gm1 = 1.0 / (p->gamma - 1);
    return m;
}

int check_common(vg_ctx *ctx, int dim, int ny, int nx, double dx, int limiter, int bc, const vg_euler_params *mp)
{
    (void)limiter;
    VG_CHECK(ctx != nullptr, "null context");
    VG_CHECK(mp != nullptr, "null material parameters");
    VG_CHECK(bc == VG_BC_TRANSMISSIVE || bc == VG_BC_PERIODIC, "unknown boundary condition");
    VG_CHECK(nx >= 2 && (dim == 1 ? ny == 1 : ny >= 2), "grid needs at least 2 cells per direction");
    VG_CHECK(dx > 0.0 && std::isfinite(dx), "cell spacing must be positive");
    return VG_OK;
}

struct DeviceBuffers {
    std::vector<double *> ptrs;
    ~DeviceBuffers()
    {
        for (double *p : ptrs) hipFree(p);
    }
    int alloc(double **p, size_t bytes)
    {
        hipError_t e = hipMalloc(p, bytes);
        if (e != hipSuccess) {
            vg::set_error(std::string("hipMalloc: ") + hipGetErrorString(e));
            return VG_ERR_HIP;
        }
        ptrs.push_back(*p);
        return VG_OK;
    }
};

// Fixed-dt forward over steps [s0, s1) through the grid slots: slot[base + k]
// is the input of sweep k of step s0 + base / per.  With skip_last the final
// sweep (whose output the adjoint never reads) is not run.
int forward_fixed(vg_ctx *ctx, EGrid &g, const double *dts_dev, int s0, int s1, double dx, double bias, int limiter,
                  int bc, const EulerPhys::Params &m, const std::vector<double *> &slots, bool skip_last = false)
{
    vg::Clock h{};
    h.it = s0;
    h.t_final = INFINITY;
    h.max_steps = s1;
    VG_HIP(hipMemcpyAsync(g.clk, &h, sizeof(h), hipMemcpyHostToDevice, ctx->stream));
    const int per = (g.dim == 2) ? 3 : 1;
    for (int s = s0; s < s1; s++) {
        vg::clock_fixed_kernel<<<1, 1, 0, ctx->stream>>>(g.clk, dts_dev, s1);
        const int base = (s - s0) * per;
        const bool last = skip_last && s == s1 - 1;
        int rc = VG_OK;
        // the grid's own sweep launch: the 4-field kernels when its last run took them
        auto sweep = [&](int dir, const double *in, double *out, double scale) {
            return g.launch_sweep<2, vg::MODE_EXACT>(dir, in, out, scale, 0, bc, dx, bias, limiter, m);
        };
        if (g.dim == 1) {
            if (!last) rc = sweep(0, slots[base], slots[base + 1], 1.0);
        } else {
            rc = sweep(0, slots[base], slots[base + 1], 0.5);
            if (!rc) rc = sweep(1, slots[base + 1], slots[base + 2], 1.0);
            if (!rc && !last) rc = sweep(0, slots[base + 2], slots[base + 3], 0.5);
        }
        if (rc) return rc;
    }
    return VG_OK;
}

// Checkpoint interval: the caller's; else every step when the run's length is
// known and one grid per step fits in free device memory (the recomputation
// then re-runs two sweeps per 2D step and none in 1D); else ceil(sqrt(steps)).
int choose_interval(vg_ctx *ctx, int checkpoint_every, int max_steps, int per, size_t gbytes)
{
    if (checkpoint_every > 0) return checkpoint_every;
    if (max_steps <= 0) return 16;
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) == hipSuccess) {
        const double avail = 0.85 * (double)(free_b + ctx->workspace_bytes());
        const double need = (double)gbytes * ((double)max_steps + 1.0 + per + 1.0 + 2.0) + 64.0 * 1024 * 1024;
        if (need < avail) return 1;
    }
    return std::max(1, (int)std::ceil(std::sqrt((double)max_steps)));
}

// The adjoint sweeps' physics for this run: the 4-field ones (AdjPlanar) when
// the forward took the 4-field sweeps, E_f = 0 and the seed (SoA, in seed_soa)
// has zero w and lambda parts; the w and lambda adjoints then stay zero, so
// planes 3 and 5 of the other adjoint buffer are cleared once and neither
// buffer's planes 3 and 5 are written again.
int choose_adjoint_physics(vg_ctx *ctx, const EGrid &g, const EulerPhys::Params &m, const double *seed_soa,
                           double *other, bool &planar)
{
    planar = false;
    if (!g.reduced || m.e_form != 0.0) return VG_OK;
    hipStream_t s = ctx->stream;
    int *flag = nullptr;
    VG_HIP(hipMalloc(&flag, sizeof(int)));
    const int one = 1;
    int ok = 0;
    hipError_t e = hipMemcpyAsync(flag, &one, sizeof(int), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) {
        vg::zero_planes_kernel<<<EGrid::blocks_for(g.cells), 256, 0, s>>>(seed_soa, g.plane, g.cells, flag);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(&ok, flag, sizeof(int), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    hipFree(flag);
    if (e != hipSuccess) {
        vg::set_error(std::string("adjoint seed check: ") + hipGetErrorString(e));
        return VG_ERR_HIP;
    }
    if (!ok) return VG_OK;
    VG_HIP(hipMemsetAsync(other + 3 * g.plane, 0, sizeof(double) * g.cells, s));
    VG_HIP(hipMemsetAsync(other + 5 * g.plane, 0, sizeof(double) * g.cells, s));
    planar = true;
    return VG_OK;
}

int adj_sweep_p(bool planar, vg_ctx *ctx, int dir, int variant, const double *w, const double *wbar_out,
                double *wbar_in, double *ghost, long long plane, int ny, int nx, int bc, double h, double bias,
                int limiter, double ts, const EulerPhys::Params &m, const double *w35 = nullptr)
{
    if (planar)
        return adj_sweep<vg::AdjPlanar>(ctx, dir, variant, w, wbar_out, wbar_in, ghost, plane, ny, nx, bc, h, bias,
                                        limiter, ts, m, w35);
    return adj_sweep<vg::AdjFull>(ctx, dir, variant, w, wbar_out, wbar_in, ghost, plane, ny, nx, bc, h, bias, limiter,
                                  ts, m, w35);
}

int tune_adjoint_p(bool planar, vg_ctx *ctx, int dir, const vg::AdjointArgs &a, int &variant)
{
    return planar ? tune_adjoint<vg::AdjPlanar>(ctx, dir, a, variant) : tune_adjoint<vg::AdjFull>(ctx, dir, a, variant);
}

// The recorded adjoint's layout for a run of max_steps steps: the number of
// steps, the last ones, that keep every sweep input (2D: their two inner sweep
// inputs too); the others keep their start only and the reverse pass
// recomputes their inner sweep inputs into two scratch buffers.  rbytes is one
// recording buffer, gbytes one full grid (the adjoint's own buffers).  -1 when
// not even the step starts fit in free device memory (or the context's
// adjoint memory limit).
int record_layout(vg_ctx *ctx, int max_steps, int dim, size_t rbytes, size_t gbytes)
{
    if (max_steps <= 0) return -1;
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return -1;
    double avail = 0.85 * (double)(free_b + ctx->workspace_bytes());
    if (ctx->adjoint_memory > 0) avail = std::min(avail, (double)ctx->adjoint_memory);
    const double left = avail - 2.0 * (double)gbytes - 64.0 * 1024 * 1024;
    const double rb = (double)rbytes;
    if (dim == 1) return (double)max_steps * rb <= left ? max_steps : -1;
    if (3.0 * max_steps * rb <= left) return max_steps;
    const double base = ((double)max_steps + 2.0) * rb;   // every start and the two scratch buffers
    if (base > left) return -1;
    return std::min(max_steps - 1, (int)((left - base) / (2.0 * rb)));
}

// The adjoint with the forward's states recorded by the forward itself: the
// CFL forward runs its steps (the solve's clock, kernels and fused wave-speed
// maximum) with its sweeps writing into recording buffers instead of the grid's
// two, so nothing is copied.  start(s) = slots[s] is the input of step s
// (slots[0] the loaded grid, slots[s + 1] its output); the last `full` steps
// (from s0 = max_steps - full on) also keep their two inner sweep inputs
// (2D) in slots of their own, the earlier ones write them into two scratch
// buffers and the reverse pass recomputes them for each step with its
// recorded dt.  When the grid takes the 4-field sweeps, their outputs leave
// planes 3 and 5 unwritten, the recording buffers hold planes 0-4 only and the
// adjoint sweeps read planes 3 and 5 (w = +0, lambda = 1) from slot 0.
int adjoint_recorded(vg_ctx *ctx, EGrid &g, int full, int dim, int ny, int nx, double dx, double cfl, double t_final,
                     int max_steps, double bias, int limiter, int bc, const EulerPhys::Params &m, const double *seed,
                     double *prim_out, double *grad, int on_device, double *dt_log, int dt_log_cap,
                     vg_run_stats *stats, std::chrono::steady_clock::time_point w0)
{
    hipStream_t s = ctx->stream;
    const size_t gbytes = sizeof(double) * 6 * g.plane;
    const bool reduced = g.reduced;   // decided by the caller's choose_physics
    const size_t rbytes = reduced ? sizeof(double) * 5 * g.plane : gbytes;
    auto wsp = [&](size_t i, size_t bytes, double **p) {
        void *v = nullptr;
        int r = ctx->workspace_buffer(i, bytes, &v);
        *p = static_cast<double *>(v);
        return r;
    };
    const int s0 = max_steps - full;
    const bool scratch = (dim == 2 && s0 > 0);
    // workspace: 0 dts, 1-2 adjoint ping-pong, 3 ghost adjoints, 4-5 scratch,
    // 6 .. the recording buffers after slot 0
    std::vector<double *> slots((size_t)max_steps + 1 + (dim == 2 ? 2 * (size_t)full : 0), nullptr);
    slots[0] = g.buf[g.cur];
    double *dts_dev = nullptr, *wb[2] = {nullptr, nullptr}, *ghost = nullptr, *T[2] = {nullptr, nullptr};
    int rc = wsp(0, sizeof(double) * max_steps, &dts_dev);
    if (!rc) rc = wsp(1, gbytes, &wb[0]);
    if (!rc) rc = wsp(2, gbytes, &wb[1]);
    if (!rc) rc = wsp(3, sizeof(double) * 4 * 6 * (size_t)std::max(nx, ny), &ghost);
    if (!rc && scratch) rc = wsp(4, rbytes, &T[0]);
    if (!rc && scratch) rc = wsp(5, rbytes, &T[1]);
    for (size_t k = 1; k < slots.size() && !rc; k++) rc = wsp(6 + k - 1, rbytes, &slots[k]);
    if (rc) return rc;
    // ---- forward: the steps of GridState::run, into the recording buffers -------------
    if ((rc = g.refresh_max(bc, m, g.fused_count(2)))) return rc;
    if (dim == 2) {
        if (!g.tune_clk) VG_HIP(hipMalloc(&g.tune_clk, sizeof(vg::Clock)));
        vg::Clock tc{};
        tc.dt = 1e-3 * dx;
        VG_HIP(hipMemcpyAsync(g.tune_clk, &tc, sizeof(vg::Clock), hipMemcpyHostToDevice, s));
        rc = reduced ? g.tune_as<vg::EulerPlanarPhys, 2, vg::MODE_EXACT>(bc, dx, bias, limiter, m)
                     : g.tune_as<EulerPhys, 2, vg::MODE_EXACT>(bc, dx, bias, limiter, m);
        if (rc) return rc;
    }
    vg::Clock h{};
    h.t_final = t_final;
    h.cfl = cfl;
    h.dx = dx;
    h.max_steps = max_steps;
    VG_HIP(hipMemcpyAsync(g.clk, &h, sizeof(vg::Clock), hipMemcpyHostToDevice, s));
    // the buffers a step reads and writes: its start, (2D) the two inner sweep
    // inputs and its end; 1D: its start and its end
    auto step_bufs = [&](int st, double *(&b)[4]) {
        b[0] = slots[st];
        if (dim == 1) {
            b[1] = slots[st + 1];
            b[2] = b[3] = nullptr;
            return;
        }
        const size_t in = (size_t)max_steps + 1 + 2 * (size_t)(st - s0);
        b[1] = st >= s0 ? slots[in] : T[0];
        b[2] = st >= s0 ? slots[in + 1] : T[1];
        b[3] = slots[st + 1];
    };
    for (int st = 0; st < max_steps && !rc; st++) {
        vg::clock_begin_kernel<<<1, vg::kReduceThreads, 0, s>>>(g.clk, g.partials, g.n_partials, dts_dev, max_steps);
        double *b[4];
        step_bufs(st, b);
        if (dim == 1) {
            rc = g.launch_sweep<2, vg::MODE_EXACT>(0, b[0], b[1], 1.0, 1, bc, dx, bias, limiter, m);
        } else {
            rc = g.launch_sweep<2, vg::MODE_EXACT>(0, b[0], b[1], 0.5, 0, bc, dx, bias, limiter, m);
            if (!rc) rc = g.launch_sweep<2, vg::MODE_EXACT>(1, b[1], b[2], 1.0, 0, bc, dx, bias, limiter, m);
            if (!rc) rc = g.launch_sweep<2, vg::MODE_EXACT>(0, b[2], b[3], 0.5, 1, bc, dx, bias, limiter, m);
        }
    }
    if (rc) return vg::launch_status("adjoint forward");
    vg::clock_finish_kernel<<<1, 1, 0, s>>>(g.clk);
    VG_HIP(hipMemcpyAsync(&h, g.clk, sizeof(vg::Clock), hipMemcpyDeviceToHost, s));
    VG_HIP(hipStreamSynchronize(s));
    if (h.stalled) {
        vg::set_error("a time step could not advance t (non-finite or zero wave speed)");
        return VG_ERR_NONFINITE;
    }
    const int steps = h.steps_this_run;
    std::vector<double> dts(std::max(steps, 1));
    if (steps > 0) VG_HIP(hipMemcpyAsync(dts.data(), dts_dev, sizeof(double) * steps, hipMemcpyDeviceToHost, s));
    // the final state; with the 4-field sweeps its four planes and slot 0's
    // constant planes go into the grid's other buffer first
    double *fin = slots[steps];
    if (reduced && steps > 0) {
        double *out = g.buf[1 - g.cur];
        const size_t pb = sizeof(double) * g.plane;
        VG_HIP(hipMemcpyAsync(out, fin, 3 * pb, hipMemcpyDeviceToDevice, s));
        VG_HIP(hipMemcpyAsync(out + 4 * g.plane, fin + 4 * g.plane, pb, hipMemcpyDeviceToDevice, s));
        for (int f = 3; f <= 5; f += 2)
            VG_HIP(hipMemcpyAsync(out + f * g.plane, slots[0] + f * g.plane, pb, hipMemcpyDeviceToDevice, s));
        fin = out;
    }
    {
        double *dst = on_device ? prim_out : g.staging;
        vg::soa_to_aos_kernel<6><<<EGrid::blocks_for(g.cells), 256, 0, s>>>(fin, dst, g.plane, g.cells);
        if ((rc = vg::launch_status("adjoint forward store"))) return rc;
        if (!on_device)
            VG_HIP(hipMemcpyAsync(prim_out, g.staging, sizeof(double) * 6 * g.cells, hipMemcpyDeviceToHost, s));
        VG_HIP(hipStreamSynchronize(s));
    }
    // ---- reverse -------------------------------------------------------------------------
    const double *seed_aos = seed;
    if (!on_device) {
        VG_HIP(hipMemcpyAsync(g.staging, seed, sizeof(double) * 6 * g.cells, hipMemcpyHostToDevice, s));
        seed_aos = g.staging;
    }
    vg::aos_to_soa_kernel<6><<<EGrid::blocks_for(g.cells), 256, 0, s>>>(seed_aos, wb[0], g.plane, g.cells);
    bool planar = false;
    if ((rc = choose_adjoint_physics(ctx, g, m, wb[0], wb[1], planar))) return rc;
    const double *w35 = reduced ? slots[0] : nullptr;
    int cw = 0;
    int xv = kAdjDefault[0], yv = kAdjDefault[1];
    if (steps > 0) {
        const vg::AdjointArgs ta =
            adjoint_args(slots[0], wb[0], wb[1], ghost, g.plane, ny, nx, bc, dx, bias, limiter, dts[0], m);
        rc = tune_adjoint_p(planar, ctx, 0, ta, xv);
        if (!rc && dim == 2) rc = tune_adjoint_p(planar, ctx, 1, ta, yv);
    }
    for (int st = steps - 1; st >= 0 && !rc; st--) {
        double *b[4];
        step_bufs(st, b);
        if (dim == 2 && st < s0) {
            // the step's two inner sweep inputs again, with its recorded dt
            const std::vector<double *> rs = {b[0], b[1], b[2], nullptr};
            rc = forward_fixed(ctx, g, dts_dev, st, st + 1, dx, bias, limiter, bc, m, rs, true);
            if (rc) break;
        }
        const double dt = dts[st];
        if (dim == 1) {
            rc = adj_sweep_p(planar, ctx, 0, xv, b[0], wb[cw], wb[1 - cw], ghost, g.plane, ny, nx, bc, dx, bias, limiter, dt, m,
                           w35);
            cw = 1 - cw;
        } else {
            rc = adj_sweep_p(planar, ctx, 0, xv, b[2], wb[cw], wb[1 - cw], ghost, g.plane, ny, nx, bc, dx, bias, limiter,
                           dt * 0.5, m, w35);
            cw = 1 - cw;
            if (!rc)
                rc = adj_sweep_p(planar, ctx, 1, yv, b[1], wb[cw], wb[1 - cw], ghost, g.plane, ny, nx, bc, dx, bias, limiter, dt,
                               m, w35);
            cw = 1 - cw;
            if (!rc)
                rc = adj_sweep_p(planar, ctx, 0, xv, b[0], wb[cw], wb[1 - cw], ghost, g.plane, ny, nx, bc, dx, bias, limiter,
                               dt * 0.5, m, w35);
            cw = 1 - cw;
        }
    }
    if (!rc) {
        double *dst = on_device ? grad : g.staging;
        vg::soa_to_aos_kernel<6><<<EGrid::blocks_for(g.cells), 256, 0, s>>>(wb[cw], dst, g.plane, g.cells);
        rc = vg::launch_status("adjoint store");
    }
    if (!rc) {
        hipError_t e = hipSuccess;
        if (!on_device) e = hipMemcpyAsync(grad, g.staging, sizeof(double) * 6 * g.cells, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) {
            vg::set_error(std::string("adjoint copy: ") + hipGetErrorString(e));
            rc = VG_ERR_HIP;
        }
    }
    if (!rc && ctx->timing) rc = ctx->resolve_timing();
    if (dt_log) std::memcpy(dt_log, dts.data(), sizeof(double) * std::min(steps, std::max(dt_log_cap, 0)));
    if (stats) {
        *stats = vg_run_stats{};
        stats->steps = steps;
        stats->total_steps = steps;
        stats->t = h.t;
        stats->dt_min = steps ? *std::min_element(dts.begin(), dts.begin() + steps) : 0.0;
        stats->dt_max = steps ? *std::max_element(dts.begin(), dts.begin() + steps) : 0.0;
        stats->wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - w0).count();
        stats->first_nonfinite = -1;
    }
    return rc;
}

int adjoint(vg_ctx *ctx, int dim, const double *prim_in, double *prim_out, const double *seed, double *grad, int ny,
            int nx, double dx, double cfl, double t_final, int max_steps, double bias, int limiter, int bc,
            const vg_euler_params *mp, int checkpoint_every, double *dt_log, int dt_log_cap, vg_run_stats *stats,
            int on_device)
{
    vg::clear_error();
    int rc = check_common(ctx, dim, ny, nx, dx, limiter, bc, mp);
    if (rc) return rc;
    VG_CHECK(prim_in && prim_out && seed && grad, "null state pointer");
    VG_CHECK(cfl > 0.0 && std::isfinite(cfl), "CFL coefficient must be positive");
    VG_CHECK(max_steps >= 0 && (max_steps > 0 || std::isfinite(t_final)), "an infinite final time needs a step cap");
    VG_CHECK(checkpoint_every >= 0, "checkpoint_every must be >= 0");
    rc = vg::bind(ctx);
    if (rc) return rc;
    auto w0 = std::chrono::steady_clock::now();
    const EulerPhys::Params m = to_params(mp);
    hipStream_t s = ctx->stream;
    EGrid g;
    rc = g.create(ctx, dim, ny, nx);
    if (!rc) rc = g.load(prim_in, on_device);
    if (!rc && !on_device) rc = g.ensure_staging();
    if (rc) {
        g.destroy();
        return rc;
    }
    const size_t gbytes = sizeof(double) * 6 * g.plane;
    const int per = (dim == 2) ? 3 : 1;
    // checkpoint_every 0: the fullest recorded layout that fits; 1: the step
    // starts only (a checkpoint every step, written by the forward itself)
    if (max_steps > 0 && checkpoint_every <= 1) {
        if ((rc = g.choose_physics())) {
            g.destroy();
            return rc;
        }
        const size_t rbytes = g.reduced ? sizeof(double) * 5 * g.plane : gbytes;
        int full = record_layout(ctx, max_steps, dim, rbytes, gbytes);
        if (checkpoint_every == 1 && full >= 0) full = (dim == 1) ? max_steps : 0;
        if (full >= 0) {
            rc = adjoint_recorded(ctx, g, full, dim, ny, nx, dx, cfl, t_final, max_steps, bias, limiter, bc, m, seed,
                                  prim_out, grad, on_device, dt_log, dt_log_cap, stats, w0);
            g.destroy();
            return rc;
        }
    }
    const int K = choose_interval(ctx, checkpoint_every, max_steps, per, gbytes);
    // Workspace buffers: 0 dts, 1-2 adjoint ping-pong, 3 ghost adjoints,
    // 4 .. 4 + per * K - 1 sweep inputs, then the checkpoints.
    const int cap = max_steps > 0 ? max_steps : 1 << 22;
    auto wsp = [&](size_t i, size_t bytes, double **p) {
        void *v = nullptr;
        int r = ctx->workspace_buffer(i, bytes, &v);
        *p = static_cast<double *>(v);
        return r;
    };
    const size_t ck0 = 4 + (size_t)per * K;
    // ---- forward with checkpoints ---------------------------------------------------
    std::vector<double> dts;
    std::vector<double *> ckpt;
    vg_run_stats fst{};
    for (;;) {
        if ((int)dts.size() >= cap) break;
        double *slot = nullptr;
        if ((rc = wsp(ck0 + ckpt.size(), gbytes, &slot))) break;
        VG_HIP(hipMemcpyAsync(slot, g.buf[g.cur], gbytes, hipMemcpyDeviceToDevice, s));
        ckpt.push_back(slot);
        std::vector<double> part(K);
        vg_run_stats st{};
        const int want = std::min(K, cap - (int)dts.size());
        rc = g.run(2, VG_MODE_EXACT, dx, cfl, t_final, want, bias, limiter, bc, m, part.data(), K, &st);
        if (rc) break;
        dts.insert(dts.end(), part.begin(), part.begin() + st.steps);
        fst = st;
        if (st.steps < want) break;   // reached t_final
    }
    if (rc) {
        g.destroy();
        return rc;
    }
    const int steps = (int)dts.size();
    rc = g.store(prim_out, on_device);
    // ---- reverse, segment by segment ----------------------------------------------------
    double *dts_dev = nullptr, *wb[2] = {nullptr, nullptr}, *ghost = nullptr;
    std::vector<double *> slots(per * K + 1, nullptr);
    if (!rc) rc = wsp(0, sizeof(double) * std::max(steps, 1), &dts_dev);
    if (!rc) rc = wsp(1, gbytes, &wb[0]);
    if (!rc) rc = wsp(2, gbytes, &wb[1]);
    if (!rc) rc = wsp(3, sizeof(double) * 4 * 6 * (size_t)std::max(nx, ny), &ghost);
    // slot 0 of a segment is its checkpoint; the last step's final sweep is not
    // recomputed, so slot per * K is never written
    for (int k = 1; k < per * K && !rc; k++) rc = wsp(4 + k - 1, gbytes, &slots[k]);
    if (rc) {
        g.destroy();
        return rc;
    }
    if (steps > 0) VG_HIP(hipMemcpyAsync(dts_dev, dts.data(), sizeof(double) * steps, hipMemcpyHostToDevice, s));
    // The 4-field sweeps leave the constant planes alone: give every recompute
    // slot those planes once (the adjoint sweeps read all six).
    if (g.reduced && !ckpt.empty()) {
        using R = vg::ReducedPhys<EulerPhys>;
        for (int k = 1; k < per * K; k++)
            for (int i = 0; i < R::kConst; i++) {
                const long long off = (long long)R::plane(i) * g.plane;
                VG_HIP(hipMemcpyAsync(slots[k] + off, ckpt[0] + off, sizeof(double) * g.cells, hipMemcpyDeviceToDevice,
                                      s));
            }
    }
    // seed: array of structures -> planes
    const double *seed_aos = seed;
    if (!on_device) {
        VG_HIP(hipMemcpyAsync(g.staging, seed, sizeof(double) * 6 * g.cells, hipMemcpyHostToDevice, s));
        seed_aos = g.staging;
    }
    vg::aos_to_soa_kernel<6><<<EGrid::blocks_for(g.cells), 256, 0, s>>>(seed_aos, wb[0], g.plane, g.cells);
    bool planar = false;
    if (!rc) rc = choose_adjoint_physics(ctx, g, m, wb[0], wb[1], planar);
    int cw = 0;
    const double dy = dx;
    int xv = kAdjDefault[0], yv = kAdjDefault[1];
    if (steps > 0 && !rc) {
        // tune on the checkpointed initial grid with the first time step
        const vg::AdjointArgs ta =
            adjoint_args(ckpt[0], wb[0], wb[1], ghost, g.plane, ny, nx, bc, dx, bias, limiter, dts[0], m);
        rc = tune_adjoint_p(planar, ctx, 0, ta, xv);
        if (!rc && dim == 2) rc = tune_adjoint_p(planar, ctx, 1, ta, yv);
    }
    for (int seg = (int)ckpt.size() - 1; seg >= 0 && !rc; seg--) {
        const int s0 = seg * K, s1 = std::min(steps, s0 + K);
        if (s0 >= s1) continue;
        slots[0] = ckpt[seg];
        rc = forward_fixed(ctx, g, dts_dev, s0, s1, dx, bias, limiter, bc, m, slots, true);
        for (int st = s1 - 1; st >= s0 && !rc; st--) {
            const int base = (st - s0) * per;
            const double dt = dts[st];
            if (dim == 1) {
                rc = adj_sweep_p(planar, ctx, 0, xv, slots[base], wb[cw], wb[1 - cw], ghost, g.plane, ny, nx, bc, dx, bias,
                               limiter, dt, m);
                cw = 1 - cw;
            } else {
                rc = adj_sweep_p(planar, ctx, 0, xv, slots[base + 2], wb[cw], wb[1 - cw], ghost, g.plane, ny, nx, bc, dx, bias,
                               limiter, dt * 0.5, m);
                cw = 1 - cw;
                if (!rc)
                    rc = adj_sweep_p(planar, ctx, 1, yv, slots[base + 1], wb[cw], wb[1 - cw], ghost, g.plane, ny, nx, bc, dy,
                                   bias, limiter, dt, m);
                cw = 1 - cw;
                if (!rc)
                    rc = adj_sweep_p(planar, ctx, 0, xv, slots[base], wb[cw], wb[1 - cw], ghost, g.plane, ny, nx, bc, dx, bias,
                                   limiter, dt * 0.5, m);
                cw = 1 - cw;
            }
        }
    }
    if (!rc) {
        double *dst = on_device ? grad : g.staging;
        vg::soa_to_aos_kernel<6><<<EGrid::blocks_for(g.cells), 256, 0, s>>>(wb[cw], dst, g.plane, g.cells);
        rc = vg::launch_status("adjoint store");
    }
    if (!rc) {
        hipError_t e = hipSuccess;
        if (!on_device) e = hipMemcpyAsync(grad, g.staging, sizeof(double) * 6 * g.cells, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) {
            vg::set_error(std::string("adjoint copy: ") + hipGetErrorString(e));
            rc = VG_ERR_HIP;
        }
    }
    if (!rc && ctx->timing) rc = ctx->resolve_timing();
    if (dt_log) std::memcpy(dt_log, dts.data(), sizeof(double) * std::min(steps, std::max(dt_log_cap, 0)));
    if (stats) {
        *stats = fst;
        stats->steps = steps;
        stats->total_steps = steps;
        stats->dt_min = steps ? *std::min_element(dts.begin(), dts.end()) : 0.0;
        stats->dt_max = steps ? *std::max_element(dts.begin(), dts.end()) : 0.0;
        stats->wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - w0).count();
        stats->nonfinite_cells = 0;
        stats->first_nonfinite = -1;
    }
    g.destroy();
    return rc;
}

int solve_fixed(vg_ctx *ctx, int dim, const double *prim_in, double *prim_out, int ny, int nx, double dx,
                const double *dts, int n_steps, double bias, int limiter, int bc, const vg_euler_params *mp)
{
    vg::clear_error();
    int rc = check_common(ctx, dim, ny, nx, dx, limiter, bc, mp);
    if (rc) return rc;
    VG_CHECK(prim_in && prim_out && (dts || n_steps == 0) && n_steps >= 0, "null argument");
    rc = vg::bind(ctx);
    if (rc) return rc;
    const EulerPhys::Params m = to_params(mp);
    EGrid g;
    rc = g.create(ctx, dim, ny, nx);
    if (!rc) rc = g.load(prim_in, 0);
    DeviceBuffers mem;
    double *dts_dev = nullptr;
    if (!rc) rc = mem.alloc(&dts_dev, sizeof(double) * std::max(n_steps, 1));
    if (!rc && n_steps > 0)
        VG_HIP(hipMemcpyAsync(dts_dev, dts, sizeof(double) * n_steps, hipMemcpyHostToDevice, ctx->stream));
    for (int s = 0; s < n_steps && !rc; s++) {
        // ping-pong through the grid's two buffers, one step at a time
        std::vector<double *> slots;
        if (dim == 1) slots = {g.buf[g.cur], g.buf[1 - g.cur]};
        else slots = {g.buf[g.cur], g.buf[1 - g.cur], g.buf[g.cur], g.buf[1 - g.cur]};
        rc = forward_fixed(ctx, g, dts_dev, s, s + 1, dx, bias, limiter, bc, m, slots);
        g.cur = 1 - g.cur;
    }
    if (!rc) rc = g.store(prim_out, 0);
    g.destroy();
    return rc;
}

}  // namespace

extern "C" {

int vg_euler_adjoint_1d(vg_ctx *ctx, const double *prim_in, double *prim_out, const double *seed, double *grad, int n,
                        double dx, double cfl, double t_final, int max_steps, double bias, int limiter, int bc,
                        const vg_euler_params *mp, int checkpoint_every, double *dt_log, int dt_log_cap,
                        vg_run_stats *stats)
{
    return adjoint(ctx, 1, prim_in, prim_out, seed, grad, 1, n, dx, cfl, t_final, max_steps, bias, limiter, bc, mp,
                   checkpoint_every, dt_log, dt_log_cap, stats, 0);
}

int vg_euler_adjoint_1d_device(vg_ctx *ctx, const double *prim_in, double *prim_out, const double *seed, double *grad,
                               int n, double dx, double cfl, double t_final, int max_steps, double bias, int limiter,
                               int bc, const vg_euler_params *mp, int checkpoint_every, double *dt_log, int dt_log_cap,
                               vg_run_stats *stats)
{
    return adjoint(ctx, 1, prim_in, prim_out, seed, grad, 1, n, dx, cfl, t_final, max_steps, bias, limiter, bc, mp,
                   checkpoint_every, dt_log, dt_log_cap, stats, 1);
}

int vg_euler_adjoint_2d(vg_ctx *ctx, const double *prim_in, double *prim_out, const double *seed, double *grad, int ny,
                        int nx, double dx, double cfl, double t_final, int max_steps, double bias, int limiter, int bc,
                        const vg_euler_params *mp, int checkpoint_every, double *dt_log, int dt_log_cap,
                        vg_run_stats *stats)
{
    return adjoint(ctx, 2, prim_in, prim_out, seed, grad, ny, nx, dx, cfl, t_final, max_steps, bias, limiter, bc, mp,
                   checkpoint_every, dt_log, dt_log_cap, stats, 0);
}

int vg_euler_adjoint_2d_device(vg_ctx *ctx, const double *prim_in, double *prim_out, const double *seed, double *grad,
                               int ny, int nx, double dx, double cfl, double t_final, int max_steps, double bias,
                               int limiter, int bc, const vg_euler_params *mp, int checkpoint_every, double *dt_log,
                               int dt_log_cap, vg_run_stats *stats)
{
    return adjoint(ctx, 2, prim_in, prim_out, seed, grad, ny, nx, dx, cfl, t_final, max_steps, bias, limiter, bc, mp,
                   checkpoint_every, dt_log, dt_log_cap, stats, 1);
}

int vg_euler_solve_fixed_dt_1d(vg_ctx *ctx, const double *prim_in, double *prim_out, int n, double dx,
                               const double *dts, int n_steps, double bias, int limiter, int bc,
                               const vg_euler_params *mp)
{
    return solve_fixed(ctx, 1, prim_in, prim_out, 1, n, dx, dts, n_steps, bias, limiter, bc, mp);
}

int vg_euler_solve_fixed_dt_2d(vg_ctx *ctx, const double *prim_in, double *prim_out, int ny, int nx, double dx,
                               const double *dts, int n_steps, double bias, int limiter, int bc,
                               const vg_euler_params *mp)
{
    return solve_fixed(ctx, 2, prim_in, prim_out, ny, nx, dx, dts, n_steps, bias, limiter, bc, mp);
}

}  // extern "C"
