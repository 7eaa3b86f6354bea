// RGFM C ABI: two-material Euler with the real ghost fluid method, 1D and 2D
// whole solves and a batch of 2D instances, all device resident.
//
// One step (RGFMSolver::solveHLLC/solveHLLC2D, rgfmsolver.cpp:745-840):
//   loop test -> ghost fill (rows, then columns) -> wave-speed maximum of both
//   materials -> dt -> level set X(dt/2) Y(dt) X(dt/2) -> SLIC X(dt/2) Y(dt) X(dt/2)
//   for material 1, then material 2.
// 1D is the same with one row, one level-set pass and one sweep per material.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "vg_grid.h"
#include "vg_rgfm.h"
#include "vg_runtime.h"

namespace vg {

// Loop test of the step about to start (secondordersolver.cpp:295 style
// `while (currentTime < finalTime)`): folds in the previous step and stops the
// clock before anything of the next step runs.
static __global__ void clock_check_kernel(Clock *c)
{
    if (c->done) return;
    double t = c->t;
    int it = c->it;
    if (c->pending) {
        t += c->dt;
        it += 1;
    }
    if (!(t < c->t_final) || (c->max_steps > 0 && it >= c->max_steps)) {
        c->t = t;
        c->it = it;
        c->pending = 0;
        c->done = 1;
    }
}

struct RGFMGrid {
    static constexpr int BSX = 256, RX = 4, BSY = 256, HY = 64;
    vg_ctx *ctx = nullptr;
    int dim = 2, ny = 1, nx = 1;
    long long cells = 0, plane = 0;
    double *m1[2] = {nullptr, nullptr}, *m2[2] = {nullptr, nullptr}, *phi[2] = {nullptr, nullptr};
    double *rec = nullptr, *staging = nullptr, *circ = nullptr, *dt_dev = nullptr;
    int dt_dev_cap = 0;
    int cur = 0, pcur = 0;
    Clock *clk = nullptr, *tune_clk = nullptr;
    unsigned long long *partials = nullptr;
    int max_blocks = 0, circ_blocks = 0;
    int y_variant = 0;                  // autotuned y-sweep shape
    hipGraphExec_t graph = nullptr;     // captured chunk of steps
    std::string graph_key;
    static constexpr int kGraphSteps = 8;   // even: buffers return to their start

    int create(vg_ctx *c, int d, int y, int x)
    {
        ctx = c;
        dim = d;
        ny = y;
        nx = x;
        cells = (long long)ny * nx;
        plane = (cells + 63) / 64 * 64;
        for (int b = 0; b < 2; b++) {
            VG_HIP(hipMalloc(&m1[b], sizeof(double) * 6 * plane));
            VG_HIP(hipMalloc(&m2[b], sizeof(double) * 6 * plane));
            VG_HIP(hipMalloc(&phi[b], sizeof(double) * cells));
        }
        VG_HIP(hipMalloc(&rec, sizeof(double) * 4 * cells));
        VG_HIP(hipMalloc(&staging, sizeof(double) * 6 * cells));
        VG_HIP(hipMalloc(&clk, sizeof(Clock)));
        max_blocks = (int)std::min<long long>((cells + 255) / 256, 2048);
        VG_HIP(hipMalloc(&partials, sizeof(unsigned long long) * 2 * max_blocks));
        circ_blocks = (int)std::min<long long>((cells + 255) / 256, 1024);
        VG_HIP(hipMalloc(&circ, sizeof(double) * circ_blocks));
        return VG_OK;
    }

    void destroy()
    {
        for (int b = 0; b < 2; b++) {
            hipFree(m1[b]);
            hipFree(m2[b]);
            hipFree(phi[b]);
        }
        hipFree(rec);
        hipFree(staging);
        hipFree(clk);
        hipFree(partials);
        hipFree(circ);
        hipFree(dt_dev);
        hipFree(tune_clk);
        if (graph) hipGraphExecDestroy(graph);
    }

    static int blocks_for(long long n) { return (int)std::min<long long>((n + 255) / 256, 4096); }

    int load(const double *a1, const double *a2, const double *ph)
    {
        hipStream_t s = ctx->stream;
        VG_HIP(hipMemcpyAsync(staging, a1, sizeof(double) * 6 * cells, hipMemcpyHostToDevice, s));
        aos_to_soa_kernel<6><<<blocks_for(cells), 256, 0, s>>>(staging, m1[0], plane, cells);
        VG_HIP(hipStreamSynchronize(s));
        VG_HIP(hipMemcpyAsync(staging, a2, sizeof(double) * 6 * cells, hipMemcpyHostToDevice, s));
        aos_to_soa_kernel<6><<<blocks_for(cells), 256, 0, s>>>(staging, m2[0], plane, cells);
        VG_HIP(hipMemcpyAsync(phi[0], ph, sizeof(double) * cells, hipMemcpyHostToDevice, s));
        VG_HIP(hipStreamSynchronize(s));
        cur = pcur = 0;
        return launch_status("rgfm load");
    }

    int store(double *a1, double *a2, double *ph)
    {
        hipStream_t s = ctx->stream;
        soa_to_aos_kernel<6><<<blocks_for(cells), 256, 0, s>>>(m1[cur], staging, plane, cells);
        VG_HIP(hipMemcpyAsync(a1, staging, sizeof(double) * 6 * cells, hipMemcpyDeviceToHost, s));
        VG_HIP(hipStreamSynchronize(s));
        soa_to_aos_kernel<6><<<blocks_for(cells), 256, 0, s>>>(m2[cur], staging, plane, cells);
        VG_HIP(hipMemcpyAsync(a2, staging, sizeof(double) * 6 * cells, hipMemcpyDeviceToHost, s));
        VG_HIP(hipMemcpyAsync(ph, phi[pcur], sizeof(double) * cells, hipMemcpyDeviceToHost, s));
        VG_HIP(hipStreamSynchronize(s));
        return launch_status("rgfm store");
    }

    LineArgs line_args(int riemann, const EP &p1, const EP &p2) const
    {
        LineArgs a;
        a.m1 = m1[cur];
        a.m2 = m2[cur];
        a.phi = phi[pcur];
        a.phi_out = phi[1 - pcur];
        a.rec = rec;
        a.plane = plane;
        a.ny = ny;
        a.nx = nx;
        a.init_flag = (dim == 2) ? 1 : 0;
        a.riemann = riemann;
        a.h = 0.0;
        a.dt_scale = 1.0;
        a.p1 = p1;
        a.p2 = p2;
        a.clk = clk;
        return a;
    }

    template <int MODE>
    int sweep(int dir, const double *in, double *out, double scale, double h, double bias, int limiter, const EP &m)
    {
        SweepArgs<EulerPhys> a;
        a.in = in;
        a.out = out;
        a.plane = plane;
        a.ny = ny;
        a.nx = nx;
        a.bc = VG_BC_TRANSMISSIVE;
        a.limiter = limiter;
        a.dx = h;
        a.bias = bias;
        a.dt_scale = scale;
        a.dim = dim;
        a.fuse_max = 0;
        a.m = m;
        a.clk = clk;
        a.partials = nullptr;
        if (dir == 0) {
            dim3 grid((nx + (BSX - 4) - 1) / (BSX - 4), (ny + RX - 1) / RX);
            return ctx->timed("rgfm_sweep_x", [&] {
                sweep_x_kernel<EulerPhys, 2, MODE, BSX, RX><<<grid, BSX, 0, ctx->stream>>>(a);
            });
        }
        return launch_y<EulerPhys, 2, MODE>(ctx, y_variant, a);
    }

    template <int MODE>
    int enqueue_step(int riemann, double dx, double dy, double bias, int limiter, const EP &p1, const EP &p2,
                     double *dt_log, int dt_cap)
    {
        hipStream_t s = ctx->stream;
        clock_check_kernel<<<1, 1, 0, s>>>(clk);
        LineArgs la = line_args(riemann, p1, p2);
        int rc = ctx->timed("rgfm_fill_x", [&] { rgfm_fill_kernel<0><<<ny, kLineBS, 0, s>>>(la); });
        if (!rc && dim == 2) rc = ctx->timed("rgfm_fill_y", [&] { rgfm_fill_kernel<1><<<nx, kLineBS, 0, s>>>(la); });
        if (rc) return launch_status("rgfm_fill");
        wave_max_kernel<EulerPhys><<<max_blocks, 256, 0, s>>>(m1[cur], plane, cells, dim, p1, partials);
        wave_max_kernel<EulerPhys><<<max_blocks, 256, 0, s>>>(m2[cur], plane, cells, dim, p2, partials + max_blocks);
        clock_begin_kernel<<<1, kReduceThreads, 0, s>>>(clk, partials, 2 * max_blocks, dt_log, dt_cap);
        rc = launch_status("rgfm clock");
        if (rc) return rc;
        if (dim == 1) {
            la.h = dx;
            la.dt_scale = 1.0;
            rc = ctx->timed("rgfm_level_set", [&] { rgfm_level_set_kernel<0><<<1, kLineBS, 0, s>>>(la); });
        } else {
            LineArgs a = la;
            a.h = dx;
            a.dt_scale = 0.5;
            a.phi = phi[pcur];
            a.phi_out = phi[1 - pcur];
            rc = ctx->timed("rgfm_level_set", [&] { rgfm_level_set_kernel<0><<<ny, kLineBS, 0, s>>>(a); });
            a.h = dy;
            a.dt_scale = 1.0;
            a.phi = phi[1 - pcur];
            a.phi_out = phi[pcur];
            if (!rc) rc = ctx->timed("rgfm_level_set", [&] { rgfm_level_set_kernel<1><<<nx, kLineBS, 0, s>>>(a); });
            a.h = dx;
            a.dt_scale = 0.5;
            a.phi = phi[pcur];
            a.phi_out = phi[1 - pcur];
            if (!rc) rc = ctx->timed("rgfm_level_set", [&] { rgfm_level_set_kernel<0><<<ny, kLineBS, 0, s>>>(a); });
        }
        if (rc) return launch_status("rgfm_level_set");
        pcur = 1 - pcur;
        double *mats[2][2] = {{m1[cur], m1[1 - cur]}, {m2[cur], m2[1 - cur]}};
        const EP *ps[2] = {&p1, &p2};
        for (int q = 0; q < 2 && !rc; q++) {
            double *A = mats[q][0], *B = mats[q][1];
            if (dim == 1) {
                rc = sweep<MODE>(0, A, B, 1.0, dx, bias, limiter, *ps[q]);
            } else {
                rc = sweep<MODE>(0, A, B, 0.5, dx, bias, limiter, *ps[q]);
                if (!rc) rc = sweep<MODE>(1, B, A, 1.0, dy, bias, limiter, *ps[q]);
                if (!rc) rc = sweep<MODE>(0, A, B, 0.5, dx, bias, limiter, *ps[q]);
            }
        }
        if (rc) return launch_status("rgfm sweep");
        cur = 1 - cur;
        return VG_OK;
    }

    // Runs to t_final or max_steps.  Each executed step moves both materials
    // and the level set to the other buffer once.
    int run(int riemann, int mode, double dx, double dy, double cfl, double t_final, int max_steps, double bias,
            int limiter, const EP &p1, const EP &p2, double *dt_log, int dt_cap, vg_run_stats *stats)
    {
        auto w0 = std::chrono::steady_clock::now();
        hipStream_t s = ctx->stream;
        Clock h{};
        h.t_final = t_final;
        h.cfl = cfl;
        h.dx = (dim == 1) ? dx : std::min(dx, dy);
        h.max_steps = max_steps;
        VG_HIP(hipMemcpyAsync(clk, &h, sizeof(Clock), hipMemcpyHostToDevice, s));
        if (dim == 2) {
            rgfm_corner_kernel<<<1, 1, 0, s>>>(p1, p2, clk);
            int rc = launch_status("rgfm_corner");
            if (rc) return rc;
        }
        if (dt_log && dt_cap > 0 && dt_dev_cap < dt_cap) {
            hipFree(dt_dev);
            dt_dev = nullptr;
            VG_HIP(hipMalloc(&dt_dev, sizeof(double) * dt_cap));
            dt_dev_cap = dt_cap;
        }
        double *dlog = (dt_log && dt_cap > 0) ? dt_dev : nullptr;
        if (dim == 2) {
            // y-sweep shape for this grid (shared with Euler grids of the same shape).
            if (!tune_clk) VG_HIP(hipMalloc(&tune_clk, sizeof(Clock)));
            Clock tc{};
            tc.dt = 1e-3 * std::min(dx, dy);
            VG_HIP(hipMemcpyAsync(tune_clk, &tc, sizeof(Clock), hipMemcpyHostToDevice, s));
            SweepArgs<EulerPhys> ta;
            ta.in = m1[cur];
            ta.out = m1[1 - cur];
            ta.plane = plane;
            ta.ny = ny;
            ta.nx = nx;
            ta.bc = VG_BC_TRANSMISSIVE;
            ta.limiter = limiter;
            ta.dx = dy;
            ta.bias = bias;
            ta.dt_scale = 1.0;
            ta.dim = 2;
            ta.fuse_max = 0;
            ta.m = p1;
            ta.clk = tune_clk;
            ta.partials = nullptr;
            int rc = (mode == VG_MODE_FAST)
                         ? tune_y<EulerPhys, 2, MODE_FAST>(ctx, tune_key("euler", 2, mode, ny, nx), ta, y_variant)
                         : tune_y<EulerPhys, 2, MODE_EXACT>(ctx, tune_key("euler", 2, mode, ny, nx), ta, y_variant);
            if (rc) return rc;
        }
        hipEvent_t e0, e1;
        VG_HIP(hipEventCreate(&e0));
        VG_HIP(hipEventCreate(&e1));
        VG_HIP(hipEventRecord(e0, s));
        const int start_cur = cur, start_pcur = pcur;
        const int chunk = (max_steps > 0 && max_steps <= 1024) ? max_steps : 128;
        auto enqueue = [&]() {
            return (mode == VG_MODE_FAST) ? enqueue_step<MODE_FAST>(riemann, dx, dy, bias, limiter, p1, p2, dlog, dt_cap)
                                          : enqueue_step<MODE_EXACT>(riemann, dx, dy, bias, limiter, p1, p2, dlog, dt_cap);
        };
        if (ctx->graphs && !ctx->timing) {
            // Capture kGraphSteps steps once per argument set and replay them;
            // every step moves both materials and the level set once, so an even
            // count returns all buffers to their start.
            std::string key(reinterpret_cast<const char *>(&p1), sizeof(p1));
            key.append(reinterpret_cast<const char *>(&p2), sizeof(p2));
            const double num[3] = {dx, dy, bias};
            key.append(reinterpret_cast<const char *>(num), sizeof(num));
            key += "|" + std::to_string(riemann) + "|" + std::to_string(mode) + "|" + std::to_string(limiter) + "|" +
                   std::to_string((unsigned long long)(uintptr_t)dlog) + "|" + std::to_string(dt_cap) + "|" +
                   std::to_string(y_variant) + "|" + std::to_string(cur) + "|" + std::to_string(pcur);
            if (!graph || key != graph_key) {
                if (graph) {
                    hipGraphExecDestroy(graph);
                    graph = nullptr;
                }
                hipGraph_t g = nullptr;
                VG_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                int rc = VG_OK;
                for (int k = 0; k < kGraphSteps && !rc; k++) rc = enqueue();
                hipError_t ec = hipStreamEndCapture(s, &g);
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
                for (int r = 0; r < reps; r++) VG_HIP(hipGraphLaunch(graph, s));
                VG_HIP(hipMemcpyAsync(&h, clk, sizeof(Clock), hipMemcpyDeviceToHost, s));
                VG_HIP(hipStreamSynchronize(s));
                if (h.done) break;
            }
        } else {
            for (;;) {
                for (int k = 0; k < chunk; k++) {
                    int rc = enqueue();
                    if (rc) return rc;
                }
                VG_HIP(hipMemcpyAsync(&h, clk, sizeof(Clock), hipMemcpyDeviceToHost, s));
                VG_HIP(hipStreamSynchronize(s));
                if (h.done) break;
            }
        }
        clock_finish_kernel<<<1, 1, 0, s>>>(clk);
        VG_HIP(hipEventRecord(e1, s));
        VG_HIP(hipMemcpyAsync(&h, clk, sizeof(Clock), hipMemcpyDeviceToHost, s));
        VG_HIP(hipStreamSynchronize(s));
        float dev_ms = 0.0f;
        VG_HIP(hipEventElapsedTime(&dev_ms, e0, e1));
        hipEventDestroy(e0);
        hipEventDestroy(e1);
        cur = start_cur ^ (h.steps_this_run & 1);
        pcur = start_pcur ^ (h.steps_this_run & 1);
        if (dlog) {
            const int n = std::min(h.steps_this_run, dt_cap);
            if (n > 0) VG_HIP(hipMemcpy(dt_log, dt_dev, sizeof(double) * n, hipMemcpyDeviceToHost));
        }
        if (ctx->timing) {
            int rc = ctx->resolve_timing();
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
        if (h.stalled) {
            set_error("RGFM: a time step could not advance t (non-finite or zero wave speed)");
            return VG_ERR_NONFINITE;
        }
        return VG_OK;
    }

    int circulation(double dx, double dy, double *out)
    {
        hipStream_t s = ctx->stream;
        rgfm_circulation_kernel<<<circ_blocks, 256, 0, s>>>(m1[cur], plane, phi[pcur], ny, nx, dx, dy, circ);
        int rc = launch_status("rgfm_circulation");
        if (rc) return rc;
        std::vector<double> part(circ_blocks);
        VG_HIP(hipMemcpyAsync(part.data(), circ, sizeof(double) * circ_blocks, hipMemcpyDeviceToHost, s));
        VG_HIP(hipStreamSynchronize(s));
        double g = 0.0;
        for (double v : part) g += v;
        *out = g;
        return VG_OK;
    }
};

}  // namespace vg

namespace {

vg::EP to_params(const vg_euler_params *p)
{
    vg::EP m;
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

int check_args(int dim, int ny, int nx, double dx, double dy, double cfl, double t_final, int max_steps, int limiter,
               int riemann, const vg_euler_params *p1, const vg_euler_params *p2, int mode)
{
    (void)limiter;
    VG_CHECK(p1 != nullptr && p2 != nullptr, "null material parameters");
    VG_CHECK(riemann == VG_RIEMANN_HLLC || riemann == VG_RIEMANN_EXACT, "riemann must be VG_RIEMANN_HLLC or _EXACT");
    VG_CHECK(mode == VG_MODE_EXACT || mode == VG_MODE_FAST, "mode must be VG_MODE_EXACT or VG_MODE_FAST");
    VG_CHECK(nx >= 8 && (dim == 1 ? ny == 1 : ny >= 8), "RGFM grids need at least 8 cells per direction");
    VG_CHECK(nx <= vg::kMaxLine && ny <= vg::kMaxLine, "RGFM grid lines are limited to 16384 cells");
    VG_CHECK(dx > 0.0 && std::isfinite(dx) && dy > 0.0 && std::isfinite(dy), "cell spacings must be positive");
    VG_CHECK(cfl > 0.0 && std::isfinite(cfl), "CFL coefficient must be positive");
    VG_CHECK(max_steps >= 0 && (max_steps > 0 || std::isfinite(t_final)), "an infinite final time needs a step cap");
    return VG_OK;
}

int solve(vg_ctx *ctx, int dim, const double *m1_in, const double *m2_in, const double *phi_in, double *m1_out,
          double *m2_out, double *phi_out, int ny, int nx, double dx, double dy, double cfl, double t_final,
          int max_steps, double bias, int limiter, int riemann, const vg_euler_params *p1,
          const vg_euler_params *p2, int mode, double *dt_log, int dt_log_cap, vg_run_stats *stats,
          double *circulation)
{
    vg::clear_error();
    VG_CHECK(ctx != nullptr, "null context");
    VG_CHECK(m1_in && m2_in && phi_in && m1_out && m2_out && phi_out, "null state pointer");
    int rc = check_args(dim, ny, nx, dx, dy, cfl, t_final, max_steps, limiter, riemann, p1, p2, mode);
    if (rc) return rc;
    rc = vg::bind(ctx);
    if (rc) return rc;
    vg::RGFMGrid g;
    rc = g.create(ctx, dim, ny, nx);
    if (!rc) rc = g.load(m1_in, m2_in, phi_in);
    vg_run_stats local{};
    int run_rc = VG_OK;
    if (!rc) {
        run_rc = g.run(riemann, mode, dx, dy, cfl, t_final, max_steps, bias, limiter, to_params(p1), to_params(p2),
                       dt_log, dt_log_cap, &local);
        if (run_rc != VG_OK && run_rc != VG_ERR_NONFINITE) rc = run_rc;
    }
    if (!rc) rc = g.store(m1_out, m2_out, phi_out);
    if (!rc && circulation) rc = g.circulation(dx, dy, circulation);
    g.destroy();
    if (stats) *stats = local;
    return rc ? rc : run_rc;
}

}  // namespace

extern "C" {

int vg_rgfm_solve_1d(vg_ctx *ctx, const double *m1_in, const double *m2_in, const double *phi_in, double *m1_out,
                     double *m2_out, double *phi_out, int n, double dx, double cfl, double t_final, int max_steps,
                     double bias, int limiter, int riemann, const vg_euler_params *p1, const vg_euler_params *p2,
                     int mode, double *dt_log, int dt_log_cap, vg_run_stats *stats)
{
    return solve(ctx, 1, m1_in, m2_in, phi_in, m1_out, m2_out, phi_out, 1, n, dx, dx, cfl, t_final, max_steps, bias,
                 limiter, riemann, p1, p2, mode, dt_log, dt_log_cap, stats, nullptr);
}

int vg_rgfm_solve_2d(vg_ctx *ctx, const double *m1_in, const double *m2_in, const double *phi_in, double *m1_out,
                     double *m2_out, double *phi_out, int ny, int nx, double dx, double dy, double cfl,
                     double t_final, int max_steps, double bias, int limiter, int riemann,
                     const vg_euler_params *p1, const vg_euler_params *p2, int mode, double *dt_log, int dt_log_cap,
                     vg_run_stats *stats, double *circulation)
{
    return solve(ctx, 2, m1_in, m2_in, phi_in, m1_out, m2_out, phi_out, ny, nx, dx, dy, cfl, t_final, max_steps,
                 bias, limiter, riemann, p1, p2, mode, dt_log, dt_log_cap, stats, circulation);
}

int vg_rgfm_solve_batch_2d(vg_ctx *ctx, int n_inst, const double *m1_in, const double *m2_in, const double *phi_in,
                           double *m1_out, double *m2_out, double *phi_out, int ny, int nx, double dx, double dy,
                           double cfl, const double *t_final, int max_steps, double bias, int limiter, int riemann,
                           const vg_euler_params *p1, const vg_euler_params *p2, int mode, int32_t *steps,
                           double *t_out, double *circulation, vg_run_stats *stats)
{
    vg::clear_error();
    VG_CHECK(ctx != nullptr && n_inst >= 1 && t_final != nullptr, "null argument");
    VG_CHECK(m1_in && m2_in && phi_in && m1_out && m2_out && phi_out, "null state pointer");
    for (int q = 0; q < n_inst; q++) {
        int rc = check_args(2, ny, nx, dx, dy, cfl, t_final[q], max_steps, limiter, riemann, p1, p2, mode);
        if (rc) return rc;
    }
    int rc = vg::bind(ctx);
    if (rc) return rc;
    vg::RGFMGrid g;
    rc = g.create(ctx, 2, ny, nx);
    const size_t cells = (size_t)ny * nx;
    vg_run_stats total{};
    double dt_min = 0.0, dt_max = 0.0;
    int run_rc = VG_OK;
    for (int q = 0; q < n_inst && !rc; q++) {
        rc = g.load(m1_in + 6 * cells * q, m2_in + 6 * cells * q, phi_in + cells * q);
        vg_run_stats st{};
        if (!rc) {
            const int r = g.run(riemann, mode, dx, dy, cfl, t_final[q], max_steps, bias, limiter, to_params(p1),
                                to_params(p2), nullptr, 0, &st);
            if (r == VG_ERR_NONFINITE) run_rc = r;
            else rc = r;
        }
        if (!rc) rc = g.store(m1_out + 6 * cells * q, m2_out + 6 * cells * q, phi_out + cells * q);
        if (!rc && circulation) rc = g.circulation(dx, dy, circulation + q);
        if (steps) steps[q] = st.steps;
        if (t_out) t_out[q] = st.t;
        total.steps += st.steps;
        total.device_ms += st.device_ms;
        total.wall_ms += st.wall_ms;
        if (q == 0 || st.dt_min < dt_min) dt_min = st.dt_min;
        if (q == 0 || st.dt_max > dt_max) dt_max = st.dt_max;
    }
    g.destroy();
    total.total_steps = total.steps;
    total.dt_min = dt_min;
    total.dt_max = dt_max;
    total.first_nonfinite = -1;
    if (stats) *stats = total;
    return rc ? rc : run_rc;
}

}  // extern "C"
