// MHD RGFM C ABI: two ideal-MHD materials with the real ghost fluid method,
// 1D and 2D whole solves and the standalone ghost fill, all device resident.
//
// One step (MHDRGFMSolver::solve / solve2D, RGFM/mhdrgfmsolver.cpp:543-693):
//   loop test -> ghost fill (rows, then columns) -> wave-speed maximum of both
//   materials -> dt -> level set (1D: one pass; 2D: X(dt/2) Y(dt) X(dt/2)) ->
//   material 1 advanced, then material 2.
// A 1D material advance is one SLIC step with the caller's parameters (c_h as
// given; the 1D solve runs no sub-cycles).  A 2D advance runs `subcycles` RK4
// GLM sub-cycles of dt / 2, the SLIC sweeps X(dt/2) Y(dt) X(dt/2), and the
// sub-cycles again, with c_h = 4 and c_p = sqrt(0.18 c_h): solve2D sets them
// after the first step's ghost fill and level set, so the first fill alone
// uses the caller's c_h (mhdrgfmsolver.cpp:618-624).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <string>

#include "vg_mhd_host.h"
#include "vg_mhd_rgfm.h"

namespace vg {
namespace {

struct MHDRGFMGrid {
    vg_ctx *ctx = nullptr;
    int dim = 2, ny = 1, nx = 1;
    long long cells = 0, plane = 0;
    GridState<MHDPhys> g[2];
    double *phi[2] = {nullptr, nullptr};
    double *rec = nullptr, *dt_dev = nullptr;
    int dt_dev_cap = 0;
    int pcur = 0;
    int max_blocks = 0;

    Clock *clk() const { return g[0].clk; }

    int create(vg_ctx *c, int d, int y, int x)
    {
        ctx = c;
        dim = d;
        ny = y;
        nx = x;
        cells = (long long)ny * nx;
        for (int q = 0; q < 2; q++) {
            int rc = g[q].create(c, d, y, x);
            if (rc) return rc;
        }
        plane = g[0].plane;
        VG_HIP(hipMalloc(&phi[0], sizeof(double) * cells));
        VG_HIP(hipMalloc(&phi[1], sizeof(double) * cells));
        VG_HIP(hipMalloc(&rec, sizeof(double) * kMHDRec * cells));
        // both materials' per-block maxima go to material 1's partials
        max_blocks = (int)std::min<long long>((cells + 255) / 256, 2048);
        return VG_OK;
    }

    void destroy()
    {
        g[0].destroy();
        g[1].destroy();
        hipFree(phi[0]);
        hipFree(phi[1]);
        hipFree(rec);
        hipFree(dt_dev);
    }

    int load(const double *a1, const double *a2, const double *ph)
    {
        int rc = g[0].load(a1, 0);
        if (!rc) rc = g[1].load(a2, 0);
        if (rc) return rc;
        VG_HIP(hipMemcpyAsync(phi[0], ph, sizeof(double) * cells, hipMemcpyHostToDevice, ctx->stream));
        VG_HIP(hipStreamSynchronize(ctx->stream));
        pcur = 0;
        return VG_OK;
    }

    int store(double *a1, double *a2, double *ph)
    {
        int rc = g[0].store(a1, 0);
        if (!rc) rc = g[1].store(a2, 0);
        if (rc) return rc;
        if (ph) {
            VG_HIP(hipMemcpyAsync(ph, phi[pcur], sizeof(double) * cells, hipMemcpyDeviceToHost, ctx->stream));
            VG_HIP(hipStreamSynchronize(ctx->stream));
        }
        return VG_OK;
    }

    MHDLineArgs fill_args(const MP &p1, const MP &p2, const Clock *c) const
    {
        MHDLineArgs a;
        a.m1 = g[0].buf[g[0].cur];
        a.m2 = g[1].buf[g[1].cur];
        a.phi = phi[pcur];
        a.rec = rec;
        a.plane = plane;
        a.ny = ny;
        a.nx = nx;
        a.init_flag = (dim == 2) ? 1 : 0;
        a.with_psi = (dim == 2) ? 1 : 0;
        a.p1 = p1;
        a.p2 = p2;
        a.clk = c;
        return a;
    }

    // applyRGFMBoundaryConditions[2D]: rows, then (2D) columns, in place.
    int fill(const MP &p1, const MP &p2, const Clock *c)
    {
        const MHDLineArgs a = fill_args(p1, p2, c);
        int rc = ctx->timed("mhd_rgfm_fill_x", [&] { mhd_rgfm_fill_kernel<0><<<ny, kLineBS, 0, ctx->stream>>>(a); });
        if (!rc && dim == 2)
            rc = ctx->timed("mhd_rgfm_fill_y", [&] { mhd_rgfm_fill_kernel<1><<<nx, kLineBS, 0, ctx->stream>>>(a); });
        return rc ? rc : launch_status("mhd_rgfm_fill");
    }

    // The level-set passes of the step (the Euler RGFM's line kernels: they
    // read only each material's velocity plane along the line).
    int level_set(double dx)
    {
        hipStream_t s = ctx->stream;
        LineArgs a{};
        a.m1 = g[0].buf[g[0].cur];
        a.m2 = g[1].buf[g[1].cur];
        a.plane = plane;
        a.ny = ny;
        a.nx = nx;
        a.init_flag = (dim == 2) ? 1 : 0;
        a.clk = clk();
        a.h = dx;
        int rc;
        if (dim == 1) {
            a.dt_scale = 1.0;
            a.phi = phi[pcur];
            a.phi_out = phi[1 - pcur];
            rc = ctx->timed("mhd_rgfm_level_set", [&] { rgfm_level_set_kernel<0><<<1, kLineBS, 0, s>>>(a); });
        } else {
            a.dt_scale = 0.5;
            a.phi = phi[pcur];
            a.phi_out = phi[1 - pcur];
            rc = ctx->timed("mhd_rgfm_level_set", [&] { rgfm_level_set_kernel<0><<<ny, kLineBS, 0, s>>>(a); });
            a.dt_scale = 1.0;
            a.phi = phi[1 - pcur];
            a.phi_out = phi[pcur];
            if (!rc)
                rc = ctx->timed("mhd_rgfm_level_set", [&] {
                    rgfm_level_set_y_kernel<<<(nx + kLsCols - 1) / kLsCols, kLsCols * kLsSegs, 0, s>>>(a);
                });
            a.dt_scale = 0.5;
            a.phi = phi[pcur];
            a.phi_out = phi[1 - pcur];
            if (!rc) rc = ctx->timed("mhd_rgfm_level_set", [&] { rgfm_level_set_kernel<0><<<ny, kLineBS, 0, s>>>(a); });
        }
        if (rc) return launch_status("mhd_rgfm_level_set");
        pcur = 1 - pcur;
        return VG_OK;
    }

    // One RK4 GLM sub-cycle of dt / (2 subcycles) (computeRungeKuttaTimeStep2D)
    // on material q with parameters as given.
    int source(int q, int subcycles, double dx, double bias, int limiter, const MP &m)
    {
        GridState<MHDPhys> &G = g[q];
        ForcingArgs<MHDPhys> a;
        a.in = G.buf[G.cur];
        a.out = G.buf[1 - G.cur];
        a.plane = plane;
        a.ny = ny;
        a.nx = nx;
        a.dim = 2;
        a.bc = VG_BC_TRANSMISSIVE;
        a.limiter = limiter;
        a.subcycles = subcycles;
        a.dx = dx;
        a.bias = bias;
        a.fuse_max = 0;
        a.m = m;
        a.clk = clk();
        a.partials = nullptr;
        a.fixed_params = 1;
        const unsigned nb = (unsigned)forcing_blocks(2, ny, nx);
        int rc = ctx->timed("mhd_rgfm_source_rk4",
                            [&] { forcing_kernel<MHDPhys, 2><<<nb, kForceBS, 0, ctx->stream>>>(a); });
        if (rc) return launch_status("mhd_rgfm_source");
        G.cur = 1 - G.cur;
        return VG_OK;
    }

    template <int MODE>
    int sweep(int q, int dir, double scale, double dx, double bias, int limiter, const MP &m)
    {
        GridState<MHDPhys> &G = g[q];
        int rc = G.launch_sweep<2, MODE>(dir, G.buf[G.cur], G.buf[1 - G.cur], scale, 0, VG_BC_TRANSMISSIVE, dx, bias,
                                         limiter, m, clk(), 1);
        if (rc) return launch_status("mhd_rgfm_sweep");
        G.cur = 1 - G.cur;
        return VG_OK;
    }

    // One step.  fp1/fp2: the ghost fill's parameters; sp1/sp2: the material
    // advances' (and the wave speeds').  Every material buffer and the level
    // set move to the other buffer an odd number of times per step.
    template <int MODE>
    int enqueue_step(const MP &fp1, const MP &fp2, const MP &sp1, const MP &sp2, double dx, double bias, int limiter,
                     int subcycles, double *dlog, int dt_cap)
    {
        hipStream_t s = ctx->stream;
        clock_check_kernel<<<1, 1, 0, s>>>(clk());
        int rc = fill(fp1, fp2, clk());
        if (rc) return rc;
        unsigned long long *partials = g[0].partials;
        wave_max_kernel<MHDPhys><<<max_blocks, 256, 0, s>>>(g[0].buf[g[0].cur], plane, cells, dim, sp1, partials);
        wave_max_kernel<MHDPhys><<<max_blocks, 256, 0, s>>>(g[1].buf[g[1].cur], plane, cells, dim, sp2,
                                                            partials + max_blocks);
        clock_begin_kernel<<<1, kReduceThreads, 0, s>>>(clk(), partials, 2 * max_blocks, dlog, dt_cap);
        rc = launch_status("mhd_rgfm clock");
        if (!rc) rc = level_set(dx);
        const MP *ps[2] = {&sp1, &sp2};
        for (int q = 0; q < 2 && !rc; q++) {
            if (dim == 1) {
                rc = sweep<MODE>(q, 0, 1.0, dx, bias, limiter, *ps[q]);
                continue;
            }
            for (int k = 0; k < subcycles && !rc; k++) rc = source(q, subcycles, dx, bias, limiter, *ps[q]);
            if (!rc) rc = sweep<MODE>(q, 0, 0.5, dx, bias, limiter, *ps[q]);
            if (!rc) rc = sweep<MODE>(q, 1, 1.0, dx, bias, limiter, *ps[q]);
            if (!rc) rc = sweep<MODE>(q, 0, 0.5, dx, bias, limiter, *ps[q]);
            for (int k = 0; k < subcycles && !rc; k++) rc = source(q, subcycles, dx, bias, limiter, *ps[q]);
        }
        return rc;
    }

    int run(int mode, double dx, double cfl, double t_final, int max_steps, double bias, int limiter, int subcycles,
            const MP &p1, const MP &p2, double *dt_log, int dt_cap, vg_run_stats *stats)
    {
        auto w0 = std::chrono::steady_clock::now();
        hipStream_t s = ctx->stream;
        Clock h{};
        h.t_final = t_final;
        h.cfl = cfl;
        h.dx = dx;
        h.max_steps = max_steps;
        VG_HIP(hipMemcpyAsync(clk(), &h, sizeof(Clock), hipMemcpyHostToDevice, s));
        if (dim == 2) {
            mhd_rgfm_corner_kernel<<<1, 1, 0, s>>>(p1, p2, clk());
            int rc = launch_status("mhd_rgfm_corner");
            if (rc) return rc;
        }
        if (dt_log && dt_cap > 0 && dt_dev_cap < dt_cap) {
            hipFree(dt_dev);
            dt_dev = nullptr;
            VG_HIP(hipMalloc(&dt_dev, sizeof(double) * dt_cap));
            dt_dev_cap = dt_cap;
        }
        double *dlog = (dt_log && dt_cap > 0) ? dt_dev : nullptr;
        // solve2D's GLM constants from the first step's advance on
        MP q1 = p1, q2 = p2;
        if (dim == 2) {
            q1.ch = 4.0;
            q2.ch = 4.0;
            q1.cp = std::sqrt(q1.ch * 0.18);
            q2.cp = std::sqrt(q2.ch * 0.18);
        }
        const int start[3] = {g[0].cur, g[1].cur, pcur};
        hipEvent_t e0, e1;
        VG_HIP(hipEventCreate(&e0));
        VG_HIP(hipEventCreate(&e1));
        VG_HIP(hipEventRecord(e0, s));
        const int chunk = (max_steps > 0 && max_steps <= 256) ? max_steps : 64;
        bool first = true;
        for (;;) {
            for (int k = 0; k < chunk; k++) {
                const MP &f1 = first ? p1 : q1, &f2 = first ? p2 : q2;
                int rc = (mode == VG_MODE_FAST)
                             ? enqueue_step<MODE_FAST>(f1, f2, q1, q2, dx, bias, limiter, subcycles, dlog, dt_cap)
                             : enqueue_step<MODE_EXACT>(f1, f2, q1, q2, dx, bias, limiter, subcycles, dlog, dt_cap);
                if (rc) return rc;
                first = false;
            }
            VG_HIP(hipMemcpyAsync(&h, clk(), sizeof(Clock), hipMemcpyDeviceToHost, s));
            VG_HIP(hipStreamSynchronize(s));
            if (h.done || (max_steps > 0 && h.steps_this_run >= max_steps)) break;
        }
        clock_finish_kernel<<<1, 1, 0, s>>>(clk());
        VG_HIP(hipEventRecord(e1, s));
        VG_HIP(hipMemcpyAsync(&h, clk(), sizeof(Clock), hipMemcpyDeviceToHost, s));
        VG_HIP(hipStreamSynchronize(s));
        float dev_ms = 0.0f;
        VG_HIP(hipEventElapsedTime(&dev_ms, e0, e1));
        hipEventDestroy(e0);
        hipEventDestroy(e1);
        // each executed step moves every buffer an odd number of times; the
        // steps enqueued after the clock stopped move nothing
        const int odd = h.steps_this_run & 1;
        g[0].cur = start[0] ^ odd;
        g[1].cur = start[1] ^ odd;
        pcur = start[2] ^ odd;
        if (dlog) {
            const int n = std::min(h.steps_this_run, dt_cap);
            if (n > 0) {
                VG_HIP(hipMemcpyAsync(dt_log, dt_dev, sizeof(double) * n, hipMemcpyDeviceToHost, s));
                VG_HIP(hipStreamSynchronize(s));
            }
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
            set_error("MHD RGFM: a time step could not advance t (non-finite or zero wave speed)");
            return VG_ERR_NONFINITE;
        }
        return VG_OK;
    }
};

int check_params(const vg_mhd_params *p1, const vg_mhd_params *p2)
{
    VG_CHECK(p1 != nullptr && p2 != nullptr, "null material parameters");
    VG_CHECK(p1->gamma != 1.0 && p2->gamma != 1.0, "gamma must differ from 1");
    return VG_OK;
}

int solve(vg_ctx *ctx, int dim, const double *m1_in, const double *m2_in, const double *phi_in, double *m1_out,
          double *m2_out, double *phi_out, int ny, int nx, double dx, double cfl, double t_final, int max_steps,
          double bias, int limiter, int subcycles, const vg_mhd_params *p1, const vg_mhd_params *p2, int mode,
          double *dt_log, int dt_log_cap, vg_run_stats *stats)
{
    clear_error();
    VG_CHECK(ctx != nullptr, "null context");
    VG_CHECK(m1_in && m2_in && phi_in && m1_out && m2_out && phi_out, "null state array");
    int rc = check_params(p1, p2);
    if (rc) return rc;
    VG_CHECK(mode == VG_MODE_EXACT || mode == VG_MODE_FAST, "unknown mode");
    VG_CHECK(nx >= 2 && ny >= 1, "grid needs at least 2 cells per line");
    VG_CHECK(dim == 1 || ny >= 2, "2D grid needs at least 2 rows");
    VG_CHECK(nx <= kMaxLine && ny <= kMaxLine, "line length above the fill kernel's limit of 16384 cells");
    VG_CHECK(dx > 0.0 && cfl > 0.0, "cell spacing and CFL must be positive");
    VG_CHECK(subcycles >= 0, "subcycles must be >= 0");
    VG_CHECK(max_steps >= 0 && (max_steps > 0 || std::isfinite(t_final)), "an infinite final time needs a step cap");
    rc = bind(ctx);
    if (rc) return rc;
    MHDRGFMGrid G;
    rc = G.create(ctx, dim, ny, nx);
    if (!rc) rc = G.load(m1_in, m2_in, phi_in);
    vg_run_stats local{};
    if (!rc)
        rc = G.run(mode, dx, cfl, t_final, max_steps, bias, limiter, dim == 2 ? subcycles : 0, mhd_params(p1),
                   mhd_params(p2), dt_log, dt_log_cap, &local);
    if (!rc) rc = G.store(m1_out, m2_out, phi_out);
    G.destroy();
    if (stats) *stats = local;
    return rc;
}

}  // namespace
}  // namespace vg

extern "C" {

int vg_mhd_rgfm_solve_1d(vg_ctx *ctx, const double *m1_in, const double *m2_in, const double *phi_in,
                         double *m1_out, double *m2_out, double *phi_out, int n, double dx, double cfl,
                         double t_final, int max_steps, double bias, int limiter, int subcycles,
                         const vg_mhd_params *p1, const vg_mhd_params *p2, int mode, double *dt_log, int dt_log_cap,
                         vg_run_stats *stats)
{
    return vg::solve(ctx, 1, m1_in, m2_in, phi_in, m1_out, m2_out, phi_out, 1, n, dx, cfl, t_final, max_steps, bias,
                     limiter, subcycles, p1, p2, mode, dt_log, dt_log_cap, stats);
}

int vg_mhd_rgfm_solve_2d(vg_ctx *ctx, const double *m1_in, const double *m2_in, const double *phi_in,
                         double *m1_out, double *m2_out, double *phi_out, int ny, int nx, double dx, double cfl,
                         double t_final, int max_steps, double bias, int limiter, int subcycles,
                         const vg_mhd_params *p1, const vg_mhd_params *p2, int mode, double *dt_log, int dt_log_cap,
                         vg_run_stats *stats)
{
    return vg::solve(ctx, 2, m1_in, m2_in, phi_in, m1_out, m2_out, phi_out, ny, nx, dx, cfl, t_final, max_steps,
                     bias, limiter, subcycles, p1, p2, mode, dt_log, dt_log_cap, stats);
}

int vg_mhd_rgfm_ghost_fill(vg_ctx *ctx, int dim, int ny, int nx, double *m1, double *m2, const double *phi,
                           const vg_mhd_params *p1, const vg_mhd_params *p2)
{
    vg::clear_error();
    VG_CHECK(ctx != nullptr && m1 && m2 && phi, "null argument");
    int rc = vg::check_params(p1, p2);
    if (rc) return rc;
    VG_CHECK(dim == 1 || dim == 2, "dim must be 1 or 2");
    if (dim == 1) ny = 1;
    VG_CHECK(nx >= 1 && ny >= 1, "empty grid");
    VG_CHECK(nx <= vg::kMaxLine && ny <= vg::kMaxLine, "line length above the fill kernel's limit of 16384 cells");
    rc = vg::bind(ctx);
    if (rc) return rc;
    vg::MHDRGFMGrid G;
    rc = G.create(ctx, dim, ny, nx);
    if (!rc) rc = G.load(m1, m2, phi);
    if (!rc) {
        // a running clock (done = 0) for the fill kernels
        vg::Clock h{};
        hipError_t e = hipMemcpyAsync(G.clk(), &h, sizeof(vg::Clock), hipMemcpyHostToDevice, ctx->stream);
        if (e != hipSuccess) {
            vg::set_error(std::string("hipMemcpyAsync: ") + hipGetErrorString(e));
            rc = VG_ERR_HIP;
        }
    }
    if (!rc) rc = G.fill(vg::mhd_params(p1), vg::mhd_params(p2), G.clk());
    if (!rc) rc = G.store(m1, m2, nullptr);
    G.destroy();
    return rc;
}

}  // extern "C"
