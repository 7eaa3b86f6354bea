// RGFM C ABI: two-material Euler with the real ghost fluid method, 1D and 2D
// whole solves and a batch of 2D instances, all device resident.
//
// One step (RGFMSolver::solveHLLC/solveHLLC2D, rgfmsolver.cpp:745-840):
//   loop test -> ghost fill (rows, then columns) -> wave-speed maximum of both
//   materials -> dt -> level set X(dt/2) Y(dt) X(dt/2) -> SLIC X(dt/2) Y(dt) X(dt/2)
//   for material 1, then material 2.
// 1D is the same with one row, one level-set pass and one sweep per material.
#include <algorithm>
#include <atomic>
#include <thread>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "vg_adjoint_host.h"
#include "vg_grid.h"
#include "vg_rgfm.h"
#include "vg_rgfm_adjoint.h"
#include "vg_runtime.h"

namespace vg {

// The inside of one step, kept for the adjoint: both materials after the x
// ghost fill (P1) and after the y fill (P2, the first x sweep's input), after
// the first x sweep (Q1, the y sweep's input) and after the y sweep (Q2, the
// last x sweep's input).  1D steps keep P1 only.
struct StepInternals {
    double *P1[2], *P2[2], *Q1[2], *Q2[2];
};

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
    bool reduced = false;               // this run sweeps with the 4-field EulerPlanarPhys
    int *flag = nullptr;                // constant-plane check result
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
        hipFree(flag);
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

    // One ghost-fill pass, HLLC, in place on (a1, a2) with level set ph, under
    // the clock c (which must not be done).
    int fill_pass(int dir, double *a1, double *a2, const double *ph, const EP &p1, const EP &p2, Clock *c)
    {
        LineArgs la = line_args(VG_RIEMANN_HLLC, p1, p2);
        la.m1 = a1;
        la.m2 = a2;
        la.phi = ph;
        la.clk = c;
        if (dir == 0) rgfm_fill_kernel<0><<<ny, kLineBS, 0, ctx->stream>>>(la);
        else rgfm_fill_kernel<1><<<nx, kLineBS, 0, ctx->stream>>>(la);
        return launch_status("rgfm fill (recomputation)");
    }

    template <class Q, int MODE>
    int sweep_as(int dir, const double *in, double *out, double scale, double h, double bias, int limiter,
                 const EP &m, Clock *use_clk)
    {
        SweepArgs<Q> a;
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
        a.clk = use_clk ? use_clk : clk;
        a.partials = nullptr;
        if (dir == 0) {
            dim3 grid((nx + (BSX - 4) - 1) / (BSX - 4), (ny + RX - 1) / RX);
            return ctx->timed("rgfm_sweep_x", [&] {
                sweep_x_kernel<Q, 2, MODE, BSX, RX><<<grid, BSX, 0, ctx->stream>>>(a);
            });
        }
        return launch_y<Q, 2, MODE>(ctx, y_variant, a);
    }

    template <int MODE>
    int sweep(int dir, const double *in, double *out, double scale, double h, double bias, int limiter, const EP &m,
              Clock *use_clk = nullptr)
    {
        if (reduced) return sweep_as<EulerPlanarPhys, MODE>(dir, in, out, scale, h, bias, limiter, m, use_clk);
        return sweep_as<EulerPhys, MODE>(dir, in, out, scale, h, bias, limiter, m, use_clk);
    }

    // The 4-field sweeps for this run when both materials hold w = +0 and
    // lambda = 1 in every cell of the current buffers.  The ghost fill writes
    // only rho, the normal velocity and p, and the 6-field sweeps keep those two
    // planes at their bits while computing the same rho, u, v, p as the 4-field
    // ones, so the results are the same bits.  The other buffers receive the two
    // planes here; the sweeps then leave them alone.
    int choose_physics()
    {
        reduced = false;
        if (!ctx->reduce_fields) return VG_OK;
        using R = ReducedPhys<EulerPhys>;
        hipStream_t s = ctx->stream;
        if (!flag) VG_HIP(hipMalloc(&flag, sizeof(int)));
        const int one = 1;
        VG_HIP(hipMemcpyAsync(flag, &one, sizeof(int), hipMemcpyHostToDevice, s));
        constant_planes_kernel<R><<<blocks_for(cells), 256, 0, s>>>(m1[cur], plane, cells, flag);
        constant_planes_kernel<R><<<blocks_for(cells), 256, 0, s>>>(m2[cur], plane, cells, flag);
        int rc = launch_status("constant_planes");
        if (rc) return rc;
        int ok = 0;
        VG_HIP(hipMemcpyAsync(&ok, flag, sizeof(int), hipMemcpyDeviceToHost, s));
        VG_HIP(hipStreamSynchronize(s));
        if (!ok) return VG_OK;
        for (int i = 0; i < R::kConst; i++) {
            const long long off = (long long)R::plane(i) * plane;
            VG_HIP(hipMemcpyAsync(m1[1 - cur] + off, m1[cur] + off, sizeof(double) * cells, hipMemcpyDeviceToDevice,
                                  s));
            VG_HIP(hipMemcpyAsync(m2[1 - cur] + off, m2[cur] + off, sizeof(double) * cells, hipMemcpyDeviceToDevice,
                                  s));
        }
        reduced = true;
        return VG_OK;
    }

    // y-sweep shape of this run's physics (shared with Euler grids of the same
    // shape and physics).
    template <class Q>
    int tune_as(int mode, double dx, double dy, double bias, int limiter, const EP &m)
    {
        if (!tune_clk) VG_HIP(hipMalloc(&tune_clk, sizeof(Clock)));
        Clock tc{};
        tc.dt = 1e-3 * std::min(dx, dy);
        VG_HIP(hipMemcpyAsync(tune_clk, &tc, sizeof(Clock), hipMemcpyHostToDevice, ctx->stream));
        SweepArgs<Q> ta;
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
        ta.m = m;
        ta.clk = tune_clk;
        ta.partials = nullptr;
        const std::string key = tune_key(Q::name, 2, mode, ny, nx);
        return (mode == VG_MODE_FAST) ? tune_y<Q, 2, MODE_FAST>(ctx, key, ta, y_variant)
                                      : tune_y<Q, 2, MODE_EXACT>(ctx, key, ta, y_variant);
    }

    // Device copies of both materials' grids: whole, or the planes the 4-field
    // sweeps carry (rho, u, v: planes 0-2; p: plane 4).
    int copy_grids(double *d1, const double *s1, double *d2, const double *s2, size_t bytes)
    {
        VG_HIP(hipMemcpyAsync(d1, s1, bytes, hipMemcpyDeviceToDevice, ctx->stream));
        VG_HIP(hipMemcpyAsync(d2, s2, bytes, hipMemcpyDeviceToDevice, ctx->stream));
        return VG_OK;
    }
    int copy_carried(double *d1, const double *s1, double *d2, const double *s2)
    {
        const size_t pb = sizeof(double) * plane;
        for (int k = 0; k < 2; k++) {
            double *d = k ? d2 : d1;
            const double *src = k ? s2 : s1;
            VG_HIP(hipMemcpyAsync(d, src, 3 * pb, hipMemcpyDeviceToDevice, ctx->stream));
            VG_HIP(hipMemcpyAsync(d + 4 * plane, src + 4 * plane, pb, hipMemcpyDeviceToDevice, ctx->stream));
        }
        return VG_OK;
    }

    // One step; with fixed_dts the time step is fixed_dts[it] (the clock's
    // step counter) instead of the CFL formula.
    template <int MODE>
    // keep (optional): where to leave the step's inside for the adjoint; the
    // arithmetic is the same either way.
    int enqueue_step(int riemann, double dx, double dy, double bias, int limiter, const EP &p1, const EP &p2,
                     double *dt_log, int dt_cap, const double *fixed_dts = nullptr, int n_fixed = 0,
                     const StepInternals *keep = nullptr)
    {
        hipStream_t s = ctx->stream;
        clock_check_kernel<<<1, 1, 0, s>>>(clk);
        LineArgs la = line_args(riemann, p1, p2);
        const size_t gbytes = sizeof(double) * 6 * plane;
        // with the 4-field sweeps only rho, u, v and p are copied: the caller
        // has set the other two planes of the destinations
        auto snapshot = [&](double *const (&dst)[2]) {
            return reduced ? copy_carried(dst[0], m1[cur], dst[1], m2[cur])
                           : copy_grids(dst[0], m1[cur], dst[1], m2[cur], gbytes);
        };
        int rc = ctx->timed("rgfm_fill_x", [&] { rgfm_fill_kernel<0><<<ny, kLineBS, 0, s>>>(la); });
        if (!rc && keep) rc = snapshot(keep->P1);
        if (!rc && dim == 2) rc = ctx->timed("rgfm_fill_y", [&] { rgfm_fill_kernel<1><<<nx, kLineBS, 0, s>>>(la); });
        if (!rc && keep && dim == 2) rc = snapshot(keep->P2);
        if (rc) return launch_status("rgfm_fill");
        if (fixed_dts) {
            clock_fixed_kernel<<<1, 1, 0, s>>>(clk, fixed_dts, n_fixed);
        } else {
            wave_max_kernel<EulerPhys><<<max_blocks, 256, 0, s>>>(m1[cur], plane, cells, dim, p1, partials);
            wave_max_kernel<EulerPhys><<<max_blocks, 256, 0, s>>>(m2[cur], plane, cells, dim, p2,
                                                                  partials + max_blocks);
            clock_begin_kernel<<<1, kReduceThreads, 0, s>>>(clk, partials, 2 * max_blocks, dt_log, dt_cap);
        }
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
            if (!rc)
                rc = ctx->timed("rgfm_level_set", [&] {
                    rgfm_level_set_y_kernel<<<(nx + kLsCols - 1) / kLsCols, kLsCols * kLsSegs, 0, s>>>(a);
                });
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
                // with keep, the two intermediate states go to the record
                double *X1 = keep ? keep->Q1[q] : B, *X2 = keep ? keep->Q2[q] : A;
                rc = sweep<MODE>(0, A, X1, 0.5, dx, bias, limiter, *ps[q]);
                if (!rc) rc = sweep<MODE>(1, X1, X2, 1.0, dy, bias, limiter, *ps[q]);
                if (!rc) rc = sweep<MODE>(0, X2, B, 0.5, dx, bias, limiter, *ps[q]);
            }
        }
        if (rc) return launch_status("rgfm sweep");
        cur = 1 - cur;
        return VG_OK;
    }

    // Runs to t_final or max_steps.  Each executed step moves both materials
    // and the level set to the other buffer once.
    // With resume the clock continues (t, step count, start-up ramp) from the
    // previous run on this grid and max_steps counts from there.
    int run(int riemann, int mode, double dx, double dy, double cfl, double t_final, int max_steps, double bias,
            int limiter, const EP &p1, const EP &p2, double *dt_log, int dt_cap, vg_run_stats *stats,
            bool resume = false)
    {
        auto w0 = std::chrono::steady_clock::now();
        hipStream_t s = ctx->stream;
        Clock h{};
        if (resume) {
            Clock prev;
            VG_HIP(hipMemcpyAsync(&prev, clk, sizeof(Clock), hipMemcpyDeviceToHost, s));
            VG_HIP(hipStreamSynchronize(s));
            h.t = prev.t;
            h.it = prev.it;
        }
        h.t_final = t_final;
        h.cfl = cfl;
        h.dx = (dim == 1) ? dx : std::min(dx, dy);
        h.max_steps = (max_steps > 0) ? h.it + max_steps : 0;
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
        int prc = choose_physics();
        if (prc) return prc;
        // every exit of this run returns to the 6-field sweeps: steps enqueued
        // outside a run (the fixed-dt forward, the adjoint's recording
        // recomputation) write all six planes of their outputs
        struct Reset {
            bool &r;
            ~Reset() { r = false; }
        } reset{reduced};
        if (dim == 2) {
            int rc = reduced ? tune_as<EulerPlanarPhys>(mode, dx, dy, bias, limiter, p1)
                             : tune_as<EulerPhys>(mode, dx, dy, bias, limiter, p1);
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
                   std::to_string(y_variant) + "|" + std::to_string(cur) + "|" + std::to_string(pcur) + "|" +
                   std::to_string((int)reduced);
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
                // a capped run is complete once its last step has started (the
                // clock folds that step in at the next check, or at the finish)
                if (h.done || (max_steps > 0 && h.steps_this_run >= max_steps)) break;
            }
        } else {
            for (;;) {
                for (int k = 0; k < chunk; k++) {
                    int rc = enqueue();
                    if (rc) return rc;
                }
                VG_HIP(hipMemcpyAsync(&h, clk, sizeof(Clock), hipMemcpyDeviceToHost, s));
                VG_HIP(hipStreamSynchronize(s));
                if (h.done || (max_steps > 0 && h.steps_this_run >= max_steps)) break;
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

// Final-state check on the host copies: in every cell the material the level
// set makes real there (material 1 where phi < 0, else material 2) must be
// finite with rho > 0 and p + p_inf > 0, and phi finite.  Counts the cells that
// fail and the first one's flat index (-1 if none).
void scan_final(const double *m1, const double *m2, const double *phi, size_t cells, const vg_euler_params *p1,
                const vg_euler_params *p2, int64_t &count, int64_t &first)
{
    count = 0;
    first = -1;
    for (size_t c = 0; c < cells; c++) {
        const bool one = phi[c] < 0.0;
        const double *w = (one ? m1 : m2) + 6 * c;
        bool bad = !std::isfinite(phi[c]) || !(w[0] > 0.0) || !((w[4] + (one ? p1 : p2)->p_inf) > 0.0);
        for (int f = 0; f < 6; f++) bad = bad || !std::isfinite(w[f]);
        if (bad) {
            if (first < 0) first = (int64_t)c;
            count += 1;
        }
    }
}

int nonfinite_error(int64_t count, int64_t first)
{
    vg::set_error("final state holds " + std::to_string(count) +
                  " non-finite or inadmissible cells (rho <= 0 or p + p_inf <= 0 in the material the level set "
                  "selects), first at flat index " + std::to_string(first));
    return VG_ERR_NONFINITE;
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
    if (!rc) scan_final(m1_out, m2_out, phi_out, (size_t)ny * nx, p1, p2, local.nonfinite_cells, local.first_nonfinite);
    if (stats) *stats = local;
    if (rc || run_rc) return rc ? rc : run_rc;
    return local.nonfinite_cells > 0 ? nonfinite_error(local.nonfinite_cells, local.first_nonfinite) : VG_OK;
}

// ---- adjoint (no reference counterpart) ---------------------------------------------

// The fixed-time-step map the adjoint differentiates: steps as `solve` takes
// them, with dts[it] for the time step.
int solve_fixed(vg_ctx *ctx, int dim, const double *m1_in, const double *m2_in, const double *phi_in,
                double *m1_out, double *m2_out, double *phi_out, int ny, int nx, double dx, double dy,
                const double *dts, int n_steps, double bias, int limiter, const vg_euler_params *p1,
                const vg_euler_params *p2)
{
    vg::clear_error();
    VG_CHECK(ctx != nullptr, "null context");
    VG_CHECK(m1_in && m2_in && phi_in && m1_out && m2_out && phi_out && (dts || n_steps == 0) && n_steps >= 0,
             "null argument");
    int rc = check_args(dim, ny, nx, dx, dy, 1.0, 1.0, 0, limiter, VG_RIEMANN_HLLC, p1, p2, VG_MODE_EXACT);
    if (rc) return rc;
    rc = vg::bind(ctx);
    if (rc) return rc;
    const vg::EP q1 = to_params(p1), q2 = to_params(p2);
    vg::RGFMGrid g;
    rc = g.create(ctx, dim, ny, nx);
    if (!rc) rc = g.load(m1_in, m2_in, phi_in);
    double *dts_dev = nullptr;
    if (!rc) {
        void *v = nullptr;
        rc = ctx->workspace_buffer(0, sizeof(double) * std::max(n_steps, 1), &v);
        dts_dev = static_cast<double *>(v);
    }
    if (!rc && n_steps > 0)
        VG_HIP(hipMemcpyAsync(dts_dev, dts, sizeof(double) * n_steps, hipMemcpyHostToDevice, ctx->stream));
    if (!rc) {
        vg::Clock h{};
        h.t_final = INFINITY;
        h.max_steps = n_steps;
        VG_HIP(hipMemcpyAsync(g.clk, &h, sizeof(h), hipMemcpyHostToDevice, ctx->stream));
    }
    for (int st = 0; st < n_steps && !rc; st++)
        rc = g.enqueue_step<vg::MODE_EXACT>(VG_RIEMANN_HLLC, dx, dy, bias, limiter, q1, q2, nullptr, 0, dts_dev,
                                            n_steps);
    if (!rc) rc = g.store(m1_out, m2_out, phi_out);
    g.destroy();
    return rc;
}

// Reverse mode through the HLLC RGFM solve, time steps frozen.  Forward with a
// checkpoint (both materials and the level set) every K steps; then, segment by
// segment from the end, the segment is recomputed from its checkpoint keeping
// each step's starting state and inside (StepInternals), and its steps, from
// the last, are reversed: the SLIC sweeps of both materials, then the ghost
// fills (y, then x).  K is ceil(sqrt(steps)) (64 for runs to t_final), or
// less when K kept steps (10 material grids and a level set each in 2D) would
// not fit in 60 % of free device memory.
int adjoint(vg_ctx *ctx, int dim, const double *m1_in, const double *m2_in, const double *phi_in, double *m1_out,
            double *m2_out, double *phi_out, const double *seed1, const double *seed2, double *grad1, double *grad2,
            int ny, int nx, double dx, double dy, double cfl, double t_final, int max_steps, double bias, int limiter,
            const vg_euler_params *p1, const vg_euler_params *p2, int checkpoint_every, double *dt_log, int dt_cap,
            vg_run_stats *stats, double mem_budget = 0.0, int n_est = 0)
{
    vg::clear_error();
    VG_CHECK(ctx != nullptr, "null context");
    VG_CHECK(m1_in && m2_in && phi_in && m1_out && m2_out && phi_out && seed1 && seed2 && grad1 && grad2,
             "null state pointer");
    VG_CHECK(checkpoint_every >= 0, "checkpoint_every must be >= 0");
    int rc = check_args(dim, ny, nx, dx, dy, cfl, t_final, max_steps, limiter, VG_RIEMANN_HLLC, p1, p2, VG_MODE_EXACT);
    if (rc) return rc;
    rc = vg::bind(ctx);
    if (rc) return rc;
    auto w0 = std::chrono::steady_clock::now();
    const vg::EP q1 = to_params(p1), q2 = to_params(p2);
    hipStream_t s = ctx->stream;
    vg::RGFMGrid g;
    rc = g.create(ctx, dim, ny, nx);
    if (!rc) rc = g.load(m1_in, m2_in, phi_in);
    // the forward takes the 4-field sweeps when both materials hold w = +0 and
    // lambda = 1 in every cell (its run decides the same way)
    if (!rc) rc = g.choose_physics();
    const bool planar_fwd = g.reduced;
    g.reduced = false;
    if (rc) {
        g.destroy();
        return rc;
    }
    const size_t gb = sizeof(double) * 6 * g.plane, pb = sizeof(double) * g.cells;
    const int n_run = max_steps > 0 ? max_steps : n_est;   // steps of the run, when known or predicted
    int K = checkpoint_every > 0 ? checkpoint_every
                                 : (n_run > 0 ? std::max(1, (int)std::ceil(std::sqrt((double)n_run))) : 64);
    constexpr int kKept = 11;   // per kept step: m1, m2, phi, P1 x2, P2 x2, Q1 x2, Q2 x2
    if (checkpoint_every == 0) {
        const double per = 10.0 * (double)gb + (double)pb;
        if (mem_budget > 0.0) {
            // a share of the device (the batch): the kept steps and the
            // checkpoints of a run of n_run steps within the budget
            const double ckb = 2.0 * (double)gb + (double)pb;
            auto need = [&](int k) { return (double)k * per + (n_run > 0 ? std::ceil((double)n_run / k) * ckb : 0.0); };
            while (K > 1 && need(K) > mem_budget) K--;
        } else {
            size_t free_b = 0, total_b = 0;
            if (hipMemGetInfo(&free_b, &total_b) == hipSuccess) {
                const double avail = 0.6 * (double)(free_b + ctx->workspace_bytes());
                K = std::max(1, std::min(K, (int)(avail / per)));
            }
        }
    }
    // workspace: 0 dts, 1 flags, 2 ghost adjoints, 3-6 adjoint ping-pong
    // (A[b][q]), 7-10 reverse temporaries (T1, T2 per material), 11 the fill
    // adjoints' VJP scratch, then K kept steps (kKept buffers each) and the
    // checkpoints (m1, m2, phi)
    auto wsp = [&](size_t i, size_t bytes, double **p) {
        void *v = nullptr;
        int r = ctx->workspace_buffer(i, bytes, &v);
        *p = static_cast<double *>(v);
        return r;
    };
    const size_t st0 = 12, ck0 = st0 + kKept * (size_t)K;
    // ---- forward with checkpoints ---------------------------------------------------
    std::vector<double> dts;
    std::vector<double *> ck;   // 3 per checkpoint
    const int cap = max_steps > 0 ? max_steps : 1 << 22;
    vg_run_stats fst{};
    bool first = true;
    for (;;) {
        if ((int)dts.size() >= cap) break;
        double *c1 = nullptr, *c2 = nullptr, *cp = nullptr;
        const size_t base = ck0 + ck.size();
        if ((rc = wsp(base, gb, &c1)) || (rc = wsp(base + 1, gb, &c2)) || (rc = wsp(base + 2, pb, &cp))) break;
        VG_HIP(hipMemcpyAsync(c1, g.m1[g.cur], gb, hipMemcpyDeviceToDevice, s));
        VG_HIP(hipMemcpyAsync(c2, g.m2[g.cur], gb, hipMemcpyDeviceToDevice, s));
        VG_HIP(hipMemcpyAsync(cp, g.phi[g.pcur], pb, hipMemcpyDeviceToDevice, s));
        ck.push_back(c1);
        ck.push_back(c2);
        ck.push_back(cp);
        std::vector<double> part(K);
        vg_run_stats st{};
        const int want = std::min(K, cap - (int)dts.size());
        rc = g.run(VG_RIEMANN_HLLC, VG_MODE_EXACT, dx, dy, cfl, t_final, want, bias, limiter, q1, q2, part.data(), K,
                   &st, !first);
        first = false;
        if (rc) break;
        dts.insert(dts.end(), part.begin(), part.begin() + st.steps);
        fst = st;
        if (st.steps < want) break;
    }
    if (rc) {
        g.destroy();
        return rc;
    }
    const int steps = (int)dts.size();
    rc = g.store(m1_out, m2_out, phi_out);
    if (!rc) {
        // A forward that ends with inadmissible real cells has no derivative:
        // report it as solve() does; the gradients are set to NaN.
        int64_t bad = 0, first_bad = -1;
        scan_final(m1_out, m2_out, phi_out, (size_t)ny * nx, p1, p2, bad, first_bad);
        if (bad > 0) {
            const size_t n6 = 6 * (size_t)ny * nx;
            std::fill(grad1, grad1 + n6, std::nan(""));
            std::fill(grad2, grad2 + n6, std::nan(""));
            if (dt_log) std::memcpy(dt_log, dts.data(), sizeof(double) * std::min(steps, std::max(dt_cap, 0)));
            if (stats) {
                *stats = fst;
                stats->steps = stats->total_steps = steps;
                stats->nonfinite_cells = bad;
                stats->first_nonfinite = first_bad;
            }
            g.destroy();
            return nonfinite_error(bad, first_bad);
        }
    }
    // ---- buffers of the reverse pass --------------------------------------------------
    double *dts_dev = nullptr, *ghost = nullptr, *A[2][2], *T1[2], *T2[2], *flag_buf = nullptr, *vjp = nullptr;
    if (!rc) rc = wsp(0, sizeof(double) * std::max(steps, 1), &dts_dev);
    if (!rc) rc = wsp(1, sizeof(double), &flag_buf);
    int *bad_step = reinterpret_cast<int *>(flag_buf);   // latest reversed step after which the adjoint is non-finite
    int *seed_flag = bad_step + 1;                       // the seeds' w and lambda parts are zero
    if (!rc) VG_HIP(hipMemsetAsync(bad_step, 0xFF, sizeof(int), s));
    if (!rc) rc = wsp(2, sizeof(double) * 4 * 6 * (size_t)std::max(nx, ny), &ghost);
    for (int b = 0; b < 2 && !rc; b++)
        for (int q = 0; q < 2 && !rc; q++) rc = wsp(3 + 2 * b + q, gb, &A[b][q]);
    for (int q = 0; q < 2 && !rc; q++) rc = wsp(7 + q, gb, &T1[q]);
    for (int q = 0; q < 2 && !rc; q++) rc = wsp(9 + q, gb, &T2[q]);
    if (!rc) rc = wsp(11, 2 * gb, &vjp);
    // kept step k: sts[3k .. 3k+2] = starting m1, m2, phi; ins[k] = its inside
    std::vector<double *> sts(3 * (size_t)K, nullptr);
    std::vector<vg::StepInternals> ins((size_t)K);
    for (int k = 0; k < K && !rc; k++) {
        const size_t b = st0 + (size_t)kKept * k;
        if ((rc = wsp(b, gb, &sts[3 * k]))) break;
        if ((rc = wsp(b + 1, gb, &sts[3 * k + 1]))) break;
        if ((rc = wsp(b + 2, pb, &sts[3 * k + 2]))) break;
        double **slots[4] = {ins[k].P1, ins[k].P2, ins[k].Q1, ins[k].Q2};
        for (int j = 0; j < 4 && !rc; j++)
            for (int q = 0; q < 2 && !rc; q++) rc = wsp(b + 3 + 2 * j + q, gb, &slots[j][q]);
    }
    if (rc) {
        g.destroy();
        return rc;
    }
    if (steps > 0) VG_HIP(hipMemcpyAsync(dts_dev, dts.data(), sizeof(double) * steps, hipMemcpyHostToDevice, s));
    // seeds: arrays of structures -> planes
    VG_HIP(hipMemcpyAsync(g.staging, seed1, gb / g.plane * g.cells, hipMemcpyHostToDevice, s));
    vg::aos_to_soa_kernel<6><<<vg::RGFMGrid::blocks_for(g.cells), 256, 0, s>>>(g.staging, A[0][0], g.plane, g.cells);
    VG_HIP(hipStreamSynchronize(s));
    VG_HIP(hipMemcpyAsync(g.staging, seed2, gb / g.plane * g.cells, hipMemcpyHostToDevice, s));
    vg::aos_to_soa_kernel<6><<<vg::RGFMGrid::blocks_for(g.cells), 256, 0, s>>>(g.staging, A[0][1], g.plane, g.cells);
    // The 4-field adjoint sweeps (AdjPlanar) when the forward took the 4-field
    // sweeps, E_f = 0 for both materials and both seeds have zero w and lambda
    // parts: the w and lambda adjoints then stay zero (the ghost fill's VJP
    // gives them exact zeros: the interface values do not depend on lambda,
    // and on w only through squares at w = +0).  The recomputation also takes
    // the 4-field sweeps and copies four planes; the kept buffers get planes 3
    // and 5 once, from the first checkpoint.
    bool planar = false;
    if (planar_fwd && q1.e_form == 0.0 && q2.e_form == 0.0) {
        const int one = 1;
        int ok = 0;
        VG_HIP(hipMemcpyAsync(seed_flag, &one, sizeof(int), hipMemcpyHostToDevice, s));
        for (int q = 0; q < 2; q++)
            vg::zero_planes_kernel<<<vg::RGFMGrid::blocks_for(g.cells), 256, 0, s>>>(A[0][q], g.plane, g.cells,
                                                                                    seed_flag);
        rc = vg::launch_status("adjoint seed check");
        if (!rc) VG_HIP(hipMemcpyAsync(&ok, seed_flag, sizeof(int), hipMemcpyDeviceToHost, s));
        if (!rc) VG_HIP(hipStreamSynchronize(s));
        planar = !rc && ok;
    }
    if (planar) {
        for (int q = 0; q < 2; q++)
            for (int f = 3; f <= 5; f += 2)
                VG_HIP(hipMemsetAsync(A[1][q] + f * g.plane, 0, sizeof(double) * g.cells, s));
        for (int k = 0; k < K; k++) {
            double *dst[6] = {sts[3 * k], sts[3 * k + 1], ins[k].P1[0], ins[k].P1[1], ins[k].P2[0], ins[k].P2[1]};
            for (int j = 0; j < (dim == 2 ? 6 : 4); j++)
                for (int f = 3; f <= 5; f += 2)
                    VG_HIP(hipMemcpyAsync(dst[j] + f * g.plane, ck[j % 2] + f * g.plane, sizeof(double) * g.cells,
                                          hipMemcpyDeviceToDevice, s));
        }
    }
    auto sweep_adj = [&](int dir, int variant, const double *w, const double *wbar_out, double *wbar_in, double h,
                         double ts, const vg::EP &m, const double *w35) {
        if (planar)
            return vg::adj_sweep<vg::AdjPlanar>(ctx, dir, variant, w, wbar_out, wbar_in, ghost, g.plane, ny, nx,
                                                VG_BC_TRANSMISSIVE, h, bias, limiter, ts, m, w35);
        return vg::adj_sweep<vg::AdjFull>(ctx, dir, variant, w, wbar_out, wbar_in, ghost, g.plane, ny, nx,
                                          VG_BC_TRANSMISSIVE, h, bias, limiter, ts, m);
    };
    int xv = vg::kAdjDefault[0], yv = vg::kAdjDefault[1];
    if (steps > 0 && !rc) {
        const vg::AdjointArgs ta =
            vg::adjoint_args(ck[0], A[0][0], A[1][0], ghost, g.plane, ny, nx, VG_BC_TRANSMISSIVE, dx, bias, limiter,
                             dts[0], q1);
        if (planar) {
            rc = vg::tune_adjoint<vg::AdjPlanar>(ctx, 0, ta, xv);
            if (!rc && dim == 2) rc = vg::tune_adjoint<vg::AdjPlanar>(ctx, 1, ta, yv);
        } else {
            rc = vg::tune_adjoint(ctx, 0, ta, xv);
            if (!rc && dim == 2) rc = vg::tune_adjoint(ctx, 1, ta, yv);
        }
    }
    const vg::EP *qs[2] = {&q1, &q2};
    int cw = 0;   // the adjoint of the state after the step being reversed is in A[cw]
    const int nseg = (int)ck.size() / 3;
    for (int seg = nseg - 1; seg >= 0 && !rc; seg--) {
        const int s0 = seg * K, s1 = std::min(steps, s0 + K);
        if (s0 >= s1) continue;
        // recompute the segment, keeping each step's starting state and inside
        g.cur = 0;
        g.pcur = 0;
        VG_HIP(hipMemcpyAsync(g.m1[0], ck[3 * seg], gb, hipMemcpyDeviceToDevice, s));
        VG_HIP(hipMemcpyAsync(g.m2[0], ck[3 * seg + 1], gb, hipMemcpyDeviceToDevice, s));
        VG_HIP(hipMemcpyAsync(g.phi[0], ck[3 * seg + 2], pb, hipMemcpyDeviceToDevice, s));
        vg::Clock h{};
        h.it = s0;
        h.t_final = INFINITY;
        h.max_steps = s1;
        VG_HIP(hipMemcpyAsync(g.clk, &h, sizeof(h), hipMemcpyHostToDevice, s));
        g.reduced = planar;
        for (int st = s0; st < s1 && !rc; st++) {
            const int k = st - s0;
            rc = planar ? g.copy_carried(sts[3 * k], g.m1[g.cur], sts[3 * k + 1], g.m2[g.cur])
                        : g.copy_grids(sts[3 * k], g.m1[g.cur], sts[3 * k + 1], g.m2[g.cur], gb);
            VG_HIP(hipMemcpyAsync(sts[3 * k + 2], g.phi[g.pcur], pb, hipMemcpyDeviceToDevice, s));
            if (!rc)
                rc = g.enqueue_step<vg::MODE_EXACT>(VG_RIEMANN_HLLC, dx, dy, bias, limiter, q1, q2, nullptr, 0, dts_dev,
                                                    s1, &ins[k]);
        }
        g.reduced = false;
        // reverse the steps of the segment
        for (int st = s1 - 1; st >= s0 && !rc; st--) {
            const int k = st - s0;
            double *S[2] = {sts[3 * k], sts[3 * k + 1]};
            const double *phs = sts[3 * k + 2];
            const double dt = dts[st];
            double *const *P1 = ins[k].P1, *const *Q1 = ins[k].Q1, *const *Q2 = ins[k].Q2;
            double *post[2] = {dim == 2 ? ins[k].P2[0] : P1[0], dim == 2 ? ins[k].P2[1] : P1[1]};
            // reverse: sweeps of each material, A[cw] -> A[1 - cw]
            for (int q = 0; q < 2 && !rc; q++) {
                const double *w35 = ck[q];   // w = +0, lambda = 1 (4-field sweeps only)
                if (dim == 1) {
                    rc = sweep_adj(0, xv, post[q], A[cw][q], A[1 - cw][q], dx, dt, *qs[q], w35);
                } else {
                    rc = sweep_adj(0, xv, Q2[q], A[cw][q], T1[q], dx, 0.5 * dt, *qs[q], w35);
                    if (!rc) rc = sweep_adj(1, yv, Q1[q], T1[q], T2[q], dy, dt, *qs[q], w35);
                    if (!rc) rc = sweep_adj(0, xv, post[q], T2[q], A[1 - cw][q], dx, 0.5 * dt, *qs[q], w35);
                }
            }
            if (rc) break;
            // reverse: ghost fills (the y pass read P1, the x pass read S), in place
            vg::FillAdjArgs fa;
            fa.phi = phs;
            fa.adj1 = A[1 - cw][0];
            fa.adj2 = A[1 - cw][1];
            fa.vjp = vjp;
            fa.plane = g.plane;
            fa.ny = ny;
            fa.nx = nx;
            fa.init_flag = (dim == 2) ? 1 : 0;
            fa.p1 = q1;
            fa.p2 = q2;
            if (dim == 2) {
                fa.m1 = P1[0];
                fa.m2 = P1[1];
                rc = ctx->timed("rgfm_fill_adjoint",
                                [&] { vg::rgfm_fill_adjoint_kernel<1><<<nx, vg::kLineBS, 0, s>>>(fa); });
            }
            if (!rc) {
                fa.m1 = S[0];
                fa.m2 = S[1];
                rc = ctx->timed("rgfm_fill_adjoint",
                                [&] { vg::rgfm_fill_adjoint_kernel<0><<<ny, vg::kLineBS, 0, s>>>(fa); });
            }
            cw = 1 - cw;
            for (int q = 0; q < 2 && !rc; q++) {
                vg::adjoint_nonfinite_kernel<<<vg::RGFMGrid::blocks_for(g.cells), 256, 0, s>>>(A[cw][q], g.plane,
                                                                                              g.cells, st, bad_step);
                rc = vg::launch_status("adjoint_nonfinite");
            }
        }
    }
    for (int q = 0; q < 2 && !rc; q++) {
        vg::soa_to_aos_kernel<6><<<vg::RGFMGrid::blocks_for(g.cells), 256, 0, s>>>(A[cw][q], g.staging, g.plane,
                                                                                    g.cells);
        hipError_t e = hipMemcpyAsync(q == 0 ? grad1 : grad2, g.staging, sizeof(double) * 6 * g.cells,
                                      hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) {
            vg::set_error(std::string("rgfm adjoint copy: ") + hipGetErrorString(e));
            rc = VG_ERR_HIP;
        }
    }
    if (!rc && ctx->timing) rc = ctx->resolve_timing();
    int first_bad_step = -1;
    if (!rc) {
        VG_HIP(hipMemcpyAsync(&first_bad_step, bad_step, sizeof(int), hipMemcpyDeviceToHost, s));
        VG_HIP(hipStreamSynchronize(s));
    }
    if (dt_log) std::memcpy(dt_log, dts.data(), sizeof(double) * std::min(steps, std::max(dt_cap, 0)));
    if (stats) {
        *stats = fst;
        stats->steps = steps;
        stats->total_steps = steps;
        stats->dt_min = steps ? *std::min_element(dts.begin(), dts.end()) : 0.0;
        stats->dt_max = steps ? *std::max_element(dts.begin(), dts.end()) : 0.0;
        stats->wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - w0).count();
    }
    g.destroy();
    if (!rc) {
        // A gradient that is not finite is reported, with the forward step whose
        // reversal first produced a non-finite adjoint.
        const size_t n6 = 6 * (size_t)ny * nx;
        int64_t bad = 0, first = -1;
        for (size_t k = 0; k < 2 * n6; k++) {
            const double v = k < n6 ? grad1[k] : grad2[k - n6];
            if (!std::isfinite(v)) {
                if (first < 0) first = (int64_t)(k % n6) / 6;
                bad += 1;
            }
        }
        if (stats) {
            stats->nonfinite_cells = bad;
            stats->first_nonfinite = first;
        }
        if (bad > 0) {
            vg::set_error("gradient holds " + std::to_string(bad) + " non-finite values (first at cell " +
                          std::to_string(first) + "); the adjoint first became non-finite in the reverse of step " +
                          std::to_string(first_bad_step) + " of " + std::to_string(steps));
            return VG_ERR_NONFINITE;
        }
    }
    return rc;
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
    const size_t cells = (size_t)ny * nx;
    const vg::EP e1 = to_params(p1), e2 = to_params(p2);
    std::vector<vg_run_stats> per(n_inst);
    std::vector<char> nonfinite(n_inst, 0);
    // One instance on grid g of context c: load, run, store, circulation.
    auto solve_one = [&](vg::RGFMGrid &g, int q) {
        int r = g.load(m1_in + 6 * cells * q, m2_in + 6 * cells * q, phi_in + cells * q);
        if (!r) {
            r = g.run(riemann, mode, dx, dy, cfl, t_final[q], max_steps, bias, limiter, e1, e2, nullptr, 0, &per[q]);
            if (r == VG_ERR_NONFINITE) {
                nonfinite[q] = 1;
                r = VG_OK;
            }
        }
        if (!r) r = g.store(m1_out + 6 * cells * q, m2_out + 6 * cells * q, phi_out + cells * q);
        if (!r && circulation) r = g.circulation(dx, dy, circulation + q);
        if (!r) {
            scan_final(m1_out + 6 * cells * q, m2_out + 6 * cells * q, phi_out + cells * q, cells, p1, p2,
                       per[q].nonfinite_cells, per[q].first_nonfinite);
            if (per[q].nonfinite_cells > 0) nonfinite[q] = 1;
        }
        return r;
    };
    const int nw = ctx->timing ? 1 : std::min(n_inst, ctx->batch_streams);
    if (nw == 1) {
        vg::RGFMGrid g;
        rc = g.create(ctx, 2, ny, nx);
        for (int q = 0; q < n_inst && !rc; q++) rc = solve_one(g, q);
        g.destroy();
    } else {
        // Longest predicted run first: an instance's step count is about
        // t_final times its largest signal speed max(|u|, |v|) + c (the time
        // step's denominator), c^2 = gamma (p + p_inf) / rho.  The order only
        // affects how well the streams stay busy.
        std::vector<std::pair<double, int>> cost(n_inst);
        for (int q = 0; q < n_inst; q++) {
            double smax = 0.0;
            const double *ms[2] = {m1_in + 6 * cells * q, m2_in + 6 * cells * q};
            const vg_euler_params *ps[2] = {p1, p2};
            for (int k = 0; k < 2; k++)
                for (size_t c = 0; c < cells; c++) {
                    const double *w = ms[k] + 6 * c;
                    const double a2 = ps[k]->gamma * (w[4] + ps[k]->p_inf) / w[0];
                    if (a2 > 0.0 && std::isfinite(a2))
                        smax = std::max(smax, std::max(std::fabs(w[1]), std::fabs(w[2])) + std::sqrt(a2));
                }
            cost[q] = {-t_final[q] * smax, q};
        }
        std::stable_sort(cost.begin(), cost.end());
        // The y-sweep shape is tuned here, alone on the device, before the
        // workers start (they take it from this context): timings taken while
        // other streams run are disturbed.
        const bool tuned = ctx->tuned.count(vg::tune_key(vg::EulerPhys::name, 2, mode, ny, nx)) &&
                           ctx->tuned.count(vg::tune_key(vg::EulerPlanarPhys::name, 2, mode, ny, nx));
        if (ctx->autotune && !tuned) {
            vg::RGFMGrid g;
            const int q = cost[0].second;
            rc = g.create(ctx, 2, ny, nx);
            if (!rc) rc = g.load(m1_in + 6 * cells * q, m2_in + 6 * cells * q, phi_in + cells * q);
            if (!rc) rc = g.choose_physics();
            if (!rc)
                rc = g.reduced ? g.tune_as<vg::EulerPlanarPhys>(mode, dx, dy, bias, limiter, e1)
                               : g.tune_as<vg::EulerPhys>(mode, dx, dy, bias, limiter, e1);
            if (ctx->stream) hipStreamSynchronize(ctx->stream);
            g.destroy();
        }
        if (!rc) rc = vg::batch_workers(ctx, nw);
        std::atomic<int> next{0};
        std::atomic<bool> stop{false};
        std::vector<int> wrc(nw, VG_OK);
        std::vector<std::string> werr(nw);
        // Each worker runs instances from the shared queue on its own stream
        // and grid until the queue is empty or another worker has failed.
        auto worker = [&](int k) {
            vg_ctx *c = ctx->workers[k];
            int r = vg::bind(c);
            vg::RGFMGrid g;
            if (!r) r = g.create(c, 2, ny, nx);
            while (!r && !stop.load()) {
                const int i = next.fetch_add(1);
                if (i >= n_inst) break;
                r = solve_one(g, cost[i].second);
            }
            if (c->stream) hipStreamSynchronize(c->stream);
            g.destroy();
            if (r) {
                stop = true;
                werr[k] = vg_last_error();
            }
            wrc[k] = r;
        };
        if (!rc) {
            std::vector<std::thread> threads;
            for (int k = 0; k < nw; k++) threads.emplace_back(worker, k);
            for (auto &t : threads) t.join();
            vg::merge_worker_tuning(ctx);
            for (int k = 0; k < nw && !rc; k++)
                if (wrc[k]) {
                    rc = wrc[k];
                    vg::set_error(werr[k]);
                }
        }
    }
    // Totals; first_nonfinite is a flat index over the concatenated instances.
    vg_run_stats total{};
    total.first_nonfinite = -1;
    double dt_min = 0.0, dt_max = 0.0;
    bool any_nonfinite = false;
    for (int q = 0; q < n_inst; q++) {
        const vg_run_stats &st = per[q];
        if (steps) steps[q] = st.steps;
        if (t_out) t_out[q] = st.t;
        total.steps += st.steps;
        total.device_ms += st.device_ms;
        total.wall_ms += st.wall_ms;
        total.nonfinite_cells += st.nonfinite_cells;
        if (total.first_nonfinite < 0 && st.nonfinite_cells > 0)
            total.first_nonfinite = (int64_t)cells * q + st.first_nonfinite;
        if (q == 0 || st.dt_min < dt_min) dt_min = st.dt_min;
        if (q == 0 || st.dt_max > dt_max) dt_max = st.dt_max;
        any_nonfinite = any_nonfinite || nonfinite[q];
    }
    total.total_steps = total.steps;
    total.dt_min = dt_min;
    total.dt_max = dt_max;
    if (stats) *stats = total;
    if (!rc && any_nonfinite) {
        vg::set_error("RGFM batch: " + std::to_string(total.nonfinite_cells) +
                      " non-finite or inadmissible final cells, or a stalled time step, in at least one instance");
        return VG_ERR_NONFINITE;
    }
    return rc;
}

int vg_rgfm_adjoint_1d(vg_ctx *ctx, const double *m1_in, const double *m2_in, const double *phi_in, double *m1_out,
                       double *m2_out, double *phi_out, const double *seed1, const double *seed2, double *grad1,
                       double *grad2, int n, double dx, double cfl, double t_final, int max_steps, double bias,
                       int limiter, const vg_euler_params *p1, const vg_euler_params *p2, int checkpoint_every,
                       double *dt_log, int dt_log_cap, vg_run_stats *stats)
{
    return adjoint(ctx, 1, m1_in, m2_in, phi_in, m1_out, m2_out, phi_out, seed1, seed2, grad1, grad2, 1, n, dx, dx,
                   cfl, t_final, max_steps, bias, limiter, p1, p2, checkpoint_every, dt_log, dt_log_cap, stats);
}

int vg_rgfm_adjoint_2d(vg_ctx *ctx, const double *m1_in, const double *m2_in, const double *phi_in, double *m1_out,
                       double *m2_out, double *phi_out, const double *seed1, const double *seed2, double *grad1,
                       double *grad2, int ny, int nx, double dx, double dy, double cfl, double t_final, int max_steps,
                       double bias, int limiter, const vg_euler_params *p1, const vg_euler_params *p2,
                       int checkpoint_every, double *dt_log, int dt_log_cap, vg_run_stats *stats)
{
    return adjoint(ctx, 2, m1_in, m2_in, phi_in, m1_out, m2_out, phi_out, seed1, seed2, grad1, grad2, ny, nx, dx, dy,
                   cfl, t_final, max_steps, bias, limiter, p1, p2, checkpoint_every, dt_log, dt_log_cap, stats);
}

int vg_rgfm_adjoint_batch_2d(vg_ctx *ctx, int n_inst, const double *m1_in, const double *m2_in, const double *phi_in,
                             double *m1_out, double *m2_out, double *phi_out, const double *seed1,
                             const double *seed2, double *grad1, double *grad2, int ny, int nx, double dx, double dy,
                             double cfl, const double *t_final, int max_steps, double bias, int limiter,
                             const vg_euler_params *p1, const vg_euler_params *p2, int checkpoint_every,
                             int32_t *steps, int32_t *status, vg_run_stats *stats)
{
    vg::clear_error();
    VG_CHECK(ctx != nullptr && n_inst >= 1 && t_final != nullptr, "null argument");
    VG_CHECK(m1_in && m2_in && phi_in && m1_out && m2_out && phi_out && seed1 && seed2 && grad1 && grad2,
             "null state pointer");
    VG_CHECK(checkpoint_every >= 0, "checkpoint_every must be >= 0");
    for (int q = 0; q < n_inst; q++) {
        int rc = check_args(2, ny, nx, dx, dy, cfl, t_final[q], max_steps, limiter, VG_RIEMANN_HLLC, p1, p2,
                            VG_MODE_EXACT);
        if (rc) return rc;
    }
    int rc = vg::bind(ctx);
    if (rc) return rc;
    const size_t cells = (size_t)ny * nx, n6 = 6 * cells;
    const vg::EP e1 = to_params(p1), e2 = to_params(p2);
    // Predicted steps (t_final times the largest signal speed over the CFL
    // spacing, with a margin for the start-up steps) order the instances,
    // longest first, and size each one's checkpoints.
    std::vector<std::pair<double, int>> order(n_inst);
    std::vector<int> n_est(n_inst, 0);
    for (int q = 0; q < n_inst; q++) {
        double smax = 0.0;
        const double *ms[2] = {m1_in + n6 * q, m2_in + n6 * q};
        const vg_euler_params *ps[2] = {p1, p2};
        for (int k = 0; k < 2; k++)
            for (size_t c = 0; c < cells; c++) {
                const double *w = ms[k] + 6 * c;
                const double a2 = ps[k]->gamma * (w[4] + ps[k]->p_inf) / w[0];
                if (a2 > 0.0 && std::isfinite(a2))
                    smax = std::max(smax, std::max(std::fabs(w[1]), std::fabs(w[2])) + std::sqrt(a2));
            }
        const double n = 1.25 * t_final[q] * smax / (cfl * std::min(dx, dy)) + 16.0;
        n_est[q] = (std::isfinite(n) && n < 1e8) ? (int)n : 0;
        order[q] = {-(double)n_est[q], q};
    }
    std::stable_sort(order.begin(), order.end());
    const int nw = std::min(n_inst, ctx->timing ? 1 : ctx->batch_streams);
    // Tile shapes are tuned here, alone on the device, before the workers
    // start (they take them from this context): the y sweeps of the forward,
    // and the adjoint sweeps in both the 4- and 6-field forms.
    if (ctx->autotune) {
        vg::RGFMGrid g;
        const int q = order[0].second;
        double *buf = nullptr;
        rc = g.create(ctx, 2, ny, nx);
        if (!rc) rc = g.load(m1_in + n6 * q, m2_in + n6 * q, phi_in + cells * q);
        if (!rc) rc = g.choose_physics();
        if (!rc)
            rc = g.reduced ? g.tune_as<vg::EulerPlanarPhys>(VG_MODE_EXACT, dx, dy, bias, limiter, e1)
                           : g.tune_as<vg::EulerPhys>(VG_MODE_EXACT, dx, dy, bias, limiter, e1);
        const size_t gb = sizeof(double) * 6 * g.plane;
        if (!rc && hipMalloc(&buf, 2 * gb + sizeof(double) * 4 * 6 * std::max(nx, ny)) != hipSuccess) {
            vg::set_error("RGFM adjoint batch: out of device memory for the tile tuning");
            rc = VG_ERR_HIP;
        }
        if (!rc && hipMemsetAsync(buf, 0, gb, ctx->stream) != hipSuccess) rc = VG_ERR_HIP;
        if (!rc) {
            const vg::AdjointArgs ta = vg::adjoint_args(g.m1[g.cur], buf, buf + 6 * g.plane, buf + 12 * g.plane,
                                                        g.plane, ny, nx, VG_BC_TRANSMISSIVE, dx, bias, limiter,
                                                        1e-3 * std::min(dx, dy), e1);
            int v = 0;
            for (int dir = 0; dir < 2 && !rc; dir++) {
                rc = vg::tune_adjoint<vg::AdjPlanar>(ctx, dir, ta, v);
                if (!rc) rc = vg::tune_adjoint<vg::AdjFull>(ctx, dir, ta, v);
            }
        }
        if (ctx->stream) hipStreamSynchronize(ctx->stream);
        if (buf) hipFree(buf);
        g.destroy();
        if (rc) return rc;
    }
    rc = vg::batch_workers(ctx, nw);
    if (rc) return rc;
    // Each worker's share of the free device memory for its instance's
    // checkpoints and kept steps; fewer workers when the longest predicted
    // run would not fit in a share (an instance needs at least K kept steps
    // and n / K checkpoints, least at K = sqrt(n checkpoint / kept step)).
    for (int k = 0; k < (int)ctx->workers.size(); k++) ctx->workers[k]->release_workspace();
    size_t free_b = 0, total_b = 0;
    VG_HIP(hipMemGetInfo(&free_b, &total_b));
    const double usable = 0.55 * (double)free_b;
    int nw_mem = nw;
    if (checkpoint_every == 0) {
        const double plane = (double)(((long long)cells + 63) / 64 * 64), gb = 48.0 * plane, pb = 8.0 * cells;
        const double kept = 10.0 * gb + pb, ckb = 2.0 * gb + pb;
        const double n = std::max(1.0, (double)(max_steps > 0 ? max_steps : -order[0].first));
        const double k_best = std::max(1.0, std::sqrt(n * ckb / kept));
        const double least = k_best * kept + std::ceil(n / k_best) * ckb;
        nw_mem = std::max(1, std::min(nw, (int)(usable / least)));
    }
    const int nwork = nw_mem;
    const double budget = usable / nwork;
    std::vector<vg_run_stats> per(n_inst);
    std::vector<int> inst_rc(n_inst, VG_OK);
    std::vector<std::string> inst_err(n_inst);
    std::atomic<int> next{0};
    std::atomic<bool> stop{false};
    std::vector<int> wrc(nwork, VG_OK);
    std::vector<std::string> werr(nwork);
    // Each worker takes instances from the shared queue, longest first, and
    // runs the single-instance adjoint on its own stream and workspace.
    auto worker = [&](int k) {
        vg_ctx *c = ctx->workers[k];
        int r = vg::bind(c);
        while (!r && !stop.load()) {
            const int i = next.fetch_add(1);
            if (i >= n_inst) break;
            const int q = order[i].second;
            const int s = adjoint(c, 2, m1_in + n6 * q, m2_in + n6 * q, phi_in + cells * q, m1_out + n6 * q,
                                  m2_out + n6 * q, phi_out + cells * q, seed1 + n6 * q, seed2 + n6 * q,
                                  grad1 + n6 * q, grad2 + n6 * q, ny, nx, dx, dy, cfl, t_final[q], max_steps, bias,
                                  limiter, p1, p2, checkpoint_every, nullptr, 0, &per[q], budget, n_est[q]);
            inst_rc[q] = s;
            if (s) inst_err[q] = vg_last_error();
            // an instance without a derivative is reported in its status
            r = (s == VG_ERR_NONFINITE) ? VG_OK : s;
        }
        if (c->stream) hipStreamSynchronize(c->stream);
        c->release_workspace();
        if (r) {
            stop = true;
            werr[k] = vg_last_error();
        }
        wrc[k] = r;
    };
    {
        std::vector<std::thread> threads;
        for (int k = 0; k < nwork; k++) threads.emplace_back(worker, k);
        for (auto &t : threads) t.join();
    }
    vg::merge_worker_tuning(ctx);
    for (int k = 0; k < nwork && !rc; k++)
        if (wrc[k]) {
            rc = wrc[k];
            vg::set_error(werr[k]);
        }
    vg_run_stats total{};
    total.first_nonfinite = -1;
    int n_bad = 0, first_bad = -1;
    for (int q = 0; q < n_inst; q++) {
        if (steps) steps[q] = per[q].steps;
        if (status) status[q] = inst_rc[q];
        total.steps += per[q].steps;
        total.wall_ms += per[q].wall_ms;
        if (inst_rc[q] == VG_ERR_NONFINITE) {
            n_bad++;
            if (first_bad < 0) first_bad = q;
            total.nonfinite_cells += per[q].nonfinite_cells;
            if (total.first_nonfinite < 0 && per[q].first_nonfinite >= 0)
                total.first_nonfinite = (int64_t)cells * q + per[q].first_nonfinite;
        }
    }
    total.total_steps = total.steps;
    if (stats) *stats = total;
    if (!rc && n_bad > 0) {
        vg::set_error("RGFM adjoint batch: " + std::to_string(n_bad) + " instance(s) without a finite derivative; " +
                      "first: instance " + std::to_string(first_bad) + ": " + inst_err[first_bad]);
        return VG_ERR_NONFINITE;
    }
    return rc;
}

int vg_rgfm_solve_fixed_dt_1d(vg_ctx *ctx, const double *m1_in, const double *m2_in, const double *phi_in,
                              double *m1_out, double *m2_out, double *phi_out, int n, double dx, const double *dts,
                              int n_steps, double bias, int limiter, const vg_euler_params *p1,
                              const vg_euler_params *p2)
{
    return solve_fixed(ctx, 1, m1_in, m2_in, phi_in, m1_out, m2_out, phi_out, 1, n, dx, dx, dts, n_steps, bias,
                       limiter, p1, p2);
}

int vg_rgfm_solve_fixed_dt_2d(vg_ctx *ctx, const double *m1_in, const double *m2_in, const double *phi_in,
                              double *m1_out, double *m2_out, double *phi_out, int ny, int nx, double dx, double dy,
                              const double *dts, int n_steps, double bias, int limiter, const vg_euler_params *p1,
                              const vg_euler_params *p2)
{
    return solve_fixed(ctx, 2, m1_in, m2_in, phi_in, m1_out, m2_out, phi_out, ny, nx, dx, dy, dts, n_steps, bias,
                       limiter, p1, p2);
}

}  // extern "C"
