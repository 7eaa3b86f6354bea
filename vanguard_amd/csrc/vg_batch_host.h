// Host side of the batched solve: uploads the instances, orders them by
// decreasing estimated work (cells^2 x initial wave speed: the step count of an
// instance grows like N * S / CFL), launches one workgroup per instance in that
// order and collects results.  The launch stops every instance at 1.25x the
// largest step count the initial wave speeds predict; an instance still short
// of its final time then continues on the grid path -- the same arithmetic
// from the same state and clock, its sweeps spread over all CUs -- so that one
// instance whose speeds grow during the run does not hold the whole launch.
#pragma once

#include <algorithm>
#include <chrono>
#include <cmath>
#include <numeric>
#include <vector>

#include "vg_batch.h"
#include "vg_grid.h"
#include "vg_runtime.h"

namespace vg {

template <class P>
struct BatchBuffers {
    long long *offsets = nullptr;
    double *dx = nullptr, *t_final = nullptr, *t_out = nullptr, *speed = nullptr, *dt_log = nullptr;
    typename P::Params *params = nullptr;
    double *buf0 = nullptr, *buf1 = nullptr, *staging = nullptr;
    int *steps = nullptr, *where = nullptr, *stalled = nullptr, *order = nullptr;
    ~BatchBuffers()
    {
        hipFree(offsets); hipFree(dx); hipFree(t_final); hipFree(t_out); hipFree(speed); hipFree(dt_log);
        hipFree(params); hipFree(buf0); hipFree(buf1); hipFree(staging); hipFree(steps);
        hipFree(where); hipFree(stalled); hipFree(order);
    }
};

template <class P, int ORDER, int MODE, int BS>
int launch_batch(vg_ctx *ctx, const BatchArgs<P> &args, int n_inst)
{
    return ctx->timed("batch_solve", [&] {
        batch_solve_kernel<P, ORDER, MODE, BS><<<n_inst, BS, 0, ctx->stream>>>(args);
    });
}

// prim_in / prim_out: all instances' cells, array of structures (NF doubles per cell).
template <class P, int BS>
int run_batch(vg_ctx *ctx, int order, int mode, int n_inst, const int64_t *offsets_h, const double *prim_in,
              double *prim_out, const double *dx_h, const std::vector<typename P::Params> &params_h,
              const double *t_final_h, double cfl, int max_steps, double bias, int limiter, int bc, double *dt_log_h,
              int dt_cap, int32_t *steps_h, double *t_out_h, vg_run_stats *stats)
{
    constexpr int NF = P::NF;
    auto w0 = std::chrono::steady_clock::now();
    VG_CHECK(n_inst >= 1, "need at least one instance");
    VG_CHECK(offsets_h && prim_in && prim_out && dx_h && t_final_h, "null argument");
    VG_CHECK(order == 1 || order == 2, "order must be 1 (FORCE) or 2 (SLIC)");
    VG_CHECK(mode == VG_MODE_EXACT || mode == VG_MODE_FAST, "unknown mode");
    VG_CHECK(bc == VG_BC_TRANSMISSIVE || bc == VG_BC_PERIODIC, "unknown boundary condition");
    VG_CHECK(max_steps >= 0, "max_steps must be >= 0");
    VG_CHECK(cfl > 0.0 && std::isfinite(cfl), "CFL coefficient must be positive");
    VG_CHECK(offsets_h[0] == 0, "offsets[0] must be 0");
    for (int q = 0; q < n_inst; q++) {
        VG_CHECK(offsets_h[q + 1] - offsets_h[q] >= 2, "every instance needs at least 2 cells");
        VG_CHECK(offsets_h[q + 1] - offsets_h[q] < (1ll << 31), "instance too large");
        VG_CHECK(dx_h[q] > 0.0 && std::isfinite(dx_h[q]), "cell spacing must be positive");
        VG_CHECK(max_steps > 0 || std::isfinite(t_final_h[q]), "an infinite final time needs a step cap");
    }
    const long long cells = offsets_h[n_inst];
    BatchBuffers<P> b;
    VG_HIP(hipMalloc(&b.offsets, sizeof(long long) * (n_inst + 1)));
    VG_HIP(hipMalloc(&b.dx, sizeof(double) * n_inst));
    VG_HIP(hipMalloc(&b.t_final, sizeof(double) * n_inst));
    VG_HIP(hipMalloc(&b.t_out, sizeof(double) * n_inst));
    VG_HIP(hipMalloc(&b.speed, sizeof(double) * n_inst));
    VG_HIP(hipMalloc(&b.params, sizeof(typename P::Params) * n_inst));
    VG_HIP(hipMalloc(&b.buf0, sizeof(double) * NF * cells));
    VG_HIP(hipMalloc(&b.buf1, sizeof(double) * NF * cells));
    VG_HIP(hipMalloc(&b.staging, sizeof(double) * NF * cells));
    VG_HIP(hipMalloc(&b.steps, sizeof(int) * n_inst));
    VG_HIP(hipMalloc(&b.where, sizeof(int) * n_inst));
    VG_HIP(hipMalloc(&b.stalled, sizeof(int) * n_inst));
    VG_HIP(hipMalloc(&b.order, sizeof(int) * n_inst));
    if (dt_log_h && dt_cap > 0) VG_HIP(hipMalloc(&b.dt_log, sizeof(double) * dt_cap));
    hipStream_t s = ctx->stream;
    std::vector<long long> off(offsets_h, offsets_h + n_inst + 1);
    VG_HIP(hipMemcpyAsync(b.offsets, off.data(), sizeof(long long) * (n_inst + 1), hipMemcpyHostToDevice, s));
    VG_HIP(hipMemcpyAsync(b.dx, dx_h, sizeof(double) * n_inst, hipMemcpyHostToDevice, s));
    VG_HIP(hipMemcpyAsync(b.t_final, t_final_h, sizeof(double) * n_inst, hipMemcpyHostToDevice, s));
    VG_HIP(hipMemcpyAsync(b.params, params_h.data(), sizeof(typename P::Params) * n_inst, hipMemcpyHostToDevice, s));
    VG_HIP(hipMemcpyAsync(b.staging, prim_in, sizeof(double) * NF * cells, hipMemcpyHostToDevice, s));
    const int gblocks = (int)std::min<long long>((cells + 255) / 256, 4096);
    batch_aos_to_planes<NF><<<gblocks, 256, 0, s>>>(b.staging, b.buf0, b.offsets, n_inst, cells);
    int rc = launch_status("batch_aos_to_planes");
    if (rc) return rc;

    // Launch order: decreasing cells^2 * initial wave speed.
    std::vector<int> order_h(n_inst);
    std::iota(order_h.begin(), order_h.end(), 0);
    if (n_inst > 1) {
        batch_speed_kernel<P><<<n_inst, 256, 0, s>>>(b.offsets, b.params, b.buf0, b.speed);
        rc = launch_status("batch_speed");
        if (rc) return rc;
        std::vector<double> sp(n_inst);
        VG_HIP(hipMemcpyAsync(sp.data(), b.speed, sizeof(double) * n_inst, hipMemcpyDeviceToHost, s));
        VG_HIP(hipStreamSynchronize(s));
        std::vector<double> est(n_inst);
        for (int q = 0; q < n_inst; q++) {
            const double n = (double)(off[q + 1] - off[q]);
            est[q] = n * n * (std::isfinite(sp[q]) && sp[q] > 0.0 ? sp[q] : 1.0);
        }
        std::stable_sort(order_h.begin(), order_h.end(), [&](int x, int y) { return est[x] > est[y]; });
    }
    // The launch's step cap: 1.25x the largest predicted step count
    // (t_final S / (CFL dx) from the initial wave speed), at least 1024.
    int cap = max_steps;
    if (n_inst > 1) {
        std::vector<double> sp(n_inst);
        VG_HIP(hipMemcpyAsync(sp.data(), b.speed, sizeof(double) * n_inst, hipMemcpyDeviceToHost, s));
        VG_HIP(hipStreamSynchronize(s));
        double most = 0.0;
        bool known = true;
        for (int q = 0; q < n_inst; q++) {
            const double speed = (std::isfinite(sp[q]) && sp[q] > 0.0) ? sp[q] : 0.0;
            if (!std::isfinite(t_final_h[q])) known = false;
            else most = std::max(most, t_final_h[q] * speed / (cfl * dx_h[q]));
        }
        if (known) {
            const double c = std::max(1024.0, std::ceil(1.25 * most));
            if (c < (double)0x7fffffff && (max_steps == 0 || c < max_steps)) cap = (int)c;
        }
    }
    VG_HIP(hipMemcpyAsync(b.order, order_h.data(), sizeof(int) * n_inst, hipMemcpyHostToDevice, s));

    BatchArgs<P> a;
    a.n_inst = n_inst;
    a.order = b.order;
    a.offsets = b.offsets;
    a.dx = b.dx;
    a.params = b.params;
    a.t_final = b.t_final;
    a.cfl = cfl;
    a.bias = bias;
    a.limiter = limiter;
    a.max_steps = cap;
    a.bc = bc;
    a.buf0 = b.buf0;
    a.buf1 = b.buf1;
    a.steps_out = b.steps;
    a.t_out = b.t_out;
    a.where = b.where;
    a.stalled = b.stalled;
    a.dt_log = (n_inst == 1) ? b.dt_log : nullptr;
    a.dt_cap = dt_cap;

    hipEvent_t e0, e1;
    VG_HIP(hipEventCreate(&e0));
    VG_HIP(hipEventCreate(&e1));
    VG_HIP(hipEventRecord(e0, s));
    if (order == 2 && mode == 0) rc = launch_batch<P, 2, 0, BS>(ctx, a, n_inst);
    else if (order == 2) rc = launch_batch<P, 2, 1, BS>(ctx, a, n_inst);
    else if (mode == 0) rc = launch_batch<P, 1, 0, BS>(ctx, a, n_inst);
    else rc = launch_batch<P, 1, 1, BS>(ctx, a, n_inst);
    if (rc) return launch_status("batch_solve");
    VG_HIP(hipEventRecord(e1, s));
    batch_planes_to_aos<NF><<<gblocks, 256, 0, s>>>(b.buf0, b.buf1, b.where, b.staging, b.offsets, n_inst, cells);
    rc = launch_status("batch_planes_to_aos");
    if (rc) return rc;
    VG_HIP(hipMemcpyAsync(prim_out, b.staging, sizeof(double) * NF * cells, hipMemcpyDeviceToHost, s));
    std::vector<int> steps(n_inst), stalled(n_inst);
    std::vector<double> tend(n_inst);
    VG_HIP(hipMemcpyAsync(steps.data(), b.steps, sizeof(int) * n_inst, hipMemcpyDeviceToHost, s));
    VG_HIP(hipMemcpyAsync(stalled.data(), b.stalled, sizeof(int) * n_inst, hipMemcpyDeviceToHost, s));
    VG_HIP(hipMemcpyAsync(tend.data(), b.t_out, sizeof(double) * n_inst, hipMemcpyDeviceToHost, s));
    VG_HIP(hipStreamSynchronize(s));
    // instances the cap stopped short of their final time: the rest of their
    // run on the grid path, from their state, time and step count
    for (int q = 0; q < n_inst && cap != max_steps; q++) {
        if (stalled[q] || steps[q] < cap || !(tend[q] < t_final_h[q])) continue;
        GridState<P> g;
        rc = g.create(ctx, 1, 1, (int)(off[q + 1] - off[q]));
        if (!rc) rc = g.load(prim_out + (size_t)NF * off[q], 0);
        if (!rc) {
            Clock h{};
            h.t = tend[q];
            h.it = steps[q];
            hipError_t e = hipMemcpyAsync(g.clk, &h, sizeof(h), hipMemcpyHostToDevice, s);
            if (e != hipSuccess) {
                set_error(std::string("batch continuation clock: ") + hipGetErrorString(e));
                rc = VG_ERR_HIP;
            }
        }
        vg_run_stats st{};
        if (!rc)
            rc = g.run(order, mode, dx_h[q], cfl, t_final_h[q], max_steps > 0 ? max_steps - steps[q] : 0, bias,
                       limiter, bc, params_h[q], nullptr, 0, &st);
        if (!rc) rc = g.store(prim_out + (size_t)NF * off[q], 0);
        g.destroy();
        if (rc) return rc;
        steps[q] += st.steps;
        tend[q] = st.t;
    }
    float dev_ms = 0.0f;
    VG_HIP(hipEventElapsedTime(&dev_ms, e0, e1));
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (ctx->timing) {
        rc = ctx->resolve_timing();
        if (rc) return rc;
    }
    if (b.dt_log && n_inst == 1 && steps[0] > 0) {
        const int n = std::min(steps[0], dt_cap);
        VG_HIP(hipMemcpy(dt_log_h, b.dt_log, sizeof(double) * n, hipMemcpyDeviceToHost));
    }
    long long total_steps = 0, stalled_count = 0;
    for (int q = 0; q < n_inst; q++) {
        if (steps_h) steps_h[q] = steps[q];
        if (t_out_h) t_out_h[q] = tend[q];
        total_steps += steps[q];
        stalled_count += stalled[q];
    }
    if (stats) {
        stats->steps = (int32_t)std::min<long long>(total_steps, 0x7fffffff);
        stats->total_steps = stats->steps;
        stats->t = tend[0];
        stats->dt_min = stats->dt_max = 0.0;
        stats->device_ms = dev_ms;
        stats->wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - w0).count();
        stats->nonfinite_cells = 0;
        stats->first_nonfinite = -1;
    }
    if (stalled_count) {
        set_error(std::to_string(stalled_count) + " instance(s) stopped because a time step could not advance t");
        return VG_ERR_NONFINITE;
    }
    return VG_OK;
}

}  // namespace vg
