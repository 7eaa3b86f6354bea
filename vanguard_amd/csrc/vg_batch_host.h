// Host side of the batched solve: uploads the instances, orders them by
// decreasing estimated work (cells^2 x initial wave speed: the step count of an
// instance grows like N * S / CFL), launches one workgroup per instance in that
// order and collects results.  The launch stops every instance at 1.25x the
// largest step count the initial wave speeds predict and, when there are more
// instances than resident workgroups, at a work budget of cells x steps (the
// predicted total over the resident workgroups); an instance still short of
// its final time then continues on the grid path -- the same arithmetic from
// the same state and clock, its sweeps spread over all CUs, several instances
// at once on the context's worker streams -- so that no single instance, one
// whose speeds grow during the run or one both large and long, holds the
// whole launch on one CU.
#pragma once

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <numeric>
#include <string>
#include <thread>
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

// Workgroups of the batch kernel resident at once on the device.
template <class P, int ORDER, int MODE, int BS>
int batch_slots(vg_ctx *ctx)
{
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, batch_solve_kernel<P, ORDER, MODE, BS>, BS, 0) !=
        hipSuccess)
        per_cu = 1;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device) != hipSuccess) cus = 1;
    return std::max(1, per_cu) * std::max(1, cus);
}

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
    // With more instances than resident workgroups, each instance also gets a
    // work budget (cells x steps): the predicted total over the resident
    // workgroups, so a large, long instance leaves the launch at about the
    // time the others finish instead of holding it alone on one CU.
    double budget = 0.0;
    if (n_inst > 1) {
        const int slots = (order == 2 && mode == 0)   ? batch_slots<P, 2, 0, BS>(ctx)
                          : (order == 2)              ? batch_slots<P, 2, 1, BS>(ctx)
                          : (mode == 0)               ? batch_slots<P, 1, 0, BS>(ctx)
                                                      : batch_slots<P, 1, 1, BS>(ctx);
        if (n_inst > slots) {
            std::vector<double> sp(n_inst);
            VG_HIP(hipMemcpyAsync(sp.data(), b.speed, sizeof(double) * n_inst, hipMemcpyDeviceToHost, s));
            VG_HIP(hipStreamSynchronize(s));
            double total = 0.0;
            bool known = true;
            for (int q = 0; q < n_inst; q++) {
                const double speed = (std::isfinite(sp[q]) && sp[q] > 0.0) ? sp[q] : 0.0;
                if (!std::isfinite(t_final_h[q])) known = false;
                else total += (double)(off[q + 1] - off[q]) * (t_final_h[q] * speed / (cfl * dx_h[q]));
            }
            if (known && total > 0.0 && std::isfinite(total)) budget = total / slots;
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
    a.work_budget = budget;

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
    // Instances a cap stopped short of their final time (and of the caller's
    // max_steps): the rest of their run on the grid path, from their state,
    // time and step count, largest remaining work first, on the context's
    // worker streams so several continue at once.
    std::vector<std::pair<double, int>> cont;
    for (int q = 0; q < n_inst; q++) {
        if (stalled[q] || !(tend[q] < t_final_h[q]) || (max_steps > 0 && steps[q] >= max_steps)) continue;
        const double left = (tend[q] > 0.0) ? steps[q] * ((t_final_h[q] - tend[q]) / tend[q]) : 1.0;
        cont.push_back({-(double)(off[q + 1] - off[q]) * left, q});
    }
    std::stable_sort(cont.begin(), cont.end());
    // Every continuation grid is allocated here, before any worker starts, and
    // freed after they all finish: no allocation or free overlaps a worker's
    // graph capture.
    std::vector<GridState<P>> grids(cont.size());
    for (size_t k = 0; k < cont.size() && !rc; k++) {
        const int q = cont[k].second;
        rc = grids[k].create(ctx, 1, 1, (int)(off[q + 1] - off[q]));
        if (!rc) rc = grids[k].ensure_staging();
    }
    auto continue_one = [&](vg_ctx *c, size_t k) {
        const int q = cont[k].second;
        GridState<P> &g = grids[k];
        g.ctx = c;   // run on the worker's stream
        int r = g.load(prim_out + (size_t)NF * off[q], 0);
        if (!r) {
            Clock h{};
            h.t = tend[q];
            h.it = steps[q];
            hipError_t e = hipMemcpyAsync(g.clk, &h, sizeof(h), hipMemcpyHostToDevice, c->stream);
            if (e != hipSuccess) {
                set_error(std::string("batch continuation clock: ") + hipGetErrorString(e));
                r = VG_ERR_HIP;
            }
        }
        vg_run_stats st{};
        if (!r)
            r = g.run(order, mode, dx_h[q], cfl, t_final_h[q], max_steps > 0 ? max_steps - steps[q] : 0, bias,
                      limiter, bc, params_h[q], nullptr, 0, &st);
        if (!r) r = g.store(prim_out + (size_t)NF * off[q], 0);
        if (c->stream) hipStreamSynchronize(c->stream);
        g.ctx = ctx;
        if (!r) {
            steps[q] += st.steps;
            tend[q] = st.t;
        }
        return r;
    };
    const int nw = ctx->timing ? 1 : std::min((int)cont.size(), ctx->batch_streams);
    if (rc) {
        // creation failed: nothing runs
    } else if (nw <= 1) {
        for (size_t k = 0; k < cont.size() && !rc; k++) rc = continue_one(ctx, k);
    } else {
        rc = batch_workers(ctx, nw);
        std::atomic<int> next{0};
        std::atomic<bool> stop{false};
        std::vector<int> wrc(nw, VG_OK);
        std::vector<std::string> werr(nw);
        auto worker = [&](int k) {
            vg_ctx *c = ctx->workers[k];
            int r = bind(c);
            while (!r && !stop.load()) {
                const int i = next.fetch_add(1);
                if (i >= (int)cont.size()) break;
                r = continue_one(c, (size_t)i);
            }
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
            merge_worker_tuning(ctx);
            for (int k = 0; k < nw && !rc; k++)
                if (wrc[k]) {
                    rc = wrc[k];
                    set_error(werr[k]);
                }
        }
    }
    for (auto &g : grids) g.destroy();
    if (rc) return rc;
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
        VG_HIP(hipMemcpyAsync(dt_log_h, b.dt_log, sizeof(double) * n, hipMemcpyDeviceToHost, s));
        VG_HIP(hipStreamSynchronize(s));
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
