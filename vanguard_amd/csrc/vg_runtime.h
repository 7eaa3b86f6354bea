// Host-side runtime shared by the solvers: the context (device, stream,
// kernel-timing table), error reporting, and launch helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <vector>

#include "../../include/vanguard.h"

namespace vg {

void set_error(const std::string &msg);
void clear_error();

#define VG_HIP(call)                                                                        \
    do {                                                                                    \
        hipError_t vg_e_ = (call);                                                          \
        if (vg_e_ != hipSuccess) {                                                          \
            ::vg::set_error(std::string(#call) + ": " + hipGetErrorString(vg_e_));          \
            return VG_ERR_HIP;                                                              \
        }                                                                                   \
    } while (0)

#define VG_CHECK(cond, msg)                                                                 \
    do {                                                                                    \
        if (!(cond)) {                                                                      \
            ::vg::set_error(msg);                                                           \
            return VG_ERR_ARG;                                                              \
        }                                                                                   \
    } while (0)

struct KernelTimer {
    int64_t launches = 0;
    double total_ms = 0.0;
};

}  // namespace vg

struct vg_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    bool timing = false;
    // Replay chunks of time steps as captured HIP graphs (only while per-kernel
    // timing is off: timing brackets each launch with events).
    bool graphs = true;
    // Time the candidate tile shapes on first use of a grid shape and keep the
    // fastest (key: physics/order/mode/ny/nx -> variant index).
    bool autotune = true;
    // Run a grid with its narrower physics (ReducedPhys) when the fields that
    // physics drops are constant; results are the same bits either way.
    bool reduce_fields = true;
    // Streams a batch call may run independent instances on at once (1: one
    // after another on this context's stream), and the worker contexts (one
    // stream each) that carry them, kept between calls.
    int batch_streams = 4;
    size_t adjoint_memory = 0;   // limit of the Euler adjoint's recording (0: 85 % of free memory)
    std::vector<vg_ctx *> workers;
    std::map<std::string, int> tuned;
    std::map<std::string, vg::KernelTimer> timers;
    // Event pairs recorded around launches and not yet resolved.
    std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> pending;
    std::vector<hipEvent_t> event_pool;

    hipEvent_t take_event();
    // Brackets a launch with events when timing is on.
    template <class F>
    int timed(const char *name, F &&launch)
    {
        if (!timing) {
            launch();
            return hipPeekAtLastError() == hipSuccess ? VG_OK : VG_ERR_HIP;
        }
        hipEvent_t a = take_event(), b = take_event();
        hipEventRecord(a, stream);
        launch();
        hipEventRecord(b, stream);
        pending.push_back({name, {a, b}});
        return hipPeekAtLastError() == hipSuccess ? VG_OK : VG_ERR_HIP;
    }
    // Synchronises the stream and folds the recorded pairs into the table.
    int resolve_timing();
    // Device buffers kept between calls (the adjoint's checkpoints and sweep
    // inputs): buffer i is reused when it is large enough, else reallocated.
    // One call at a time per context uses them.
    std::vector<std::pair<void *, size_t>> workspace;
    int workspace_buffer(size_t i, size_t bytes, void **out);
    size_t workspace_bytes() const;
    void release_workspace();
};

namespace vg {
// Makes the context's device current; returns VG_ERR_HIP on failure.
int bind(vg_ctx *ctx);
// Ensures n worker contexts on ctx's device and gives them ctx's switches and
// tuned shapes; `graphs` false makes them launch eagerly (no stream capture,
// for workers that allocate and free grids while others run).
int batch_workers(vg_ctx *ctx, int n, bool graphs = true);
// Copies the shapes the workers tuned into ctx's table (existing keys kept).
void merge_worker_tuning(vg_ctx *ctx);
int launch_status(const char *what);
}  // namespace vg
