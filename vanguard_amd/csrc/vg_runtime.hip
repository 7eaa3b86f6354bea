// Context, errors and kernel timing.
#include <string>

#include "vg_runtime.h"

#include <cstring>

namespace vg {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }
void clear_error() { g_last_error.clear(); }

int bind(vg_ctx *ctx)
{
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) {
        set_error(std::string("hipSetDevice: ") + hipGetErrorString(e));
        return VG_ERR_HIP;
    }
    return VG_OK;
}

int batch_workers(vg_ctx *ctx, int n, bool graphs)
{
    while ((int)ctx->workers.size() < n) {
        vg_ctx *w = nullptr;
        const int rc = vg_ctx_create(&w, ctx->device);
        if (rc) return rc;
        ctx->workers.push_back(w);
    }
    for (vg_ctx *w : ctx->workers) {
        w->graphs = ctx->graphs && graphs;
        w->autotune = ctx->autotune;
        w->reduce_fields = ctx->reduce_fields;
        w->timing = false;
        for (const auto &kv : ctx->tuned) w->tuned[kv.first] = kv.second;
    }
    return VG_OK;
}

void merge_worker_tuning(vg_ctx *ctx)
{
    for (vg_ctx *w : ctx->workers)
        for (const auto &kv : w->tuned) ctx->tuned.insert(kv);
}

int launch_status(const char *what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error(std::string(what) + ": " + hipGetErrorString(e));
        return VG_ERR_HIP;
    }
    return VG_OK;
}

}  // namespace vg

hipEvent_t vg_ctx::take_event()
{
    if (!event_pool.empty()) {
        hipEvent_t e = event_pool.back();
        event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    hipEventCreate(&e);
    return e;
}

int vg_ctx::resolve_timing()
{
    VG_HIP(hipStreamSynchronize(stream));
    for (auto &p : pending) {
        float ms = 0.0f;
        VG_HIP(hipEventElapsedTime(&ms, p.second.first, p.second.second));
        auto &t = timers[p.first];
        t.launches += 1;
        t.total_ms += ms;
        event_pool.push_back(p.second.first);
        event_pool.push_back(p.second.second);
    }
    pending.clear();
    return VG_OK;
}

int vg_ctx::workspace_buffer(size_t i, size_t bytes, void **out)
{
    if (workspace.size() <= i) workspace.resize(i + 1, {nullptr, 0});
    auto &b = workspace[i];
    if (b.second < bytes) {
        if (b.first) {
            VG_HIP(hipStreamSynchronize(stream));
            VG_HIP(hipFree(b.first));
        }
        b = {nullptr, 0};
        VG_HIP(hipMalloc(&b.first, bytes));
        b.second = bytes;
    }
    *out = b.first;
    return VG_OK;
}

size_t vg_ctx::workspace_bytes() const
{
    size_t n = 0;
    for (auto &b : workspace) n += b.second;
    return n;
}

void vg_ctx::release_workspace()
{
    for (auto &b : workspace)
        if (b.first) hipFree(b.first);
    workspace.clear();
}

extern "C" {

const char *vg_last_error(void) { return vg::g_last_error.c_str(); }

const char *vg_version(void) { return "vanguard-mi355x 0.1 (gfx950)"; }

int vg_ctx_create(vg_ctx **out, int device)
{
    vg::clear_error();
    VG_CHECK(out != nullptr, "vg_ctx_create: null output pointer");
    int count = 0;
    VG_HIP(hipGetDeviceCount(&count));
    VG_CHECK(device >= 0 && device < count, "vg_ctx_create: no such device");
    vg_ctx *ctx = new vg_ctx();
    ctx->device = device;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        vg::set_error(std::string("vg_ctx_create: ") + hipGetErrorString(e));
        delete ctx;
        return VG_ERR_HIP;
    }
    *out = ctx;
    return VG_OK;
}

void vg_ctx_destroy(vg_ctx *ctx)
{
    if (!ctx) return;
    hipSetDevice(ctx->device);
    if (ctx->stream) hipStreamSynchronize(ctx->stream);
    ctx->release_workspace();
    for (auto &p : ctx->pending) {
        hipEventDestroy(p.second.first);
        hipEventDestroy(p.second.second);
    }
    for (hipEvent_t e : ctx->event_pool) hipEventDestroy(e);
    if (ctx->stream) hipStreamDestroy(ctx->stream);
    for (vg_ctx *w : ctx->workers) vg_ctx_destroy(w);
    delete ctx;
}

int vg_ctx_set_batch_streams(vg_ctx *ctx, int streams)
{
    VG_CHECK(ctx != nullptr, "vg_ctx_set_batch_streams: null context");
    VG_CHECK(streams >= 1 && streams <= 16, "vg_ctx_set_batch_streams: streams must be 1..16");
    ctx->batch_streams = streams;
    return VG_OK;
}

int vg_ctx_set_adjoint_memory(vg_ctx *ctx, int64_t bytes)
{
    VG_CHECK(ctx != nullptr, "vg_ctx_set_adjoint_memory: null context");
    VG_CHECK(bytes >= 0, "vg_ctx_set_adjoint_memory: bytes must be >= 0");
    ctx->adjoint_memory = (size_t)bytes;
    return VG_OK;
}

int vg_ctx_set_timing(vg_ctx *ctx, int enabled)
{
    VG_CHECK(ctx != nullptr, "vg_ctx_set_timing: null context");
    ctx->timing = enabled != 0;
    return VG_OK;
}

int vg_ctx_kernel_time(vg_ctx *ctx, const char *kernel, int64_t *launches, double *total_ms)
{
    VG_CHECK(ctx != nullptr && kernel != nullptr, "vg_ctx_kernel_time: null argument");
    int rc = ctx->resolve_timing();
    if (rc) return rc;
    auto it = ctx->timers.find(kernel);
    if (launches) *launches = (it == ctx->timers.end()) ? 0 : it->second.launches;
    if (total_ms) *total_ms = (it == ctx->timers.end()) ? 0.0 : it->second.total_ms;
    return VG_OK;
}

int vg_ctx_set_graphs(vg_ctx *ctx, int enabled)
{
    VG_CHECK(ctx != nullptr, "vg_ctx_set_graphs: null context");
    ctx->graphs = enabled != 0;
    return VG_OK;
}

int vg_ctx_set_autotune(vg_ctx *ctx, int enabled)
{
    VG_CHECK(ctx != nullptr, "vg_ctx_set_autotune: null context");
    ctx->autotune = enabled != 0;
    return VG_OK;
}

int vg_ctx_set_reduced_fields(vg_ctx *ctx, int enabled)
{
    VG_CHECK(ctx != nullptr, "vg_ctx_set_reduced_fields: null context");
    ctx->reduce_fields = enabled != 0;
    return VG_OK;
}

int vg_ctx_tuned_shape(vg_ctx *ctx, const char *key, int *variant)
{
    VG_CHECK(ctx != nullptr && key != nullptr && variant != nullptr, "vg_ctx_tuned_shape: null argument");
    auto it = ctx->tuned.find(key);
    *variant = (it == ctx->tuned.end()) ? -1 : it->second;
    return VG_OK;
}

int vg_ctx_release_workspace(vg_ctx *ctx)
{
    VG_CHECK(ctx != nullptr, "vg_ctx_release_workspace: null context");
    int rc = vg::bind(ctx);
    if (rc) return rc;
    VG_HIP(hipStreamSynchronize(ctx->stream));
    ctx->release_workspace();
    return VG_OK;
}

int vg_ctx_set_tuned_shape(vg_ctx *ctx, const char *key, int variant)
{
    VG_CHECK(ctx != nullptr && key != nullptr, "vg_ctx_set_tuned_shape: null argument");
    const bool adjoint = std::string(key).find("adjoint") != std::string::npos;
    VG_CHECK(variant >= -1 && variant < (adjoint ? 6 : 8),
             adjoint ? "vg_ctx_set_tuned_shape: adjoint variant must be -1..5"
                     : "vg_ctx_set_tuned_shape: y-sweep variant must be -1..7");
    if (variant < 0) ctx->tuned.erase(key);
    else ctx->tuned[key] = variant;
    return VG_OK;
}

int vg_ctx_reset_timing(vg_ctx *ctx)
{
    VG_CHECK(ctx != nullptr, "vg_ctx_reset_timing: null context");
    int rc = ctx->resolve_timing();
    if (rc) return rc;
    ctx->timers.clear();
    return VG_OK;
}

}  // extern "C"
