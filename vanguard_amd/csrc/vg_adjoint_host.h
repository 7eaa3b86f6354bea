// Host side of the Euler adjoint sweeps: tile variants, their launch (sweep +
// ghost fold) and the per-shape autotune.  Shared by the Euler adjoint and the
// RGFM adjoint (which advances each material with the same sweeps).
#pragma once

#include <string>

#include "vg_adjoint.h"
#include "vg_grid.h"

namespace vg {

template <class A, int DIR, int BS, int LW, int WPE>
inline void launch_adjoint_tile(vg_ctx *ctx, const AdjointArgs &a)
{
    const int n = (DIR == 0) ? a.nx : a.ny, lines = (DIR == 0) ? a.ny : a.nx;
    using T = AdjTile<BS, LW>;
    dim3 grid((n + 4 + T::out - 1) / T::out, (lines + LW - 1) / LW);
    euler_adjoint_sweep_kernel<A, DIR, BS, LW, WPE><<<grid, BS, 0, ctx->stream>>>(a);
}

// Tile variants per direction (threads, lines per block, waves per SIMD).  X:
// one line per block; Y: LW adjacent columns per block, so each row read is LW
// consecutive doubles.  Variants 4-5 trade halo for tile length: a tile of
// BS / LW positions keeps BS / LW - 8 outputs.
constexpr int kAdjVariants = 6;
constexpr int kAdjDefault[2] = {0, 0};

template <class A = AdjFull>
inline void launch_adjoint_dir(vg_ctx *ctx, int dir, int variant, const AdjointArgs &a)
{
    if (dir == 0) {
        switch (variant) {
        case 1: launch_adjoint_tile<A, 0, 128, 1, 3>(ctx, a); break;
        case 2: launch_adjoint_tile<A, 0, 256, 1, 2>(ctx, a); break;
        case 3: launch_adjoint_tile<A, 0, 64, 1, 2>(ctx, a); break;
        case 4: launch_adjoint_tile<A, 0, 256, 1, 3>(ctx, a); break;
        case 5: launch_adjoint_tile<A, 0, 512, 1, 3>(ctx, a); break;
        default: launch_adjoint_tile<A, 0, 128, 1, 2>(ctx, a); break;
        }
        return;
    }
    switch (variant) {
    case 1: launch_adjoint_tile<A, 1, 256, 8, 3>(ctx, a); break;
    case 2: launch_adjoint_tile<A, 1, 256, 16, 2>(ctx, a); break;
    case 3: launch_adjoint_tile<A, 1, 256, 4, 2>(ctx, a); break;
    case 4: launch_adjoint_tile<A, 1, 512, 8, 3>(ctx, a); break;
    case 5: launch_adjoint_tile<A, 1, 512, 16, 3>(ctx, a); break;
    default: launch_adjoint_tile<A, 1, 256, 8, 2>(ctx, a); break;
    }
}

inline AdjointArgs adjoint_args(const double *w, const double *wbar_out, double *wbar_in, double *ghost,
                             long long plane, int ny, int nx, int bc, double h, double bias, int limiter, double ts,
                             const EulerPhys::Params &m)
{
    AdjointArgs a;
    a.w = w;
    a.w35 = w;
    a.wbar_out = wbar_out;
    a.wbar_in = wbar_in;
    a.ghost = ghost;
    a.plane = plane;
    a.ny = ny;
    a.nx = nx;
    a.bc = bc;
    a.limiter = limiter;
    a.dx = h;
    a.bias = bias;
    a.ts = ts;
    a.m = m;
    return a;
}

// One adjoint sweep: wbar_in = (d sweep / d w)^T wbar_out at input w.  Every
// tile variant computes each output with the same arithmetic (the halo is
// recomputed, not shared), so the variant changes nothing in the result.
template <class A = AdjFull>
inline int adj_sweep(vg_ctx *ctx, int dir, int variant, const double *w, const double *wbar_out, double *wbar_in,
              double *ghost, long long plane, int ny, int nx, int bc, double h, double bias, int limiter, double ts,
              const EulerPhys::Params &m, const double *w35 = nullptr)
{
    AdjointArgs a = adjoint_args(w, wbar_out, wbar_in, ghost, plane, ny, nx, bc, h, bias, limiter, ts, m);
    if (w35) a.w35 = w35;
    const int lines = (dir == 0) ? ny : nx;
    return ctx->timed(dir == 0 ? "adjoint_sweep_x" : "adjoint_sweep_y", [&] {
        launch_adjoint_dir<A>(ctx, dir, variant, a);
        if (dir == 0)
            euler_adjoint_fold_kernel<A, 0><<<(lines + 255) / 256, 256, 0, ctx->stream>>>(wbar_in, ghost, plane, ny,
                                                                                               nx, bc);
        else
            euler_adjoint_fold_kernel<A, 1><<<(lines + 255) / 256, 256, 0, ctx->stream>>>(wbar_in, ghost, plane, ny,
                                                                                               nx, bc);
    });
}

inline std::string adjoint_key(const char *phys, int dir, int ny, int nx)
{
    return std::string(phys) + (dir == 0 ? "_x/" : "_y/") + std::to_string(ny) + "/" + std::to_string(nx);
}

// Tile variant of one direction for an ny x nx grid: the context's table, else
// each variant runs once to warm up and twice timed on the caller's buffers
// (the real sweep overwrites what they write), and the fastest is recorded.
template <class A = AdjFull>
inline int tune_adjoint(vg_ctx *ctx, int dir, const AdjointArgs &a, int &variant)
{
    const std::string key = adjoint_key(A::key, dir, a.ny, a.nx);
    auto it = ctx->tuned.find(key);
    if (it != ctx->tuned.end()) {
        variant = it->second;
        return VG_OK;
    }
    variant = kAdjDefault[dir];
    if (!ctx->autotune || ctx->timing) return VG_OK;
    hipEvent_t e0, e1;
    VG_HIP(hipEventCreate(&e0));
    VG_HIP(hipEventCreate(&e1));
    float best = 0.0f;
    for (int v = 0; v < kAdjVariants; v++) {
        launch_adjoint_dir<A>(ctx, dir, v, a);
        VG_HIP(hipEventRecord(e0, ctx->stream));
        for (int r = 0; r < 2; r++) launch_adjoint_dir<A>(ctx, dir, v, a);
        VG_HIP(hipEventRecord(e1, ctx->stream));
        VG_HIP(hipEventSynchronize(e1));
        int rc = launch_status("adjoint autotune");
        if (rc) return rc;
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

}  // namespace vg
