// Adjoint C ABI for the HPR second-order solve (stiffened gas, not plastic),
// and the fixed-time-step forward it differentiates.
//
// vg_hpr_adjoint_{1d,2d}: forward solve (the solve's own kernels and
// arithmetic) recording every dt and keeping a checkpoint every K steps
// (K = ceil(sqrt(steps)) unless given); then, segment by segment from the end,
// the segment is recomputed from its checkpoint with the recorded dts, storing
// the input of every operation of every step (sub-cycles, sweeps, sub-cycles),
// and the operations' adjoints run backwards through it.  The gradient is
// returned for the initial primitive state and for the two relaxation times
// (tau1: strain dissipation, tau2: thermal impulse); a parameter such as the
// viscosity reaches them through the caller's tau(mu).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "vg_hpr_adjoint.h"
#include "vg_hpr_host.h"

namespace {

using vg::HPRGrid;
using vg::HPRPhys;
namespace ha = vg::hpradj;
constexpr int NF = 17;

int check_adjoint_args(vg_ctx *ctx, int dim, int ny, int nx, double dx, int bc, int subcycles, const vg_hpr_params *mp)
{
    VG_CHECK(ctx != nullptr, "null context");
    int rc = vg::hpr_check_params(mp, subcycles);
    if (rc) return rc;
    rc = vg::hpr_check_adjoint_params(mp);
    if (rc) return rc;
    VG_CHECK(bc == VG_BC_TRANSMISSIVE || bc == VG_BC_PERIODIC, "unknown boundary condition");
    VG_CHECK(dim == 1 || dim == 2, "dim must be 1 or 2");
    VG_CHECK(nx >= 3 && (dim == 1 ? ny == 1 : ny >= 3), "the adjoint needs at least 3 cells per direction");
    VG_CHECK(ny <= 65535 && nx <= 65535, "the adjoint's launches take at most 65535 cells per direction");
    VG_CHECK((long long)ny * nx <= (1ll << 30), "grid too large for one state");
    VG_CHECK(dx > 0.0 && std::isfinite(dx), "cell spacing must be positive");
    return VG_OK;
}

// One step's operations: sub-cycles, sweeps (X in 1D; X(1/2) Y X(1/2) in 2D),
// sub-cycles.  kind 0 = sub-cycle, 1 = x sweep, 2 = y sweep; scale = the
// sweep's fraction of dt.
struct Op {
    int kind;
    double scale;
};

std::vector<Op> step_ops(int dim, int nsub)
{
    std::vector<Op> ops;
    for (int s = 0; s < nsub; s++) ops.push_back({0, 0.0});
    if (dim == 1) {
        ops.push_back({1, 1.0});
    } else {
        ops.push_back({1, 0.5});
        ops.push_back({2, 1.0});
        ops.push_back({1, 0.5});
    }
    for (int s = 0; s < nsub; s++) ops.push_back({0, 0.0});
    return ops;
}

// Fixed-dt forward over steps [s0, s1): states[(s - s0) * nops + o] receives
// the input of operation o of step s, states[(s1 - s0) * nops] the output.
int forward_fixed(vg_ctx *ctx, HPRGrid &g, const double *dts_dev, int s0, int s1, double dx, double bias, int limiter,
                  int bc, const HPRPhys::Params &m, const std::vector<Op> &ops, const std::vector<double *> &states)
{
    vg::Clock h{};
    h.it = s0;
    h.t_final = INFINITY;
    h.max_steps = s1;
    VG_HIP(hipMemcpyAsync(g.clk, &h, sizeof(h), hipMemcpyHostToDevice, ctx->stream));
    const int nops = (int)ops.size();
    for (int s = s0; s < s1; s++) {
        vg::clock_fixed_kernel<<<1, 1, 0, ctx->stream>>>(g.clk, dts_dev, s1);
        int rc = vg::launch_status("clock_fixed");
        for (int o = 0; o < nops && !rc; o++) {
            const int k = (s - s0) * nops + o;
            const double *in = states[k];
            double *out = states[k + 1];
            if (ops[o].kind == 0)
                rc = vg::SourceStage<HPRPhys>::launch(g, in, out, 0, bc, dx, bias, limiter, m);
            else
                rc = g.launch_sweep<2, vg::MODE_EXACT>(ops[o].kind - 1, in, out, ops[o].scale, 0, bc, dx, bias,
                                                       limiter, m);
        }
        if (rc) return rc;
    }
    return VG_OK;
}

// Planes of the sub-cycle adjoint's scratch: rt, um (9 x 17), ue (3 x 17), nb,
// umb, kb (4 x kMaxSlots), four neighbour accumulators.
constexpr int kForcePlanes = NF + 9 * NF + 3 * NF + NF + NF + 4 * vg::ForcingTraits<HPRPhys>::kMaxSlots + 4 * NF;

struct Work {
    double *wb[2] = {nullptr, nullptr};
    double *nbar = nullptr;
    double *ext[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    double *fs = nullptr;   // sub-cycle adjoint scratch, kForcePlanes planes
    double *tau_part = nullptr;
    long long eplane = 0;
    int force_blocks = 0;
};

ha::SweepAdjArgs sweep_args(const HPRGrid &g, const Work &w, const double *in, const double *out, const double *wb_out,
                            double *wb_in, int bc, int limiter, double dx, double bias, double ts,
                            const HPRPhys::Params &m)
{
    ha::SweepAdjArgs a;
    a.w = in;
    a.w_out = out;
    a.wb_out = wb_out;
    a.wb_in = wb_in;
    a.nbar = w.nbar;
    a.xlb = w.ext[0];
    a.xrb = w.ext[1];
    a.own = w.ext[2];
    a.tol = w.ext[3];
    a.tor = w.ext[4];
    a.plane = g.plane;
    a.eplane = w.eplane;
    a.ny = g.ny;
    a.nx = g.nx;
    a.bc = bc;
    a.limiter = limiter;
    a.dx = dx;
    a.bias = bias;
    a.ts = ts;
    a.m = m;
    return a;
}

// Launch shape over `len` positions per line: x is the fastest index in both
// directions (positions along x for DIR 0, lines along x for DIR 1).
dim3 line_grid(int dir, int len, int ny, int nx)
{
    if (dir == 0) return dim3((unsigned)((len + ha::kBS - 1) / ha::kBS), (unsigned)ny);
    return dim3((unsigned)((nx + ha::kBS - 1) / ha::kBS), (unsigned)len);
}

template <int DIR>
int sweep_adjoint_dir(vg_ctx *ctx, const ha::SweepAdjArgs &a)
{
    const int n = (DIR == 0) ? a.nx : a.ny;
    const bool periodic = (a.bc == VG_BC_PERIODIC);
    hipStream_t s = ctx->stream;
    return ctx->timed(DIR == 0 ? "hpr_adjoint_sweep_x" : "hpr_adjoint_sweep_y", [&] {
        ha::hpr_adj_nbar_kernel<DIR><<<line_grid(DIR, n, a.ny, a.nx), ha::kBS, 0, s>>>(a);
        ha::hpr_adj_face_kernel<DIR><<<line_grid(DIR, periodic ? n : n + 1, a.ny, a.nx), ha::kBS, 0, s>>>(a);
        ha::hpr_adj_evolve_kernel<DIR><<<line_grid(DIR, periodic ? n : n + 2, a.ny, a.nx), ha::kBS, 0, s>>>(a);
        ha::hpr_adj_gather_kernel<DIR><<<line_grid(DIR, n, a.ny, a.nx), ha::kBS, 0, s>>>(a);
    });
}

int forcing_adjoint(vg_ctx *ctx, const HPRGrid &g, const Work &w, const double *in, const double *wb_out,
                    double *wb_in, int bc, int limiter, double dx, double bias, double ts, const HPRPhys::Params &m)
{
    constexpr int NS = vg::ForcingTraits<HPRPhys>::kMaxSlots;
    ha::ForcingAdjArgs a;
    a.w = in;
    a.wb_out = wb_out;
    a.wb_in = wb_in;
    a.own = w.ext[0];
    a.tl = w.ext[1];
    a.tr = w.ext[2];
    a.tt = w.ext[3];
    a.tb = w.ext[4];
    double *f = w.fs;
    const long long P = g.plane;
    a.rt = f;
    a.um = f + NF * P;
    a.ue = f + 10 * NF * P;
    a.nb = f + 13 * NF * P;
    a.umb = f + 14 * NF * P;
    a.kb = f + 15 * NF * P;
    a.ulb = a.kb + 4 * NS * P;
    a.urb = a.ulb + NF * P;
    a.utb = a.urb + NF * P;
    a.ubb = a.utb + NF * P;
    a.tau_part = w.tau_part;
    a.plane = g.plane;
    a.ny = g.ny;
    a.nx = g.nx;
    a.dim = g.dim;
    a.bc = bc;
    a.limiter = limiter;
    a.dx = dx;
    a.bias = bias;
    a.ts = ts;
    a.m = m;
    const unsigned nb = (unsigned)w.force_blocks;
    hipStream_t s = ctx->stream;
    const double cx = ts / dx;
    return ctx->timed("hpr_adjoint_source", [&] {
        ha::hpr_adj_force_rt_kernel<<<nb, ha::kBS, 0, s>>>(a);
        const unsigned tiles = (unsigned)vg::forcing_blocks(g.dim, g.ny, g.nx);
        if (g.dim == 1) ha::hpr_adj_force_record_kernel<1><<<tiles, vg::kForceBS, 0, s>>>(a);
        else ha::hpr_adj_force_record_kernel<2><<<tiles, vg::kForceBS, 0, s>>>(a);
        for (int st = 2; st >= 0; st--) {
            const double coef = (st < 2) ? 0.5 : 1.0;
            ha::hpr_adj_force_source_kernel<<<nb, ha::kBS, 0, s>>>(a, st);
            if (g.dim == 1) {
                ha::hpr_adj_force_frac_kernel<0><<<nb, ha::kBS, 0, s>>>(a, st, 0, coef * cx, coef);
            } else {
                const double fx = (st < 2) ? 0.25 : 0.5, fy = (st < 2) ? 0.5 : 1.0;
                ha::hpr_adj_force_frac_kernel<0><<<nb, ha::kBS, 0, s>>>(a, st, 2, fx * cx, coef);
                ha::hpr_adj_force_frac_kernel<1><<<nb, ha::kBS, 0, s>>>(a, st, 1, fy * cx, coef);
                ha::hpr_adj_force_frac_kernel<0><<<nb, ha::kBS, 0, s>>>(a, st, 0, fx * cx, coef);
            }
        }
        if (g.dim == 1) ha::hpr_adj_force_final_kernel<1><<<nb, ha::kBS, 0, s>>>(a);
        else ha::hpr_adj_force_final_kernel<2><<<nb, ha::kBS, 0, s>>>(a);
        ha::hpr_adj_forcing_gather_kernel<<<nb, ha::kBS, 0, s>>>(a);
    });
}

int adjoint(vg_ctx *ctx, int dim, const double *prim_in, double *prim_out, const double *seed, double *grad,
            double *grad_tau, int ny, int nx, double dx, double cfl, double t_final, int max_steps, double bias,
            int limiter, int bc, int subcycles, const vg_hpr_params *mp, int checkpoint_every, double *dt_log,
            int dt_log_cap, vg_run_stats *stats)
{
    vg::clear_error();
    int rc = check_adjoint_args(ctx, dim, ny, nx, dx, bc, subcycles, mp);
    if (rc) return rc;
    VG_CHECK(prim_in && prim_out && seed && grad, "null state pointer");
    VG_CHECK(cfl > 0.0 && std::isfinite(cfl), "CFL coefficient must be positive");
    VG_CHECK(max_steps >= 0 && (max_steps > 0 || std::isfinite(t_final)), "an infinite final time needs a step cap");
    VG_CHECK(checkpoint_every >= 0, "checkpoint_every must be >= 0");
    rc = vg::bind(ctx);
    if (rc) return rc;
    auto w0 = std::chrono::steady_clock::now();
    const HPRPhys::Params m = vg::hpr_to_params(mp);
    hipStream_t s = ctx->stream;
    HPRGrid g;
    rc = g.create(ctx, dim, ny, nx);
    if (!rc) rc = g.load(prim_in, 0);
    if (!rc) rc = g.ensure_staging();
    if (rc) {
        g.destroy();
        return rc;
    }
    g.subcycles = subcycles;
    const size_t gbytes = sizeof(double) * NF * g.plane;
    const std::vector<Op> ops = step_ops(dim, subcycles);
    const int nops = (int)ops.size();
    const int K = checkpoint_every > 0 ? checkpoint_every
                                       : (max_steps > 0 ? std::max(1, (int)std::ceil(std::sqrt((double)max_steps))) : 16);
    const int cap = max_steps > 0 ? max_steps : 1 << 22;
    auto wsp = [&](size_t i, size_t bytes, double **p) {
        void *v = nullptr;
        int r = ctx->workspace_buffer(i, bytes, &v);
        *p = static_cast<double *>(v);
        return r;
    };
    // Workspace: 0 dts, 1-2 adjoint ping-pong, 3 N-bar, 4-8 extended work planes,
    // 9 relaxation-time partial sums, 10 sub-cycle adjoint scratch, 11 .. 11 +
    // K * nops - 1 stored states (the segment's start is its checkpoint), then
    // the checkpoints.
    const size_t st0 = 11, ck0 = st0 + (size_t)K * nops;
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
    rc = g.store(prim_out, 0);
    // ---- reverse ----------------------------------------------------------------------
    Work w;
    w.eplane = (std::max<long long>((long long)ny * (nx + 2), (long long)(ny + 2) * nx) + 63) / 64 * 64;
    w.force_blocks = (int)((g.cells + ha::kBS - 1) / ha::kBS);
    double *dts_dev = nullptr;
    if (!rc) rc = wsp(0, sizeof(double) * std::max(steps, 1), &dts_dev);
    if (!rc) rc = wsp(1, gbytes, &w.wb[0]);
    if (!rc) rc = wsp(2, gbytes, &w.wb[1]);
    if (!rc) rc = wsp(3, gbytes, &w.nbar);
    for (int i = 0; i < 5 && !rc; i++) rc = wsp(4 + i, sizeof(double) * NF * w.eplane, &w.ext[i]);
    if (!rc) rc = wsp(9, sizeof(double) * 2 * w.force_blocks, &w.tau_part);
    if (!rc && subcycles > 0) rc = wsp(10, sizeof(double) * kForcePlanes * g.plane, &w.fs);
    std::vector<double *> states((size_t)K * nops + 1, nullptr);
    for (int k = 1; k <= K * nops && !rc; k++) rc = wsp(st0 + k - 1, gbytes, &states[k]);
    if (rc) {
        g.destroy();
        return rc;
    }
    if (steps > 0) VG_HIP(hipMemcpyAsync(dts_dev, dts.data(), sizeof(double) * steps, hipMemcpyHostToDevice, s));
    VG_HIP(hipMemsetAsync(w.tau_part, 0, sizeof(double) * 2 * w.force_blocks, s));
    VG_HIP(hipMemcpyAsync(g.staging, seed, gbytes / g.plane * g.cells, hipMemcpyHostToDevice, s));
    vg::aos_to_soa_kernel<NF><<<HPRGrid::blocks_for(g.cells), 256, 0, s>>>(g.staging, w.wb[0], g.plane, g.cells);
    int cw = 0;
    for (int seg = (int)ckpt.size() - 1; seg >= 0 && !rc; seg--) {
        const int s0 = seg * K, s1 = std::min(steps, s0 + K);
        if (s0 >= s1) continue;
        states[0] = ckpt[seg];
        // the last stored slot of a short segment is its output
        std::vector<double *> used(states.begin(), states.begin() + (size_t)(s1 - s0) * nops + 1);
        rc = forward_fixed(ctx, g, dts_dev, s0, s1, dx, bias, limiter, bc, m, ops, used);
        for (int st = s1 - 1; st >= s0 && !rc; st--) {
            const double dt = dts[st];
            for (int o = nops - 1; o >= 0 && !rc; o--) {
                const int k = (st - s0) * nops + o;
                if (ops[o].kind == 0) {
                    rc = forcing_adjoint(ctx, g, w, used[k], w.wb[cw], w.wb[1 - cw], bc, limiter, dx, bias,
                                         0.5 * (dt / subcycles), m);
                } else {
                    const ha::SweepAdjArgs a = sweep_args(g, w, used[k], used[k + 1], w.wb[cw], w.wb[1 - cw], bc,
                                                          limiter, dx, bias, dt * ops[o].scale, m);
                    rc = (ops[o].kind == 1) ? sweep_adjoint_dir<0>(ctx, a) : sweep_adjoint_dir<1>(ctx, a);
                }
                cw = 1 - cw;
            }
        }
    }
    if (!rc) {
        vg::soa_to_aos_kernel<NF><<<HPRGrid::blocks_for(g.cells), 256, 0, s>>>(w.wb[cw], g.staging, g.plane, g.cells);
        rc = vg::launch_status("adjoint store");
    }
    std::vector<double> part(2 * (size_t)w.force_blocks);
    if (!rc) {
        hipError_t e = hipMemcpyAsync(grad, g.staging, gbytes / g.plane * g.cells, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess)
            e = hipMemcpyAsync(part.data(), w.tau_part, sizeof(double) * part.size(), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) {
            vg::set_error(std::string("adjoint copy: ") + hipGetErrorString(e));
            rc = VG_ERR_HIP;
        }
    }
    if (!rc && grad_tau) {
        double t1 = 0.0, t2 = 0.0;
        for (int b = 0; b < w.force_blocks; b++) {
            t1 += part[2 * b];
            t2 += part[2 * b + 1];
        }
        grad_tau[0] = t1;
        grad_tau[1] = t2;
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
                const double *dts, int n_steps, double bias, int limiter, int bc, int subcycles,
                const vg_hpr_params *mp)
{
    vg::clear_error();
    int rc = check_adjoint_args(ctx, dim, ny, nx, dx, bc, subcycles, mp);
    if (rc) return rc;
    VG_CHECK(prim_in && prim_out && (dts || n_steps == 0) && n_steps >= 0, "null argument");
    rc = vg::bind(ctx);
    if (rc) return rc;
    const HPRPhys::Params m = vg::hpr_to_params(mp);
    HPRGrid g;
    rc = g.create(ctx, dim, ny, nx);
    if (!rc) rc = g.load(prim_in, 0);
    g.subcycles = subcycles;
    const std::vector<Op> ops = step_ops(dim, subcycles);
    double *dts_dev = nullptr;
    void *v = nullptr;
    if (!rc) rc = ctx->workspace_buffer(0, sizeof(double) * std::max(n_steps, 1), &v);
    dts_dev = static_cast<double *>(v);
    if (!rc && n_steps > 0)
        VG_HIP(hipMemcpyAsync(dts_dev, dts, sizeof(double) * n_steps, hipMemcpyHostToDevice, ctx->stream));
    for (int s = 0; s < n_steps && !rc; s++) {
        // ping-pong through the grid's two buffers, one step at a time
        std::vector<double *> slots(ops.size() + 1);
        for (size_t k = 0; k <= ops.size(); k++) slots[k] = g.buf[(g.cur + k) & 1];
        rc = forward_fixed(ctx, g, dts_dev, s, s + 1, dx, bias, limiter, bc, m, ops, slots);
        g.cur = (g.cur + (int)ops.size()) & 1;
    }
    if (!rc) rc = g.store(prim_out, 0);
    g.destroy();
    return rc;
}

}  // namespace

extern "C" {

int vg_hpr_adjoint_1d(vg_ctx *ctx, const double *prim_in, double *prim_out, const double *seed, double *grad,
                      double *grad_tau, int n, double dx, double cfl, double t_final, int max_steps, double bias,
                      int limiter, int bc, int subcycles, const vg_hpr_params *mp, int checkpoint_every,
                      double *dt_log, int dt_log_cap, vg_run_stats *stats)
{
    return adjoint(ctx, 1, prim_in, prim_out, seed, grad, grad_tau, 1, n, dx, cfl, t_final, max_steps, bias, limiter,
                   bc, subcycles, mp, checkpoint_every, dt_log, dt_log_cap, stats);
}

int vg_hpr_adjoint_2d(vg_ctx *ctx, const double *prim_in, double *prim_out, const double *seed, double *grad,
                      double *grad_tau, int ny, int nx, double dx, double cfl, double t_final, int max_steps,
                      double bias, int limiter, int bc, int subcycles, const vg_hpr_params *mp, int checkpoint_every,
                      double *dt_log, int dt_log_cap, vg_run_stats *stats)
{
    return adjoint(ctx, 2, prim_in, prim_out, seed, grad, grad_tau, ny, nx, dx, cfl, t_final, max_steps, bias, limiter,
                   bc, subcycles, mp, checkpoint_every, dt_log, dt_log_cap, stats);
}

int vg_hpr_solve_fixed_dt_1d(vg_ctx *ctx, const double *prim_in, double *prim_out, int n, double dx, const double *dts,
                             int n_steps, double bias, int limiter, int bc, int subcycles, const vg_hpr_params *mp)
{
    return solve_fixed(ctx, 1, prim_in, prim_out, 1, n, dx, dts, n_steps, bias, limiter, bc, subcycles, mp);
}

int vg_hpr_solve_fixed_dt_2d(vg_ctx *ctx, const double *prim_in, double *prim_out, int ny, int nx, double dx,
                             const double *dts, int n_steps, double bias, int limiter, int bc, int subcycles,
                             const vg_hpr_params *mp)
{
    return solve_fixed(ctx, 2, prim_in, prim_out, ny, nx, dx, dts, n_steps, bias, limiter, bc, subcycles, mp);
}

}  // extern "C"
