// Two-material diffuse-interface Euler C ABI: the reference's
// SecondOrderSolver / FirstOrderSolver solve[2D] and per-sweep overloads for
// EulerMultiphysicsStateVector and EulerReducedStateVector
// (Solvers/secondordersolver.h:34-69, firstordersolver.h:65-99) on the generic
// device grid, and MultiphysicsSolvers::reinitialiseVolumeFraction
// (Solvers/multiphysicssolvers.h:12-17) as the post-step stage of the
// second-order solves and as a call of its own.
#include <cmath>
#include <cstring>
#include <string>

#include "vg_grid.h"
#include "vg_phys_twomat.h"

namespace vg {

// Plane indices of the fields the reinitialisation touches.
template <class P>
struct ReinitFields;
template <>
struct ReinitFields<EulerMPPhys> {
    static constexpr int rho1 = 4, rho2 = 6, p1 = 5, p2 = 7;
    static constexpr bool pressures = true;
};
template <>
struct ReinitFields<EulerReducedPhys> {
    static constexpr int rho1 = 4, rho2 = 5, p1 = 0, p2 = 0;
    static constexpr bool pressures = false;
};

// relaxTotalDensity (and, multiphysics, relaxTotalPressure), then the new
// volume fraction (eulermultiphysicsstatevector.cpp:334-365,
// eulerreducedstatevector.cpp:337-350): rho_k := alpha1 rho1 + alpha2 rho2 with
// the stored alpha1, likewise p_k; the reference's unused mixture energy has no
// effect.
template <class P>
__device__ __forceinline__ void relax_cell(double *w, long long plane, long long c, double alpha)
{
    using F = ReinitFields<P>;
    const double a = w[c], a2 = 1.0 - a;
    const double rho = (a * w[F::rho1 * plane + c]) + (a2 * w[F::rho2 * plane + c]);
    w[F::rho1 * plane + c] = rho;
    w[F::rho2 * plane + c] = rho;
    if constexpr (F::pressures) {
        const double p = (a * w[F::p1 * plane + c]) + (a2 * w[F::p2 * plane + c]);
        w[F::p1 * plane + c] = p;
        w[F::p2 * plane + c] = p;
    }
    w[c] = alpha;
}

// Whether this step reinitialises: it ran (the clock is not done and the step
// is pending) and its number, counted from 1, is a multiple of k.  A null clock
// (the standalone call) always does.
__device__ __forceinline__ bool reinit_due(const Clock *c, int k)
{
    if (!c) return true;
    return !c->done && c->pending && ((c->it + 1) % k) == 0;
}

// 1D: the last falling crossing i >= 1 (alpha_i < 0.5 <= alpha_{i-1}), or 0.
__global__ void reinit_1d_init_kernel(const Clock *clk, int k, int *loc)
{
    if (reinit_due(clk, k)) *loc = 0;
}
__global__ __launch_bounds__(256) void reinit_1d_find_kernel(const Clock *clk, int k, const double *__restrict__ w,
                                                             int n, int *loc)
{
    if (!reinit_due(clk, k)) return;
    for (int i = blockIdx.x * 256 + threadIdx.x + 1; i < n; i += gridDim.x * 256)
        if (w[i] < 0.5 && w[i - 1] >= 0.5) atomicMax(loc, i);
}
// The window i-2 .. i+1 (0.999 before i, 0.001 from i on); cells outside the
// line are skipped (the reference writes cells -2 and -1 there).
template <class P>
__global__ void reinit_1d_window_kernel(const Clock *clk, int k, double *w, long long plane, int n, const int *loc)
{
    if (!reinit_due(clk, k)) return;
    const int l = *loc, i = l - 2 + (int)threadIdx.x;
    if (threadIdx.x < 4 && i >= 0 && i < n) relax_cell<P>(w, plane, i, i < l ? 0.999 : 0.001);
}

// 2D: the crossings of every line from the volume fractions before any of
// them changes: code 1 falling (alpha_j < 0.5 <= alpha_{j-1}), 2 rising
// (alpha_j > 0.5 >= alpha_{j-1}) along x (DIR 0) or y (DIR 1).
template <int DIR>
__global__ __launch_bounds__(256) void reinit_codes_kernel(const Clock *clk, int k, const double *__restrict__ w,
                                                           int ny, int nx, unsigned char *__restrict__ code)
{
    if (!reinit_due(clk, k)) return;
    const long long cells = (long long)ny * nx;
    for (long long c = (long long)blockIdx.x * 256 + threadIdx.x; c < cells; c += (long long)gridDim.x * 256) {
        const long long pos = (DIR == 0) ? c % nx : c / nx;
        unsigned char v = 0;
        if (pos > 0) {
            const double a = w[c], ap = w[c - (DIR == 0 ? 1 : nx)];
            if (a < 0.5 && ap >= 0.5) v = 1;
            else if (a > 0.5 && ap <= 0.5) v = 2;
        }
        code[c] = v;
    }
}

// One thread per line (a row for DIR 0, a column for DIR 1): every falling
// crossing's window in order, then every rising one's
// (multiphysicssolvers.cpp:37-161, 193-313).  Window cells outside the line are
// skipped (the reference indexes cells -1 and n there), and a line may hold any
// number of crossings (the reference's 100 per kind).
template <class P, int DIR>
__global__ __launch_bounds__(64) void reinit_lines_kernel(const Clock *clk, int k, double *w, long long plane, int ny,
                                                          int nx, const unsigned char *__restrict__ code)
{
    if (!reinit_due(clk, k)) return;
    const int line = blockIdx.x * 64 + threadIdx.x;
    const int lines = (DIR == 0) ? ny : nx, n = (DIR == 0) ? nx : ny;
    if (line >= lines) return;
    const long long base = (DIR == 0) ? (long long)line * nx : line;
    const long long st = (DIR == 0) ? 1 : nx;
    for (int pass = 1; pass <= 2; pass++)
        for (int j = 1; j < n; j++) {
            if (code[base + j * st] != pass) continue;
            for (int q = j - 2; q < j + 2; q++) {
                if (q < 0 || q >= n) continue;
                const double alpha = (pass == 1) ? (q < j ? 0.999 : 0.001) : (q < j ? 0.001 : 0.999);
                relax_cell<P>(w, plane, base + q * st, alpha);
            }
        }
}

// Launches one reinitialisation of the grid `w` (clock-gated when clk is set).
template <class P>
int launch_reinit(hipStream_t s, const Clock *clk, int k, double *w, long long plane, int dim, int ny, int nx,
                  void *scratch)
{
    const long long cells = (long long)ny * nx;
    const int blocks = (int)std::min<long long>((cells + 255) / 256, 4096);
    if (dim == 1) {
        int *loc = reinterpret_cast<int *>(scratch);
        reinit_1d_init_kernel<<<1, 1, 0, s>>>(clk, k, loc);
        reinit_1d_find_kernel<<<blocks, 256, 0, s>>>(clk, k, w, nx, loc);
        reinit_1d_window_kernel<P><<<1, 64, 0, s>>>(clk, k, w, plane, nx, loc);
        return launch_status("reinit_1d");
    }
    unsigned char *code = reinterpret_cast<unsigned char *>(scratch);
    reinit_codes_kernel<0><<<blocks, 256, 0, s>>>(clk, k, w, ny, nx, code);
    reinit_lines_kernel<P, 0><<<(ny + 63) / 64, 64, 0, s>>>(clk, k, w, plane, ny, nx, code);
    reinit_codes_kernel<1><<<blocks, 256, 0, s>>>(clk, k, w, ny, nx, code);
    reinit_lines_kernel<P, 1><<<(nx + 63) / 64, 64, 0, s>>>(clk, k, w, plane, ny, nx, code);
    return launch_status("reinit_2d");
}

inline size_t reinit_scratch_bytes(long long cells) { return (size_t)std::max<long long>(cells, 4) + 256; }

// After every reinit_every-th step of a second-order solve
// (secondordersolver.cpp:343-349, 452-458).
template <class P>
struct TwoMaterialPostStep {
    static bool active(const TwoMaterialParams &m) { return m.reinit_every > 0; }
    template <class GS>
    static int prepare(GS &g, const TwoMaterialParams &)
    {
        if (!g.post_buf) VG_HIP(hipMalloc(&g.post_buf, reinit_scratch_bytes(g.cells)));
        return VG_OK;
    }
    template <class GS>
    static int launch(GS &g, double *w, const TwoMaterialParams &m)
    {
        return launch_reinit<P>(g.ctx->stream, g.clk, m.reinit_every, w, g.plane, g.dim, g.ny, g.nx, g.post_buf);
    }
};
template <>
struct PostStep<EulerMPPhys> : TwoMaterialPostStep<EulerMPPhys> {};
template <>
struct PostStep<EulerReducedPhys> : TwoMaterialPostStep<EulerReducedPhys> {};

}  // namespace vg

namespace {

vg::EulerPhys::Params euler_params(const vg_euler_params *p)
{
    vg::EulerPhys::Params m{};
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

vg::TwoMaterialParams to_params(const vg_euler_params *p1, const vg_euler_params *p2, int reinit_every)
{
    vg::TwoMaterialParams m{};
    m.m1 = euler_params(p1);
    m.m2 = euler_params(p2);
    m.reinit_every = reinit_every;
    return m;
}

int check_run_args(int dim, int ny, int nx, double dx, double cfl, double t_final, int max_steps, int order, int bc,
                   int reinit_every, const vg_euler_params *p1, const vg_euler_params *p2)
{
    VG_CHECK(p1 != nullptr && p2 != nullptr, "null material parameters");
    VG_CHECK(order == 1 || order == 2, "order must be 1 (FORCE) or 2 (SLIC)");
    VG_CHECK(bc == VG_BC_TRANSMISSIVE || bc == VG_BC_PERIODIC, "unknown boundary condition");
    VG_CHECK(dim == 1 || dim == 2, "dim must be 1 or 2");
    VG_CHECK(nx >= 2 && ny >= 1 && (dim == 1 ? ny == 1 : ny >= 2), "grid needs at least 2 cells per direction");
    VG_CHECK((long long)ny * nx <= (1ll << 31), "grid too large for one state");
    VG_CHECK(dx > 0.0 && std::isfinite(dx), "cell spacing must be positive");
    VG_CHECK(cfl > 0.0 && std::isfinite(cfl), "CFL coefficient must be positive");
    VG_CHECK(max_steps >= 0, "max_steps must be >= 0");
    VG_CHECK(max_steps > 0 || std::isfinite(t_final), "an infinite final time needs a step cap");
    VG_CHECK(reinit_every >= 0, "reinitialisation frequency must be >= 0");
    return VG_OK;
}

template <class P>
int solve(vg_ctx *ctx, int dim, const double *prim_in, double *prim_out, int ny, int nx, double dx, double cfl,
          double t_final, int max_steps, double bias, int limiter, int order, int bc, int reinit_every,
          const vg_euler_params *p1, const vg_euler_params *p2, double *dt_log, int dt_log_cap, vg_run_stats *stats)
{
    vg::clear_error();
    VG_CHECK(ctx != nullptr, "null context");
    VG_CHECK(prim_in != nullptr && prim_out != nullptr, "null state pointer");
    int rc = check_run_args(dim, ny, nx, dx, cfl, t_final, max_steps, order, bc, reinit_every, p1, p2);
    if (rc) return rc;
    rc = vg::bind(ctx);
    if (rc) return rc;
    const vg::TwoMaterialParams m = to_params(p1, p2, reinit_every);
    vg::GridState<P> g;
    rc = g.create(ctx, dim, ny, nx);
    if (!rc) rc = g.load(prim_in, 0);
    vg_run_stats local{};
    // One arithmetic: EXACT.
    if (!rc) rc = g.run(order, VG_MODE_EXACT, dx, cfl, t_final, max_steps, bias, limiter, bc, m, dt_log, dt_log_cap,
                        &local);
    if (!rc) rc = g.store(prim_out, 0);
    if (!rc) rc = g.scan_nonfinite(m, &local.nonfinite_cells, &local.first_nonfinite);
    g.destroy();
    if (stats) *stats = local;
    if (!rc && local.nonfinite_cells > 0) {
        vg::set_error("final state holds " + std::to_string(local.nonfinite_cells) +
                      " non-finite cells, first at flat index " +
                      std::to_string(local.first_nonfinite));
        return VG_ERR_NONFINITE;
    }
    return rc;
}

template <class P>
int time_step(vg_ctx *ctx, int dim, double *cells, const double *with_boundary, int ny, int nx, int dir, double dx,
              double dt, double bias, int limiter, int order, const vg_euler_params *p1, const vg_euler_params *p2)
{
    vg::clear_error();
    VG_CHECK(ctx != nullptr, "null context");
    VG_CHECK(p2 != nullptr, "null material parameters");
    int rc = vg::check_time_step_args(dim, ny, nx, dir, dx, dt, order, VG_MODE_EXACT, cells, with_boundary, p1);
    if (rc) return rc;
    rc = vg::bind(ctx);
    if (rc) return rc;
    return vg::ghosted_time_step<P>(ctx, dim, cells, with_boundary, ny, nx, dir, dx, dt, bias, limiter, order,
                                    VG_MODE_EXACT, to_params(p1, p2, 0));
}

template <class P>
int reinitialise(vg_ctx *ctx, int dim, double *prim, int ny, int nx)
{
    vg::clear_error();
    VG_CHECK(ctx != nullptr && prim != nullptr, "null argument");
    VG_CHECK(dim == 1 || dim == 2, "dim must be 1 or 2");
    VG_CHECK(nx >= 1 && ny >= 1 && (dim == 1 ? ny == 1 : true), "grid needs at least 1 cell");
    VG_CHECK((long long)ny * nx <= (1ll << 31), "grid too large for one state");
    int rc = vg::bind(ctx);
    if (rc) return rc;
    vg::GridState<P> g;
    rc = g.create(ctx, dim, ny, nx);
    if (!rc) rc = g.load(prim, 0);
    void *scratch = nullptr;
    if (!rc && hipMalloc(&scratch, vg::reinit_scratch_bytes(g.cells)) != hipSuccess) {
        vg::set_error("hipMalloc of the reinitialisation scratch failed");
        rc = VG_ERR_HIP;
    }
    if (!rc) rc = vg::launch_reinit<P>(ctx->stream, nullptr, 1, g.buf[g.cur], g.plane, dim, ny, nx, scratch);
    if (!rc) rc = g.store(prim, 0);
    hipStreamSynchronize(ctx->stream);
    hipFree(scratch);
    g.destroy();
    return rc;
}

}  // namespace

extern "C" {

int vg_euler_multiphysics_solve_1d(vg_ctx *ctx, const double *prim_in, double *prim_out, int n, double dx, double cfl,
                                   double t_final, int max_steps, double bias, int limiter, int order, int bc,
                                   int reinit_every, const vg_euler_params *p1, const vg_euler_params *p2,
                                   double *dt_log, int dt_log_cap, vg_run_stats *stats)
{
    return solve<vg::EulerMPPhys>(ctx, 1, prim_in, prim_out, 1, n, dx, cfl, t_final, max_steps, bias, limiter, order,
                                  bc, reinit_every, p1, p2, dt_log, dt_log_cap, stats);
}

int vg_euler_multiphysics_solve_2d(vg_ctx *ctx, const double *prim_in, double *prim_out, int ny, int nx, double dx,
                                   double cfl, double t_final, int max_steps, double bias, int limiter, int order,
                                   int bc, int reinit_every, const vg_euler_params *p1, const vg_euler_params *p2,
                                   double *dt_log, int dt_log_cap, vg_run_stats *stats)
{
    return solve<vg::EulerMPPhys>(ctx, 2, prim_in, prim_out, ny, nx, dx, cfl, t_final, max_steps, bias, limiter,
                                  order, bc, reinit_every, p1, p2, dt_log, dt_log_cap, stats);
}

int vg_euler_reduced_solve_1d(vg_ctx *ctx, const double *prim_in, double *prim_out, int n, double dx, double cfl,
                              double t_final, int max_steps, double bias, int limiter, int order, int bc,
                              int reinit_every, const vg_euler_params *p1, const vg_euler_params *p2, double *dt_log,
                              int dt_log_cap, vg_run_stats *stats)
{
    return solve<vg::EulerReducedPhys>(ctx, 1, prim_in, prim_out, 1, n, dx, cfl, t_final, max_steps, bias, limiter,
                                       order, bc, reinit_every, p1, p2, dt_log, dt_log_cap, stats);
}

int vg_euler_reduced_solve_2d(vg_ctx *ctx, const double *prim_in, double *prim_out, int ny, int nx, double dx,
                              double cfl, double t_final, int max_steps, double bias, int limiter, int order, int bc,
                              int reinit_every, const vg_euler_params *p1, const vg_euler_params *p2, double *dt_log,
                              int dt_log_cap, vg_run_stats *stats)
{
    return solve<vg::EulerReducedPhys>(ctx, 2, prim_in, prim_out, ny, nx, dx, cfl, t_final, max_steps, bias, limiter,
                                       order, bc, reinit_every, p1, p2, dt_log, dt_log_cap, stats);
}

int vg_euler_multiphysics_time_step_1d(vg_ctx *ctx, double *cells, const double *cells_with_boundary, int n,
                                       double dx, double dt, double bias, int limiter, int order,
                                       const vg_euler_params *p1, const vg_euler_params *p2)
{
    return time_step<vg::EulerMPPhys>(ctx, 1, cells, cells_with_boundary, 1, n, 0, dx, dt, bias, limiter, order, p1,
                                      p2);
}

int vg_euler_multiphysics_time_step_2d(vg_ctx *ctx, double *cells, const double *cells_with_boundary, int ny, int nx,
                                       int dir, double dx, double dt, double bias, int limiter, int order,
                                       const vg_euler_params *p1, const vg_euler_params *p2)
{
    return time_step<vg::EulerMPPhys>(ctx, 2, cells, cells_with_boundary, ny, nx, dir, dx, dt, bias, limiter, order,
                                      p1, p2);
}

int vg_euler_reduced_time_step_1d(vg_ctx *ctx, double *cells, const double *cells_with_boundary, int n, double dx,
                                  double dt, double bias, int limiter, int order, const vg_euler_params *p1,
                                  const vg_euler_params *p2)
{
    return time_step<vg::EulerReducedPhys>(ctx, 1, cells, cells_with_boundary, 1, n, 0, dx, dt, bias, limiter, order,
                                           p1, p2);
}

int vg_euler_reduced_time_step_2d(vg_ctx *ctx, double *cells, const double *cells_with_boundary, int ny, int nx,
                                  int dir, double dx, double dt, double bias, int limiter, int order,
                                  const vg_euler_params *p1, const vg_euler_params *p2)
{
    return time_step<vg::EulerReducedPhys>(ctx, 2, cells, cells_with_boundary, ny, nx, dir, dx, dt, bias, limiter,
                                           order, p1, p2);
}

int vg_euler_multiphysics_reinitialise(vg_ctx *ctx, double *prim, int dim, int ny, int nx)
{
    return reinitialise<vg::EulerMPPhys>(ctx, dim, prim, ny, nx);
}

int vg_euler_reduced_reinitialise(vg_ctx *ctx, double *prim, int dim, int ny, int nx)
{
    return reinitialise<vg::EulerReducedPhys>(ctx, dim, prim, ny, nx);
}

}  // extern "C"
