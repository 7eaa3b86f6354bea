// Euler C ABI: whole-solve drop-ins and the device-resident state.
#include <cmath>
#include <cstring>

#include <vector>

#include "vg_batch_host.h"
#include "vg_grid.h"
#include "vg_phys_euler.h"

using EulerGrid = vg::GridState<vg::EulerPhys>;

struct vg_euler_state {
    EulerGrid g;
};

namespace {

// 1D grids up to this size run as one persistent workgroup.
constexpr int kPersistent1D = 32768;

vg::EulerPhys::Params to_params(const vg_euler_params *p)
{
    vg::EulerPhys::Params m;
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

int check_run_args(int dim, int ny, int nx, double dx, double cfl, double t_final, int max_steps, int limiter,
                   int order, int bc, const vg_euler_params *mp, int mode)
{
    (void)limiter;
    VG_CHECK(mp != nullptr, "null material parameters");
    VG_CHECK(order == 1 || order == 2, "order must be 1 (FORCE) or 2 (SLIC)");
    VG_CHECK(mode == VG_MODE_EXACT || mode == VG_MODE_FAST, "mode must be VG_MODE_EXACT or VG_MODE_FAST");
    VG_CHECK(bc == VG_BC_TRANSMISSIVE || bc == VG_BC_PERIODIC, "unknown boundary condition");
    VG_CHECK(dim == 1 || dim == 2, "dim must be 1 or 2");
    VG_CHECK(nx >= 2 && ny >= 1 && (dim == 1 ? ny == 1 : ny >= 2), "grid needs at least 2 cells per direction");
    VG_CHECK((long long)ny * nx <= (1ll << 31), "grid too large for one state");
    VG_CHECK(dx > 0.0 && std::isfinite(dx), "cell spacing must be positive");
    VG_CHECK(cfl > 0.0 && std::isfinite(cfl), "CFL coefficient must be positive");
    VG_CHECK(max_steps >= 0, "max_steps must be >= 0");
    VG_CHECK(max_steps > 0 || std::isfinite(t_final), "an infinite final time needs a step cap");
    return VG_OK;
}

int solve(vg_ctx *ctx, int dim, const double *prim_in, double *prim_out, int ny, int nx, double dx, double cfl,
          double t_final, int max_steps, double bias, int limiter, int order, int bc, const vg_euler_params *mp,
          int mode, double *dt_log, int dt_log_cap, vg_run_stats *stats)
{
    vg::clear_error();
    VG_CHECK(ctx != nullptr, "null context");
    VG_CHECK(prim_in != nullptr && prim_out != nullptr, "null state pointer");
    int rc = check_run_args(dim, ny, nx, dx, cfl, t_final, max_steps, limiter, order, bc, mp, mode);
    if (rc) return rc;
    rc = vg::bind(ctx);
    if (rc) return rc;
    if (dim == 1 && nx <= kPersistent1D) {
        // A small 1D grid runs its whole time loop in one persistent workgroup:
        // one launch instead of a launch per step.
        const int64_t offsets[2] = {0, nx};
        std::vector<vg::EulerPhys::Params> params(1, to_params(mp));
        int32_t steps = 0;
        double tend = 0.0;
        vg_run_stats local{};
        rc = vg::run_batch<vg::EulerPhys, 256>(ctx, order, mode, 1, offsets, prim_in, prim_out, &dx, params, &t_final,
                                               cfl, max_steps, bias, limiter, bc, dt_log, dt_log_cap, &steps, &tend,
                                               &local);
        local.steps = local.total_steps = steps;
        local.t = tend;
        local.first_nonfinite = -1;
        for (int j = 0; j < nx && rc == VG_OK; j++) {
            const double *w = prim_out + 6 * (size_t)j;
            bool bad = !(w[0] > 0.0) || !((w[4] + mp->p_inf) > 0.0);
            for (int f = 0; f < 6; f++) bad = bad || !std::isfinite(w[f]);
            if (bad) {
                if (local.first_nonfinite < 0) local.first_nonfinite = j;
                local.nonfinite_cells += 1;
            }
        }
        if (stats) *stats = local;
        if (!rc && local.nonfinite_cells > 0) {
            vg::set_error("final state holds " + std::to_string(local.nonfinite_cells) +
                          " non-finite or inadmissible cells, first at flat index " +
                          std::to_string(local.first_nonfinite));
            return VG_ERR_NONFINITE;
        }
        return rc;
    }
    EulerGrid g;
    rc = g.create(ctx, dim, ny, nx);
    if (!rc) rc = g.load(prim_in, 0);
    vg_run_stats local{};
    if (!rc) rc = g.run(order, mode, dx, cfl, t_final, max_steps, bias, limiter, bc, to_params(mp), dt_log, dt_log_cap, &local);
    if (!rc) rc = g.store(prim_out, 0);
    if (!rc) rc = g.scan_nonfinite(to_params(mp), &local.nonfinite_cells, &local.first_nonfinite);
    g.destroy();
    if (stats) *stats = local;
    if (!rc && local.nonfinite_cells > 0) {
        vg::set_error("final state holds " + std::to_string(local.nonfinite_cells) +
                      " non-finite or inadmissible cells (rho <= 0 or p + p_inf <= 0), first at flat index " +
                      std::to_string(local.first_nonfinite));
        return VG_ERR_NONFINITE;
    }
    return rc;
}

}  // namespace

extern "C" {

int vg_euler_solve_1d(vg_ctx *ctx, const double *prim_in, double *prim_out, int n, double dx, double cfl,
                      double t_final, int max_steps, double bias, int limiter, int order, int bc,
                      const vg_euler_params *mp, int mode, double *dt_log, int dt_log_cap, vg_run_stats *stats)
{
    return solve(ctx, 1, prim_in, prim_out, 1, n, dx, cfl, t_final, max_steps, bias, limiter, order, bc, mp, mode,
                 dt_log, dt_log_cap, stats);
}

int vg_euler_solve_2d(vg_ctx *ctx, const double *prim_in, double *prim_out, int ny, int nx, double dx, double cfl,
                      double t_final, int max_steps, double bias, int limiter, int order, int bc,
                      const vg_euler_params *mp, int mode, double *dt_log, int dt_log_cap, vg_run_stats *stats)
{
    return solve(ctx, 2, prim_in, prim_out, ny, nx, dx, cfl, t_final, max_steps, bias, limiter, order, bc, mp, mode,
                 dt_log, dt_log_cap, stats);
}

int vg_euler_state_create(vg_ctx *ctx, int dim, int ny, int nx, vg_euler_state **st)
{
    vg::clear_error();
    VG_CHECK(ctx != nullptr && st != nullptr, "null argument");
    VG_CHECK(dim == 1 || dim == 2, "dim must be 1 or 2");
    VG_CHECK(nx >= 2 && (dim == 1 ? ny == 1 : ny >= 2), "grid needs at least 2 cells per direction");
    int rc = vg::bind(ctx);
    if (rc) return rc;
    vg_euler_state *s = new vg_euler_state();
    rc = s->g.create(ctx, dim, ny, nx);
    if (rc) {
        s->g.destroy();
        delete s;
        return rc;
    }
    *st = s;
    return VG_OK;
}

void vg_euler_state_destroy(vg_euler_state *st)
{
    if (!st) return;
    vg::bind(st->g.ctx);
    hipStreamSynchronize(st->g.ctx->stream);
    st->g.destroy();
    delete st;
}

int vg_euler_state_load(vg_euler_state *st, const double *prim, int on_device)
{
    vg::clear_error();
    VG_CHECK(st != nullptr && prim != nullptr, "null argument");
    int rc = vg::bind(st->g.ctx);
    return rc ? rc : st->g.load(prim, on_device);
}

int vg_euler_state_store(vg_euler_state *st, double *prim, int on_device)
{
    vg::clear_error();
    VG_CHECK(st != nullptr && prim != nullptr, "null argument");
    int rc = vg::bind(st->g.ctx);
    return rc ? rc : st->g.store(prim, on_device);
}

int vg_euler_state_planes(vg_euler_state *st, double **planes, int64_t *plane)
{
    VG_CHECK(st != nullptr && planes != nullptr && plane != nullptr, "null argument");
    *planes = st->g.buf[st->g.cur];
    *plane = st->g.plane;
    st->g.planar_known = false;   // the caller may write any plane
    return VG_OK;
}

int vg_euler_state_run(vg_euler_state *st, int order, double dx, double cfl, double t_final, int max_steps,
                       double bias, int limiter, int bc, const vg_euler_params *mp, int mode, double *dt_log,
                       int dt_log_cap, vg_run_stats *stats)
{
    vg::clear_error();
    VG_CHECK(st != nullptr, "null state");
    int rc = check_run_args(st->g.dim, st->g.ny, st->g.nx, dx, cfl, t_final, max_steps, limiter, order, bc, mp, mode);
    if (rc) return rc;
    rc = vg::bind(st->g.ctx);
    if (rc) return rc;
    return st->g.run(order, mode, dx, cfl, t_final, max_steps, bias, limiter, bc, to_params(mp), dt_log, dt_log_cap,
                     stats);
}

int vg_euler_state_max_wave_speed(vg_euler_state *st, int bc, const vg_euler_params *mp, double *smax)
{
    vg::clear_error();
    VG_CHECK(st != nullptr && mp != nullptr && smax != nullptr, "null argument");
    int rc = vg::bind(st->g.ctx);
    if (rc) return rc;
    rc = st->g.refresh_max(bc, to_params(mp), st->g.fused_count(2));
    if (rc) return rc;
    vg::Clock h;
    VG_HIP(hipMemcpyAsync(&h, st->g.clk, sizeof(h), hipMemcpyDeviceToHost, st->g.ctx->stream));
    VG_HIP(hipStreamSynchronize(st->g.ctx->stream));
    double s;
    std::memcpy(&s, &h.smax, sizeof(s));
    if (h.corner > s) s = h.corner;
    *smax = s;
    return VG_OK;
}

int vg_euler_state_sweep(vg_euler_state *st, int dir, int order, double dx, double dt, double bias, int limiter, int bc,
                         const vg_euler_params *mp, int mode)
{
    vg::clear_error();
    VG_CHECK(st != nullptr && mp != nullptr, "null argument");
    VG_CHECK(order == 1 || order == 2, "order must be 1 (FORCE) or 2 (SLIC)");
    VG_CHECK(mode == VG_MODE_EXACT || mode == VG_MODE_FAST, "mode must be VG_MODE_EXACT or VG_MODE_FAST");
    VG_CHECK(bc == VG_BC_TRANSMISSIVE || bc == VG_BC_PERIODIC, "unknown boundary condition");
    VG_CHECK(dir == 0 || (dir == 1 && st->g.dim == 2), "dir must be 0 (x) or 1 (y; 2D only)");
    VG_CHECK(dx > 0.0 && std::isfinite(dx) && std::isfinite(dt), "cell spacing must be positive, dt finite");
    int rc = vg::bind(st->g.ctx);
    if (rc) return rc;
    return st->g.sweep_once(order, mode, dir, dt, bc, dx, bias, limiter, to_params(mp));
}

int vg_euler_time_step_1d(vg_ctx *ctx, double *cells, const double *cells_with_boundary, int n, double dx, double dt,
                          double bias, int limiter, int order, const vg_euler_params *mp, int mode)
{
    vg::clear_error();
    VG_CHECK(ctx != nullptr, "null context");
    int rc = vg::check_time_step_args(1, 1, n, 0, dx, dt, order, mode, cells, cells_with_boundary, mp);
    if (rc) return rc;
    rc = vg::bind(ctx);
    if (rc) return rc;
    return vg::ghosted_time_step<vg::EulerPhys>(ctx, 1, cells, cells_with_boundary, 1, n, 0, dx, dt, bias, limiter,
                                                order, mode, to_params(mp));
}

int vg_euler_time_step_2d(vg_ctx *ctx, double *cells, const double *cells_with_boundary, int ny, int nx, int dir,
                          double dx, double dt, double bias, int limiter, int order, const vg_euler_params *mp,
                          int mode)
{
    vg::clear_error();
    VG_CHECK(ctx != nullptr, "null context");
    int rc = vg::check_time_step_args(2, ny, nx, dir, dx, dt, order, mode, cells, cells_with_boundary, mp);
    if (rc) return rc;
    rc = vg::bind(ctx);
    if (rc) return rc;
    return vg::ghosted_time_step<vg::EulerPhys>(ctx, 2, cells, cells_with_boundary, ny, nx, dir, dx, dt, bias, limiter,
                                                order, mode, to_params(mp));
}

}  // extern "C"

extern "C" int vg_euler_solve_batch(vg_ctx *ctx, int n_inst, const int64_t *offsets, const double *prim_in,
                                    double *prim_out, const double *dx, const vg_euler_params *mp,
                                    const double *t_final, double cfl, int max_steps, double bias, int limiter,
                                    int order, int mode, int32_t *steps, double *t_out, vg_run_stats *stats)
{
    vg::clear_error();
    VG_CHECK(ctx != nullptr && mp != nullptr && n_inst >= 1, "null argument");
    int rc = vg::bind(ctx);
    if (rc) return rc;
    std::vector<vg::EulerPhys::Params> params(n_inst);
    for (int q = 0; q < n_inst; q++) params[q] = to_params(&mp[q]);
    return vg::run_batch<vg::EulerPhys, 256>(ctx, order, mode, n_inst, offsets, prim_in, prim_out, dx, params, t_final,
                                             cfl, max_steps, bias, limiter, VG_BC_TRANSMISSIVE, nullptr, 0, steps,
                                             t_out, stats);
}
