// HPR C ABI: 1D/2D whole-solve drop-ins and the device-resident state, with the
// RK4 relaxation sub-cycles run as a source stage around the sweeps.
#include <cmath>
#include <cstring>
#include <string>

#include "vg_hpr_host.h"

using HPRGrid = vg::HPRGrid;

struct vg_hpr_state {
    HPRGrid g;
};

namespace {

using vg::HPRPhys;

HPRPhys::Params to_params(const vg_hpr_params *p) { return vg::hpr_to_params(p); }

int check_params(const vg_hpr_params *mp, int subcycles) { return vg::hpr_check_params(mp, subcycles); }

int check_run_args(int dim, int ny, int nx, double dx, double cfl, double t_final, int max_steps, int order, int bc,
                   int subcycles, const vg_hpr_params *mp, int mode)
{
    int rc = check_params(mp, subcycles);
    if (rc) return rc;
    VG_CHECK(order == 1 || order == 2, "order must be 1 (FORCE) or 2 (SLIC)");
    VG_CHECK(mode == VG_MODE_EXACT || mode == VG_MODE_FAST, "mode must be VG_MODE_EXACT or VG_MODE_FAST");
    VG_CHECK(bc == VG_BC_TRANSMISSIVE || bc == VG_BC_PERIODIC, "unknown boundary condition");
    VG_CHECK(dim == 1 || dim == 2, "dim must be 1 or 2");
    VG_CHECK(nx >= 2 && ny >= 1 && (dim == 1 ? ny == 1 : ny >= 2), "grid needs at least 2 cells per direction");
    VG_CHECK((long long)ny * nx <= (1ll << 30), "grid too large for one state");
    VG_CHECK(dx > 0.0 && std::isfinite(dx), "cell spacing must be positive");
    VG_CHECK(cfl > 0.0 && std::isfinite(cfl), "CFL coefficient must be positive");
    VG_CHECK(max_steps >= 0, "max_steps must be >= 0");
    VG_CHECK(max_steps > 0 || std::isfinite(t_final), "an infinite final time needs a step cap");
    return VG_OK;
}

int run_grid(HPRGrid &g, int order, double dx, double cfl, double t_final, int max_steps, double bias, int limiter,
             int bc, int subcycles, const vg_hpr_params *mp, int mode, double *dt_log, int dt_log_cap,
             vg_run_stats *stats)
{
    g.subcycles = subcycles;
    // HPR has no separate FAST arithmetic: both modes run the exact kernels.
    (void)mode;
    return g.run(order, VG_MODE_EXACT, dx, cfl, t_final, max_steps, bias, limiter, bc, to_params(mp), dt_log,
                 dt_log_cap, stats);
}

int solve(vg_ctx *ctx, int dim, const double *prim_in, double *prim_out, int ny, int nx, double dx, double cfl,
          double t_final, int max_steps, double bias, int limiter, int order, int bc, int subcycles,
          const vg_hpr_params *mp, int mode, double *dt_log, int dt_log_cap, vg_run_stats *stats)
{
    vg::clear_error();
    VG_CHECK(ctx != nullptr, "null context");
    VG_CHECK(prim_in != nullptr && prim_out != nullptr, "null state pointer");
    int rc = check_run_args(dim, ny, nx, dx, cfl, t_final, max_steps, order, bc, subcycles, mp, mode);
    if (rc) return rc;
    rc = vg::bind(ctx);
    if (rc) return rc;
    HPRGrid g;
    rc = g.create(ctx, dim, ny, nx);
    if (!rc) rc = g.load(prim_in, 0);
    vg_run_stats local{};
    if (!rc) rc = run_grid(g, order, dx, cfl, t_final, max_steps, bias, limiter, bc, subcycles, mp, mode, dt_log,
                           dt_log_cap, &local);
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

int vg_hpr_solve_1d(vg_ctx *ctx, const double *prim_in, double *prim_out, int n, double dx, double cfl,
                    double t_final, int max_steps, double bias, int limiter, int order, int bc, int subcycles,
                    const vg_hpr_params *mp, int mode, double *dt_log, int dt_log_cap, vg_run_stats *stats)
{
    return solve(ctx, 1, prim_in, prim_out, 1, n, dx, cfl, t_final, max_steps, bias, limiter, order, bc, subcycles,
                 mp, mode, dt_log, dt_log_cap, stats);
}

int vg_hpr_solve_2d(vg_ctx *ctx, const double *prim_in, double *prim_out, int ny, int nx, double dx, double cfl,
                    double t_final, int max_steps, double bias, int limiter, int order, int bc, int subcycles,
                    const vg_hpr_params *mp, int mode, double *dt_log, int dt_log_cap, vg_run_stats *stats)
{
    return solve(ctx, 2, prim_in, prim_out, ny, nx, dx, cfl, t_final, max_steps, bias, limiter, order, bc, subcycles,
                 mp, mode, dt_log, dt_log_cap, stats);
}

int vg_hpr_state_create(vg_ctx *ctx, int dim, int ny, int nx, vg_hpr_state **st)
{
    vg::clear_error();
    VG_CHECK(ctx != nullptr && st != nullptr, "null argument");
    VG_CHECK(dim == 1 || dim == 2, "dim must be 1 or 2");
    VG_CHECK(nx >= 2 && (dim == 1 ? ny == 1 : ny >= 2), "grid needs at least 2 cells per direction");
    VG_CHECK((long long)ny * nx <= (1ll << 30), "grid too large for one state");
    int rc = vg::bind(ctx);
    if (rc) return rc;
    vg_hpr_state *s = new vg_hpr_state();
    rc = s->g.create(ctx, dim, ny, nx);
    if (rc) {
        s->g.destroy();
        delete s;
        return rc;
    }
    *st = s;
    return VG_OK;
}

void vg_hpr_state_destroy(vg_hpr_state *st)
{
    if (!st) return;
    vg::bind(st->g.ctx);
    hipStreamSynchronize(st->g.ctx->stream);
    st->g.destroy();
    delete st;
}

int vg_hpr_state_load(vg_hpr_state *st, const double *prim, int on_device)
{
    vg::clear_error();
    VG_CHECK(st != nullptr && prim != nullptr, "null argument");
    int rc = vg::bind(st->g.ctx);
    return rc ? rc : st->g.load(prim, on_device);
}

int vg_hpr_state_store(vg_hpr_state *st, double *prim, int on_device)
{
    vg::clear_error();
    VG_CHECK(st != nullptr && prim != nullptr, "null argument");
    int rc = vg::bind(st->g.ctx);
    return rc ? rc : st->g.store(prim, on_device);
}

int vg_hpr_state_planes(vg_hpr_state *st, double **planes, int64_t *plane)
{
    vg::clear_error();
    VG_CHECK(st != nullptr && planes != nullptr && plane != nullptr, "null argument");
    *planes = st->g.buf[st->g.cur];
    *plane = st->g.plane;
    return VG_OK;
}

int vg_hpr_state_run(vg_hpr_state *st, int order, double dx, double cfl, double t_final, int max_steps, double bias,
                     int limiter, int bc, int subcycles, const vg_hpr_params *mp, int mode, double *dt_log,
                     int dt_log_cap, vg_run_stats *stats)
{
    vg::clear_error();
    VG_CHECK(st != nullptr, "null state");
    int rc = check_run_args(st->g.dim, st->g.ny, st->g.nx, dx, cfl, t_final, max_steps, order, bc, subcycles, mp,
                            mode);
    if (rc) return rc;
    rc = vg::bind(st->g.ctx);
    if (rc) return rc;
    return run_grid(st->g, order, dx, cfl, t_final, max_steps, bias, limiter, bc, subcycles, mp, mode, dt_log,
                    dt_log_cap, stats);
}

int vg_hpr_state_max_wave_speed(vg_hpr_state *st, int bc, const vg_hpr_params *mp, double *smax)
{
    vg::clear_error();
    VG_CHECK(st != nullptr && smax != nullptr, "null argument");
    int rc = check_params(mp, 0);
    if (rc) return rc;
    rc = vg::bind(st->g.ctx);
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

int vg_hpr_state_sweep(vg_hpr_state *st, int dir, int order, double dx, double dt, double bias, int limiter, int bc,
                       const vg_hpr_params *mp)
{
    vg::clear_error();
    VG_CHECK(st != nullptr, "null state");
    int rc = check_params(mp, 0);
    if (rc) return rc;
    VG_CHECK(order == 1 || order == 2, "order must be 1 (FORCE) or 2 (SLIC)");
    VG_CHECK(bc == VG_BC_TRANSMISSIVE || bc == VG_BC_PERIODIC, "unknown boundary condition");
    VG_CHECK(dir == 0 || (dir == 1 && st->g.dim == 2), "dir must be 0 (x) or 1 (y; 2D only)");
    VG_CHECK(dx > 0.0 && std::isfinite(dx) && std::isfinite(dt), "cell spacing must be positive, dt finite");
    rc = vg::bind(st->g.ctx);
    if (rc) return rc;
    return st->g.sweep_once(order, VG_MODE_EXACT, dir, dt, bc, dx, bias, limiter, to_params(mp));
}

int vg_hpr_time_step_1d(vg_ctx *ctx, double *cells, const double *cells_with_boundary, int n, double dx, double dt,
                        double bias, int limiter, int order, const vg_hpr_params *mp)
{
    vg::clear_error();
    VG_CHECK(ctx != nullptr, "null context");
    int rc = check_params(mp, 0);
    if (rc) return rc;
    rc = vg::check_time_step_args(1, 1, n, 0, dx, dt, order, VG_MODE_EXACT, cells, cells_with_boundary, mp);
    if (rc) return rc;
    rc = vg::bind(ctx);
    if (rc) return rc;
    return vg::ghosted_time_step<HPRPhys>(ctx, 1, cells, cells_with_boundary, 1, n, 0, dx, dt, bias, limiter, order,
                                          VG_MODE_EXACT, to_params(mp));
}

int vg_hpr_time_step_2d(vg_ctx *ctx, double *cells, const double *cells_with_boundary, int ny, int nx, int dir,
                        double dx, double dt, double bias, int limiter, int order, const vg_hpr_params *mp)
{
    vg::clear_error();
    VG_CHECK(ctx != nullptr, "null context");
    int rc = check_params(mp, 0);
    if (rc) return rc;
    rc = vg::check_time_step_args(2, ny, nx, dir, dx, dt, order, VG_MODE_EXACT, cells, cells_with_boundary, mp);
    if (rc) return rc;
    rc = vg::bind(ctx);
    if (rc) return rc;
    return vg::ghosted_time_step<HPRPhys>(ctx, 2, cells, cells_with_boundary, ny, nx, dir, dx, dt, bias, limiter,
                                          order, VG_MODE_EXACT, to_params(mp));
}

}  // extern "C"
