// MHD C ABI: 1D and 2D solve drop-ins (with the RK4 GLM sub-cycles of
// MHDForcingSolver run as a source stage around the sweeps) and the batched
// persistent solve.
#include <cmath>
#include <vector>

#include "vg_batch_host.h"
#include "vg_mhd_host.h"

namespace {

using vg::MHDPhys;

MHDPhys::Params to_params(const vg_mhd_params *p) { return vg::mhd_params(p); }

constexpr int kBatchBS = 128;


// GridState path: 2D, or 1D with sub-cycles.
int grid_solve(vg_ctx *ctx, int dim, const double *prim_in, double *prim_out, int ny, int nx, double dx, double cfl,
               double t_final, int max_steps, double bias, int limiter, int order, int bc, int subcycles,
               const vg_mhd_params *mp, int mode, double *dt_log, int dt_log_cap, vg_run_stats *stats)
{
    VG_CHECK(ctx != nullptr && prim_in != nullptr && prim_out != nullptr, "null argument");
    VG_CHECK(mp != nullptr, "null material parameters");
    VG_CHECK(order == 1 || order == 2, "order must be 1 (FORCE) or 2 (SLIC)");
    VG_CHECK(mode == VG_MODE_EXACT || mode == VG_MODE_FAST, "unknown mode");
    VG_CHECK(bc == VG_BC_TRANSMISSIVE || bc == VG_BC_PERIODIC, "unknown boundary condition");
    VG_CHECK(subcycles >= 0, "subcycles must be >= 0");
    VG_CHECK(dx > 0.0 && cfl > 0.0, "cell spacing and CFL must be positive");
    VG_CHECK(max_steps >= 0 && (max_steps > 0 || std::isfinite(t_final)), "an infinite final time needs a step cap");
    int rc = vg::bind(ctx);
    if (rc) return rc;
    vg::GridState<MHDPhys> g;
    rc = g.create(ctx, dim, ny, nx);
    g.subcycles = subcycles;
    if (!rc) rc = g.load(prim_in, 0);
    vg_run_stats local{};
    if (!rc) rc = g.run(order, mode, dx, cfl, t_final, max_steps, bias, limiter, bc, to_params(mp), dt_log, dt_log_cap, &local);
    if (!rc) rc = g.store(prim_out, 0);
    if (!rc) rc = g.scan_nonfinite(to_params(mp), &local.nonfinite_cells, &local.first_nonfinite);
    g.destroy();
    if (stats) *stats = local;
    if (!rc && local.nonfinite_cells > 0) {
        vg::set_error("final state holds non-finite or inadmissible cells, first at flat index " +
                      std::to_string(local.first_nonfinite));
        return VG_ERR_NONFINITE;
    }
    return rc;
}

}  // namespace

extern "C" {

int vg_mhd_solve_1d(vg_ctx *ctx, const double *prim_in, double *prim_out, int n, double dx, double cfl,
                    double t_final, int max_steps, double bias, int limiter, int order, int bc, int subcycles,
                    const vg_mhd_params *mp, int mode, double *dt_log, int dt_log_cap, vg_run_stats *stats)
{
    vg::clear_error();
    VG_CHECK(ctx != nullptr, "null context");
    VG_CHECK(mp != nullptr, "null material parameters");
    VG_CHECK(subcycles >= 0, "subcycles must be >= 0");
    if (order == 2 && subcycles > 0) {
        // the source stage runs on the grid path; without it the one-instance
        // persistent batch launch is the faster route
        VG_CHECK(n >= 2, "grid needs at least 2 cells");
        return grid_solve(ctx, 1, prim_in, prim_out, 1, n, dx, cfl, t_final, max_steps, bias, limiter, order, bc,
                          subcycles, mp, mode, dt_log, dt_log_cap, stats);
    }
    int rc = vg::bind(ctx);
    if (rc) return rc;
    const int64_t offsets[2] = {0, n};
    std::vector<MHDPhys::Params> params(1, to_params(mp));
    int32_t steps = 0;
    double tend = 0.0;
    rc = vg::run_batch<MHDPhys, kBatchBS>(ctx, order, mode, 1, offsets, prim_in, prim_out, &dx, params, &t_final, cfl,
                                          max_steps, bias, limiter, bc, dt_log, dt_log_cap, &steps, &tend, stats);
    if (stats) {
        stats->steps = steps;
        stats->total_steps = steps;
        stats->t = tend;
    }
    return rc;
}

int vg_mhd_solve_2d(vg_ctx *ctx, const double *prim_in, double *prim_out, int ny, int nx, double dx, double cfl,
                    double t_final, int max_steps, double bias, int limiter, int order, int bc, int subcycles,
                    const vg_mhd_params *mp, int mode, double *dt_log, int dt_log_cap, vg_run_stats *stats)
{
    vg::clear_error();
    VG_CHECK(nx >= 2 && ny >= 2, "grid needs at least 2 cells per direction");
    return grid_solve(ctx, 2, prim_in, prim_out, ny, nx, dx, cfl, t_final, max_steps, bias, limiter, order, bc,
                      subcycles, mp, mode, dt_log, dt_log_cap, stats);
}

int vg_mhd_solve_batch(vg_ctx *ctx, int n_inst, const int64_t *offsets, const double *prim_in, double *prim_out,
                       const double *dx, const vg_mhd_params *mp, const double *t_final, double cfl, int max_steps,
                       double bias, int limiter, int order, int mode, int32_t *steps, double *t_out,
                       vg_run_stats *stats)
{
    vg::clear_error();
    VG_CHECK(ctx != nullptr && mp != nullptr && n_inst >= 1, "null argument");
    int rc = vg::bind(ctx);
    if (rc) return rc;
    std::vector<MHDPhys::Params> params(n_inst);
    for (int q = 0; q < n_inst; q++) params[q] = to_params(&mp[q]);
    return vg::run_batch<MHDPhys, kBatchBS>(ctx, order, mode, n_inst, offsets, prim_in, prim_out, dx, params, t_final,
                                            cfl, max_steps, bias, limiter, VG_BC_TRANSMISSIVE, nullptr, 0, steps,
                                            t_out, stats);
}

int vg_mhd_time_step_1d(vg_ctx *ctx, double *cells, const double *cells_with_boundary, int n, double dx, double dt,
                        double bias, int limiter, int order, const vg_mhd_params *mp, int mode)
{
    vg::clear_error();
    VG_CHECK(ctx != nullptr, "null context");
    int rc = vg::check_time_step_args(1, 1, n, 0, dx, dt, order, mode, cells, cells_with_boundary, mp);
    if (rc) return rc;
    rc = vg::bind(ctx);
    if (rc) return rc;
    return vg::ghosted_time_step<MHDPhys>(ctx, 1, cells, cells_with_boundary, 1, n, 0, dx, dt, bias, limiter, order,
                                          mode, to_params(mp));
}

int vg_mhd_time_step_2d(vg_ctx *ctx, double *cells, const double *cells_with_boundary, int ny, int nx, int dir,
                        double dx, double dt, double bias, int limiter, int order, const vg_mhd_params *mp, int mode)
{
    vg::clear_error();
    VG_CHECK(ctx != nullptr, "null context");
    int rc = vg::check_time_step_args(2, ny, nx, dir, dx, dt, order, mode, cells, cells_with_boundary, mp);
    if (rc) return rc;
    rc = vg::bind(ctx);
    if (rc) return rc;
    return vg::ghosted_time_step<MHDPhys>(ctx, 2, cells, cells_with_boundary, ny, nx, dir, dx, dt, bias, limiter,
                                          order, mode, to_params(mp));
}

}  // extern "C"
