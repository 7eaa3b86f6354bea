// HPR C ABI: 1D/2D whole-solve drop-ins and the device-resident state, with the
// RK4 relaxation sub-cycles run as a source stage around the sweeps.  Each
// equation of state has its own kernels (HPRPhysT<EOS>); a call picks them from
// the material parameters.
#include <cmath>
#include <cstring>
#include <string>
#include <type_traits>

#include "vg_hpr_host.h"

// The device state: one grid per equation of state, of which the one used last
// holds the storage (it moves to another when a call names another EOS).
struct vg_hpr_state {
    vg::HPRGridT<0> g0;
    vg::HPRGridT<1> g1;
    vg::HPRGridT<2> g2;
    int eos = 0;
};

namespace {

vg::HPRParams to_params(const vg_hpr_params *p) { return vg::hpr_to_params(p); }

int check_params(const vg_hpr_params *mp, int subcycles) { return vg::hpr_check_params(mp, subcycles); }

// f(grid) on the grid of equation of state `eos` (a run-time value) .
template <class F>
int with_eos(int eos, F &&f)
{
    if (eos == VG_HPR_EOS_SHOCK_MIE_GRUNEISEN) return f(std::integral_constant<int, 1>{});
    if (eos == VG_HPR_EOS_GODUNOV_ROMENSKI) return f(std::integral_constant<int, 2>{});
    return f(std::integral_constant<int, 0>{});
}

template <int E>
vg::HPRGridT<E> &grid_of(vg_hpr_state *st)
{
    if constexpr (E == 0) return st->g0;
    else if constexpr (E == 1) return st->g1;
    else return st->g2;
}

// Moves the state's storage to the grid of equation of state E.
template <int E>
vg::HPRGridT<E> &select(vg_hpr_state *st)
{
    if (st->eos != E) {
        with_eos(st->eos, [&](auto k) {
            grid_of<E>(st).adopt(grid_of<decltype(k)::value>(st));
            return 0;
        });
        st->eos = E;
    }
    return grid_of<E>(st);
}

// f(grid) on whichever grid holds the storage.
template <class F>
auto with_active(vg_hpr_state *st, F &&f)
{
    return with_eos(st->eos, [&](auto k) { return f(grid_of<decltype(k)::value>(st)); });
}

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

template <class GS>
int run_grid(GS &g, int order, double dx, double cfl, double t_final, int max_steps, double bias, int limiter, int bc,
             int subcycles, const vg_hpr_params *mp, int mode, double *dt_log, int dt_log_cap, vg_run_stats *stats)
{
    g.subcycles = subcycles;
    // HPR has no separate FAST arithmetic: both modes run the exact kernels.
    (void)mode;
    return g.run(order, VG_MODE_EXACT, dx, cfl, t_final, max_steps, bias, limiter, bc, to_params(mp), dt_log,
                 dt_log_cap, stats);
}

template <int E>
int solve_eos(vg_ctx *ctx, int dim, const double *prim_in, double *prim_out, int ny, int nx, double dx, double cfl,
              double t_final, int max_steps, double bias, int limiter, int order, int bc, int subcycles,
              const vg_hpr_params *mp, int mode, double *dt_log, int dt_log_cap, vg_run_stats *stats)
{
    vg::HPRGridT<E> g;
    int rc = g.create(ctx, dim, ny, nx);
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
                      " non-finite or inadmissible cells (rho <= 0, or p + p_inf <= 0 for the stiffened gas), first at flat index " +
                      std::to_string(local.first_nonfinite));
        return VG_ERR_NONFINITE;
    }
    return rc;
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
    return with_eos(mp->eos, [&](auto k) -> int {
        return solve_eos<decltype(k)::value>(ctx, dim, prim_in, prim_out, ny, nx, dx, cfl, t_final, max_steps, bias,
                                             limiter, order, bc, subcycles, mp, mode, dt_log, dt_log_cap, stats);
    });
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
    rc = s->g0.create(ctx, dim, ny, nx);
    if (rc) {
        s->g0.destroy();
        delete s;
        return rc;
    }
    *st = s;
    return VG_OK;
}

void vg_hpr_state_destroy(vg_hpr_state *st)
{
    if (!st) return;
    with_active(st, [](auto &g) {
        vg::bind(g.ctx);
        hipStreamSynchronize(g.ctx->stream);
        g.destroy();
        return 0;
    });
    delete st;
}

int vg_hpr_state_load(vg_hpr_state *st, const double *prim, int on_device)
{
    vg::clear_error();
    VG_CHECK(st != nullptr && prim != nullptr, "null argument");
    return with_active(st, [&](auto &g) -> int {
        int rc = vg::bind(g.ctx);
        return rc ? rc : g.load(prim, on_device);
    });
}

int vg_hpr_state_store(vg_hpr_state *st, double *prim, int on_device)
{
    vg::clear_error();
    VG_CHECK(st != nullptr && prim != nullptr, "null argument");
    return with_active(st, [&](auto &g) -> int {
        int rc = vg::bind(g.ctx);
        return rc ? rc : g.store(prim, on_device);
    });
}

int vg_hpr_state_planes(vg_hpr_state *st, double **planes, int64_t *plane)
{
    vg::clear_error();
    VG_CHECK(st != nullptr && planes != nullptr && plane != nullptr, "null argument");
    return with_active(st, [&](auto &g) -> int {
        *planes = g.buf[g.cur];
        *plane = g.plane;
        return VG_OK;
    });
}

int vg_hpr_state_run(vg_hpr_state *st, int order, double dx, double cfl, double t_final, int max_steps, double bias,
                     int limiter, int bc, int subcycles, const vg_hpr_params *mp, int mode, double *dt_log,
                     int dt_log_cap, vg_run_stats *stats)
{
    vg::clear_error();
    VG_CHECK(st != nullptr, "null state");
    VG_CHECK(mp != nullptr, "null material parameters");
    int rc = with_active(st, [&](auto &g) {
        return check_run_args(g.dim, g.ny, g.nx, dx, cfl, t_final, max_steps, order, bc, subcycles, mp, mode);
    });
    if (rc) return rc;
    return with_eos(mp->eos, [&](auto k) -> int {
        auto &g = select<decltype(k)::value>(st);
        int r = vg::bind(g.ctx);
        return r ? r : run_grid(g, order, dx, cfl, t_final, max_steps, bias, limiter, bc, subcycles, mp, mode, dt_log,
                                dt_log_cap, stats);
    });
}

int vg_hpr_state_max_wave_speed(vg_hpr_state *st, int bc, const vg_hpr_params *mp, double *smax)
{
    vg::clear_error();
    VG_CHECK(st != nullptr && smax != nullptr, "null argument");
    int rc = check_params(mp, 0);
    if (rc) return rc;
    return with_eos(mp->eos, [&](auto k) -> int {
        auto &g = select<decltype(k)::value>(st);
        int r = vg::bind(g.ctx);
        if (r) return r;
        r = g.refresh_max(bc, to_params(mp), g.fused_count(2));
        if (r) return r;
        vg::Clock h;
        VG_HIP(hipMemcpyAsync(&h, g.clk, sizeof(h), hipMemcpyDeviceToHost, g.ctx->stream));
        VG_HIP(hipStreamSynchronize(g.ctx->stream));
        double s;
        std::memcpy(&s, &h.smax, sizeof(s));
        if (h.corner > s) s = h.corner;
        *smax = s;
        return VG_OK;
    });
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
    const int dim = with_active(st, [](auto &g) { return g.dim; });
    VG_CHECK(dir == 0 || (dir == 1 && dim == 2), "dir must be 0 (x) or 1 (y; 2D only)");
    VG_CHECK(dx > 0.0 && std::isfinite(dx) && std::isfinite(dt), "cell spacing must be positive, dt finite");
    return with_eos(mp->eos, [&](auto k) -> int {
        auto &g = select<decltype(k)::value>(st);
        int r = vg::bind(g.ctx);
        return r ? r : g.sweep_once(order, VG_MODE_EXACT, dir, dt, bc, dx, bias, limiter, to_params(mp));
    });
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
    return with_eos(mp->eos, [&](auto k) -> int {
        return vg::ghosted_time_step<vg::HPRPhysT<decltype(k)::value>>(ctx, 1, cells, cells_with_boundary, 1, n, 0, dx,
                                                                       dt, bias, limiter, order, VG_MODE_EXACT,
                                                                       to_params(mp));
    });
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
    return with_eos(mp->eos, [&](auto k) -> int {
        return vg::ghosted_time_step<vg::HPRPhysT<decltype(k)::value>>(ctx, 2, cells, cells_with_boundary, ny, nx, dir,
                                                                       dx, dt, bias, limiter, order, VG_MODE_EXACT,
                                                                       to_params(mp));
    });
}

}  // extern "C"
