/* Generic SLIC / FORCE scheme -- TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * The reference implements the same scheme once per state type (Euler, MHD,
 * HPR) with identical arithmetic; this file restates it once, over a physics
 * table.  Each face flux is evaluated once per adjacent cell as the reference
 * does, and ghost cells are materialised as insertBoundaryCells[2D] builds them.
 */
#include "scheme_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* std::min / std::max as the reference's libstdc++ defines them. */
double orc_smin(double a, double b) { return (b < a) ? b : a; }
double orc_smax(double a, double b) { return (a < b) ? b : a; }

/* SlopeLimiters::computeRCoefficient (Solvers/slopelimiters.cpp:7-10). */
double orc_r_coefficient(double r, double bias)
{
    return 2.0 / ((1.0 - bias) + ((1.0 + bias) * r));
}

/* SlopeLimiters::computeSlopeLimiter and the three limiters
 * (Solvers/slopelimiters.cpp:12-74): 0 superbee, 1 van Leer, else minbee. */
double orc_limiter_value(double r, double bias, int limiter)
{
    if (limiter == 0) {
        if (r < 0.0) return 0.0;
        if (r < 0.5) return 2.0 * r;
        if (r < 1.0) return 1.0;
        return orc_smin(orc_smin(r, orc_r_coefficient(r, bias)), 2.0);
    }
    if (limiter == 1) {
        if (r < 0.0) return 0.0;
        return orc_smin((2.0 * r) / (1.0 + r), orc_r_coefficient(r, bias));
    }
    if (r < 0.0) return 0.0;
    if (r < 1.0) return r;
    return orc_smin(1.0, orc_r_coefficient(r, bias));
}

/* SlopeLimiters::computeSlopeVector on conserved vectors
 * (Solvers/slopelimiters.cpp:217-246); a magnitude below 1e-5 becomes +1e-5. */
void orc_slope_vector(int nf, const double *UL, const double *UM, const double *UR, double bias, int limiter,
                      double *d)
{
    for (int k = 0; k < nf; k++) {
        double dl = UM[k] - UL[k];
        double dr = UR[k] - UM[k];
        d[k] = ((0.5 * (1.0 + bias)) * dl) + ((0.5 * (1.0 - bias)) * dr);
        double num = dl, den = dr;
        if (fabs(num) < 1e-5) num = 1e-5;
        if (fabs(den) < 1e-5) den = 1e-5;
        d[k] *= orc_limiter_value(num / den, bias, limiter);
    }
}

/* Solvers::evolveStateByHalfX/YTimeStep + evolveStateByHalfTimeStep +
 * computeEvolutionVector (Solvers/solvers.cpp:454-466, 502-514, 550-565,
 * 601-609); MHDSolvers and HPRSolvers repeat the same arithmetic
 * (mhdsolvers.cpp:499-511, hprsolvers.cpp:493). */
void orc_evolve_half(const orc_phys *ph, int dir, const double *wl, const double *wm, const double *wr, double dx,
                     double ts, double bias, int limiter, int side, double *out)
{
    const int nf = ph->nf;
    double Ul[ORC_MAX_NF], Um[ORC_MAX_NF], Ur[ORC_MAX_NF], d[ORC_MAX_NF], xl[ORC_MAX_NF], xr[ORC_MAX_NF];
    double Fl[ORC_MAX_NF], Fr[ORC_MAX_NF], U[ORC_MAX_NF];
    ph->prim_to_cons(ph, wl, Ul);
    ph->prim_to_cons(ph, wm, Um);
    ph->prim_to_cons(ph, wr, Ur);
    orc_slope_vector(nf, Ul, Um, Ur, bias, limiter, d);
    for (int k = 0; k < nf; k++) {
        xl[k] = Um[k] - (0.5 * d[k]);
        xr[k] = Um[k] + (0.5 * d[k]);
    }
    ph->flux(ph, dir, xl, Fl);
    ph->flux(ph, dir, xr, Fr);
    double c = 0.5 * (ts / dx);
    for (int k = 0; k < nf; k++) {
        double evo = c * (Fl[k] - Fr[k]);
        U[k] = (side == 0 ? xl[k] : xr[k]) + evo;
    }
    ph->cons_to_prim(ph, U, out);
}

/* FirstOrderSolver::computeLaxFriedrichsFlux (Solvers/firstordersolver.cpp:79-85). */
void orc_lax_friedrichs(int nf, const double *Ul, const double *Ur, const double *Fl, const double *Fr, double dx,
                        double ts, double *out)
{
    double c = dx / ts;
    for (int k = 0; k < nf; k++) out[k] = 0.5 * ((Fl[k] + Fr[k]) + (c * (Ul[k] - Ur[k])));
}

/* FirstOrderSolver::computeRichtmyerFlux (Solvers/firstordersolver.cpp:165-171). */
void orc_richtmyer_state(int nf, const double *Ul, const double *Ur, const double *Fl, const double *Fr, double dx,
                         double ts, double *out)
{
    double c = ts / dx;
    for (int k = 0; k < nf; k++) out[k] = 0.5 * ((Ul[k] + Ur[k]) + (c * (Fl[k] - Fr[k])));
}

/* FirstOrderSolver::computeX/YFORCEFlux (Solvers/firstordersolver.cpp:173-180,
 * 200-207; MHD: mhdfirstordersolver.cpp:189-196). */
void orc_force_flux(const orc_phys *ph, int dir, const double *wl, const double *wr, double dx, double ts,
                    double *F)
{
    const int nf = ph->nf;
    double Ul[ORC_MAX_NF], Ur[ORC_MAX_NF], Fl[ORC_MAX_NF], Fr[ORC_MAX_NF], lf[ORC_MAX_NF], us[ORC_MAX_NF];
    double ri[ORC_MAX_NF];
    ph->prim_to_cons(ph, wl, Ul);
    ph->prim_to_cons(ph, wr, Ur);
    ph->flux(ph, dir, Ul, Fl);   /* state overloads: flux(computeConservedVariableVector()) */
    ph->flux(ph, dir, Ur, Fr);
    orc_lax_friedrichs(nf, Ul, Ur, Fl, Fr, dx, ts, lf);
    orc_richtmyer_state(nf, Ul, Ur, Fl, Fr, dx, ts, us);
    ph->flux(ph, dir, us, ri);
    for (int k = 0; k < nf; k++) F[k] = 0.5 * (lf[k] + ri[k]);
}

/* SecondOrderSolver::computeX/YSLICFlux (Solvers/secondordersolver.cpp:7-17, 43-53). */
void orc_slic_flux(const orc_phys *ph, int dir, const double *wa, const double *wb, const double *wc,
                   const double *wd, double dx, double ts, double bias, int limiter, double *F)
{
    double er[ORC_MAX_NF], el[ORC_MAX_NF];
    orc_evolve_half(ph, dir, wa, wb, wc, dx, ts, bias, limiter, 1, er);
    orc_evolve_half(ph, dir, wb, wc, wd, dx, ts, bias, limiter, 0, el);
    orc_force_flux(ph, dir, er, el, dx, ts, F);
}

/* FirstOrderSolver::computeFORCEUpdate (firstordersolver.cpp:428-433) then
 * setConservedVariableVector. */
static void force_update(const orc_phys *ph, const double *w, const double *Fm, const double *Fp, double dx,
                         double ts, double *out)
{
    const int nf = ph->nf;
    double U[ORC_MAX_NF], N[ORC_MAX_NF];
    ph->prim_to_cons(ph, w, U);
    double c = ts / dx;
    for (int k = 0; k < nf; k++) N[k] = U[k] + (c * (Fm[k] - Fp[k]));
    ph->cons_to_prim(ph, N, out);
}

/* Transmissive ghosts are mirror-ordered copies for two layers and copies for
 * one (Solvers/solvers.cpp:12-23, 94-131); both reduce to this reflection.
 * bc 1 is periodic (not in the reference). */
static int ghost_index(int k, int n, int bc)
{
    if (k >= 0 && k < n) return k;
    if (bc == 1) return k < 0 ? k + n : k - n;
    return k < 0 ? -1 - k : 2 * n - 1 - k;
}

static double *ghosted_2d(const orc_phys *ph, int ny, int nx, int g, int bc, const double *w)
{
    const int nf = orc_np(ph);
    int GY = ny + 2 * g, GX = nx + 2 * g;
    double *out = (double *)malloc(sizeof(double) * nf * (size_t)GY * GX);
    for (int i = 0; i < GY; i++) {
        for (int j = 0; j < GX; j++) {
            int ii = i - g, jj = j - g;
            int rin = ii >= 0 && ii < ny, cin = jj >= 0 && jj < nx;
            double *dst = out + nf * ((size_t)i * GX + j);
            if (!rin && !cin && bc == 0) {
                memcpy(dst, ph->default_state, sizeof(double) * nf);
            } else {
                int si = ghost_index(ii, ny, bc), sj = ghost_index(jj, nx, bc);
                memcpy(dst, w + nf * ((size_t)si * nx + sj), sizeof(double) * nf);
            }
        }
    }
    return out;
}

static double *ghosted_1d(const orc_phys *ph, int nx, int g, int bc, const double *w)
{
    const int nf = orc_np(ph);
    double *out = (double *)malloc(sizeof(double) * nf * (size_t)(nx + 2 * g));
    for (int j = 0; j < nx + 2 * g; j++)
        memcpy(out + nf * (size_t)j, w + nf * (size_t)ghost_index(j - g, nx, bc), sizeof(double) * nf);
    return out;
}

/* computeMaximumWaveSpeed[2D] over the ghosted grid: starts at 0.0 and is raised
 * by `waveSpeed > max`, so NaN is skipped (Solvers/solvers.cpp:259-339); ghosts
 * repeat interior cells, so only the 2D transmissive corners add a candidate. */
double orc_ghosted_max(const orc_phys *ph, int dim, int ny, int nx, int bc, const double *prim)
{
    const int nf = orc_np(ph);
    double m = 0.0;
    size_t cells = (size_t)(dim == 1 ? 1 : ny) * nx;
    for (size_t c = 0; c < cells; c++) {
        double s = ph->wave_speed(ph, prim + nf * c, dim);
        if (s > m) m = s;
    }
    if (dim == 2 && bc == 0) {
        double s = ph->wave_speed(ph, ph->default_state, dim);
        if (s > m) m = s;
    }
    return m;
}

void orc_sweep(const orc_phys *ph, int order, int dim, int dir, int ny, int nx, double dx, double ts, double bias,
               int limiter, int bc, const double *prim_in, double *prim_out)
{
    const int nf = orc_np(ph);   /* primitive stride */
    int g = (order == 1) ? 1 : 2;
    if (dim == 1) {
        double *wb = ghosted_1d(ph, nx, g, bc, prim_in);
        for (int j = 0; j < nx; j++) {
            double Fm[ORC_MAX_NF], Fp[ORC_MAX_NF];
            const double *c = wb + nf * (size_t)(j + g);
            if (order == 1) {
                orc_force_flux(ph, 0, c - nf, c, dx, ts, Fm);
                orc_force_flux(ph, 0, c, c + nf, dx, ts, Fp);
            } else {
                orc_slic_flux(ph, 0, c - 2 * nf, c - nf, c, c + nf, dx, ts, bias, limiter, Fm);
                orc_slic_flux(ph, 0, c - nf, c, c + nf, c + 2 * nf, dx, ts, bias, limiter, Fp);
            }
            force_update(ph, prim_in + nf * (size_t)j, Fm, Fp, dx, ts, prim_out + nf * (size_t)j);
        }
        free(wb);
        return;
    }
    int GX = nx + 2 * g;
    double *wb = ghosted_2d(ph, ny, nx, g, bc, prim_in);
    size_t s = (dir == 0) ? (size_t)nf : (size_t)nf * GX;
    for (int i = 0; i < ny; i++) {
        for (int j = 0; j < nx; j++) {
            const double *c = wb + nf * ((size_t)(i + g) * GX + (j + g));
            double Fm[ORC_MAX_NF], Fp[ORC_MAX_NF];
            if (order == 1) {
                orc_force_flux(ph, dir, c - s, c, dx, ts, Fm);
                orc_force_flux(ph, dir, c, c + s, dx, ts, Fp);
            } else {
                orc_slic_flux(ph, dir, c - 2 * s, c - s, c, c + s, dx, ts, bias, limiter, Fm);
                orc_slic_flux(ph, dir, c - s, c, c + s, c + 2 * s, dx, ts, bias, limiter, Fp);
            }
            size_t o = nf * ((size_t)i * nx + j);
            force_update(ph, prim_in + o, Fm, Fp, dx, ts, prim_out + o);
        }
    }
    free(wb);
}

/* Solvers::computeStableTimeStep (Solvers/solvers.cpp:389-395, 437-452). */
double orc_stable_time_step(double S, double dx, double cfl, double t, double T, int it)
{
    double dt = cfl * (dx / S);
    if (it <= 5) dt *= 0.2;
    if ((t + dt) > T) dt = T - t;
    return dt;
}

/* Solvers::evolveStateByFractionalX/YTimeStep (HPR: hprsolvers.cpp:606-620,
 * 673-687; MHD: mhdsolvers.cpp:612-626 and the Y twin): half-slope
 * extrapolations, flux difference scaled by frac * (ts / dx) added to the middle
 * conserved vector (computeFractionalEvolutionVector), then
 * evolveStateByFractionalTimeStep: the new primitive state. */
static void evolve_fractional(const orc_phys *ph, int dir, double frac, const double *wl, const double *wm,
                              const double *wr, double dx, double ts, double bias, int limiter, double *out)
{
    const int nf = ph->nf;
    double Ul[ORC_MAX_NF], Um[ORC_MAX_NF], Ur[ORC_MAX_NF], d[ORC_MAX_NF], xl[ORC_MAX_NF], xr[ORC_MAX_NF];
    double Fl[ORC_MAX_NF], Fr[ORC_MAX_NF], N[ORC_MAX_NF];
    ph->prim_to_cons(ph, wl, Ul);
    ph->prim_to_cons(ph, wm, Um);
    ph->prim_to_cons(ph, wr, Ur);
    orc_slope_vector(nf, Ul, Um, Ur, bias, limiter, d);
    for (int k = 0; k < nf; k++) {
        xl[k] = Um[k] - (0.5 * d[k]);
        xr[k] = Um[k] + (0.5 * d[k]);
    }
    ph->flux(ph, dir, xl, Fl);
    ph->flux(ph, dir, xr, Fr);
    double c = frac * (ts / dx);
    for (int k = 0; k < nf; k++) N[k] = Um[k] + (c * (Fl[k] - Fr[k]));
    ph->cons_to_prim(ph, N, out);
}

/* HPRForcingSolver / MHDForcingSolver::evolveConservedVariableVector[2D]
 * (hprforcingsolver.cpp:7-44, 134-192; mhdforcingsolver.cpp:7-44, 134-194) and
 * HPRForcingSolver::computeEvolvedConservedVariableVector (:332-339): RK4 on the
 * source, each stage predicting the middle state with fractional MUSCL steps
 * (X(1/2) in 1D; X(1/4) Y(1/2) X(1/4), then X(1/2) Y(1) X(1/2), in 2D) from
 * fixed neighbours. */
void orc_forcing_cell(const orc_phys *ph, int dim, const double *Ul_, const double *Um, const double *Ur_,
                      const double *Ut_, const double *Ub_, double dx, double ts, double bias, int limiter,
                      double *Unew)
{
    const int nf = ph->nf;
    double wl[ORC_MAX_NF], wr[ORC_MAX_NF], wt[ORC_MAX_NF], wb[ORC_MAX_NF], wm[ORC_MAX_NF];
    double e1[ORC_MAX_NF], e2[ORC_MAX_NF], e3[ORC_MAX_NF];
    double k1[ORC_MAX_NF], k2[ORC_MAX_NF], k3[ORC_MAX_NF], k4[ORC_MAX_NF], S[ORC_MAX_NF], Us[ORC_MAX_NF];
    double Ue[ORC_MAX_NF];
    ph->cons_to_prim(ph, Ul_, wl);
    ph->cons_to_prim(ph, Ur_, wr);
    if (dim == 2) {
        ph->cons_to_prim(ph, Ut_, wt);
        ph->cons_to_prim(ph, Ub_, wb);
    }
    static const double fx2[3] = {0.25, 0.25, 0.5}, fy2[3] = {0.5, 0.5, 1.0}, fx1[3] = {0.5, 0.5, 1.0};
    double *ks[4] = {k1, k2, k3, k4};
    ph->source(ph, Um, S);
    for (int k = 0; k < nf; k++) k1[k] = ts * S[k];
    for (int st = 0; st < 3; st++) {
        const double *kp = ks[st];
        for (int k = 0; k < nf; k++) Us[k] = Um[k] + (st < 2 ? (0.5 * kp[k]) : kp[k]);
        ph->cons_to_prim(ph, Us, wm);
        if (dim == 1) {
            evolve_fractional(ph, 0, fx1[st], wl, wm, wr, dx, ts, bias, limiter, e3);
        } else {
            evolve_fractional(ph, 0, fx2[st], wl, wm, wr, dx, ts, bias, limiter, e1);
            evolve_fractional(ph, 1, fy2[st], wt, e1, wb, dx, ts, bias, limiter, e2);
            evolve_fractional(ph, 0, fx2[st], wl, e2, wr, dx, ts, bias, limiter, e3);
        }
        ph->prim_to_cons(ph, e3, Ue);
        ph->source(ph, Ue, S);
        for (int k = 0; k < nf; k++) ks[st + 1][k] = ts * S[k];
    }
    for (int k = 0; k < nf; k++)
        Unew[k] = Um[k] + ((1.0 / 6.0) * (k1[k] + ((2.0 * k2[k]) + ((2.0 * k3[k]) + k4[k]))));
}

/* HPRForcingSolver / MHDForcingSolver::computeRungeKuttaTimeStep[2D]
 * (hprforcingsolver.cpp:341-356, 394-417; mhdforcingsolver.cpp:341-356, 394-417)
 * over a one-layer ghosted grid: every cell's conserved vector from its
 * primitive state, the RK4 update, setConservedVariableVector. */
void orc_forcing_step(const orc_phys *ph, int dim, int ny, int nx, int bc, double dx, double ts, double bias,
                      int limiter, const double *prim_in, double *prim_out)
{
    const int nf = ph->nf;
    size_t cells = (size_t)(dim == 1 ? 1 : ny) * nx;
    double *U = (double *)malloc(sizeof(double) * nf * cells);
    for (size_t c = 0; c < cells; c++) ph->prim_to_cons(ph, prim_in + nf * c, U + nf * c);
    for (int i = 0; i < (dim == 1 ? 1 : ny); i++) {
        for (int j = 0; j < nx; j++) {
            size_t c = (size_t)i * nx + j;
            const double *L = U + nf * ((size_t)i * nx + ghost_index(j - 1, nx, bc));
            const double *R = U + nf * ((size_t)i * nx + ghost_index(j + 1, nx, bc));
            const double *T = NULL, *B = NULL;
            if (dim == 2) {
                T = U + nf * ((size_t)ghost_index(i - 1, ny, bc) * nx + j);
                B = U + nf * ((size_t)ghost_index(i + 1, ny, bc) * nx + j);
            }
            double N[ORC_MAX_NF];
            orc_forcing_cell(ph, dim, L, U + nf * c, R, T, B, dx, ts, bias, limiter, N);
            ph->cons_to_prim(ph, N, prim_out + nf * c);
            if (ph->forcing_finish) ph->forcing_finish(ph, prim_in + nf * c, prim_out + nf * c);
        }
    }
    free(U);
}

/* The solve loops of SecondOrderSolver/FirstOrderSolver::solve[2D]
 * (secondordersolver.cpp:288-314, 392-424; firstordersolver.cpp:435-461,
 * 521-553), MHDSecondOrderSolver::solve[2D] (mhdsecondordersolver.cpp:318-360,
 * 501-546: c_h set from the step's maximum, sub-cycles around the SLIC step),
 * HPRSecondOrderSolver::solve[2D] (hprsecondordersolver.cpp:314-352, 502-546)
 * and the first-order MHD/HPR solvers, which neither set c_h nor sub-cycle
 * (mhdfirstordersolver.cpp:457-477, 552-578; hprfirstordersolver.cpp:455-475),
 * with an optional step cap. */
int orc_solve_sub(orc_phys *ph, int order, int dim, int ny, int nx, double dx, double cfl, double t_final,
                  int max_steps, double bias, int limiter, int bc, int subcycles, const double *prim_in,
                  double *prim_out, double *dt_log, int dt_cap)
{
    const int nf = orc_np(ph);   /* primitive stride */
    const int nsub = (order == 2 && ph->source) ? subcycles : 0;
    size_t cells = (size_t)(dim == 1 ? 1 : ny) * nx;
    double *a = (double *)malloc(sizeof(double) * nf * cells);
    double *b = (double *)malloc(sizeof(double) * nf * cells);
    memcpy(a, prim_in, sizeof(double) * nf * cells);
    double t = 0.0;
    int it = 0;
    while (t < t_final && (max_steps == 0 || it < max_steps)) {
        double S = orc_ghosted_max(ph, dim, ny, nx, bc, a);
        double dt = orc_stable_time_step(S, dx, cfl, t, t_final, it);
        if (order == 2 && ph->begin_step) ph->begin_step(ph, S);
        for (int s = 0; s < nsub; s++) {
            orc_forcing_step(ph, dim, ny, nx, bc, dx, 0.5 * (dt / subcycles), bias, limiter, a, b);
            double *tmp = a; a = b; b = tmp;
        }
        if (dim == 1) {
            orc_sweep(ph, order, 1, 0, 1, nx, dx, dt, bias, limiter, bc, a, b);
        } else {
            double half = 0.5 * dt;
            orc_sweep(ph, order, 2, 0, ny, nx, dx, half, bias, limiter, bc, a, b);
            orc_sweep(ph, order, 2, 1, ny, nx, dx, dt, bias, limiter, bc, b, a);
            orc_sweep(ph, order, 2, 0, ny, nx, dx, half, bias, limiter, bc, a, b);
        }
        { double *tmp = a; a = b; b = tmp; }
        for (int s = 0; s < nsub; s++) {
            orc_forcing_step(ph, dim, ny, nx, bc, dx, 0.5 * (dt / subcycles), bias, limiter, a, b);
            double *tmp = a; a = b; b = tmp;
        }
        if (dt_log && it < dt_cap) dt_log[it] = dt;
        t += dt;
        it += 1;
        if (ph->after_step) ph->after_step(ph, order, dim, ny, nx, it, a);
    }
    memcpy(prim_out, a, sizeof(double) * nf * cells);
    free(a);
    free(b);
    return it;
}

int orc_solve(orc_phys *ph, int order, int dim, int ny, int nx, double dx, double cfl, double t_final,
              int max_steps, double bias, int limiter, int bc, const double *prim_in, double *prim_out,
              double *dt_log, int dt_cap)
{
    return orc_solve_sub(ph, order, dim, ny, nx, dx, cfl, t_final, max_steps, bias, limiter, bc, 0, prim_in,
                         prim_out, dt_log, dt_cap);
}
