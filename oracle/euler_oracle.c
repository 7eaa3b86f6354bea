/* Euler SLIC / FORCE restatement -- TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * Plain C, one cell at a time, in the reference's operation order.  Every face
 * flux is evaluated once per adjacent cell, as the reference does
 * (Solvers/secondordersolver.cpp:90-93), and ghost cells are materialised as the
 * reference's insertBoundaryCells[2D] does (Solvers/solvers.cpp:7-32, 88-143),
 * including the default-state corners that enter the 2D wave-speed maximum.
 * Pinned bit for bit against the compiled reference by tests/test_oracle_euler.py.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define NF 6

/* std::min / std::max as the reference's libstdc++ defines them. */
static double smin(double a, double b) { return (b < a) ? b : a; }
static double smax(double a, double b) { return (a < b) ? b : a; }

/* Euler/eulerequationofstate.cpp:7-13 */
static double eos_e(double rho, double p, const orc_euler_params *mp)
{
    return (p + (mp->gamma * mp->p_inf)) / ((mp->gamma - 1) * rho);
}

/* Euler/eulerequationofstate.cpp:15-21 */
static double eos_p(double rho, double e, const orc_euler_params *mp)
{
    return (e * (mp->gamma - 1) * rho) - (mp->gamma * mp->p_inf);
}

/* Euler/eulerequationofstate.cpp:23-29 */
static double eos_c(double rho, double p, const orc_euler_params *mp)
{
    return sqrt((mp->gamma * (p + mp->p_inf)) / rho);
}

/* EulerStateVector::computeConservedVariableVector + computeTotalEnergy
 * (Euler/eulerstatevector.cpp:78-91, 192-198). */
static void prim_to_cons(const double *w, double *U, const orc_euler_params *mp)
{
    double rho = w[0], u = w[1], v = w[2], z = w[3], p = w[4], lam = w[5];
    double vsq = (u * u) + (v * v) + (z * z);
    U[0] = rho;
    U[1] = rho * u;
    U[2] = rho * v;
    U[3] = rho * z;
    U[4] = rho * ((0.5 * vsq) + eos_e(rho, p, mp) + (mp->e_form * (lam - 1.0)));
    U[5] = rho * lam;
}

/* EulerStateVector::setConservedVariableVector (Euler/eulerstatevector.cpp:47-61). */
static void cons_to_prim(const double *U, double *w, const orc_euler_params *mp)
{
    double rho = U[0];
    double u = U[1] / rho, v = U[2] / rho, z = U[3] / rho;
    double lam = U[5] / rho;
    double vsq = (u * u) + (v * v) + (z * z);
    double e = (U[4] / rho) - (0.5 * vsq) - (mp->e_form * (lam - 1.0));
    w[0] = rho;
    w[1] = u;
    w[2] = v;
    w[3] = z;
    w[4] = eos_p(rho, e, mp);
    w[5] = lam;
}

/* EulerStateVector::computeXFluxVector / computeYFluxVector on a conserved
 * vector (Euler/eulerstatevector.cpp:93-120, 127-154). */
static void flux(int dir, const double *U, double *F, const orc_euler_params *mp)
{
    double rho = U[0];
    double u = U[1] / rho, v = U[2] / rho, z = U[3] / rho;
    double lam = U[5] / rho;
    double vsq = (u * u) + (v * v) + (z * z);
    double et = U[4] / rho;
    double e = et - (0.5 * vsq) - (mp->e_form * (lam - 1.0));
    double p = eos_p(rho, e, mp);
    if (dir == 0) {
        F[0] = rho * u;
        F[1] = (rho * (u * u)) + p;
        F[2] = rho * (u * v);
        F[3] = rho * (u * z);
        F[4] = (rho * (u * et)) + (u * p);
        F[5] = rho * (u * lam);
    } else {
        F[0] = rho * v;
        F[1] = rho * (v * u);
        F[2] = (rho * (v * v)) + p;
        F[3] = rho * (v * z);
        F[4] = (rho * (v * et)) + (v * p);
        F[5] = rho * (v * lam);
    }
}

/* Flux of a primitive state: the state overloads round-trip through the
 * conserved vector first (Euler/eulerstatevector.cpp:122-125, 156-159). */
static void flux_of_prim(int dir, const double *w, double *F, const orc_euler_params *mp)
{
    double U[NF];
    prim_to_cons(w, U, mp);
    flux(dir, U, F, mp);
}

/* SlopeLimiters::computeRCoefficient (Solvers/slopelimiters.cpp:7-10). */
static double r_coefficient(double r, double bias)
{
    return 2.0 / ((1.0 - bias) + ((1.0 + bias) * r));
}

/* SlopeLimiters::computeSlopeLimiter and the three limiters
 * (Solvers/slopelimiters.cpp:12-74): 0 superbee, 1 van Leer, else minbee. */
static double limiter_value(double r, double bias, int limiter)
{
    if (limiter == 0) {
        if (r < 0.0) return 0.0;
        if (r < 0.5) return 2.0 * r;
        if (r < 1.0) return 1.0;
        return smin(smin(r, r_coefficient(r, bias)), 2.0);
    }
    if (limiter == 1) {
        if (r < 0.0) return 0.0;
        return smin((2.0 * r) / (1.0 + r), r_coefficient(r, bias));
    }
    if (r < 0.0) return 0.0;
    if (r < 1.0) return r;
    return smin(1.0, r_coefficient(r, bias));
}

/* SlopeLimiters::computeSlopeVector on conserved vectors
 * (Solvers/slopelimiters.cpp:217-246).  The small-difference clamp replaces a
 * magnitude below 1e-5 with +1e-5 (sign dropped), as the reference does. */
static void slope_vector(int nf, const double *UL, const double *UM, const double *UR, double bias,
                         int limiter, double *d)
{
    for (int k = 0; k < nf; k++) {
        double dl = UM[k] - UL[k];
        double dr = UR[k] - UM[k];
        d[k] = ((0.5 * (1.0 + bias)) * dl) + ((0.5 * (1.0 - bias)) * dr);
        double num = dl, den = dr;
        if (fabs(num) < 1e-5) num = 1e-5;
        if (fabs(den) < 1e-5) den = 1e-5;
        d[k] *= limiter_value(num / den, bias, limiter);
    }
}

/* Solvers::evolveStateByHalfX/YTimeStep + evolveStateByHalfTimeStep +
 * computeEvolutionVector (Solvers/solvers.cpp:454-466, 502-514, 550-565,
 * 601-609).  side 0 returns the evolved left extrapolation, side 1 the right. */
static void evolve_half(int dir, const double *wl, const double *wm, const double *wr, double dx,
                        double ts, double bias, int limiter, int side, const orc_euler_params *mp,
                        double *out)
{
    double Ul[NF], Um[NF], Ur[NF], d[NF], xl[NF], xr[NF], Fl[NF], Fr[NF], U[NF];
    prim_to_cons(wl, Ul, mp);
    prim_to_cons(wm, Um, mp);
    prim_to_cons(wr, Ur, mp);
    slope_vector(NF, Ul, Um, Ur, bias, limiter, d);
    for (int k = 0; k < NF; k++) {
        xl[k] = Um[k] - (0.5 * d[k]);
        xr[k] = Um[k] + (0.5 * d[k]);
    }
    flux(dir, xl, Fl, mp);
    flux(dir, xr, Fr, mp);
    double c = 0.5 * (ts / dx);
    for (int k = 0; k < NF; k++) {
        double evo = c * (Fl[k] - Fr[k]);
        U[k] = (side == 0 ? xl[k] : xr[k]) + evo;
    }
    cons_to_prim(U, out, mp);
}

/* FirstOrderSolver::computeLaxFriedrichsFlux (Solvers/firstordersolver.cpp:79-85). */
static void lax_friedrichs(const double *Ul, const double *Ur, const double *Fl, const double *Fr,
                           double dx, double ts, double *out)
{
    double c = dx / ts;
    for (int k = 0; k < NF; k++) out[k] = 0.5 * ((Fl[k] + Fr[k]) + (c * (Ul[k] - Ur[k])));
}

/* FirstOrderSolver::computeRichtmyerFlux (Solvers/firstordersolver.cpp:165-171):
 * returns the intermediate conserved state. */
static void richtmyer_state(const double *Ul, const double *Ur, const double *Fl, const double *Fr,
                            double dx, double ts, double *out)
{
    double c = ts / dx;
    for (int k = 0; k < NF; k++) out[k] = 0.5 * ((Ul[k] + Ur[k]) + (c * (Fl[k] - Fr[k])));
}

/* FirstOrderSolver::computeX/YFORCEFlux (Solvers/firstordersolver.cpp:173-180,
 * 200-207) with computeX/YLaxFriedrichsFlux (:7-17, 43-53), computeX/YRichtmyerFlux
 * (:87-98, 126-137) and computeFORCEFlux (:227-230). */
static void force_flux(int dir, const double *wl, const double *wr, double dx, double ts,
                       const orc_euler_params *mp, double *F)
{
    double Ul[NF], Ur[NF], Fl[NF], Fr[NF], lf[NF], us[NF], ri[NF];
    prim_to_cons(wl, Ul, mp);
    prim_to_cons(wr, Ur, mp);
    flux_of_prim(dir, wl, Fl, mp);
    flux_of_prim(dir, wr, Fr, mp);
    lax_friedrichs(Ul, Ur, Fl, Fr, dx, ts, lf);
    richtmyer_state(Ul, Ur, Fl, Fr, dx, ts, us);
    flux(dir, us, ri, mp);
    for (int k = 0; k < NF; k++) F[k] = 0.5 * (lf[k] + ri[k]);
}

/* SecondOrderSolver::computeX/YSLICFlux (Solvers/secondordersolver.cpp:7-17, 43-53). */
static void slic_flux(int dir, const double *wa, const double *wb, const double *wc, const double *wd,
                      double dx, double ts, double bias, int limiter, const orc_euler_params *mp,
                      double *F)
{
    double er[NF], el[NF];
    evolve_half(dir, wa, wb, wc, dx, ts, bias, limiter, 1, mp, er);
    evolve_half(dir, wb, wc, wd, dx, ts, bias, limiter, 0, mp, el);
    force_flux(dir, er, el, dx, ts, mp, F);
}

/* FirstOrderSolver::computeFORCEUpdate (Solvers/firstordersolver.cpp:428-433)
 * followed by setConservedVariableVector. */
static void force_update(const double *w, const double *Fm, const double *Fp, double dx, double ts,
                         const orc_euler_params *mp, double *out)
{
    double U[NF], N[NF];
    prim_to_cons(w, U, mp);
    double c = ts / dx;
    for (int k = 0; k < NF; k++) N[k] = U[k] + (c * (Fm[k] - Fp[k]));
    cons_to_prim(N, out, mp);
}

/* Wave speeds of Solvers::computeMaximumWaveSpeed (Solvers/solvers.cpp:259-276)
 * and computeMaximumWaveSpeed2D (:318-339). */
static double wave_speed_1d(const double *w, const orc_euler_params *mp)
{
    return fabs(w[1]) + eos_c(w[0], w[4], mp);
}

static double wave_speed_2d(const double *w, const orc_euler_params *mp)
{
    return smax(fabs(w[1]), fabs(w[2])) + eos_c(w[0], w[4], mp);
}

/* EulerStateVector default constructor (Euler/eulerstatevector.cpp:3-12). */
static const double DEFAULT_STATE[NF] = {1.0, 0.0, 0.0, 0.0, 1.0, 1.0};

/* Ghost index for boundary conditions: transmissive ghosts are mirror-ordered
 * for two layers and copies for one (Solvers/solvers.cpp:12-23, 94-131); both
 * reduce to the reflection below.  bc 1 is periodic (not in the reference). */
static int ghost_index(int k, int n, int bc)
{
    if (k >= 0 && k < n) return k;
    if (bc == 1) return k < 0 ? k + n : k - n;
    return k < 0 ? -1 - k : 2 * n - 1 - k;
}

/* Ghosted 2D copy as Solvers::insertBoundaryCells2D builds it: interior, the
 * four edge strips, and default-state corners (transmissive) or wrapped corners
 * (periodic). */
static double *ghosted_2d(int ny, int nx, int g, int bc, const double *w)
{
    int GY = ny + 2 * g, GX = nx + 2 * g;
    double *out = (double *)malloc(sizeof(double) * NF * (size_t)GY * GX);
    for (int i = 0; i < GY; i++) {
        for (int j = 0; j < GX; j++) {
            int ii = i - g, jj = j - g;
            int rin = ii >= 0 && ii < ny, cin = jj >= 0 && jj < nx;
            double *dst = out + NF * ((size_t)i * GX + j);
            if (!rin && !cin && bc == 0) {
                memcpy(dst, DEFAULT_STATE, sizeof(DEFAULT_STATE));
            } else {
                int si = ghost_index(ii, ny, bc), sj = ghost_index(jj, nx, bc);
                memcpy(dst, w + NF * ((size_t)si * nx + sj), sizeof(double) * NF);
            }
        }
    }
    return out;
}

static double *ghosted_1d(int nx, int g, int bc, const double *w)
{
    double *out = (double *)malloc(sizeof(double) * NF * (size_t)(nx + 2 * g));
    for (int j = 0; j < nx + 2 * g; j++)
        memcpy(out + NF * (size_t)j, w + NF * (size_t)ghost_index(j - g, nx, bc), sizeof(double) * NF);
    return out;
}

double orc_euler_max_wave_speed(int dim, int ny, int nx, const orc_euler_params *mp, const double *prim)
{
    /* maximumWaveSpeed starts at 0.0 and is raised by `waveSpeed > max`, so NaN
     * wave speeds are skipped (Solvers/solvers.cpp:261-273). */
    double m = 0.0;
    if (dim == 1) {
        for (int j = 0; j < nx; j++) {
            double s = wave_speed_1d(prim + NF * (size_t)j, mp);
            if (s > m) m = s;
        }
    } else {
        for (size_t c = 0; c < (size_t)ny * nx; c++) {
            double s = wave_speed_2d(prim + NF * c, mp);
            if (s > m) m = s;
        }
    }
    return m;
}

/* The ghosted grid's maximum: ghosts repeat interior cells, so only the 2D
 * transmissive corners (default states) add a candidate. */
static double ghosted_max_wave_speed(int dim, int ny, int nx, int bc, const orc_euler_params *mp,
                                     const double *prim)
{
    double m = orc_euler_max_wave_speed(dim, ny, nx, mp, prim);
    if (dim == 2 && bc == 0) {
        double s = wave_speed_2d(DEFAULT_STATE, mp);
        if (s > m) m = s;
    }
    return m;
}

void orc_euler_sweep(int order, int dim, int dir, int ny, int nx, double dx, double ts, double bias,
                     int limiter, int bc, const orc_euler_params *mp, const double *prim_in,
                     double *prim_out)
{
    int g = (order == 1) ? 1 : 2;
    if (dim == 1) {
        double *wb = ghosted_1d(nx, g, bc, prim_in);
        for (int j = 0; j < nx; j++) {
            double Fm[NF], Fp[NF];
            const double *c = wb + NF * (size_t)(j + g);
            if (order == 1) {
                /* FirstOrderSolver::computeFORCETimeStep (firstordersolver.cpp:232-246) */
                force_flux(0, c - NF, c, dx, ts, mp, Fm);
                force_flux(0, c, c + NF, dx, ts, mp, Fp);
            } else {
                /* SecondOrderSolver::computeSLICTimeStep (secondordersolver.cpp:80-98) */
                slic_flux(0, c - 2 * NF, c - NF, c, c + NF, dx, ts, bias, limiter, mp, Fm);
                slic_flux(0, c - NF, c, c + NF, c + 2 * NF, dx, ts, bias, limiter, mp, Fp);
            }
            force_update(prim_in + NF * (size_t)j, Fm, Fp, dx, ts, mp, prim_out + NF * (size_t)j);
        }
        free(wb);
        return;
    }
    int GX = nx + 2 * g;
    double *wb = ghosted_2d(ny, nx, g, bc, prim_in);
    /* X sweeps read row i+g, columns j..j+2g; Y sweeps column j+g, rows i..i+2g
     * (secondordersolver.cpp:140-162, 214-236; firstordersolver.cpp:286-378). */
    size_t s = (dir == 0) ? (size_t)NF : (size_t)NF * GX;
    for (int i = 0; i < ny; i++) {
        for (int j = 0; j < nx; j++) {
            const double *c = wb + NF * ((size_t)(i + g) * GX + (j + g));
            double Fm[NF], Fp[NF];
            if (order == 1) {
                force_flux(dir, c - s, c, dx, ts, mp, Fm);
                force_flux(dir, c, c + s, dx, ts, mp, Fp);
            } else {
                slic_flux(dir, c - 2 * s, c - s, c, c + s, dx, ts, bias, limiter, mp, Fm);
                slic_flux(dir, c - s, c, c + s, c + 2 * s, dx, ts, bias, limiter, mp, Fp);
            }
            size_t o = NF * ((size_t)i * nx + j);
            force_update(prim_in + o, Fm, Fp, dx, ts, mp, prim_out + o);
        }
    }
    free(wb);
}

/* Solvers::computeStableTimeStep(timeStep, t, T, it) (Solvers/solvers.cpp:437-452)
 * applied to CFL*(dx/S) (:389-395, 413-419). */
static double stable_time_step(double S, double dx, double cfl, double t, double T, int it)
{
    double dt = cfl * (dx / S);
    if (it <= 5) dt *= 0.2;
    if ((t + dt) > T) dt = T - t;
    return dt;
}

/* SecondOrderSolver::solve/solve2D (secondordersolver.cpp:288-314, 392-424) and
 * FirstOrderSolver::solve/solve2D (firstordersolver.cpp:435-461, 521-553), with
 * an optional step cap. */
int orc_euler_solve(int order, int dim, int ny, int nx, double dx, double cfl, double t_final,
                    int max_steps, double bias, int limiter, int bc, const orc_euler_params *mp,
                    const double *prim_in, double *prim_out, double *dt_log, int dt_cap)
{
    size_t cells = (size_t)(dim == 1 ? 1 : ny) * nx;
    double *a = (double *)malloc(sizeof(double) * NF * cells);
    double *b = (double *)malloc(sizeof(double) * NF * cells);
    memcpy(a, prim_in, sizeof(double) * NF * cells);
    double t = 0.0;
    int it = 0;
    while (t < t_final && (max_steps == 0 || it < max_steps)) {
        double S = ghosted_max_wave_speed(dim, ny, nx, bc, mp, a);
        double dt = stable_time_step(S, dx, cfl, t, t_final, it);
        if (dim == 1) {
            orc_euler_sweep(order, 1, 0, 1, nx, dx, dt, bias, limiter, bc, mp, a, b);
            double *tmp = a; a = b; b = tmp;
        } else {
            double half = 0.5 * dt;
            orc_euler_sweep(order, 2, 0, ny, nx, dx, half, bias, limiter, bc, mp, a, b);
            orc_euler_sweep(order, 2, 1, ny, nx, dx, dt, bias, limiter, bc, mp, b, a);
            orc_euler_sweep(order, 2, 0, ny, nx, dx, half, bias, limiter, bc, mp, a, b);
            double *tmp = a; a = b; b = tmp;
        }
        if (dt_log && it < dt_cap) dt_log[it] = dt;
        t += dt;
        it += 1;
    }
    memcpy(prim_out, a, sizeof(double) * NF * cells);
    free(a);
    free(b);
    return it;
}

void orc_euler_units(int n, double bias, int limiter, double dx, double ts, const orc_euler_params *mp,
                     const double *states, double *out)
{
    for (int r = 0; r < n; r++) {
        const double *a = states + 30 * (size_t)r, *b = a + 6, *c = a + 12, *d = a + 18;
        double *o = out + 92 * (size_t)r;
        double Ua[NF], Ub[NF], Uc[NF];
        prim_to_cons(a, Ua, mp);
        prim_to_cons(b, Ub, mp);
        prim_to_cons(c, Uc, mp);
        slope_vector(NF, Ua, Ub, Uc, bias, limiter, o); o += 6;
        flux(0, Ub, o, mp); o += 6;
        flux(1, Ub, o, mp); o += 6;
        evolve_half(0, a, b, c, dx, ts, bias, limiter, 0, mp, o); o += 6;
        evolve_half(0, a, b, c, dx, ts, bias, limiter, 1, mp, o); o += 6;
        evolve_half(1, a, b, c, dx, ts, bias, limiter, 0, mp, o); o += 6;
        evolve_half(1, a, b, c, dx, ts, bias, limiter, 1, mp, o); o += 6;
        force_flux(0, b, c, dx, ts, mp, o); o += 6;
        force_flux(1, b, c, dx, ts, mp, o); o += 6;
        slic_flux(0, a, b, c, d, dx, ts, bias, limiter, mp, o); o += 6;
        slic_flux(1, a, b, c, d, dx, ts, bias, limiter, mp, o); o += 6;
        {
            double s1 = wave_speed_1d(b, mp), s2 = wave_speed_2d(b, mp);
            *o++ = (s1 > 0.0) ? s1 : 0.0;
            *o++ = (s2 > 0.0) ? s2 : 0.0;
        }
        memcpy(o, Ub, sizeof(Ub)); o += 6;
        cons_to_prim(Ub, o, mp); o += 6;
        {
            double Fb[NF], Fc[NF];
            flux_of_prim(0, b, Fb, mp);
            flux_of_prim(0, c, Fc, mp);
            lax_friedrichs(Ub, Uc, Fb, Fc, dx, ts, o); o += 6;
            double us[NF];
            richtmyer_state(Ub, Uc, Fb, Fc, dx, ts, us);
            flux(0, us, o, mp); o += 6;
        }
    }
}

void orc_limiter_units(int n, double bias, const double *r, double *out)
{
    for (int k = 0; k < n; k++) {
        out[5 * (size_t)k + 0] = limiter_value(r[k], bias, 0);
        out[5 * (size_t)k + 1] = limiter_value(r[k], bias, 1);
        out[5 * (size_t)k + 2] = limiter_value(r[k], bias, 2);
        out[5 * (size_t)k + 3] = r_coefficient(r[k], bias);
        out[5 * (size_t)k + 4] = limiter_value(r[k], bias, 7);
    }
}
