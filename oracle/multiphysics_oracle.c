/* Two-material diffuse-interface Euler models of the oracle -- TEST
 * INFRASTRUCTURE ONLY.
 *
 * EulerMultiphysicsStateVector (Euler/Multiphysics/eulermultiphysicsstatevector.cpp:
 * 8 primitives alpha1, u, v, w, rho1, p1, rho2, p2; 9 conserved variables) and
 * EulerReducedStateVector (eulerreducedstatevector.cpp: 7 primitives alpha1, u,
 * v, w, rho1, rho2, p; 8 conserved variables), restated in the reference's
 * operation order and plugged into the generic scheme (scheme_oracle.c: the
 * SLIC and FORCE overloads of Solvers/secondordersolver.cpp:19-97, 100-160,
 * 164-312 and Solvers/firstordersolver.cpp:19-247 are the Euler scheme with
 * these state vectors), plus the volume-fraction reinitialisation of
 * Solvers/multiphysicssolvers.cpp:7-313 and the solve loops of
 * secondordersolver.cpp:316-390, 426-517 and firstordersolver.cpp:463-640.
 * Pinned bit for bit against the compiled reference by
 * tests/test_oracle_multiphysics.py.
 */
#include "oracle.h"
#include "scheme_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    orc_euler_params m1, m2;
    int reinit_every;
} mp_ctx;

/* Euler/eulerequationofstate.cpp:7-29 */
static double eos_e(double rho, double p, const orc_euler_params *mp)
{
    return (p + (mp->gamma * mp->p_inf)) / ((mp->gamma - 1) * rho);
}
static double eos_p(double rho, double e, const orc_euler_params *mp)
{
    return (e * (mp->gamma - 1) * rho) - (mp->gamma * mp->p_inf);
}
static double eos_c(double rho, double p, const orc_euler_params *mp)
{
    return sqrt((mp->gamma * (p + mp->p_inf)) / rho);
}

/* The clamp of the volume fraction to [0.001, 0.999] in every conversion
 * (eulermultiphysicsstatevector.cpp:47-58, 117-124; the reduced twin). */
static double clamp_alpha(double a)
{
    if (a < 0.001) return 0.001;
    if (a > 0.999) return 0.999;
    return a;
}

/* ---- EulerMultiphysicsStateVector ------------------------------------------ */

/* computeConservedVariableVector (eulermultiphysicsstatevector.cpp:113-140) with
 * computeMaterialkTotalEnergy (:278-283, :300-305). */
static void mp_prim_to_cons(const orc_phys *ph, const double *w, double *U)
{
    const mp_ctx *c = (const mp_ctx *)ph->params;
    double a = clamp_alpha(w[0]);
    double a2 = 1.0 - a;
    double u = w[1], v = w[2], z = w[3], r1 = w[4], p1 = w[5], r2 = w[6], p2 = w[7];
    double rho = (a * r1) + (a2 * r2);
    double vsq = (u * u) + (v * v) + (z * z);
    U[0] = rho;
    U[1] = rho * a;
    U[2] = rho * u;
    U[3] = rho * v;
    U[4] = rho * z;
    U[5] = a * r1;
    U[6] = a * (r1 * ((0.5 * vsq) + eos_e(r1, p1, &c->m1)));
    U[7] = a2 * r2;
    U[8] = a2 * (r2 * ((0.5 * vsq) + eos_e(r2, p2, &c->m2)));
}

/* setConservedVariableVector (eulermultiphysicsstatevector.cpp:43-75).  Its
 * velocity square adds w + w, not w * w (:67); kept. */
static void mp_cons_to_prim(const orc_phys *ph, const double *U, double *w)
{
    const mp_ctx *c = (const mp_ctx *)ph->params;
    double rho = U[0];
    double a = clamp_alpha(U[1] / rho);
    double a2 = 1.0 - a;
    double u = U[2] / rho, v = U[3] / rho, z = U[4] / rho;
    double r1 = U[5] / a;
    double vsq = (u * u) + (v * v) + (z + z);
    double e1 = (U[6] / (r1 * a)) - (0.5 * vsq);
    double r2 = U[7] / a2;
    double e2 = (U[8] / (r2 * a2)) - (0.5 * vsq);
    w[0] = a;
    w[1] = u;
    w[2] = v;
    w[3] = z;
    w[4] = r1;
    w[5] = eos_p(r1, e1, &c->m1);
    w[6] = r2;
    w[7] = eos_p(r2, e2, &c->m2);
}

/* computeX/YFluxVector (eulermultiphysicsstatevector.cpp:142-257). */
static void mp_flux(const orc_phys *ph, int dir, const double *U, double *F)
{
    const mp_ctx *c = (const mp_ctx *)ph->params;
    double rho = U[0];
    double a = clamp_alpha(U[1] / rho);
    double a2 = 1.0 - a;
    double u = U[2] / rho, v = U[3] / rho, z = U[4] / rho;
    double r1 = U[5] / a;
    double vsq = (u * u) + (v * v) + (z * z);
    double E1 = U[6] / (a * r1);
    double p1 = eos_p(r1, E1 - (0.5 * vsq), &c->m1);
    double r2 = U[7] / a2;
    double E2 = U[8] / (a2 * r2);
    double p2 = eos_p(r2, E2 - (0.5 * vsq), &c->m2);
    double q = (dir == 0) ? u : v;
    F[0] = rho * q;
    F[1] = rho * (q * a);
    F[2] = rho * (q * u);
    F[3] = rho * (q * v);
    F[4] = rho * (q * z);
    F[2 + dir] = (rho * (q * q)) + ((a * p1) + (a2 * p2));
    F[5] = a * (r1 * q);
    F[6] = a * ((r1 * (q * E1)) + (q * p1));
    F[7] = a2 * (r2 * q);
    F[8] = a2 * ((r2 * (q * E2)) + (q * p2));
}

/* computeMaximumWaveSpeed[2D] per cell (Solvers/solvers.cpp:278-296, 341-363):
 * |u| (or max(|u|, |v|)) + max(c1, c2). */
static double mp_wave_speed(const orc_phys *ph, const double *w, int dim)
{
    const mp_ctx *c = (const mp_ctx *)ph->params;
    double s = (dim == 1) ? fabs(w[1]) : orc_smax(fabs(w[1]), fabs(w[2]));
    return s + orc_smax(eos_c(w[4], w[5], &c->m1), eos_c(w[6], w[7], &c->m2));
}

/* EulerMultiphysicsStateVector() (eulermultiphysicsstatevector.cpp:3-15). */
static const double MP_DEFAULT[8] = {0.999, 0.0, 0.0, 0.0, 1.0, 1.0, 1.0, 1.0};

/* ---- EulerReducedStateVector ------------------------------------------------ */

/* computeConservedVariableVector (eulerreducedstatevector.cpp:94-126): the
 * material energies from the total density and the interface pressure. */
static void rd_prim_to_cons(const orc_phys *ph, const double *w, double *U)
{
    const mp_ctx *c = (const mp_ctx *)ph->params;
    double a = clamp_alpha(w[0]);
    double a2 = 1.0 - a;
    double u = w[1], v = w[2], z = w[3], r1 = w[4], r2 = w[5], p = w[6];
    double rho = (a * r1) + (a2 * r2);
    U[0] = rho;
    U[1] = rho * a;
    U[2] = rho * u;
    U[3] = rho * v;
    U[4] = rho * z;
    U[5] = a * r1;
    U[6] = a2 * r2;
    double vsq = (u * u) + (v * v) + (z * z);
    double E1 = eos_e(rho, p, &c->m1) + (0.5 * vsq);
    double E2 = eos_e(rho, p, &c->m2) + (0.5 * vsq);
    U[7] = rho * ((a * E1) + (a2 * E2));
}

/* The reduced model's mixture pressure from (rho, e):
 * alpha1 p1(rho, e) + alpha2 p2(rho, e). */
static double rd_pressure(const mp_ctx *c, double a, double a2, double rho, double e)
{
    return (a * eos_p(rho, e, &c->m1)) + (a2 * eos_p(rho, e, &c->m2));
}

/* setConservedVariableVector (eulerreducedstatevector.cpp:43-75). */
static void rd_cons_to_prim(const orc_phys *ph, const double *U, double *w)
{
    const mp_ctx *c = (const mp_ctx *)ph->params;
    double rho = U[0];
    double a = clamp_alpha(U[1] / rho);
    double a2 = 1.0 - a;
    double u = U[2] / rho, v = U[3] / rho, z = U[4] / rho;
    double vsq = (u * u) + (v * v) + (z * z);
    double e = (U[7] / rho) - (0.5 * vsq);
    w[0] = a;
    w[1] = u;
    w[2] = v;
    w[3] = z;
    w[4] = U[5] / a;
    w[5] = U[6] / a2;
    w[6] = rd_pressure(c, a, a2, rho, e);
}

/* computeX/YFluxVector (eulerreducedstatevector.cpp:128-231). */
static void rd_flux(const orc_phys *ph, int dir, const double *U, double *F)
{
    const mp_ctx *c = (const mp_ctx *)ph->params;
    double rho = U[0];
    double a = clamp_alpha(U[1] / rho);
    double a2 = 1.0 - a;
    double u = U[2] / rho, v = U[3] / rho, z = U[4] / rho;
    double r1 = U[5] / a;
    double r2 = U[6] / a2;
    double E = U[7] / rho;
    double vsq = (u * u) + (v * v) + (z * z);
    double p = rd_pressure(c, a, a2, rho, E - (0.5 * vsq));
    double q = (dir == 0) ? u : v;
    F[0] = rho * q;
    F[1] = rho * (q * a);
    F[2] = rho * (q * u);
    F[3] = rho * (q * v);
    F[4] = rho * (q * z);
    F[2 + dir] = (rho * (q * q)) + p;
    F[5] = a * (r1 * q);
    F[6] = a2 * (r2 * q);
    F[7] = (rho * (q * E)) + (q * p);
}

/* computeMaximumWaveSpeed[2D] per cell (Solvers/solvers.cpp:298-316, 365-387):
 * the material sound speeds at the interface pressure. */
static double rd_wave_speed(const orc_phys *ph, const double *w, int dim)
{
    const mp_ctx *c = (const mp_ctx *)ph->params;
    double s = (dim == 1) ? fabs(w[1]) : orc_smax(fabs(w[1]), fabs(w[2]));
    return s + orc_smax(eos_c(w[4], w[6], &c->m1), eos_c(w[5], w[6], &c->m2));
}

/* EulerReducedStateVector() (eulerreducedstatevector.cpp:3-14). */
static const double RD_DEFAULT[7] = {0.999, 0.0, 0.0, 0.0, 1.0, 1.0, 1.0};

/* ---- volume-fraction reinitialisation (Solvers/multiphysicssolvers.cpp) ------ */

/* relaxTotalDensity / relaxTotalPressure on one cell (both rho_k := alpha1 rho1
 * + alpha2 rho2; multiphysics also p_k := alpha1 p1 + alpha2 p2; the reference's
 * unused mixture energy has no effect, eulermultiphysicsstatevector.cpp:341-365,
 * eulerreducedstatevector.cpp:344-350), then the new volume fraction. */
static void relax_cell(int kind, double *w, double alpha)
{
    double a = w[0], a2 = 1.0 - a;
    if (kind == 0) {
        double rho = (a * w[4]) + (a2 * w[6]);
        w[4] = rho;
        w[6] = rho;
        double p = (a * w[5]) + (a2 * w[7]);
        w[5] = p;
        w[7] = p;
    } else {
        double rho = (a * w[4]) + (a2 * w[5]);
        w[4] = rho;
        w[5] = rho;
    }
    w[0] = alpha;
}

/* The windows of one line of n cells at stride `st` (in cells): every cell j with
 * alpha_j < 0.5 <= alpha_{j-1} (a falling crossing) relaxes cells j-2 .. j+1 to
 * 0.999 before j and 0.001 from j on; then every rising crossing (alpha_j > 0.5
 * >= alpha_{j-1}) the mirror image (multiphysicssolvers.cpp:37-161, 193-313).
 * Crossings are found before any cell changes.  The reference writes cells -1
 * and n of the line when a crossing sits at j = 1 or j = n - 1 (outside its
 * vectors) and holds at most 100 crossings of each kind per line; here window
 * cells outside the line are skipped and crossings are unbounded. */
static void reinit_line(int kind, int np, double *w, int n, size_t st, int *loc1, int *loc2)
{
    int n1 = 0, n2 = 0;
    for (int j = 1; j < n; j++) {
        double aj = w[np * st * (size_t)j], ap = w[np * st * (size_t)(j - 1)];
        if (aj < 0.5 && ap >= 0.5) loc1[n1++] = j;
        if (aj > 0.5 && ap <= 0.5) loc2[n2++] = j;
    }
    for (int k = 0; k < n1; k++)
        for (int j = loc1[k] - 2; j < loc1[k] + 2; j++)
            if (j >= 0 && j < n) relax_cell(kind, w + np * st * (size_t)j, j < loc1[k] ? 0.999 : 0.001);
    for (int k = 0; k < n2; k++)
        for (int j = loc2[k] - 2; j < loc2[k] + 2; j++)
            if (j >= 0 && j < n) relax_cell(kind, w + np * st * (size_t)j, j < loc2[k] ? 0.001 : 0.999);
}

/* reinitialiseVolumeFraction, 1D (multiphysicssolvers.cpp:7-35, 164-191): the
 * last falling crossing i >= 1 (the reference starts at i = 0 and then reads
 * cell -1; that read only matters when alpha_0 < 0.5), 0 when there is none,
 * and the window i-2 .. i+1 with cells outside the line skipped (the reference
 * writes cells -2, -1 when i < 2). */
static void reinit_1d(int kind, int np, double *w, int n)
{
    int loc = 0;
    for (int i = 1; i < n; i++)
        if (w[np * (size_t)i] < 0.5 && w[np * (size_t)(i - 1)] >= 0.5) loc = i;
    for (int i = loc - 2; i < loc + 2; i++)
        if (i >= 0 && i < n) relax_cell(kind, w + np * (size_t)i, i < loc ? 0.999 : 0.001);
}

/* reinitialiseVolumeFraction, 2D: every row, then every column. */
static void reinit_2d(int kind, int np, double *w, int ny, int nx)
{
    int m = ny > nx ? ny : nx;
    int *l1 = (int *)malloc(sizeof(int) * (size_t)m), *l2 = (int *)malloc(sizeof(int) * (size_t)m);
    for (int i = 0; i < ny; i++) reinit_line(kind, np, w + np * (size_t)i * nx, nx, 1, l1, l2);
    for (int j = 0; j < nx; j++) reinit_line(kind, np, w + np * (size_t)j, ny, (size_t)nx, l1, l2);
    free(l1);
    free(l2);
}

static void reinit_after_step(const orc_phys *ph, int kind, int order, int dim, int ny, int nx, int it, double *prim)
{
    const mp_ctx *c = (const mp_ctx *)ph->params;
    /* only the second-order solvers reinitialise (secondordersolver.cpp:343-349);
     * the first-order ones never do (firstordersolver.cpp:463-490) */
    if (order != 2 || c->reinit_every == 0 || (it % c->reinit_every) != 0) return;
    if (dim == 1) reinit_1d(kind, orc_np(ph), prim, nx);
    else reinit_2d(kind, orc_np(ph), prim, ny, nx);
}
static void mp_after_step(const orc_phys *ph, int order, int dim, int ny, int nx, int it, double *prim)
{
    reinit_after_step(ph, 0, order, dim, ny, nx, it, prim);
}
static void rd_after_step(const orc_phys *ph, int order, int dim, int ny, int nx, int it, double *prim)
{
    reinit_after_step(ph, 1, order, dim, ny, nx, it, prim);
}

static orc_phys mp_phys(int kind, mp_ctx *c)
{
    orc_phys ph;
    memset(&ph, 0, sizeof(ph));
    if (kind == 0) {
        ph.nf = 9;
        ph.np = 8;
        ph.prim_to_cons = mp_prim_to_cons;
        ph.cons_to_prim = mp_cons_to_prim;
        ph.flux = mp_flux;
        ph.wave_speed = mp_wave_speed;
        ph.default_state = MP_DEFAULT;
        ph.after_step = mp_after_step;
    } else {
        ph.nf = 8;
        ph.np = 7;
        ph.prim_to_cons = rd_prim_to_cons;
        ph.cons_to_prim = rd_cons_to_prim;
        ph.flux = rd_flux;
        ph.wave_speed = rd_wave_speed;
        ph.default_state = RD_DEFAULT;
        ph.after_step = rd_after_step;
    }
    ph.params = c;
    return ph;
}

int orc_mp_solve(int kind, int order, int dim, int ny, int nx, double dx, double cfl, double t_final, int max_steps,
                 double bias, int limiter, int bc, int reinit_every, const orc_euler_params *m1,
                 const orc_euler_params *m2, const double *prim_in, double *prim_out, double *dt_log, int dt_cap)
{
    mp_ctx c;
    c.m1 = *m1;
    c.m2 = *m2;
    c.reinit_every = reinit_every;
    orc_phys ph = mp_phys(kind, &c);
    return orc_solve(&ph, order, dim, ny, nx, dx, cfl, t_final, max_steps, bias, limiter, bc, prim_in, prim_out,
                     dt_log, dt_cap);
}

void orc_mp_sweep(int kind, int order, int dim, int dir, int ny, int nx, double dx, double ts, double bias,
                  int limiter, int bc, const orc_euler_params *m1, const orc_euler_params *m2, const double *prim_in,
                  double *prim_out)
{
    mp_ctx c;
    c.m1 = *m1;
    c.m2 = *m2;
    c.reinit_every = 0;
    orc_phys ph = mp_phys(kind, &c);
    orc_sweep(&ph, order, dim, dir, ny, nx, dx, ts, bias, limiter, bc, prim_in, prim_out);
}

void orc_mp_reinit(int kind, int dim, int ny, int nx, double *prim)
{
    int np = kind == 0 ? 8 : 7;
    if (dim == 1) reinit_1d(kind, np, prim, nx);
    else reinit_2d(kind, np, prim, ny, nx);
}

/* Per-function records, the layout of the reference harness's mp_units mode.
 * Input: four primitive states a, b, c, d per record (4 np doubles).  Output,
 * in order: cons(b) [nf] | prim(cons(b)) [np] | Fx(cons(b)) [nf] | Fy(cons(b))
 * [nf] | wave speed 1D, 2D of b [2] | slope(a, b, c) [nf] | evolveX(a, b, c)
 * side 0, side 1 [2 np] | FORCEx(b, c) [nf] | SLICx(a, b, c, d) [nf] |
 * SLICy(a, b, c, d) [nf]. */
void orc_mp_units(int kind, int n, const orc_euler_params *m1, const orc_euler_params *m2, double bias, int limiter,
                  double dx, double ts, const double *states, double *out)
{
    mp_ctx c;
    c.m1 = *m1;
    c.m2 = *m2;
    c.reinit_every = 0;
    orc_phys ph = mp_phys(kind, &c);
    const int nf = ph.nf, np = ph.np;
    const int rec = 7 * nf + 3 * np + 2;
    for (int r = 0; r < n; r++) {
        const double *a = states + (size_t)r * 4 * np, *b = a + np, *cc = b + np, *d = cc + np;
        double *o = out + (size_t)r * rec;
        double U[ORC_MAX_NF], Ua[ORC_MAX_NF], Uc[ORC_MAX_NF];
        ph.prim_to_cons(&ph, b, U);
        memcpy(o, U, sizeof(double) * nf);
        o += nf;
        ph.cons_to_prim(&ph, U, o);
        o += np;
        ph.flux(&ph, 0, U, o);
        o += nf;
        ph.flux(&ph, 1, U, o);
        o += nf;
        *o++ = ph.wave_speed(&ph, b, 1);
        *o++ = ph.wave_speed(&ph, b, 2);
        ph.prim_to_cons(&ph, a, Ua);
        ph.prim_to_cons(&ph, cc, Uc);
        orc_slope_vector(nf, Ua, U, Uc, bias, limiter, o);
        o += nf;
        orc_evolve_half(&ph, 0, a, b, cc, dx, ts, bias, limiter, 0, o);
        o += np;
        orc_evolve_half(&ph, 0, a, b, cc, dx, ts, bias, limiter, 1, o);
        o += np;
        orc_force_flux(&ph, 0, b, cc, dx, ts, o);
        o += nf;
        orc_slic_flux(&ph, 0, a, b, cc, d, dx, ts, bias, limiter, o);
        o += nf;
        orc_slic_flux(&ph, 1, a, b, cc, d, dx, ts, bias, limiter, o);
    }
}
