/* Two-material ideal MHD with the real ghost fluid method -- TEST
 * INFRASTRUCTURE ONLY.
 *
 * MHDHLLCSolver (RGFM/mhdhllcsolver.cpp: the HLLC interface state of two MHD
 * materials, its star densities and the HLL averages it builds on) and
 * MHDRGFMSolver (RGFM/mhdrgfmsolver.cpp: ghost fill, level-set transport,
 * solve / solve2D) restated in the reference's operation order, on the MHD
 * physics table of mhd_oracle.c and the generic scheme.  Primitive states are
 * (rho, u, v, w, p, Bx, By, Bz, psi).  Pinned bit for bit against the compiled
 * reference by tests/test_oracle_mhd_rgfm.py.
 */
#include "oracle.h"
#include "scheme_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define NF 9

orc_phys orc_mhd_phys_table(orc_mhd_params *mp);
double orc_mhd_fast_speed(const orc_mhd_params *mp, const double *w, int dir);

typedef const orc_mhd_params *mpar;

/* std::min / std::max with the argument order of the call. */
static double smin(double a, double b) { return (b < a) ? b : a; }
static double smax(double a, double b) { return (a < b) ? b : a; }

/* MHDStateVector::computeTotalEnergy and setConservedVariableVector through the
 * physics table. */
static double total_energy(const double *w, mpar m)
{
    orc_mhd_params local = *m;
    orc_phys ph = orc_mhd_phys_table(&local);
    double U[NF];
    ph.prim_to_cons(&ph, w, U);
    return U[4];
}
static void cons_to_prim(const double *U, double *w, mpar m)
{
    orc_mhd_params local = *m;
    orc_phys ph = orc_mhd_phys_table(&local);
    ph.cons_to_prim(&ph, U, w);
}

/* computeTotalPressure (mhdhllcsolver.cpp:268-279). */
static double total_pressure(const double *w)
{
    double b2 = (w[5] * w[5]) + (w[6] * w[6]) + (w[7] * w[7]);
    return w[4] + (0.5 * b2);
}

/* The interface problem between a (left / top, material m1) and b (right /
 * bottom, material m2) in direction d (0: x, 1: y).  n = normal, t = the
 * in-plane tangential slot, as the reference's X and Y function families use
 * them. */
typedef struct {
    int d;
    const double *a, *b;
    mpar m1, m2;
} prob;

static int vn(int d) { return 1 + d; }   /* normal velocity slot */
static int bn(int d) { return 5 + d; }   /* normal field slot */

/* computeLeft/Right/Top/BottomWaveSpeed (:1495-1537). */
static double wave_a(const prob *P)
{
    return smin(P->a[vn(P->d)], P->b[vn(P->d)]) -
           smax(orc_mhd_fast_speed(P->m1, P->a, P->d), orc_mhd_fast_speed(P->m2, P->b, P->d));
}
static double wave_b(const prob *P)
{
    return smax(P->a[vn(P->d)], P->b[vn(P->d)]) +
           smax(orc_mhd_fast_speed(P->m1, P->a, P->d), orc_mhd_fast_speed(P->m2, P->b, P->d));
}

/* computeStarRegionX/YVelocity (:281-333). */
static double star_speed(const prob *P)
{
    const double *a = P->a, *b = P->b;
    int n = vn(P->d), f = bn(P->d);
    double Sa = wave_a(P), Sb = wave_b(P);
    double ca = a[0] * (Sa - a[n]);
    double cb = b[0] * (Sb - b[n]);
    double num = (cb * b[n]) - (ca * a[n]) + total_pressure(a) - total_pressure(b) - (a[f] * a[f]) + (b[f] * b[f]);
    return num / (cb - ca);
}

/* computeStarRegionDensity and its four sides (:7-58). */
static double star_rho(double rho, double S, double u, double Ss) { return rho * ((S - u) / (S - Ss)); }
static double star_rho_a(const prob *P) { return star_rho(P->a[0], wave_a(P), P->a[vn(P->d)], star_speed(P)); }
static double star_rho_b(const prob *P) { return star_rho(P->b[0], wave_b(P), P->b[vn(P->d)], star_speed(P)); }

/* The HLL averages computeX/YStarRegion* (:547-922): (S_b q_b - S_a q_a -
 * (f_b - f_a)) / (S_b - S_a) for each conserved quantity and its flux. */
static double hll_avg(double Sa, double Sb, double qa, double qb, double fa, double fb)
{
    return ((Sb * qb) - (Sa * qa) - (fb - fa)) / (Sb - Sa);
}
/* The momentum averages write S q as (S rho) vel, left to right. */
static double hll_mom(double Sa, double Sb, const double *a, const double *b, int k, double fa, double fb)
{
    return ((Sb * b[0] * b[k]) - (Sa * a[0] * a[k]) - (fb - fa)) / (Sb - Sa);
}

/* The HLL value of field `q` (0 rho, 1-3 momentum x/y/z, 5-7 B, 8 psi):
 * computeX/YHLL* (:923-1334): the left state's value when 0 <= S_a, the right
 * state's when S_b <= 0, the HLL average otherwise. */
static double hll_field(const prob *P, int q)
{
    const double *a = P->a, *b = P->b;
    int d = P->d, n = vn(d);
    double Sa = wave_a(P), Sb = wave_b(P);
    double avg, va, vb;
    if (q == 0) {
        avg = hll_avg(Sa, Sb, a[0], b[0], a[0] * a[n], b[0] * b[n]);
        va = a[0];
        vb = b[0];
    } else if (q >= 1 && q <= 3) {
        int k = q;   /* momentum component k (velocity slot k) */
        double fa, fb;
        if (k == n) {
            double ba2 = (a[5] * a[5]) + (a[6] * a[6]) + (a[7] * a[7]);
            double bb2 = (b[5] * b[5]) + (b[6] * b[6]) + (b[7] * b[7]);
            fa = (a[0] * (a[n] * a[n])) + a[4] + (0.5 * ba2) - (a[bn(d)] * a[bn(d)]);
            fb = (b[0] * (b[n] * b[n])) + b[4] + (0.5 * bb2) - (b[bn(d)] * b[bn(d)]);
        } else {
            fa = (a[0] * (a[n] * a[k])) - (a[bn(d)] * a[4 + k]);
            fb = (b[0] * (b[n] * b[k])) - (b[bn(d)] * b[4 + k]);
        }
        avg = hll_mom(Sa, Sb, a, b, k, fa, fb);
        va = a[0] * a[k];
        vb = b[0] * b[k];
    } else if (q >= 5 && q <= 7) {
        double fa, fb;
        if (q == bn(d)) {
            fa = a[8];
            fb = b[8];
        } else {
            /* induction flux B_q u_n - B_n u_q (x: (By u) - (Bx v); y: (Bx v) - (By u)) */
            int vq = q - 4;
            fa = (a[q] * a[n]) - (a[bn(d)] * a[vq]);
            fb = (b[q] * b[n]) - (b[bn(d)] * b[vq]);
        }
        avg = hll_avg(Sa, Sb, a[q], b[q], fa, fb);
        va = a[q];
        vb = b[q];
    } else {
        double fa = (P->m1->ch * P->m1->ch) * a[bn(d)];
        double fb = (P->m2->ch * P->m2->ch) * b[bn(d)];
        avg = hll_avg(Sa, Sb, a[8], b[8], fa, fb);
        va = a[8];
        vb = b[8];
    }
    if (0.0 <= Sa) return va;
    if (Sb <= 0.0) return vb;
    return avg;
}

/* computeLeft/Right/Top/BottomStarRegionTotalPressure (:335-401). */
static double star_ptot(const prob *P, int side)
{
    const double *w = side ? P->b : P->a;
    int n = vn(P->d), f = bn(P->d);
    double S = side ? wave_b(P) : wave_a(P);
    double Ss = star_speed(P);
    double Bn = hll_field(P, f);
    return (w[0] * (S - w[n]) * (Ss - w[n])) + total_pressure(w) - (w[f] * w[f]) + (Bn * Bn);
}

/* The tangential star momenta (:403-545): rho* vel_t - (Bn* Bt* - Bn Bt) /
 * (S - S*). */
static double star_mom_t(const prob *P, int side, int k)
{
    const double *w = side ? P->b : P->a;
    int f = bn(P->d);
    double rs = side ? star_rho_b(P) : star_rho_a(P);
    double Bn = hll_field(P, f), Bt = hll_field(P, 4 + k);
    double Ss = star_speed(P);
    double S = side ? wave_b(P) : wave_a(P);
    return (rs * w[k]) - (((Bn * Bt) - (w[f] * w[4 + k])) / (S - Ss));
}

/* computeLeft/Right/Top/BottomStarRegionTotalEnergy (:1335-1493). */
static double star_energy(const prob *P, int side)
{
    const double *w = side ? P->b : P->a;
    mpar m = side ? P->m2 : P->m1;
    int n = vn(P->d), f = bn(P->d);
    double Ss = star_speed(P);
    double S = side ? wave_b(P) : wave_a(P);
    double Bx = hll_field(P, 5), By = hll_field(P, 6), Bz = hll_field(P, 7);
    double rh = hll_field(P, 0);
    double ux = hll_field(P, 1) / rh, uy = hll_field(P, 2) / rh, uz = hll_field(P, 3) / rh;
    double bv_s = (Bx * ux) + (By * uy) + (Bz * uz);
    double bv = (w[5] * w[1]) + (w[6] * w[2]) + (w[7] * w[3]);
    double pt = total_pressure(w);
    double pts = star_ptot(P, side);
    double E = total_energy(w, m);
    double Bns = (f == 5) ? Bx : By;
    double first = E * ((S - w[n]) / (S - Ss));
    double second = ((pts * Ss) - (pt * w[n]) - ((Bns * bv_s) - (w[f] * bv))) / (S - Ss);
    return first + second;
}

/* computeLeft/Right/Top/BottomStarRegionConservedVariableVector (:60-198). */
static void star_cons(const prob *P, int side, double *U)
{
    int n = vn(P->d);
    double rs = side ? star_rho_b(P) : star_rho_a(P);
    U[0] = rs;
    for (int k = 1; k <= 3; k++) U[k] = (k == n) ? rs * star_speed(P) : star_mom_t(P, side, k);
    U[4] = star_energy(P, side);
    U[5] = hll_field(P, 5);
    U[6] = hll_field(P, 6);
    U[7] = hll_field(P, 7);
    U[8] = hll_field(P, 8);
}

/* MHDHLLCSolver::solveX / solveY (:200-266). */
void orc_mhd_hllc_solve(int d, const double *a, const double *b, const orc_mhd_params *m1,
                        const orc_mhd_params *m2, double *out)
{
    prob P = {d, a, b, m1, m2};
    double Ss = star_speed(&P), Sa = wave_a(&P), Sb = wave_b(&P);
    if (0.0 <= Sa) {
        memcpy(out, a, sizeof(double) * NF);
    } else if (Sb <= 0.0) {
        memcpy(out, b, sizeof(double) * NF);
    } else if (0.0 <= Ss) {
        double U[NF];
        star_cons(&P, 0, U);
        cons_to_prim(U, out, m1);
    } else {
        double U[NF];
        star_cons(&P, 1, U);
        cons_to_prim(U, out, m2);
    }
}

/* ---- ghost fill (mhdrgfmsolver.cpp:7-378) ------------------------------------ */

/* Writes a ghost: density, velocities, the floored pressure and B from the
 * interface solution, and (2D fills only) its psi. */
static void set_ghost(double *w, double rho, const double *s, double p, int with_psi)
{
    w[0] = rho;
    w[1] = s[1];
    w[2] = s[2];
    w[3] = s[3];
    w[4] = p;
    w[5] = s[5];
    w[6] = s[6];
    w[7] = s[7];
    if (with_psi) w[8] = s[8];
}

static void fill_line(int d, int n, size_t st, const double *phi, size_t pst, const double *o1, const double *o2,
                      double *n1, double *n2, mpar p1, mpar p2, int in_ghost, int with_psi, double pfloor)
{
    double prev = phi[0];
    for (int k = 1; k < n; k++) {
        double cur = phi[(size_t)k * pst];
        if ((cur * prev) <= 0.0 && k > 2 && k < (n - 3)) {
            const double *a, *b;
            mpar pa, pb;
            double *na, *nb;
            if (!in_ghost) {
                a = o1 + (size_t)(k - 1) * st * NF; b = o2 + (size_t)k * st * NF; pa = p1; pb = p2; na = n1; nb = n2;
            } else {
                a = o2 + (size_t)(k - 1) * st * NF; b = o1 + (size_t)k * st * NF; pa = p2; pb = p1; na = n2; nb = n1;
            }
            prob P = {d, a, b, pa, pb};
            double s[NF];
            orc_mhd_hllc_solve(d, a, b, pa, pb, s);
            double p = s[4];
            if (p < pfloor) p = pfloor;
            double ra = star_rho_a(&P), rb = star_rho_b(&P);
            for (int j = 0; j < 4; j++) {
                set_ghost(na + (size_t)(k + j - 1) * st * NF, ra, s, p, with_psi);
                set_ghost(nb + (size_t)(k - j) * st * NF, rb, s, p, with_psi);
            }
            in_ghost = !in_ghost;
        }
        prev = cur;
    }
}

void orc_mhd_rgfm_ghost_fill(int dim, int ny, int nx, const orc_mhd_params *p1, const orc_mhd_params *p2, double *m1,
                             double *m2, const double *phi)
{
    const double pfloor = pow(10.0, -8.0);
    size_t cells = (size_t)ny * nx;
    double *o1 = malloc(sizeof(double) * NF * cells), *o2 = malloc(sizeof(double) * NF * cells);
    memcpy(o1, m1, sizeof(double) * NF * cells);
    memcpy(o2, m2, sizeof(double) * NF * cells);
    if (dim == 1) {
        /* the 1D fill leaves psi alone and starts outside a ghost region */
        fill_line(0, nx, 1, phi, 1, o1, o2, m1, m2, p1, p2, 0, 0, pfloor);
    } else {
        for (int i = 0; i < ny; i++)
            fill_line(0, nx, 1, phi + (size_t)i * nx, 1, o1 + (size_t)i * nx * NF, o2 + (size_t)i * nx * NF,
                      m1 + (size_t)i * nx * NF, m2 + (size_t)i * nx * NF, p1, p2, 1, 1, pfloor);
        memcpy(o1, m1, sizeof(double) * NF * cells);
        memcpy(o2, m2, sizeof(double) * NF * cells);
        for (int j = 0; j < nx; j++)
            fill_line(1, ny, (size_t)nx, phi + j, (size_t)nx, o1 + (size_t)j * NF, o2 + (size_t)j * NF,
                      m1 + (size_t)j * NF, m2 + (size_t)j * NF, p1, p2, 1, 1, pfloor);
    }
    free(o1);
    free(o2);
}

/* ---- level set (mhdrgfmsolver.cpp:380-520): the Euler RGFM's upwind update
 * with the MHD materials' velocities ---------------------------------------- */

static int mirror(int k, int n) { return (k < 0) ? -1 - k : (k >= n ? 2 * n - 1 - k : k); }

static void level_set_line(int d, int n, const double *phi, size_t ps, const double *m1, const double *m2, size_t ms,
                           double h, double dt, double *out, int in_ghost)
{
    double prev = phi[0];
    for (int k = 0; k < n; k++) {
        double c = phi[(size_t)k * ps];
        if ((prev * c) <= 0.0) in_ghost = !in_ghost;
        double v = (!in_ghost ? m1 : m2)[(size_t)k * ms * NF + 1 + d];
        double up;
        if (v < 0.0) {
            double a = phi[(size_t)mirror(k + 2, n) * ps], b = phi[(size_t)mirror(k + 1, n) * ps];
            up = (-a + (4.0 * b) - (3.0 * c)) / (2.0 * h);
        } else {
            double a = phi[(size_t)mirror(k - 1, n) * ps], b = phi[(size_t)mirror(k - 2, n) * ps];
            up = ((3.0 * c) - (4.0 * a) + b) / (2.0 * h);
        }
        out[(size_t)k * ps] = c - (dt * (v * up));
        prev = c;
    }
}

static void level_set(int dim, int d, int ny, int nx, double h, double dt, const double *m1, const double *m2,
                      const double *in, double *out)
{
    if (dim == 1) {
        level_set_line(0, nx, in, 1, m1, m2, 1, h, dt, out, 0);
    } else if (d == 0) {
        for (int i = 0; i < ny; i++)
            level_set_line(0, nx, in + (size_t)i * nx, 1, m1 + (size_t)i * nx * NF, m2 + (size_t)i * nx * NF, 1, h, dt,
                           out + (size_t)i * nx, 1);
    } else {
        for (int j = 0; j < nx; j++)
            level_set_line(1, ny, in + j, (size_t)nx, m1 + (size_t)j * NF, m2 + (size_t)j * NF, (size_t)nx, h, dt,
                           out + j, 1);
    }
}

/* ---- driver: MHDRGFMSolver::solve / solve2D (mhdrgfmsolver.cpp:522-693) ------ */

static void advance_material(orc_phys *ph, int dim, int ny, int nx, double dx, double dt, double bias, int limiter,
                             int subcycles, double *a, double *b)
{
    size_t cells = (size_t)ny * nx;
    if (dim == 1) {
        /* the 1D solve ignores its sub-cycle count */
        orc_sweep(ph, 2, 1, 0, 1, nx, dx, dt, bias, limiter, 0, a, b);
        memcpy(a, b, sizeof(double) * NF * cells);
        return;
    }
    for (int s = 0; s < subcycles; s++) {
        orc_forcing_step(ph, 2, ny, nx, 0, dx, 0.5 * (dt / subcycles), bias, limiter, a, b);
        memcpy(a, b, sizeof(double) * NF * cells);
    }
    double half = 0.5 * dt;
    orc_sweep(ph, 2, 2, 0, ny, nx, dx, half, bias, limiter, 0, a, b);
    orc_sweep(ph, 2, 2, 1, ny, nx, dx, dt, bias, limiter, 0, b, a);
    orc_sweep(ph, 2, 2, 0, ny, nx, dx, half, bias, limiter, 0, a, b);
    memcpy(a, b, sizeof(double) * NF * cells);
    for (int s = 0; s < subcycles; s++) {
        orc_forcing_step(ph, 2, ny, nx, 0, dx, 0.5 * (dt / subcycles), bias, limiter, a, b);
        memcpy(a, b, sizeof(double) * NF * cells);
    }
}

int orc_mhd_rgfm_solve(int dim, int ny, int nx, double dx, double cfl, double t_final, int max_steps, double bias,
                       int limiter, int subcycles, const orc_mhd_params *p1_in, const orc_mhd_params *p2_in,
                       const double *m1_in, const double *m2_in, const double *phi_in, double *m1_out,
                       double *m2_out, double *phi_out, double *dt_log, int dt_cap)
{
    if (dim == 1) ny = 1;
    size_t cells = (size_t)ny * nx;
    orc_mhd_params p1 = *p1_in, p2 = *p2_in;   /* solve2D mutates its copies' c_h and c_p */
    orc_phys ph1 = orc_mhd_phys_table(&p1), ph2 = orc_mhd_phys_table(&p2);
    double *a1 = malloc(sizeof(double) * NF * cells), *b1 = malloc(sizeof(double) * NF * cells);
    double *a2 = malloc(sizeof(double) * NF * cells), *b2 = malloc(sizeof(double) * NF * cells);
    double *phi = malloc(sizeof(double) * cells), *phi2 = malloc(sizeof(double) * cells);
    memcpy(a1, m1_in, sizeof(double) * NF * cells);
    memcpy(a2, m2_in, sizeof(double) * NF * cells);
    memcpy(phi, phi_in, sizeof(double) * cells);
    double t = 0.0;
    int it = 0;
    while (t < t_final && (max_steps == 0 || it < max_steps)) {
        orc_mhd_rgfm_ghost_fill(dim, ny, nx, &p1, &p2, a1, a2, phi);
        double d1 = orc_stable_time_step(orc_ghosted_max(&ph1, dim, ny, nx, 0, a1), dx, cfl, t, t_final, it);
        double d2 = orc_stable_time_step(orc_ghosted_max(&ph2, dim, ny, nx, 0, a2), dx, cfl, t, t_final, it);
        double dt = smin(d1, d2);
        if (dim == 1) {
            level_set(1, 0, 1, nx, dx, dt, a1, a2, phi, phi2);
            memcpy(phi, phi2, sizeof(double) * cells);
        } else {
            level_set(2, 0, ny, nx, dx, 0.5 * dt, a1, a2, phi, phi2);
            level_set(2, 1, ny, nx, dx, dt, a1, a2, phi2, phi);
            level_set(2, 0, ny, nx, dx, 0.5 * dt, a1, a2, phi, phi2);
            memcpy(phi, phi2, sizeof(double) * cells);
            /* c_h = 4 and configureParabolicDamping for both materials (:638-645) */
            p1.ch = 4.0;
            p2.ch = 4.0;
            p1.cp = sqrt(p1.ch * 0.18);
            p2.cp = sqrt(p2.ch * 0.18);
        }
        advance_material(&ph1, dim, ny, nx, dx, dt, bias, limiter, subcycles, a1, b1);
        advance_material(&ph2, dim, ny, nx, dx, dt, bias, limiter, subcycles, a2, b2);
        if (dt_log && it < dt_cap) dt_log[it] = dt;
        t += dt;
        it += 1;
    }
    memcpy(m1_out, a1, sizeof(double) * NF * cells);
    memcpy(m2_out, a2, sizeof(double) * NF * cells);
    memcpy(phi_out, phi, sizeof(double) * cells);
    free(a1); free(b1); free(a2); free(b2); free(phi); free(phi2);
    return it;
}

/* Per-function records: states a, b (18 doubles) -> solveX(a, b) [9],
 * solveY(a, b) [9], star densities left, right, top, bottom [4], star speeds x,
 * y [2] = 24 doubles, the layout of the harness's mhd_hllc_units mode. */
void orc_mhd_hllc_units(int n, const orc_mhd_params *p1, const orc_mhd_params *p2, const double *states, double *out)
{
    for (int r = 0; r < n; r++) {
        const double *a = states + 18 * (size_t)r, *b = a + 9;
        double *o = out + 24 * (size_t)r;
        prob X = {0, a, b, p1, p2}, Y = {1, a, b, p1, p2};
        orc_mhd_hllc_solve(0, a, b, p1, p2, o);
        orc_mhd_hllc_solve(1, a, b, p1, p2, o + 9);
        o[18] = star_rho_a(&X);
        o[19] = star_rho_b(&X);
        o[20] = star_rho_a(&Y);
        o[21] = star_rho_b(&Y);
        o[22] = star_speed(&X);
        o[23] = star_speed(&Y);
    }
}
