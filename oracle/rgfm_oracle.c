/* Two-material Euler with the real ghost fluid method (RGFM) -- TEST
 * INFRASTRUCTURE ONLY (oracle).
 *
 * HLLCSolver (RGFM/hllcsolver.cpp), ExactSolver (RGFM/exactsolver.cpp), the
 * RGFM ghost fill, the level-set advection and the solveHLLC / solveExact
 * drivers (RGFM/rgfmsolver.cpp) restated in the reference's operation order on
 * top of the Euler oracle.  States are primitive (rho, u, v, w, p, lambda).
 *
 * Pinning (tests/test_oracle_rgfm.py): the HLLC path bit for bit against golden
 * vectors of the compiled reference (per-function records and whole 1D/2D
 * solves).  The reference's exact solver does not compile
 * (exactsolver.cpp:295,331,388,424: a state built from a vector), so the
 * exact-solver path is pinned to Toro's published exact star states and to the
 * Rankine-Hugoniot / isentrope relations instead; its rarefaction-fan sampling
 * follows Toro's formulas where the reference's would not compile.
 * The anisotropic spacing (dy != dx) is not in the reference: with dy == dx
 * every function reduces to the reference's arithmetic.
 */
#include "oracle.h"
#include "scheme_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define NF 6

typedef const orc_euler_params *mpar;

static double snd(const double *w, mpar m) { return orc_euler_sound_speed(w[0], w[4], m); }
static double vel(const double *w, int dir) { return w[1 + dir]; }

/* EulerStateVector::computeTotalEnergy = the energy slot of the conserved vector. */
static double total_energy(const double *w, mpar m)
{
    orc_phys ph = orc_euler_phys_table(m);
    double U[NF];
    ph.prim_to_cons(&ph, w, U);
    return U[4];
}

static void cons_to_prim(const double *U, double *w, mpar m)
{
    orc_phys ph = orc_euler_phys_table(m);
    ph.cons_to_prim(&ph, U, w);
}

/* ---- HLLC (hllcsolver.cpp) ------------------------------------------------ */

/* computeStarRegionDensity (:7-10) */
static double hllc_star_rho(double rho, double S, double u, double Ss) { return rho * ((S - u) / (S - Ss)); }

/* computePVRSStarRegionX/YVelocity and -Pressure (:182-254). */
static double pvrs_prod(const double *a, const double *b, mpar m1, mpar m2)
{
    return 0.25 * (a[0] + b[0]) * (snd(a, m1) + snd(b, m2));
}
static double pvrs_vel(int dir, const double *a, const double *b, mpar m1, mpar m2)
{
    double prod = pvrs_prod(a, b, m1, m2);
    return (0.5 * (vel(a, dir) + vel(b, dir))) + (0.5 * (a[4] - b[4]) / prod);
}
static double pvrs_pres(int dir, const double *a, const double *b, mpar m1, mpar m2)
{
    double prod = pvrs_prod(a, b, m1, m2);
    return (0.5 * (a[4] + b[4])) + (0.5 * (vel(a, dir) - vel(b, dir)) * prod);
}

/* computeTwoRarefactionStarRegionX/YVelocity and -Pressure (:256-348). */
static double two_rare_vel(int dir, const double *a, const double *b, mpar m1, mpar m2)
{
    double ca = snd(a, m1), cb = snd(b, m2);
    double g = 0.5 * (m1->gamma + m2->gamma);
    double ex = (g - 1.0) / (2.0 * g);
    double pr = pow((a[4] / b[4]), ex);
    double co = 2.0 / (g - 1.0);
    return ((pr * (vel(a, dir) / ca)) + (vel(b, dir) / cb) + (co * (pr - 1.0))) / ((pr / ca) + (1.0 / cb));
}
static double two_rare_pres(int dir, const double *a, const double *b, mpar m1, mpar m2)
{
    double ca = snd(a, m1), cb = snd(b, m2);
    double g = 0.5 * (m1->gamma + m2->gamma);
    double co = 0.5 * (g - 1.0);
    double us = two_rare_vel(dir, a, b, m1, m2);
    double ex = 2.0 * (g / (g - 1.0));
    double ld = 1.0 + (co * ((vel(a, dir) - us) / ca));
    double rd = 1.0 + (co * ((us - vel(b, dir)) / cb));
    return 0.5 * ((a[4] * pow(ld, ex)) + (b[4] * pow(rd, ex)));
}

/* computeTwoShockStarRegionX/YVelocity and -Pressure (:350-462). */
static void two_shock_coeffs(int dir, const double *a, const double *b, mpar m1, mpar m2, double *ga, double *gb)
{
    double g = 0.5 * (m1->gamma + m2->gamma);
    double ppv = pvrs_pres(dir, a, b, m1, m2);
    double A = 2.0 / (g + 1.0);
    double B = (g - 1.0) / (g + 1.0);
    *ga = sqrt((A / a[0]) / ((B * a[4]) + ppv));
    *gb = sqrt((A / b[0]) / ((B * b[4]) + ppv));
}
static double two_shock_pres(int dir, const double *a, const double *b, mpar m1, mpar m2)
{
    double ga, gb;
    two_shock_coeffs(dir, a, b, m1, m2, &ga, &gb);
    return ((ga * a[4]) + (gb * b[4]) - (vel(b, dir) - vel(a, dir))) / (ga + gb);
}
static double two_shock_vel(int dir, const double *a, const double *b, mpar m1, mpar m2)
{
    double ps = two_shock_pres(dir, a, b, m1, m2);
    double ga, gb;
    two_shock_coeffs(dir, a, b, m1, m2, &ga, &gb);
    return (0.5 * (vel(a, dir) + vel(b, dir))) + (0.5 * ((gb * (ps - b[4])) - (ga * (ps - a[4]))));
}

/* The adaptive choice of computeStarRegionX/YVelocity and -Pressure (:464-606).
 * PVRS whenever the pressure ratio is moderate and the PVRS guess lies between
 * the two pressures, or the materials differ.  In computeStarRegionYVelocity the
 * "PVRS pressure" of that test is the PVRS *velocity* (hllcsolver.cpp:552);
 * kept as the reference has it. */
static int use_pvrs(double pa, double pb, double ppv, mpar m1, mpar m2)
{
    double pmin = orc_smin(pa, pb), pmax = orc_smax(pa, pb);
    double ratio = pmax / pmin;
    return ((ratio <= 2.0) && (pmin <= ppv) && (ppv <= pmax)) || fabs(m1->gamma - m2->gamma) > 1e-8 ||
           m1->p_inf > 0.0 || m2->p_inf > 0.0;
}
double orc_hllc_star_vel(int dir, const double *a, const double *b, mpar m1, mpar m2)
{
    double ppv = (dir == 0) ? pvrs_pres(0, a, b, m1, m2) : pvrs_vel(1, a, b, m1, m2);
    if (use_pvrs(a[4], b[4], ppv, m1, m2)) return pvrs_vel(dir, a, b, m1, m2);
    if (ppv < orc_smin(a[4], b[4])) return two_rare_vel(dir, a, b, m1, m2);
    return two_shock_vel(dir, a, b, m1, m2);
}
double orc_hllc_star_pres(int dir, const double *a, const double *b, mpar m1, mpar m2)
{
    double ppv = pvrs_pres(dir, a, b, m1, m2);
    if (use_pvrs(a[4], b[4], ppv, m1, m2)) return ppv;
    if (ppv < orc_smin(a[4], b[4])) return two_rare_pres(dir, a, b, m1, m2);
    return two_shock_pres(dir, a, b, m1, m2);
}

/* computeWaveSpeedWeighting (:608-623) and the side wave speeds (:625-667). */
static double weighting(double ps, const double *w, mpar m)
{
    double p = w[4];
    if (ps <= p) return 1.0;
    return sqrt(1.0 + ((m->gamma + 1.0) / (2.0 * m->gamma)) * ((ps + m->p_inf) / (p + m->p_inf) - 1.0));
}
static double wave_a(int dir, const double *a, const double *b, mpar m1, mpar m2)
{
    double ps = orc_hllc_star_pres(dir, a, b, m1, m2);
    return vel(a, dir) - (weighting(ps, a, m1) * snd(a, m1));
}
static double wave_b(int dir, const double *a, const double *b, mpar m1, mpar m2)
{
    double ps = orc_hllc_star_pres(dir, a, b, m1, m2);
    return vel(b, dir) + (weighting(ps, b, m2) * snd(b, m2));
}

/* computeLeft/Right/Top/BottomStarRegionDensity (:12-58). */
static double star_rho_a(int dir, const double *a, const double *b, mpar m1, mpar m2)
{
    double Ss = orc_hllc_star_vel(dir, a, b, m1, m2);
    double S = wave_a(dir, a, b, m1, m2);
    return hllc_star_rho(a[0], S, vel(a, dir), Ss);
}
static double star_rho_b(int dir, const double *a, const double *b, mpar m1, mpar m2)
{
    double Ss = orc_hllc_star_vel(dir, a, b, m1, m2);
    double S = wave_b(dir, a, b, m1, m2);
    return hllc_star_rho(b[0], S, vel(b, dir), Ss);
}

/* computeStarRegionX/YConservedVariableVector (:60-112). */
static void star_cons(int dir, const double *w, double S, double Ss, mpar m, double *U)
{
    double rho = w[0], p = w[4], un = vel(w, dir);
    double E = total_energy(w, m);
    double rs = hllc_star_rho(rho, S, un, Ss);
    double et = (Ss - un) * (Ss + (p / (rho * (S - un))));
    U[0] = rs;
    U[1] = rs * (dir == 0 ? Ss : w[1]);
    U[2] = rs * (dir == 1 ? Ss : w[2]);
    U[3] = rs * w[3];
    U[4] = rs * ((E / rho) + et);
    U[5] = rs;
}

/* HLLCSolver::solveX / solveY (:114-180): the state at x/t = 0. */
void orc_hllc_solve(int dir, const double *a, const double *b, mpar m1, mpar m2, double *out)
{
    double Ss = orc_hllc_star_vel(dir, a, b, m1, m2);
    double Sa = wave_a(dir, a, b, m1, m2);
    double Sb = wave_b(dir, a, b, m1, m2);
    double U[NF];
    if (0.0 <= Sa) {
        memcpy(out, a, sizeof(double) * NF);
    } else if (Sb <= 0.0) {
        memcpy(out, b, sizeof(double) * NF);
    } else if (0.0 <= Ss) {
        star_cons(dir, a, Sa, Ss, m1, U);
        cons_to_prim(U, out, m1);
    } else {
        star_cons(dir, b, Sb, Ss, m2, U);
        cons_to_prim(U, out, m2);
    }
}

/* ---- exact Riemann solver (exactsolver.cpp, Toro ch. 4 for stiffened gas) ---- */

static double ex_A(double rho, double g) { return 2.0 / (rho * (g + 1.0)); }
static double ex_B(double p, double g, double pinf) { return (p * ((g - 1.0) / (g + 1.0))) + ((2.0 * g * pinf) / (g + 1.0)); }

/* computeWaveJumpFunctionComponent (:165-186) and its derivative (:202-221). */
static double ex_f(double ps, const double *w, mpar m)
{
    double rho = w[0], p = w[4], c = snd(w, m), g = m->gamma, pi = m->p_inf;
    if (ps > p) {
        double co = sqrt(ex_A(rho, g) / (ps + ex_B(p, g, pi)));
        return (ps - p) * co;
    }
    double co = pow((ps + pi) / (p + pi), (g - 1.0) / (2.0 * g));
    return 2.0 * (c / (g - 1.0)) * (co - 1.0);
}
static double ex_df(double ps, const double *w, mpar m)
{
    double rho = w[0], p = w[4], c = snd(w, m), g = m->gamma, pi = m->p_inf;
    if (ps > p) {
        double co = sqrt(ex_A(rho, g) / (ps + ex_B(p, g, pi)));
        return (1.0 - ((ps - p) / (2.0 * (ps + ex_B(p, g, pi))))) * co;
    }
    double co = pow((ps + pi) / (p + pi), -(g + 1.0) / (2.0 * g));
    return (co * c) / (g * (p + pi));
}

/* computeStarRegionX/YPressure (:115-163): Newton from 3 x the mean pressure,
 * relative change 2|a-b|/|a+b| < 1e-8, at most 10^4 iterations, floor 1e-8. */
double orc_exact_star_pres(int dir, const double *a, const double *b, mpar m1, mpar m2)
{
    double cur = 0.5 * (a[4] + b[4]);
    double nw = 3.0 * cur;
    int it = 0;
    while ((2.0 * (fabs(nw - cur) / fabs(nw + cur))) >= 1e-8 && it < 10000) {
        cur = nw;
        double f = ex_f(cur, a, m1) + ex_f(cur, b, m2) + vel(b, dir) - vel(a, dir);
        double df = ex_df(cur, a, m1) + ex_df(cur, b, m2);
        nw = cur - (f / df);
        if (nw < 1e-8) nw = 1e-8;
        it += 1;
    }
    return nw;
}
/* computeStarRegionX/YVelocity (:105-113). */
double orc_exact_star_vel(int dir, double ps, const double *a, const double *b, mpar m1, mpar m2)
{
    return 0.5 * (vel(a, dir) + vel(b, dir) + ex_f(ps, b, m2) - ex_f(ps, a, m1));
}
/* computeStarRegionShockDensity / RarefactionDensity / StarRegionDensity (:68-103, 236-260). */
static double ex_shock_rho(double ps, const double *w, mpar m)
{
    double g = m->gamma, pi = m->p_inf, q = (ps + pi) / (w[4] + pi), k = (g - 1.0) / (g + 1.0);
    double num = q + k;
    double den = k * q + 1.0;
    return w[0] * (num / den);
}
static double ex_rare_rho(double ps, const double *w, mpar m)
{
    return w[0] * pow((ps + m->p_inf) / (w[4] + m->p_inf), 1.0 / m->gamma);
}
double orc_exact_star_rho(double ps, const double *w, mpar m)
{
    if (ps < w[4]) return ex_rare_rho(ps, w, m);
    return ex_shock_rho(ps, w, m);
}

/* ExactSolver::solveX / solveY (:262-446) sampled at x/t = s.  Inside a
 * rarefaction fan (a branch whose reference code does not compile) the state is
 * Toro's fan solution with the tangential velocities set to zero and lambda = 1. */
void orc_exact_solve(int dir, double s, const double *a, const double *b, mpar m1, mpar m2, double *out)
{
    double ps = orc_exact_star_pres(dir, a, b, m1, m2);
    double us = orc_exact_star_vel(dir, ps, a, b, m1, m2);
    const double *w;
    mpar m;
    double sign;
    if (s < us) { w = a; m = m1; sign = -1.0; } else { w = b; m = m2; sign = 1.0; }
    double rho = w[0], un = vel(w, dir), p = w[4], c = snd(w, m), g = m->gamma, pi = m->p_inf;
    int star = 0;
    if (ps < p) {
        double head = un + sign * c;
        double tail = us + sign * (c * pow((ps + pi) / (p + pi), (g - 1.0) / (2.0 * g)));
        if ((sign < 0 && s < head) || (sign > 0 && s > head)) {
            memcpy(out, w, sizeof(double) * NF);
            return;
        }
        if ((sign < 0 && s < tail) || (sign > 0 && s > tail)) {
            /* inside the fan (Toro eq. 4.56 / 4.63 for a stiffened gas) */
            double cs = -sign * c;   /* +c for the left fan, -c for the right (exactsolver.cpp:299, 335) */
            double k = (2.0 / (g + 1.0)) + (g - 1.0) * ((un - s) / ((g + 1.0) * cs));
            memset(out, 0, sizeof(double) * NF);
            out[0] = rho * pow(k, 2.0 / (g - 1.0));
            out[1 + dir] = 2.0 * (cs + (g - 1.0) * (un / 2.0) + s) / (g + 1.0);
            out[4] = ((p + pi) * pow(k, (2.0 * g) / (g - 1.0))) - pi;
            out[5] = 1.0;
            return;
        }
        star = 1;
    } else {
        double Q = sqrt((ps + ex_B(p, g, pi)) / ex_A(rho, g));
        double shock = un + sign * (Q / rho);
        if ((sign < 0 && s < shock) || (sign > 0 && s > shock)) {
            memcpy(out, w, sizeof(double) * NF);
            return;
        }
        star = 1;
    }
    if (star) {
        memset(out, 0, sizeof(double) * NF);
        out[0] = (ps < p) ? ex_rare_rho(ps, w, m) : ex_shock_rho(ps, w, m);
        out[1 + dir] = us;
        out[4] = ps;
        out[5] = 1.0;
    }
}

/* ---- ghost fill (rgfmsolver.cpp:92-491) ----------------------------------- */

/* Interface Riemann problem between real states a (left/top, material pa) and b
 * (material pb): star pressure, normal velocity, and the two star densities. */
static void interface_state(int riemann, int dir, const double *a, const double *b, mpar pa, mpar pb, double *p,
                            double *u, double *ra, double *rb)
{
    double s[NF];
    if (riemann == 0) {
        orc_hllc_solve(dir, a, b, pa, pb, s);
        *p = s[4];
        *u = s[1 + dir];
        *ra = star_rho_a(dir, a, b, pa, pb);
        *rb = star_rho_b(dir, a, b, pa, pb);
    } else {
        orc_exact_solve(dir, 0.0, a, b, pa, pb, s);
        *p = s[4];
        *u = s[1 + dir];
        *ra = orc_exact_star_rho(*p, a, pa);
        *rb = orc_exact_star_rho(*p, b, pb);
    }
}

static void set_ghost(double *w, int dir, double rho, double u, double p)
{
    w[0] = rho;
    w[1 + dir] = u;
    w[4] = p;
}

/* One line of cells (stride between consecutive cells): the sequential scan of
 * applyRGFMBoundaryConditions*: a sign change phi_k phi_{k-1} <= 0 with
 * 2 < k < n - 3 solves the interface problem from the ORIGINAL states and
 * overwrites 4 cells of each material; `in_ghost` toggles per crossing. */
static void fill_line(int riemann, int dir, int n, size_t stride, const double *phi, size_t pstride,
                      const double *o1, const double *o2, double *n1, double *n2, mpar p1, mpar p2, int in_ghost)
{
    double prev = phi[0];
    for (int k = 1; k < n; k++) {
        double cur = phi[(size_t)k * pstride];
        if ((cur * prev) <= 0.0 && k > 2 && k < (n - 3)) {
            double p, u, ra, rb;
            if (!in_ghost) {
                interface_state(riemann, dir, o1 + (size_t)(k - 1) * stride * NF, o2 + (size_t)k * stride * NF, p1, p2,
                                &p, &u, &ra, &rb);
                for (int j = 0; j < 4; j++) {
                    set_ghost(n1 + (size_t)(k + j - 1) * stride * NF, dir, ra, u, p);
                    set_ghost(n2 + (size_t)(k - j) * stride * NF, dir, rb, u, p);
                }
            } else {
                interface_state(riemann, dir, o2 + (size_t)(k - 1) * stride * NF, o1 + (size_t)k * stride * NF, p2, p1,
                                &p, &u, &ra, &rb);
                for (int j = 0; j < 4; j++) {
                    set_ghost(n2 + (size_t)(k + j - 1) * stride * NF, dir, ra, u, p);
                    set_ghost(n1 + (size_t)(k - j) * stride * NF, dir, rb, u, p);
                }
            }
            in_ghost = !in_ghost;
        }
        prev = cur;
    }
}

void orc_rgfm_ghost_fill(int riemann, int dim, int ny, int nx, mpar p1, mpar p2, double *m1, double *m2,
                         const double *phi)
{
    size_t cells = (size_t)ny * nx;
    double *o1 = malloc(sizeof(double) * NF * cells), *o2 = malloc(sizeof(double) * NF * cells);
    memcpy(o1, m1, sizeof(double) * NF * cells);
    memcpy(o2, m2, sizeof(double) * NF * cells);
    if (dim == 1) {
        fill_line(riemann, 0, nx, 1, phi, 1, o1, o2, m1, m2, p1, p2, 0);
    } else {
        /* rows (x), in_ghost starts true (rgfmsolver.cpp:374-379) */
        for (int i = 0; i < ny; i++)
            fill_line(riemann, 0, nx, 1, phi + (size_t)i * nx, 1, o1 + (size_t)i * nx * NF, o2 + (size_t)i * nx * NF,
                      m1 + (size_t)i * nx * NF, m2 + (size_t)i * nx * NF, p1, p2, 1);
        /* columns (y) read the states after the row pass (:430-431) */
        memcpy(o1, m1, sizeof(double) * NF * cells);
        memcpy(o2, m2, sizeof(double) * NF * cells);
        for (int j = 0; j < nx; j++)
            fill_line(riemann, 1, ny, (size_t)nx, phi + j, (size_t)nx, o1 + (size_t)j * NF, o2 + (size_t)j * NF,
                      m1 + (size_t)j * NF, m2 + (size_t)j * NF, p1, p2, 1);
    }
    free(o1);
    free(o2);
}

/* ---- level set (rgfmsolver.cpp:493-646) ------------------------------------ */

static int mirror(int k, int n) { return (k < 0) ? -1 - k : (k >= n ? 2 * n - 1 - k : k); }

/* One line: second-order one-sided upwind differences over mirror-ordered
 * ghosts, velocity component `dir` of material 1 or 2 by the crossing toggle. */
static void level_set_line(int dir, int n, const double *phi, size_t ps, const double *m1, const double *m2,
                           size_t ms, double h, double dt, double *out, int in_ghost)
{
    double prev = phi[0];
    for (int k = 0; k < n; k++) {
        double c = phi[(size_t)k * ps];
        if ((prev * c) <= 0.0) in_ghost = !in_ghost;
        double v = (!in_ghost ? m1 : m2)[(size_t)k * ms * NF + 1 + dir];
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

void orc_rgfm_level_set(int dim, int dir, int ny, int nx, double h, double dt, const double *m1, const double *m2,
                        const double *phi_in, double *phi_out)
{
    if (dim == 1) {
        level_set_line(0, nx, phi_in, 1, m1, m2, 1, h, dt, phi_out, 0);
    } else if (dir == 0) {
        for (int i = 0; i < ny; i++)
            level_set_line(0, nx, phi_in + (size_t)i * nx, 1, m1 + (size_t)i * nx * NF, m2 + (size_t)i * nx * NF, 1, h,
                           dt, phi_out + (size_t)i * nx, 1);
    } else {
        for (int j = 0; j < nx; j++)
            level_set_line(1, ny, phi_in + j, (size_t)nx, m1 + (size_t)j * NF, m2 + (size_t)j * NF, (size_t)nx, h, dt,
                           phi_out + j, 1);
    }
}

/* ---- driver (rgfmsolver.cpp:648-840) --------------------------------------- */

int orc_rgfm_solve(int riemann, int dim, int ny, int nx, double dx, double dy, double cfl, double t_final,
                   int max_steps, double bias, int limiter, const orc_euler_params *p1, const orc_euler_params *p2,
                   const double *m1_in, const double *m2_in, const double *phi_in, double *m1_out, double *m2_out,
                   double *phi_out, double *dt_log, int dt_cap)
{
    if (dim == 1) ny = 1;
    size_t cells = (size_t)ny * nx;
    orc_phys ph1 = orc_euler_phys_table(p1), ph2 = orc_euler_phys_table(p2);
    double *a1 = malloc(sizeof(double) * NF * cells), *b1 = malloc(sizeof(double) * NF * cells);
    double *a2 = malloc(sizeof(double) * NF * cells), *b2 = malloc(sizeof(double) * NF * cells);
    double *phi = malloc(sizeof(double) * cells), *phi2 = malloc(sizeof(double) * cells);
    memcpy(a1, m1_in, sizeof(double) * NF * cells);
    memcpy(a2, m2_in, sizeof(double) * NF * cells);
    memcpy(phi, phi_in, sizeof(double) * cells);
    double h = (dim == 1) ? dx : orc_smin(dx, dy);
    double t = 0.0;
    int it = 0;
    while (t < t_final && (max_steps == 0 || it < max_steps)) {
        orc_rgfm_ghost_fill(riemann, dim, ny, nx, p1, p2, a1, a2, phi);
        double d1 = orc_stable_time_step(orc_ghosted_max(&ph1, dim, ny, nx, 0, a1), h, cfl, t, t_final, it);
        double d2 = orc_stable_time_step(orc_ghosted_max(&ph2, dim, ny, nx, 0, a2), h, cfl, t, t_final, it);
        double dt = orc_smin(d1, d2);
        if (dim == 1) {
            orc_rgfm_level_set(1, 0, 1, nx, dx, dt, a1, a2, phi, phi2);
            memcpy(phi, phi2, sizeof(double) * cells);
            orc_sweep(&ph1, 2, 1, 0, 1, nx, dx, dt, bias, limiter, 0, a1, b1);
            orc_sweep(&ph2, 2, 1, 0, 1, nx, dx, dt, bias, limiter, 0, a2, b2);
            memcpy(a1, b1, sizeof(double) * NF * cells);
            memcpy(a2, b2, sizeof(double) * NF * cells);
        } else {
            double half = 0.5 * dt;
            orc_rgfm_level_set(2, 0, ny, nx, dx, half, a1, a2, phi, phi2);
            orc_rgfm_level_set(2, 1, ny, nx, dy, dt, a1, a2, phi2, phi);
            orc_rgfm_level_set(2, 0, ny, nx, dx, half, a1, a2, phi, phi2);
            memcpy(phi, phi2, sizeof(double) * cells);
            orc_sweep(&ph1, 2, 2, 0, ny, nx, dx, half, bias, limiter, 0, a1, b1);
            orc_sweep(&ph1, 2, 2, 1, ny, nx, dy, dt, bias, limiter, 0, b1, a1);
            orc_sweep(&ph1, 2, 2, 0, ny, nx, dx, half, bias, limiter, 0, a1, b1);
            orc_sweep(&ph2, 2, 2, 0, ny, nx, dx, half, bias, limiter, 0, a2, b2);
            orc_sweep(&ph2, 2, 2, 1, ny, nx, dy, dt, bias, limiter, 0, b2, a2);
            orc_sweep(&ph2, 2, 2, 0, ny, nx, dx, half, bias, limiter, 0, a2, b2);
            memcpy(a1, b1, sizeof(double) * NF * cells);
            memcpy(a2, b2, sizeof(double) * NF * cells);
        }
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

/* Per-function HLLC records: two states (12 doubles) -> 24 doubles, laid out as
 * the reference harness's hllc_units mode writes them. */
void orc_hllc_units(int n, const orc_euler_params *p1, const orc_euler_params *p2, const double *states, double *out)
{
    for (int r = 0; r < n; r++) {
        const double *a = states + 12 * (size_t)r, *b = a + 6;
        double *o = out + 24 * (size_t)r;
        for (int dir = 0; dir < 2; dir++) {
            double *q = o + 12 * dir;
            orc_hllc_solve(dir, a, b, p1, p2, q);
            q[6] = star_rho_a(dir, a, b, p1, p2);
            q[7] = star_rho_b(dir, a, b, p1, p2);
            q[8] = orc_hllc_star_vel(dir, a, b, p1, p2);
            q[9] = orc_hllc_star_pres(dir, a, b, p1, p2);
            q[10] = wave_a(dir, a, b, p1, p2);
            q[11] = wave_b(dir, a, b, p1, p2);
        }
    }
}

/* Exact-solver records: two states -> p*, u*, rho*_left, rho*_right, and the
 * sampled state at x/t = s (6). */
void orc_exact_units(int n, int dir, double s, const orc_euler_params *p1, const orc_euler_params *p2,
                     const double *states, double *out)
{
    for (int r = 0; r < n; r++) {
        const double *a = states + 12 * (size_t)r, *b = a + 6;
        double *o = out + 10 * (size_t)r;
        o[0] = orc_exact_star_pres(dir, a, b, p1, p2);
        o[1] = orc_exact_star_vel(dir, o[0], a, b, p1, p2);
        o[2] = orc_exact_star_rho(o[0], a, p1);
        o[3] = orc_exact_star_rho(o[0], b, p2);
        orc_exact_solve(dir, s, a, b, p1, p2, o + 4);
    }
}

/* Interface circulation: sum over cells with phi < 0 (interior cells only) of
 * (dv/dx - du/dy) dx dy by central differences of material 1's velocity. */
double orc_rgfm_circulation(int ny, int nx, double dx, double dy, const double *m1, const double *phi)
{
    double g = 0.0;
    for (int i = 1; i < ny - 1; i++)
        for (int j = 1; j < nx - 1; j++) {
            size_t c = (size_t)i * nx + j;
            if (!(phi[c] < 0.0)) continue;
            double dvdx = (m1[(c + 1) * NF + 2] - m1[(c - 1) * NF + 2]) / (2.0 * dx);
            double dudy = (m1[(c + nx) * NF + 1] - m1[(c - nx) * NF + 1]) / (2.0 * dy);
            g += (dvdx - dudy) * (dx * dy);
        }
    return g;
}
