/* Euler physics of the oracle -- TEST INFRASTRUCTURE ONLY.
 *
 * EulerStateVector and EulerEquationOfState restated in the reference's
 * operation order, plugged into the generic scheme (scheme_oracle.c).  Pinned bit
 * for bit against the compiled reference by tests/test_oracle_euler.py.
 */
#include "oracle.h"
#include "scheme_oracle.h"

#include <math.h>
#include <string.h>

#define NF 6

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
static void prim_to_cons(const orc_phys *ph, const double *w, double *U)
{
    const orc_euler_params *mp = (const orc_euler_params *)ph->params;
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
static void cons_to_prim(const orc_phys *ph, const double *U, double *w)
{
    const orc_euler_params *mp = (const orc_euler_params *)ph->params;
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
static void flux(const orc_phys *ph, int dir, const double *U, double *F)
{
    const orc_euler_params *mp = (const orc_euler_params *)ph->params;
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

/* Wave speeds of Solvers::computeMaximumWaveSpeed (|u| + c, solvers.cpp:267) and
 * computeMaximumWaveSpeed2D (max(|u|, |v|) + c, solvers.cpp:329). */
static double wave_speed(const orc_phys *ph, const double *w, int dim)
{
    const orc_euler_params *mp = (const orc_euler_params *)ph->params;
    double a = (dim == 1) ? fabs(w[1]) : orc_smax(fabs(w[1]), fabs(w[2]));
    return a + eos_c(w[0], w[4], mp);
}

/* EulerStateVector default constructor (Euler/eulerstatevector.cpp:3-12). */
static const double DEFAULT_STATE[NF] = {1.0, 0.0, 0.0, 0.0, 1.0, 1.0};

static orc_phys euler_phys(const orc_euler_params *mp)
{
    orc_phys ph;
    memset(&ph, 0, sizeof(ph));
    ph.nf = NF;
    ph.prim_to_cons = prim_to_cons;
    ph.cons_to_prim = cons_to_prim;
    ph.flux = flux;
    ph.wave_speed = wave_speed;
    ph.default_state = DEFAULT_STATE;
    ph.params = (void *)mp;
    return ph;
}

double orc_euler_max_wave_speed(int dim, int ny, int nx, const orc_euler_params *mp, const double *prim)
{
    orc_phys ph = euler_phys(mp);
    return orc_ghosted_max(&ph, dim, ny, nx, 1, prim);   /* bc 1: no corner candidate */
}

void orc_euler_sweep(int order, int dim, int dir, int ny, int nx, double dx, double ts, double bias,
                     int limiter, int bc, const orc_euler_params *mp, const double *prim_in,
                     double *prim_out)
{
    orc_phys ph = euler_phys(mp);
    orc_sweep(&ph, order, dim, dir, ny, nx, dx, ts, bias, limiter, bc, prim_in, prim_out);
}

int orc_euler_solve(int order, int dim, int ny, int nx, double dx, double cfl, double t_final,
                    int max_steps, double bias, int limiter, int bc, const orc_euler_params *mp,
                    const double *prim_in, double *prim_out, double *dt_log, int dt_cap)
{
    orc_phys ph = euler_phys(mp);
    return orc_solve(&ph, order, dim, ny, nx, dx, cfl, t_final, max_steps, bias, limiter, bc, prim_in, prim_out,
                     dt_log, dt_cap);
}

void orc_euler_units(int n, double bias, int limiter, double dx, double ts, const orc_euler_params *mp,
                     const double *states, double *out)
{
    orc_phys ph = euler_phys(mp);
    for (int r = 0; r < n; r++) {
        const double *a = states + 30 * (size_t)r, *b = a + 6, *c = a + 12, *d = a + 18;
        double *o = out + 92 * (size_t)r;
        double Ua[NF], Ub[NF], Uc[NF];
        prim_to_cons(&ph, a, Ua);
        prim_to_cons(&ph, b, Ub);
        prim_to_cons(&ph, c, Uc);
        orc_slope_vector(NF, Ua, Ub, Uc, bias, limiter, o); o += 6;
        flux(&ph, 0, Ub, o); o += 6;
        flux(&ph, 1, Ub, o); o += 6;
        orc_evolve_half(&ph, 0, a, b, c, dx, ts, bias, limiter, 0, o); o += 6;
        orc_evolve_half(&ph, 0, a, b, c, dx, ts, bias, limiter, 1, o); o += 6;
        orc_evolve_half(&ph, 1, a, b, c, dx, ts, bias, limiter, 0, o); o += 6;
        orc_evolve_half(&ph, 1, a, b, c, dx, ts, bias, limiter, 1, o); o += 6;
        orc_force_flux(&ph, 0, b, c, dx, ts, o); o += 6;
        orc_force_flux(&ph, 1, b, c, dx, ts, o); o += 6;
        orc_slic_flux(&ph, 0, a, b, c, d, dx, ts, bias, limiter, o); o += 6;
        orc_slic_flux(&ph, 1, a, b, c, d, dx, ts, bias, limiter, o); o += 6;
        {
            double s1 = wave_speed(&ph, b, 1), s2 = wave_speed(&ph, b, 2);
            *o++ = (s1 > 0.0) ? s1 : 0.0;
            *o++ = (s2 > 0.0) ? s2 : 0.0;
        }
        memcpy(o, Ub, sizeof(Ub)); o += 6;
        cons_to_prim(&ph, Ub, o); o += 6;
        {
            double Fb[NF], Fc[NF], us[NF];
            flux(&ph, 0, Ub, Fb);
            flux(&ph, 0, Uc, Fc);
            orc_lax_friedrichs(NF, Ub, Uc, Fb, Fc, dx, ts, o); o += 6;
            orc_richtmyer_state(NF, Ub, Uc, Fb, Fc, dx, ts, us);
            flux(&ph, 0, us, o); o += 6;
        }
    }
}

void orc_limiter_units(int n, double bias, const double *r, double *out)
{
    for (int k = 0; k < n; k++) {
        out[5 * (size_t)k + 0] = orc_limiter_value(r[k], bias, 0);
        out[5 * (size_t)k + 1] = orc_limiter_value(r[k], bias, 1);
        out[5 * (size_t)k + 2] = orc_limiter_value(r[k], bias, 2);
        out[5 * (size_t)k + 3] = orc_r_coefficient(r[k], bias);
        out[5 * (size_t)k + 4] = orc_limiter_value(r[k], bias, 7);
    }
}

/* Entry points for the RGFM oracle (rgfm_oracle.c). */
orc_phys orc_euler_phys_table(const orc_euler_params *mp) { return euler_phys(mp); }
double orc_euler_sound_speed(double rho, double p, const orc_euler_params *mp) { return eos_c(rho, p, mp); }
