/* Ideal MHD with GLM divergence cleaning -- TEST INFRASTRUCTURE ONLY (oracle).
 *
 * MHDStateVector, MHDEquationOfState and MHDWaveSpeeds restated in the
 * reference's operation order, on the generic scheme (scheme_oracle.c).  The
 * solve loop sets the GLM speed c_h to the step's maximum wave speed
 * (MHDSecondOrderSolver::solve, mhdsecondordersolver.cpp:330-332).
 * Pinned bit for bit against the compiled reference by tests/test_oracle_mhd.py.
 */
#include "oracle.h"
#include "scheme_oracle.h"

#include <math.h>
#include <string.h>

#define NF 9

/* MHD/mhdequationofstate.cpp (same stiffened-gas forms as Euler) */
static double eos_e(double rho, double p, const orc_mhd_params *mp)
{
    return (p + (mp->gamma * mp->p_inf)) / ((mp->gamma - 1) * rho);
}

static double eos_p(double rho, double e, const orc_mhd_params *mp)
{
    return (e * (mp->gamma - 1) * rho) - (mp->gamma * mp->p_inf);
}

static double eos_c(double rho, double p, const orc_mhd_params *mp)
{
    return sqrt((mp->gamma * (p + mp->p_inf)) / rho);
}

/* VectorAlgebra::computeDotProduct (Mathematics/vectoralgebra.cpp:38-49): the sum
 * starts from 0.0, which matters for the sign of a zero result. */
static double dot3(double a0, double a1, double a2, double b0, double b1, double b2)
{
    double d = 0.0;
    d += a0 * b0;
    d += a1 * b1;
    d += a2 * b2;
    return d;
}

/* MHDStateVector::computeConservedVariableVector + computeTotalEnergy
 * (MHD/mhdstatevector.cpp:89-106, 245-252). */
static void prim_to_cons(const orc_phys *ph, const double *w, double *U)
{
    const orc_mhd_params *mp = (const orc_mhd_params *)ph->params;
    double rho = w[0], u = w[1], v = w[2], z = w[3], p = w[4];
    double vsq = (u * u) + (v * v) + (z * z);
    double hydro = rho * ((0.5 * vsq) + eos_e(rho, p, mp));
    double b2 = (w[5] * w[5]) + (w[6] * w[6]) + (w[7] * w[7]);
    U[0] = rho;
    U[1] = rho * u;
    U[2] = rho * v;
    U[3] = rho * z;
    U[4] = hydro + (0.5 * b2);
    U[5] = w[5];
    U[6] = w[6];
    U[7] = w[7];
    U[8] = w[8];
}

/* MHDStateVector::setConservedVariableVector (MHD/mhdstatevector.cpp:49-68). */
static void cons_to_prim(const orc_phys *ph, const double *U, double *w)
{
    const orc_mhd_params *mp = (const orc_mhd_params *)ph->params;
    double rho = U[0];
    double u = U[1] / rho, v = U[2] / rho, z = U[3] / rho;
    double bx = U[5], by = U[6], bz = U[7];
    double b2 = (bx * bx) + (by * by) + (bz * bz);
    double et = (U[4] - (0.5 * b2)) / rho;
    double vsq = (u * u) + (v * v) + (z * z);
    double e = et - (0.5 * vsq);
    w[0] = rho;
    w[1] = u;
    w[2] = v;
    w[3] = z;
    w[4] = eos_p(rho, e, mp);
    w[5] = bx;
    w[6] = by;
    w[7] = bz;
    w[8] = U[8];
}

/* MHDStateVector::computeXFluxVector / computeYFluxVector
 * (MHD/mhdstatevector.cpp:108-158, 165-215). */
static void flux(const orc_phys *ph, int dir, const double *U, double *F)
{
    const orc_mhd_params *mp = (const orc_mhd_params *)ph->params;
    double rho = U[0];
    double u = U[1] / rho, v = U[2] / rho, z = U[3] / rho;
    double bx = U[5], by = U[6], bz = U[7], psi = U[8];
    double b2 = (bx * bx) + (by * by) + (bz * bz);
    double E = U[4];
    double het = (E - (0.5 * b2)) / rho;
    double vsq = (u * u) + (v * v) + (z * z);
    double e = het - (0.5 * vsq);
    double p = eos_p(rho, e, mp);
    double ch2 = mp->ch * mp->ch;
    double ub = dot3(u, v, z, bx, by, bz);
    if (dir == 0) {
        F[0] = rho * u;
        F[1] = (rho * (u * u)) + p + (0.5 * b2) - (bx * bx);
        F[2] = (rho * (u * v)) - (bx * by);
        F[3] = (rho * (u * z)) - (bx * bz);
        F[4] = ((E + p + (0.5 * b2)) * u) - (ub * bx);
        F[5] = psi;
        F[6] = (by * u) - (bx * v);
        F[7] = (bz * u) - (bx * z);
        F[8] = ch2 * bx;
    } else {
        F[0] = rho * v;
        F[1] = (rho * (v * u)) - (by * bx);
        F[2] = (rho * (v * v)) + p + (0.5 * b2) - (by * by);
        F[3] = (rho * (v * z)) - (by * bz);
        F[4] = ((E + p + (0.5 * b2)) * v) - (ub * by);
        F[5] = (bx * v) - (by * u);
        F[6] = psi;
        F[7] = (bz * v) - (by * z);
        F[8] = ch2 * by;
    }
}

/* MHDWaveSpeeds::computeAlfvenWaveSpeed and computeX/YFastMagnetoAcousticSpeed
 * (MHD/mhdwavespeeds.cpp:7-15, 41-63). */
static double alfven(double rho, double bx, double by, double bz)
{
    return sqrt(dot3(bx, by, bz, bx, by, bz)) / sqrt(rho);
}

static double fast_speed(const orc_mhd_params *mp, const double *w, int dir)
{
    double rho = w[0], p = w[4];
    double c = eos_c(rho, p, mp);
    double a = alfven(rho, w[5], w[6], w[7]);
    double c2 = (c * c), a2 = (a * a);
    double bn = (dir == 0) ? w[5] : w[6];
    return sqrt(0.5 * (c2 + a2 + sqrt(((c2 + a2) * (c2 + a2)) - 4.0 * ((c2 * (bn * bn)) / rho))));
}

/* MHDSolvers::computeMaximumWaveSpeed (|u| + |c_f,x|, mhdsolvers.cpp:286-303) and
 * computeMaximumWaveSpeed2D (max(|u|,|v|) + max(|c_f,x|,|c_f,y|), :365-387). */
static double wave_speed(const orc_phys *ph, const double *w, int dim)
{
    const orc_mhd_params *mp = (const orc_mhd_params *)ph->params;
    if (dim == 1) return fabs(w[1]) + fabs(fast_speed(mp, w, 0));
    return orc_smax(fabs(w[1]), fabs(w[2])) + orc_smax(fabs(fast_speed(mp, w, 0)), fabs(fast_speed(mp, w, 1)));
}

/* setHyperbolicWaveSpeed(max) + configureParabolicDamping
 * (mhdsecondordersolver.cpp:330-332, mhdmaterialparameters.cpp:31-34). */
static void begin_step(orc_phys *ph, double smax)
{
    orc_mhd_params *mp = (orc_mhd_params *)ph->params;
    mp->ch = smax;
    mp->cp = sqrt(mp->ch * 0.18);
}

/* MHDStateVector::computeSourceTermVector (MHD/mhdstatevector.cpp:222-234): the
 * GLM damping of psi, -(c_h^2 / c_p^2) psi; every other slot 0. */
static void source(const orc_phys *ph, const double *U, double *S)
{
    const orc_mhd_params *mp = (const orc_mhd_params *)ph->params;
    for (int k = 0; k < NF; k++) S[k] = 0.0;
    double chs = mp->ch * mp->ch, cps = mp->cp * mp->cp;
    S[8] = -(chs / cps) * U[8];
}

/* MHDStateVector default constructor (MHD/mhdstatevector.cpp:3-16). */
static const double DEFAULT_STATE[NF] = {1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 0.0};

static orc_phys mhd_phys(orc_mhd_params *mp)
{
    orc_phys ph;
    memset(&ph, 0, sizeof(ph));
    ph.nf = NF;
    ph.prim_to_cons = prim_to_cons;
    ph.cons_to_prim = cons_to_prim;
    ph.flux = flux;
    ph.wave_speed = wave_speed;
    ph.default_state = DEFAULT_STATE;
    ph.begin_step = begin_step;
    ph.source = source;
    ph.params = mp;
    return ph;
}

int orc_mhd_solve(int order, int dim, int ny, int nx, double dx, double cfl, double t_final, int max_steps,
                  double bias, int limiter, int bc, int subcycles, const orc_mhd_params *mp, const double *prim_in,
                  double *prim_out, double *dt_log, int dt_cap)
{
    orc_mhd_params local = *mp;   /* solve takes the parameters by value and mutates its copy */
    orc_phys ph = mhd_phys(&local);
    return orc_solve_sub(&ph, order, dim, ny, nx, dx, cfl, t_final, max_steps, bias, limiter, bc, subcycles, prim_in,
                         prim_out, dt_log, dt_cap);
}

/* One sweep with the parameters as given (the per-sweep calls
 * MHDSecondOrderSolver::computeSLICTimeStep / computeX|YSLICTimeStep2D,
 * mhdsecondordersolver.cpp:91-108, and the FORCE twins: c_h and c_p are the
 * caller's, not the step's maximum). */
void orc_mhd_sweep(int order, int dim, int dir, int ny, int nx, double dx, double ts, double bias, int limiter, int bc,
                   const orc_mhd_params *mp, const double *prim_in, double *prim_out)
{
    orc_mhd_params local = *mp;
    orc_phys ph = mhd_phys(&local);
    orc_sweep(&ph, order, dim, dir, ny, nx, dx, ts, bias, limiter, bc, prim_in, prim_out);
}

void orc_mhd_units(int n, double bias, int limiter, double dx, double ts, const orc_mhd_params *mp,
                   const double *states, double *out)
{
    orc_mhd_params local = *mp;
    orc_phys ph = mhd_phys(&local);
    for (int r = 0; r < n; r++) {
        const double *a = states + 45 * (size_t)r, *b = a + 9, *c = a + 18, *d = a + 27;
        double *o = out + 85 * (size_t)r;
        double Ua[NF], Ub[NF], Uc[NF];
        prim_to_cons(&ph, a, Ua);
        prim_to_cons(&ph, b, Ub);
        prim_to_cons(&ph, c, Uc);
        orc_slope_vector(NF, Ua, Ub, Uc, bias, limiter, o); o += 9;
        flux(&ph, 0, Ub, o); o += 9;
        flux(&ph, 1, Ub, o); o += 9;
        orc_evolve_half(&ph, 0, a, b, c, dx, ts, bias, limiter, 0, o); o += 9;
        orc_evolve_half(&ph, 0, a, b, c, dx, ts, bias, limiter, 1, o); o += 9;
        orc_force_flux(&ph, 0, b, c, dx, ts, o); o += 9;
        orc_slic_flux(&ph, 0, a, b, c, d, dx, ts, bias, limiter, o); o += 9;
        {
            double s1 = wave_speed(&ph, b, 1), s2 = wave_speed(&ph, b, 2);
            *o++ = (s1 > 0.0) ? s1 : 0.0;
            *o++ = (s2 > 0.0) ? s2 : 0.0;
        }
        memcpy(o, Ub, sizeof(Ub)); o += 9;
        cons_to_prim(&ph, Ub, o); o += 9;
        *o++ = fast_speed(&local, b, 0);
        *o++ = alfven(b[0], b[5], b[6], b[7]);
    }
}

/* Exposed for the MHD RGFM oracle (mhd_rgfm_oracle.c). */
orc_phys orc_mhd_phys_table(orc_mhd_params *mp) { return mhd_phys(mp); }
double orc_mhd_fast_speed(const orc_mhd_params *mp, const double *w, int dir) { return fast_speed(mp, w, dir); }
