/* Godunov-Peshkov-Romenski (HPR) continuum -- TEST INFRASTRUCTURE ONLY (oracle).
 *
 * HPRStateVector, HPREquationOfState, HPRMieGruneisen (stiffened-gas branch),
 * HPRDerivatives, HPRSourceTerms, HPRAcousticTensor and the Mathematics/ matrix
 * helpers restated in the reference's operation order, on the generic scheme
 * (scheme_oracle.c), plus the RK4 source sub-cycles of HPRForcingSolver.
 *
 * Layout (17 doubles).  Primitive: rho, p, A (3x3 row-major), u, v, w, J (3).
 * Conserved: rho, rho E, A, rho u, rho v, rho w, rho J.  Without heat
 * conduction the reference leaves the conserved J slots at zero and never
 * touches a cell's J (hprstatevector.cpp:59-76, 78-103); here those slots carry
 * J itself with a zero flux, which reproduces that behaviour exactly.
 *
 * Pinned against the compiled reference by tests/test_oracle_hpr.py: bit for bit
 * without sources; the stiff source uses pow(det A, 5/3) (hprsourceterms.cpp:21),
 * so runs with sub-cycles are pinned to a tolerance (glibc pow vs. the GPU's).
 */
#include "oracle.h"
#include "scheme_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define NF 17

typedef double mat3[3][3];

static void load_A(const double *src, mat3 A)
{
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) A[i][j] = src[2 + 3 * i + j];
}

/* MatrixAlgebra::multiplyMatrices (matrixalgebra.cpp:62-79): sums start at 0.0. */
static void matmul(const mat3 a, const mat3 b, mat3 c)
{
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double s = 0.0;
            for (int k = 0; k < 3; k++) s += a[i][k] * b[k][j];
            c[i][j] = s;
        }
}

static void transpose(const mat3 a, mat3 t)
{
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) t[j][i] = a[i][j];
}

/* MatrixAlgebra::computeDeterminant (matrixalgebra.cpp:130-134). */
static double det3(const mat3 m)
{
    return (m[0][0] * ((m[1][1] * m[2][2]) - (m[2][1] * m[1][2]))) - (m[0][1] * ((m[1][0] * m[2][2]) - (m[1][2] * m[2][0]))) +
           (m[0][2] * ((m[1][0] * m[2][1]) - (m[1][1] * m[2][0])));
}

/* TensorAlgebra::computeGramianMatrix (A^T A) and computeDeviator
 * (tensoralgebra.cpp:22-30): G - (tr G / 3) I via subtractMatrices. */
static void gram_dev(const mat3 A, mat3 dev)
{
    mat3 At, G;
    transpose(A, At);
    matmul(At, A, G);
    double tr = 0.0;
    for (int i = 0; i < 3; i++) tr += G[i][i];
    double s = tr / 3.0;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) dev[i][j] = G[i][j] + (-1.0 * (s * (i == j ? 1.0 : 0.0)));
}

static const orc_hpr_params *P_(const orc_phys *ph) { return (const orc_hpr_params *)ph->params; }

/* HPRMieGruneisen (hprmiegruneisen.cpp:7-263): the Grueneisen coefficient, the
 * reference pressure and reference internal energy of the three equations of
 * state, and their density derivatives.  StiffenedGas: Gamma = gamma - 1,
 * p_ref = -p_inf, e_ref = p_inf / rho.  The reference's
 * computeGruneisenCoefficientDerivative has no GodunovRomenski branch (it ends
 * without a return, :110-125); Gamma = gamma is constant there, so its
 * derivative is 0 here. */
static double mg_gamma_of(const orc_hpr_params *m, double rho)
{
    if (m->eos == 1) return m->gamma0 * (m->rho0 / rho);
    if (m->eos == 2) return m->gamma;
    return m->gamma - 1.0;
}
static double mg_pref(const orc_hpr_params *m, double rho)
{
    if (m->eos == 1) {
        double c2 = m->c0 * m->c0;
        if (rho > m->rho0)
            return c2 * (((1.0 / m->rho0) - (1.0 / rho)) /
                         pow(((1.0 / m->rho0) - (m->s_h * ((1.0 / m->rho0) - (1.0 / rho)))), 2.0));
        return c2 * (rho - m->rho0);
    }
    if (m->eos == 2) {
        double c2 = m->c0 * m->c0, co = pow((rho / m->rho0), m->alpha_gr);
        return c2 * (rho / m->alpha_gr) * (co - 1.0) * co;
    }
    return -m->p_inf;
}
static double mg_eref(const orc_hpr_params *m, double rho)
{
    if (m->eos == 1) {
        if (rho > m->rho0) return 0.5 * m->p_ref * ((1.0 / m->rho0) - (1.0 / rho));
        return 0.0;
    }
    if (m->eos == 2) {
        double c2 = m->c0 * m->c0, co = pow((rho / m->rho0), m->alpha_gr);
        return (c2 / (2.0 * (m->alpha_gr * m->alpha_gr))) * ((co - 1.0) * (co - 1.0));
    }
    return m->p_inf / rho;
}
static double mg_gamma_d(const orc_hpr_params *m, double rho)
{
    if (m->eos == 1) return -m->gamma0 * (m->rho0 / (rho * rho));
    return 0.0;
}
static double mg_pref_d(const orc_hpr_params *m, double rho)
{
    if (m->eos == 1) {
        double c2 = m->c0 * m->c0;
        if (rho > m->rho0)
            return c2 * (m->rho0 * m->rho0) * (((m->s_h * (m->rho0 - rho)) - rho) / pow((m->s_h * (rho - m->rho0)) - rho, 3.0));
        return c2;
    }
    if (m->eos == 2) {
        double c2 = m->c0 * m->c0, co = pow((rho / m->rho0), m->alpha_gr);
        return (c2 / m->alpha_gr) * co * (((1.0 + m->alpha_gr) * (co - 1.0)) + (m->alpha_gr * co));
    }
    return 0.0;
}
static double mg_eref_d(const orc_hpr_params *m, double rho)
{
    if (m->eos == 1) {
        double c2 = m->c0 * m->c0;
        if (rho > m->rho0) return -((rho - m->rho0) * m->rho0 * c2) / pow((m->s_h * (rho - m->rho0)) - rho, 3.0);
        return 0.0;
    }
    if (m->eos == 2) {
        double c2 = m->c0 * m->c0, co = pow((rho / m->rho0), m->alpha_gr);
        return (c2 / (rho * m->alpha_gr)) * (co - 1.0) * co;
    }
    return -(m->p_inf / (rho * rho));
}
/* computeInternalEnergy, computePressure, computeTemperature and the derivatives
 * (hprmiegruneisen.cpp:192-263). */
static double mg_e(const orc_hpr_params *m, double rho, double p)
{
    double G = mg_gamma_of(m, rho), pr = mg_pref(m, rho), er = mg_eref(m, rho);
    return er + ((p - pr) / (rho * G));
}
static double mg_p(const orc_hpr_params *m, double rho, double e)
{
    double G = mg_gamma_of(m, rho), pr = mg_pref(m, rho), er = mg_eref(m, rho);
    return ((e - er) * rho * G) + pr;
}
static double mg_T(const orc_hpr_params *m, double rho, double p)
{
    double G = mg_gamma_of(m, rho), pr = mg_pref(m, rho);
    return m->t_ref + ((p - pr) / (rho * G * m->cv));
}
static double mg_e_rho(const orc_hpr_params *m, double rho, double p)
{
    double G = mg_gamma_of(m, rho), Gd = mg_gamma_d(m, rho), pr = mg_pref(m, rho), prd = mg_pref_d(m, rho),
           erd = mg_eref_d(m, rho);
    return erd - (((prd * rho * G) + ((G + (rho * Gd)) * (p - pr))) / ((rho * G) * (rho * G)));
}
static double mg_e_p(const orc_hpr_params *m, double rho) { return 1.0 / (rho * mg_gamma_of(m, rho)); }
static double mg_T_rho(const orc_hpr_params *m, double rho, double p)
{
    double G = mg_gamma_of(m, rho), Gd = mg_gamma_d(m, rho), pr = mg_pref(m, rho), prd = mg_pref_d(m, rho);
    return -((prd * rho * G) + ((G + (rho * Gd)) * (p - pr))) / (((rho * G) * (rho * G)) / m->cv);
}
static double mg_T_p(const orc_hpr_params *m, double rho) { return mg_e_p(m, rho) / m->cv; }

/* HPREquationOfState energies (hprequationofstate.cpp:7-48). */
static double e_dist(const orc_hpr_params *m, const mat3 A)
{
    mat3 dev;
    gram_dev(A, dev);
    double total = 0.0;
    for (int i = 0; i < 3; i++) {
        double row = 0.0;
        for (int j = 0; j < 3; j++) row += dev[i][j] * dev[i][j];
        total += row;
    }
    return (0.25 * (m->cs * m->cs)) * total;
}
static double e_therm(const orc_hpr_params *m, double jx, double jy, double jz)
{
    return (0.5 * (m->alpha * m->alpha)) * ((jx * jx) + (jy * jy) + (jz * jz));
}
static double e_macro(double u, double v, double w) { return 0.5 * ((u * u) + (v * v) + (w * w)); }

static double total_energy(const orc_hpr_params *m, double rho, double p, double u, double v, double w, const mat3 A,
                           double jx, double jy, double jz)
{
    double E = mg_e(m, rho, p) + e_dist(m, A) + e_macro(u, v, w);
    if (m->is_thermal) E += e_therm(m, jx, jy, jz);
    return E;
}

static double pressure_of(const orc_hpr_params *m, double rho, double E, double u, double v, double w, const mat3 A,
                          double jx, double jy, double jz)
{
    double e = E - (e_macro(u, v, w) + e_dist(m, A));
    if (m->is_thermal) e -= e_therm(m, jx, jy, jz);
    return mg_p(m, rho, e);
}

/* HPRDerivatives::computeTotalEnergyDerivativeDistortionTensor: c_s^2 (A dev(A^T A)). */
static void e_A(const orc_hpr_params *m, const mat3 A, mat3 out)
{
    mat3 dev, Ad;
    gram_dev(A, dev);
    matmul(A, dev, Ad);
    double cs2 = m->cs * m->cs;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) out[i][j] = cs2 * Ad[i][j];
}

/* computeShearStressTensor: -rho A^T E_A (hprequationofstate.cpp:55-60) and its
 * density derivative -1 * A^T E_A (:76-81). */
static void shear(const orc_hpr_params *m, double rho, const mat3 A, mat3 sig, mat3 dsig_rho)
{
    mat3 EA, At, P;
    e_A(m, A, EA);
    transpose(A, At);
    matmul(At, EA, P);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            sig[i][j] = (-rho) * P[i][j];
            if (dsig_rho) dsig_rho[i][j] = -1.0 * P[i][j];
        }
}

/* HPRStateVector::computeConservedVariableVector (hprstatevector.cpp:105-132). */
static void prim_to_cons(const orc_phys *ph, const double *w, double *U)
{
    const orc_hpr_params *m = P_(ph);
    mat3 A;
    load_A(w, A);
    double rho = w[0];
    U[0] = rho;
    U[1] = rho * total_energy(m, rho, w[1], w[11], w[12], w[13], A, w[14], w[15], w[16]);
    for (int k = 2; k < 11; k++) U[k] = w[k];
    U[11] = rho * w[11];
    U[12] = rho * w[12];
    U[13] = rho * w[13];
    if (m->is_thermal) {
        U[14] = rho * w[14];
        U[15] = rho * w[15];
        U[16] = rho * w[16];
    } else {
        U[14] = w[14];
        U[15] = w[15];
        U[16] = w[16];
    }
}

/* HPRStateVector::setConservedVariableVector (hprstatevector.cpp:59-76). */
static void cons_to_prim(const orc_phys *ph, const double *U, double *w)
{
    const orc_hpr_params *m = P_(ph);
    mat3 A;
    load_A(U, A);
    double rho = U[0];
    double u = U[11] / rho, v = U[12] / rho, z = U[13] / rho;
    double jx, jy, jz;
    if (m->is_thermal) {
        jx = U[14] / rho;
        jy = U[15] / rho;
        jz = U[16] / rho;
    } else {
        jx = U[14];
        jy = U[15];
        jz = U[16];
    }
    double E = U[1] / rho;
    w[0] = rho;
    w[1] = pressure_of(m, rho, E, u, v, z, A, jx, jy, jz);
    for (int k = 2; k < 11; k++) w[k] = U[k];
    w[11] = u;
    w[12] = v;
    w[13] = z;
    w[14] = jx;
    w[15] = jy;
    w[16] = jz;
}

/* HPRStateVector::computeX/YFluxVector (hprstatevector.cpp:134-231, 238-305). */
static void flux(const orc_phys *ph, int dir, const double *U, double *F)
{
    const orc_hpr_params *m = P_(ph);
    mat3 A, sig;
    load_A(U, A);
    double rho = U[0];
    shear(m, rho, A, sig, NULL);
    double u = U[11] / rho, v = U[12] / rho, z = U[13] / rho;
    double jx = m->is_thermal ? U[14] / rho : 0.0;
    double jy = m->is_thermal ? U[15] / rho : 0.0;
    double jz = m->is_thermal ? U[16] / rho : 0.0;
    double E = U[1] / rho;
    double p = pressure_of(m, rho, E, u, v, z, A, jx, jy, jz);
    for (int k = 0; k < NF; k++) F[k] = 0.0;
    double vn = (dir == 0) ? u : v;
    F[0] = rho * vn;
    F[1] = (rho * (vn * E)) + (p * vn);
    F[11] = rho * (vn * u);
    F[12] = rho * (vn * v);
    F[13] = rho * (vn * z);
    F[11 + dir] = (rho * (vn * vn)) + p;
    double Av[3];
    for (int i = 0; i < 3; i++) {
        double s = 0.0;
        s += A[i][0] * u;
        s += A[i][1] * v;
        s += A[i][2] * z;
        Av[i] = s;
    }
    double sv = 0.0;
    sv += sig[dir][0] * u;
    sv += sig[dir][1] * v;
    sv += sig[dir][2] * z;
    F[1] -= sv;
    for (int i = 0; i < 3; i++) F[11 + i] -= sig[dir][i];
    for (int i = 0; i < 3; i++) F[2 + 3 * i + dir] = Av[i];
    if (m->is_thermal) {
        double T = mg_T(m, rho, p);
        double a2 = m->alpha * m->alpha;
        double q = T * (a2 * (dir == 0 ? jx : jy));
        F[1] += q;
        F[14] = rho * (vn * jx);
        F[15] = rho * (vn * jy);
        F[16] = rho * (vn * jz);
        F[14 + dir] = (rho * (vn * (dir == 0 ? jx : jy))) + T;
    }
}

/* HPRAcousticTensor::computeMaximumWaveSpeed (hpracoustictensor.cpp:7-97) with
 * MatrixAlgebra::computeLargestEigenvalue (matrixalgebra.cpp:182-199): ten power
 * iterations from (1, 1, 1, 1), then |M v|. */
static double acoustic_speed(const orc_hpr_params *m, const double *w, int d)
{
    mat3 A, dev, Ad, sig, dsr, EA, At, P;
    load_A(w, A);
    double rho = w[0], p = w[1];
    shear(m, rho, A, sig, dsr);
    /* d sigma / d A at [d][i][j][d] (hprequationofstate.cpp:62-74): */
    transpose(A, At);
    mat3 G;
    matmul(At, A, G);
    gram_dev(A, dev);
    matmul(A, dev, Ad);             /* B[j][k] = (A dev)[k][j] */
    (void)EA;
    (void)P;
    double cs2 = m->cs * m->cs;
    double C1[4][5], C2[5][4];
    memset(C1, 0, sizeof(C1));
    memset(C2, 0, sizeof(C2));
    double inv_rho = 1.0 / rho;
    C1[0][1] = inv_rho;
    for (int i = 0; i < 3; i++) C1[i][0] = -(1.0 / rho) * dsr[d][i];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double pa = 0.0 + (G[d][i] * A[j][d]);
            double pb = 0.0 + (G[d][d] * A[j][i]);
            double D = (pa + pb) + (-1.0 * ((2.0 / 3.0) * pa));
            D += Ad[j][i];                       /* B[i][j] */
            if (i == d) D += Ad[j][d];           /* B[d][j], second update of the same entry */
            double dsA = (-(rho * cs2)) * D;
            C1[i][2 + j] = -(1.0 / rho) * dsA;
        }
    if (m->is_thermal) {
        C1[3][0] = mg_T_rho(m, rho, p) / rho;
        C1[3][1] = mg_T_p(m, rho) / rho;
    }
    double ca = sqrt(((p / (rho * rho)) - mg_e_rho(m, rho, p)) / mg_e_p(m, rho));
    C2[0][0] = rho;
    C2[1][d] = rho * (ca * ca);
    for (int i = 0; i < 3; i++) C2[1][i] += sig[d][i] - (rho * dsr[d][i]);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) C2[2 + i][j] = A[i][j];
    if (m->is_thermal) {
        double T = mg_T(m, rho, p);
        double Tp = mg_T_p(m, rho);
        double ch = (sqrt((m->alpha * m->alpha) * (T / m->cv))) / rho;
        C2[1][3] = rho * ((ch * ch) / Tp);
    }
    double M[4][4];
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            double s = 0.0;
            for (int k = 0; k < 5; k++) s += C1[i][k] * C2[k][j];
            M[i][j] = s;
        }
    double v[4] = {1.0, 1.0, 1.0, 1.0};
    for (int it = 0; it < 10; it++) {
        double nv[4];
        for (int i = 0; i < 4; i++) {
            double s = 0.0;
            for (int k = 0; k < 4; k++) s += M[i][k] * v[k];
            nv[i] = s;
        }
        double nn = 0.0;
        for (int i = 0; i < 4; i++) nn += nv[i] * nv[i];
        double c = 1.0 / sqrt(nn);
        for (int i = 0; i < 4; i++) v[i] = c * nv[i];
    }
    double mv[4], nn = 0.0;
    for (int i = 0; i < 4; i++) {
        double s = 0.0;
        for (int k = 0; k < 4; k++) s += M[i][k] * v[k];
        mv[i] = s;
    }
    for (int i = 0; i < 4; i++) nn += mv[i] * mv[i];
    return sqrt(sqrt(nn));
}

/* HPRSolvers::computeMaximumWaveSpeed[2D] (hprsolvers.cpp:286-303, 362-385). */
static double wave_speed(const orc_phys *ph, const double *w, int dim)
{
    const orc_hpr_params *m = P_(ph);
    if (dim == 1) return fabs(w[11]) + acoustic_speed(m, w, 0);
    return orc_smax(fabs(w[11]), fabs(w[12])) + orc_smax(acoustic_speed(m, w, 0), acoustic_speed(m, w, 1));
}

/* HPRStateVector default constructor (hprstatevector.cpp:3-16): rho 1, p 1/1.4,
 * A = rho^(1/3) I = I, u = J = 0. */
static const double DEFAULT_STATE[NF] = {1.0, 1.0 / 1.4, 1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0};

/* HPRStateVector::computeSourceTermVector (hprstatevector.cpp:312-366) with
 * HPRSourceTerms::computeTheta1Reciprocal / computeTheta2Reciprocal
 * (hprsourceterms.cpp:7-39), non-plastic branch. */
/* HPRSourceTerms::computeTheta1Reciprocal (hprsourceterms.cpp:7-30): 3 det(A)^(5/3)
 * / (c_s^2 tau1), times (min(|sigma|, 1e8) / sigma_Y)^n when plastic, with
 * |sigma| TensorAlgebra::computeSigmaNorm (tensoralgebra.cpp:276-283) of the shear
 * stress -rho A^T E_A. */
static double theta1_rcp(const orc_hpr_params *m, double rho, const mat3 A)
{
    double cs2 = m->cs * m->cs;
    double base = 3.0 * (pow(det3(A), (5.0 / 3.0)) / (cs2 * m->tau1));
    if (!m->is_plastic) return base;
    mat3 sig;
    shear(m, rho, A, sig, NULL);
    double c1 = ((sig[0][0] - sig[1][1]) * (sig[0][0] - sig[1][1])) + ((sig[1][1] - sig[2][2]) * (sig[1][1] - sig[2][2])) +
                ((sig[2][2] - sig[0][0]) * (sig[2][2] - sig[0][0]));
    double c2 = (sig[0][1] * sig[0][1]) + (sig[1][2] * sig[1][2]) + (sig[2][0] * sig[2][0]);
    double norm = sqrt((0.5 * c1) + (3.0 * c2));
    norm = orc_smin(norm, pow(10.0, 8.0));
    return base * pow(norm / m->sigma_y, m->n_pow);
}

static void source(const orc_hpr_params *m, const double *U, double *S)
{
    mat3 A, EA;
    load_A(U, A);
    double rho = U[0];
    for (int k = 0; k < NF; k++) S[k] = 0.0;
    e_A(m, A, EA);
    double th1 = theta1_rcp(m, rho, A);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) S[2 + 3 * i + j] = -EA[i][j] * th1;
    if (m->is_thermal) {
        double E = U[1] / rho;
        double u = U[11] / rho, v = U[12] / rho, z = U[13] / rho;
        double jx = U[14] / rho, jy = U[15] / rho, jz = U[16] / rho;
        double p = pressure_of(m, rho, E, u, v, z, A, jx, jy, jz);
        double T = mg_T(m, rho, p);
        double a2 = m->alpha * m->alpha;
        double ej[3] = {a2 * jx, a2 * jy, a2 * jz};
        double th2 = 1.0 / (a2 * m->tau2 * ((rho / m->rho0) * (m->t0 / T)));
        for (int i = 0; i < 3; i++) S[14 + i] = -rho * th2 * ej[i];
    }
}

static void source_of(const orc_phys *ph, const double *U, double *S) { source(P_(ph), U, S); }

/* computeRungeKuttaTimeStep[2D] leaves a cell's J alone without heat conduction. */
static void keep_j(const orc_phys *ph, const double *w_in, double *w_out)
{
    if (!P_(ph)->is_thermal) {
        w_out[14] = w_in[14];
        w_out[15] = w_in[15];
        w_out[16] = w_in[16];
    }
}

static orc_phys hpr_phys(const orc_hpr_params *m)
{
    orc_phys ph;
    memset(&ph, 0, sizeof(ph));
    ph.nf = NF;
    ph.prim_to_cons = prim_to_cons;
    ph.cons_to_prim = cons_to_prim;
    ph.flux = flux;
    ph.wave_speed = wave_speed;
    ph.default_state = DEFAULT_STATE;
    ph.source = source_of;
    ph.forcing_finish = keep_j;
    ph.params = (void *)m;
    return ph;
}

/* HPRSecondOrderSolver::solve / solve2D (hprsecondordersolver.cpp:314-352,
 * 502-546): dt, [n sub-cycles of the source at 0.5 dt / n], SLIC (X in 1D;
 * X(1/2) Y X(1/2) in 2D), [n sub-cycles]; HPRFirstOrderSolver::solve never
 * sub-cycles. */
int orc_hpr_solve(int order, int dim, int ny, int nx, double dx, double cfl, double t_final, int max_steps,
                  double bias, int limiter, int bc, int subcycles, const orc_hpr_params *mp, const double *prim_in,
                  double *prim_out, double *dt_log, int dt_cap)
{
    orc_phys ph = hpr_phys(mp);
    return orc_solve_sub(&ph, order, dim, ny, nx, dx, cfl, t_final, max_steps, bias, limiter, bc, subcycles, prim_in,
                         prim_out, dt_log, dt_cap);
}

/* Per-function records of five input states (85 doubles) -> 17*9 + 3 = 156:
 * cons(b) | prim(cons b) | Fx(cons b) | Fy(cons b) | S(cons b) | evolveX side0 |
 * FORCEx(b,c) | SLICx(a,b,c,d) | forcing 1D cell (a,b,c) | speed_x | speed_y | wave 2D */
void orc_hpr_units(int n, double bias, int limiter, double dx, double ts, const orc_hpr_params *mp,
                   const double *states, double *out)
{
    orc_phys ph = hpr_phys(mp);
    for (int r = 0; r < n; r++) {
        const double *a = states + 85 * (size_t)r, *b = a + 17, *c = a + 34, *d = a + 51;
        double *o = out + 156 * (size_t)r;
        double Ub[NF], Ua[NF], Uc[NF];
        prim_to_cons(&ph, b, Ub);
        prim_to_cons(&ph, a, Ua);
        prim_to_cons(&ph, c, Uc);
        memcpy(o, Ub, sizeof(Ub)); o += NF;
        cons_to_prim(&ph, Ub, o); o += NF;
        flux(&ph, 0, Ub, o); o += NF;
        flux(&ph, 1, Ub, o); o += NF;
        source(mp, Ub, o); o += NF;
        orc_evolve_half(&ph, 0, a, b, c, dx, ts, bias, limiter, 0, o); o += NF;
        orc_force_flux(&ph, 0, b, c, dx, ts, o); o += NF;
        orc_slic_flux(&ph, 0, a, b, c, d, dx, ts, bias, limiter, o); o += NF;
        {
            double N[NF];
            orc_forcing_cell(&ph, 1, Ua, Ub, Uc, NULL, NULL, dx, ts, bias, limiter, N);
            cons_to_prim(&ph, N, o); o += NF;
        }
        *o++ = acoustic_speed(mp, b, 0);
        *o++ = acoustic_speed(mp, b, 1);
        *o++ = wave_speed(&ph, b, 2);
    }
}

void orc_hpr_eos_units(int n, const orc_hpr_params *mp, const double *states, double *out)
{
    for (int r = 0; r < n; r++) {
        const double *st = states + 11 * (size_t)r;
        double *o = out + 8 * (size_t)r;
        double rho = st[0], p = st[1];
        mat3 A;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) A[i][j] = st[2 + 3 * i + j];
        o[0] = mg_e(mp, rho, p);
        o[1] = mg_p(mp, rho, o[0]);
        o[2] = mg_T(mp, rho, p);
        o[3] = mg_e_rho(mp, rho, p);
        o[4] = mg_e_p(mp, rho);
        o[5] = mg_T_rho(mp, rho, p);
        o[6] = mg_T_p(mp, rho);
        o[7] = theta1_rcp(mp, rho, A);
    }
}
