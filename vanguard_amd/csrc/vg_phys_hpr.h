// Godunov-Peshkov-Romenski (HPR) continuum with the Mie-Grueneisen EOS family: the physics of
// the reference's HPRStateVector / HPREquationOfState / HPRMieGruneisen /
// HPRAcousticTensor / HPRSourceTerms on register arrays, in the reference's
// operation order (every matrix product sums from 0.0 as MatrixAlgebra does).
//
// Primitive (17): rho, p, A (3x3 row-major), u, v, w, J (3).
// Conserved (17): rho, rho E, A, rho u, rho v, rho w, rho J.  Without heat
// conduction the reference keeps the conserved J slots at zero and never
// changes a cell's J (hprstatevector.cpp:59-76, 105-132); here those slots carry
// J itself with zero flux, which leaves J unchanged exactly as the reference does.
#pragma once

#include "vg_device.h"
#include "vg_quot.h"

namespace vg {

// HPRMaterialParameters (HPR/hprmaterialparameters.h): the fields the solver
// reads, plus derived constants filled on the host (G = gamma - 1.0,
// cs2 = cs^2, a2 = alpha^2, c02 = c0^2: the same IEEE values the reference
// recomputes).  eos: 0 StiffenedGas, 1 ShockMieGruneisen, 2 GodunovRomenski;
// p_ref, c0, gamma0, s_h, alpha_gr are the reference pressure, reference sound
// speed, reference Grueneisen coefficient, Hugoniot slope and alpha parameter
// those branches read; sigma_y, n_pow the plastic relaxation's
// elastic-plastic transition parameter and power-law index.
struct HPRParams {
    int is_thermal, is_plastic;
    double rho0, t_ref, t0, cv, gamma, p_inf, cs, tau1, tau2, alpha;
    double G, cs2, a2;
    int eos;
    double p_ref, c0, gamma0, s_h, alpha_gr, sigma_y, n_pow, c02;
};

// EOS: the equation of state, fixed per kernel (0 StiffenedGas, 1
// ShockMieGruneisen, 2 GodunovRomenski) so that the stiffened-gas kernels carry
// none of the other curves' code; Params::eos names the same choice at run time.
template <int EOS>
struct HPRPhysT {
    static constexpr int NF = 17;
    static constexpr const char *name = EOS == 0 ? "hpr" : (EOS == 1 ? "hpr_smg" : "hpr_gr");
    using Params = HPRParams;

    __device__ static __forceinline__ void begin_step(Params &, double) {}

    // ---- HPRMieGruneisen (hprmiegruneisen.cpp:7-263) ------------------------------
    // The Grueneisen coefficient, reference pressure and reference internal energy
    // of the three equations of state and their density derivatives.  The
    // reference's computeGruneisenCoefficientDerivative has no GodunovRomenski
    // branch (it ends without a return, :110-125); Gamma = gamma is constant
    // there, so its derivative is 0.
    __device__ static __forceinline__ double mg_gamma_of(const Params &m, double rho)
    {
        if constexpr (EOS == 1) return m.gamma0 * (m.rho0 / rho);
        else if constexpr (EOS == 2) return m.gamma;
        else return m.G;
    }
    __device__ static __forceinline__ double mg_pref(const Params &m, double rho)
    {
        if constexpr (EOS == 1) {
            if (rho > m.rho0)
                return m.c02 * (((1.0 / m.rho0) - (1.0 / rho)) /
                                pow(((1.0 / m.rho0) - (m.s_h * ((1.0 / m.rho0) - (1.0 / rho)))), 2.0));
            return m.c02 * (rho - m.rho0);
        } else if constexpr (EOS == 2) {
            const double co = pow((rho / m.rho0), m.alpha_gr);
            return m.c02 * (rho / m.alpha_gr) * (co - 1.0) * co;
        } else {
            return -m.p_inf;
        }
    }
    __device__ static __forceinline__ double mg_eref(const Params &m, double rho)
    {
        if constexpr (EOS == 1) {
            return (rho > m.rho0) ? 0.5 * m.p_ref * ((1.0 / m.rho0) - (1.0 / rho)) : 0.0;
        } else if constexpr (EOS == 2) {
            const double co = pow((rho / m.rho0), m.alpha_gr);
            return (m.c02 / (2.0 * (m.alpha_gr * m.alpha_gr))) * ((co - 1.0) * (co - 1.0));
        } else {
            return m.p_inf / rho;
        }
    }
    __device__ static __forceinline__ double mg_gamma_d(const Params &m, double rho)
    {
        if constexpr (EOS == 1) return -m.gamma0 * (m.rho0 / (rho * rho));
        else return 0.0;
    }
    __device__ static __forceinline__ double mg_pref_d(const Params &m, double rho)
    {
        if constexpr (EOS == 1) {
            if (rho > m.rho0)
                return m.c02 * (m.rho0 * m.rho0) *
                       (((m.s_h * (m.rho0 - rho)) - rho) / pow((m.s_h * (rho - m.rho0)) - rho, 3.0));
            return m.c02;
        } else if constexpr (EOS == 2) {
            const double co = pow((rho / m.rho0), m.alpha_gr);
            return (m.c02 / m.alpha_gr) * co * (((1.0 + m.alpha_gr) * (co - 1.0)) + (m.alpha_gr * co));
        } else {
            return 0.0;
        }
    }
    __device__ static __forceinline__ double mg_eref_d(const Params &m, double rho)
    {
        if constexpr (EOS == 1) {
            return (rho > m.rho0) ? -((rho - m.rho0) * m.rho0 * m.c02) / pow((m.s_h * (rho - m.rho0)) - rho, 3.0)
                                  : 0.0;
        } else if constexpr (EOS == 2) {
            const double co = pow((rho / m.rho0), m.alpha_gr);
            return (m.c02 / (rho * m.alpha_gr)) * (co - 1.0) * co;
        } else {
            return -(m.p_inf / (rho * rho));
        }
    }
    // computeInternalEnergy / computePressure / computeTemperature and derivatives
    // (hprmiegruneisen.cpp:192-263).  For the stiffened gas these are
    // e_ref + (p - p_ref) / (rho Gamma) with e_ref = p_inf / rho, p_ref = -p_inf.
    __device__ static __forceinline__ double mg_e(const Params &m, double rho, double p)
    {
        return mg_eref(m, rho) + ((p - mg_pref(m, rho)) / (rho * mg_gamma_of(m, rho)));
    }
    __device__ static __forceinline__ double mg_p(const Params &m, double rho, double e)
    {
        return ((e - mg_eref(m, rho)) * rho * mg_gamma_of(m, rho)) + mg_pref(m, rho);
    }
    __device__ static __forceinline__ double mg_T(const Params &m, double rho, double p)
    {
        return m.t_ref + ((p - mg_pref(m, rho)) / (rho * mg_gamma_of(m, rho) * m.cv));
    }
    __device__ static __forceinline__ double mg_e_rho(const Params &m, double rho, double p)
    {
        const double G = mg_gamma_of(m, rho), Gd = mg_gamma_d(m, rho), pr = mg_pref(m, rho), prd = mg_pref_d(m, rho),
                     erd = mg_eref_d(m, rho);
        return erd - (((prd * rho * G) + ((G + (rho * Gd)) * (p - pr))) / ((rho * G) * (rho * G)));
    }
    __device__ static __forceinline__ double mg_e_p(const Params &m, double rho)
    {
        return 1.0 / (rho * mg_gamma_of(m, rho));
    }
    __device__ static __forceinline__ double mg_T_rho(const Params &m, double rho, double p)
    {
        const double G = mg_gamma_of(m, rho), Gd = mg_gamma_d(m, rho), pr = mg_pref(m, rho), prd = mg_pref_d(m, rho);
        return -((prd * rho * G) + ((G + (rho * Gd)) * (p - pr))) / (((rho * G) * (rho * G)) / m.cv);
    }
    __device__ static __forceinline__ double mg_T_p(const Params &m, double rho) { return mg_e_p(m, rho) / m.cv; }

    // ---- 3x3 helpers (Mathematics/matrixalgebra.cpp, tensoralgebra.cpp) ------------
    template <int NN>
    __device__ static __forceinline__ void load_A(const double (&s)[NN], double (&A)[3][3])
    {
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) A[i][j] = s[2 + 3 * i + j];
    }
    // A^T B with sums from 0.0
    __device__ static __forceinline__ void mul_tn(const double (&a)[3][3], const double (&b)[3][3], double (&c)[3][3])
    {
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) {
                double s = 0.0;
                s = s + a[0][i] * b[0][j];
                s = s + a[1][i] * b[1][j];
                s = s + a[2][i] * b[2][j];
                c[i][j] = s;
            }
    }
    __device__ static __forceinline__ void mul_nn(const double (&a)[3][3], const double (&b)[3][3], double (&c)[3][3])
    {
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) {
                double s = 0.0;
                s = s + a[i][0] * b[0][j];
                s = s + a[i][1] * b[1][j];
                s = s + a[i][2] * b[2][j];
                c[i][j] = s;
            }
    }
    // Gramian A^T A and its deviator G - (tr G / 3) I (tensoralgebra.cpp:22-30).
    __device__ static __forceinline__ void gram_dev(const double (&A)[3][3], double (&G)[3][3], double (&dev)[3][3])
    {
        mul_tn(A, A, G);
        double tr = 0.0;
        tr = tr + G[0][0];
        tr = tr + G[1][1];
        tr = tr + G[2][2];
        const double s = tr / 3.0;
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) dev[i][j] = G[i][j] + (-1.0 * (s * (i == j ? 1.0 : 0.0)));
    }
    __device__ static __forceinline__ double det3(const double (&M)[3][3])
    {
        return (M[0][0] * ((M[1][1] * M[2][2]) - (M[2][1] * M[1][2]))) - (M[0][1] * ((M[1][0] * M[2][2]) - (M[1][2] * M[2][0]))) +
               (M[0][2] * ((M[1][0] * M[2][1]) - (M[1][1] * M[2][0])));
    }

    // ---- energies (hprequationofstate.cpp:7-48) ---------------------------------
    __device__ static __forceinline__ double e_dist(const Params &m, const double (&A)[3][3])
    {
        double G[3][3], dev[3][3];
        gram_dev(A, G, dev);
        double total = 0.0;
#pragma unroll
        for (int i = 0; i < 3; i++) {
            double row = 0.0;
            row = row + dev[i][0] * dev[i][0];
            row = row + dev[i][1] * dev[i][1];
            row = row + dev[i][2] * dev[i][2];
            total = total + row;
        }
        return (0.25 * m.cs2) * total;
    }
    __device__ static __forceinline__ double e_therm(const Params &m, double jx, double jy, double jz)
    {
        return (0.5 * m.a2) * ((jx * jx) + (jy * jy) + (jz * jz));
    }
    __device__ static __forceinline__ double e_macro(double u, double v, double w)
    {
        return 0.5 * ((u * u) + (v * v) + (w * w));
    }
    __device__ static __forceinline__ double pressure_of(const Params &m, double rho, double E, double u, double v, double w,
                                                         const double (&A)[3][3], double jx, double jy, double jz)
    {
        double e = E - (e_macro(u, v, w) + e_dist(m, A));
        if (m.is_thermal) e -= e_therm(m, jx, jy, jz);
        return mg_p(m, rho, e);
    }
    // c_s^2 (A dev(A^T A)) (hprderivatives.cpp:17-24) and the shear stress
    // -rho A^T E_A with its density derivative -1 * A^T E_A (hprequationofstate.cpp:55-81).
    __device__ static __forceinline__ void shear(const Params &m, double rho, const double (&A)[3][3], double (&sig)[3][3],
                                                 double (&dsr)[3][3])
    {
        double G[3][3], dev[3][3], Ad[3][3], EA[3][3], P[3][3];
        gram_dev(A, G, dev);
        mul_nn(A, dev, Ad);
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) EA[i][j] = m.cs2 * Ad[i][j];
        mul_tn(A, EA, P);
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) {
                sig[i][j] = (-rho) * P[i][j];
                dsr[i][j] = -1.0 * P[i][j];
            }
    }

    // computeConservedVariableVector (hprstatevector.cpp:105-132).
    __device__ static __forceinline__ void prim_to_cons(const Params &m, const double (&w)[NF], double (&U)[NF])
    {
        double A[3][3];
        load_A(w, A);
        const double rho = w[0];
        double E = mg_e(m, rho, w[1]) + e_dist(m, A) + e_macro(w[11], w[12], w[13]);
        if (m.is_thermal) E += e_therm(m, w[14], w[15], w[16]);
        U[0] = rho;
        U[1] = rho * E;
#pragma unroll
        for (int k = 2; k < 11; k++) U[k] = w[k];
        U[11] = rho * w[11];
        U[12] = rho * w[12];
        U[13] = rho * w[13];
        if (m.is_thermal) {
            U[14] = rho * w[14];
            U[15] = rho * w[15];
            U[16] = rho * w[16];
        } else {
            U[14] = w[14];
            U[15] = w[15];
            U[16] = w[16];
        }
    }

    // The quotients by rho of a conserved vector (u, v, w, J and E), each the
    // correctly rounded IEEE quotient, from one shared reciprocal of rho.
    __device__ static __forceinline__ void by_rho(const double (&U)[NF], double (&q)[7])
    {
        const double a[7] = {U[11], U[12], U[13], U[14], U[15], U[16], U[1]};
        quot_group<7>(a, shared_divisor(U[0]), q);
    }

    // setConservedVariableVector (hprstatevector.cpp:59-76).
    __device__ static __forceinline__ void cons_to_prim(const Params &m, const double (&U)[NF], double (&w)[NF])
    {
        double A[3][3];
        load_A(U, A);
        const double rho = U[0];
        double q[7];
        by_rho(U, q);
        const double u = q[0], v = q[1], z = q[2];
        double jx, jy, jz;
        if (m.is_thermal) {
            jx = q[3];
            jy = q[4];
            jz = q[5];
        } else {
            jx = U[14];
            jy = U[15];
            jz = U[16];
        }
        const double E = q[6];
        w[0] = rho;
        w[1] = pressure_of(m, rho, E, u, v, z, A, jx, jy, jz);
#pragma unroll
        for (int k = 2; k < 11; k++) w[k] = U[k];
        w[11] = u;
        w[12] = v;
        w[13] = z;
        w[14] = jx;
        w[15] = jy;
        w[16] = jz;
    }

    // computeX/YFluxVector (hprstatevector.cpp:134-231, 238-305).
    template <int DIR>
    __device__ static __forceinline__ void flux(const Params &m, const double (&U)[NF], double (&F)[NF])
    {
        double A[3][3], sig[3][3], dsr[3][3];
        load_A(U, A);
        const double rho = U[0];
        shear(m, rho, A, sig, dsr);
        double q[7];
        by_rho(U, q);
        const double u = q[0], v = q[1], z = q[2];
        const double jx = m.is_thermal ? q[3] : 0.0;
        const double jy = m.is_thermal ? q[4] : 0.0;
        const double jz = m.is_thermal ? q[5] : 0.0;
        const double E = q[6];
        const double p = pressure_of(m, rho, E, u, v, z, A, jx, jy, jz);
#pragma unroll
        for (int k = 0; k < NF; k++) F[k] = 0.0;
        const double vn = (DIR == 0) ? u : v;
        F[0] = rho * vn;
        F[1] = (rho * (vn * E)) + (p * vn);
        F[11] = rho * (vn * u);
        F[12] = rho * (vn * v);
        F[13] = rho * (vn * z);
        F[11 + DIR] = (rho * (vn * vn)) + p;
        double sv = 0.0;
        sv = sv + sig[DIR][0] * u;
        sv = sv + sig[DIR][1] * v;
        sv = sv + sig[DIR][2] * z;
        F[1] = F[1] - sv;
#pragma unroll
        for (int i = 0; i < 3; i++) F[11 + i] = F[11 + i] - sig[DIR][i];
#pragma unroll
        for (int i = 0; i < 3; i++) {
            double s = 0.0;
            s = s + A[i][0] * u;
            s = s + A[i][1] * v;
            s = s + A[i][2] * z;
            F[2 + 3 * i + DIR] = s;
        }
        if (m.is_thermal) {
            const double T = mg_T(m, rho, p);
            F[1] = F[1] + (T * (m.a2 * (DIR == 0 ? jx : jy)));
            F[14] = rho * (vn * jx);
            F[15] = rho * (vn * jy);
            F[16] = rho * (vn * jz);
            F[14 + DIR] = (rho * (vn * (DIR == 0 ? jx : jy))) + T;
        }
    }

    // HPRAcousticTensor::computeMaximumWaveSpeed (hpracoustictensor.cpp:7-97) and
    // MatrixAlgebra::computeLargestEigenvalue (matrixalgebra.cpp:182-199).  The
    // entries of d sigma / d A the tensor reads ([d][i][j][d],
    // hprequationofstate.cpp:62-74) are formed directly, in the reference's order.
    __device__ static __forceinline__ double acoustic_speed(const Params &m, const double (&w)[NF], int d)
    {
        double A[3][3], G[3][3], dev[3][3], Ad[3][3], sig[3][3], dsr[3][3];
        load_A(w, A);
        const double rho = w[0], p = w[1];
        shear(m, rho, A, sig, dsr);
        gram_dev(A, G, dev);
        mul_nn(A, dev, Ad);
        double C1[4][5], C2[5][4];
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int k = 0; k < 5; k++) {
                C1[i][k] = 0.0;
                C2[k][i] = 0.0;
            }
        C1[0][1] = 1.0 / rho;
#pragma unroll
        for (int i = 0; i < 3; i++) C1[i][0] = -(1.0 / rho) * dsr[d][i];
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) {
                const double pa = 0.0 + (G[d][i] * A[j][d]);
                const double pb = 0.0 + (G[d][d] * A[j][i]);
                double D = (pa + pb) + (-1.0 * ((2.0 / 3.0) * pa));
                D = D + Ad[j][i];
                if (i == d) D = D + Ad[j][d];
                const double dsA = (-(rho * m.cs2)) * D;
                C1[i][2 + j] = -(1.0 / rho) * dsA;
            }
        if (m.is_thermal) {
            C1[3][0] = mg_T_rho(m, rho, p) / rho;
            C1[3][1] = mg_T_p(m, rho) / rho;
        }
        const double ca = sqrt(((p / (rho * rho)) - mg_e_rho(m, rho, p)) / mg_e_p(m, rho));
        C2[0][0] = rho;
        C2[1][d] = rho * (ca * ca);
#pragma unroll
        for (int i = 0; i < 3; i++) C2[1][i] = C2[1][i] + (sig[d][i] - (rho * dsr[d][i]));
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) C2[2 + i][j] = A[i][j];
        if (m.is_thermal) {
            const double T = mg_T(m, rho, p);
            const double Tp = mg_T_p(m, rho);
            const double ch = (sqrt(m.a2 * (T / m.cv))) / rho;
            C2[1][3] = rho * ((ch * ch) / Tp);
        }
        double M[4][4];
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 4; j++) {
                double s = 0.0;
#pragma unroll
                for (int k = 0; k < 5; k++) s = s + C1[i][k] * C2[k][j];
                M[i][j] = s;
            }
        double v[4] = {1.0, 1.0, 1.0, 1.0};
        for (int it = 0; it < 10; it++) {
            double nv[4];
#pragma unroll
            for (int i = 0; i < 4; i++) {
                double s = 0.0;
#pragma unroll
                for (int k = 0; k < 4; k++) s = s + M[i][k] * v[k];
                nv[i] = s;
            }
            double nn = 0.0;
#pragma unroll
            for (int i = 0; i < 4; i++) nn = nn + nv[i] * nv[i];
            const double c = 1.0 / sqrt(nn);
#pragma unroll
            for (int i = 0; i < 4; i++) v[i] = c * nv[i];
        }
        double nn = 0.0;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < 4; k++) s = s + M[i][k] * v[k];
            nn = nn + s * s;
        }
        return sqrt(sqrt(nn));
    }

    // HPRSolvers::computeMaximumWaveSpeed[2D] (hprsolvers.cpp:286-303, 362-385).
    __device__ static __forceinline__ double wave_speed(const Params &m, const double (&w)[NF], int dim)
    {
        if (dim == 1) return fabs(w[11]) + acoustic_speed(m, w, 0);
        return smax(fabs(w[11]), fabs(w[12])) + smax(acoustic_speed(m, w, 0), acoustic_speed(m, w, 1));
    }

    // The stiffened gas needs p + p_inf > 0 for a real sound speed; the
    // Mie-Grueneisen solids may carry tension (p < 0) and only need rho > 0.
    __device__ static __forceinline__ bool admissible(const Params &m, const double (&w)[NF])
    {
        if constexpr (EOS == 0) return w[0] > 0.0 && (w[1] + m.p_inf) > 0.0;
        else return w[0] > 0.0;
    }

    // HPRStateVector() (hprstatevector.cpp:3-16): rho 1, p 1/1.4, A = I.
    __device__ static __forceinline__ void default_state(double (&w)[NF])
    {
#pragma unroll
        for (int k = 0; k < NF; k++) w[k] = 0.0;
        w[0] = 1.0;
        w[1] = 1.0 / 1.4;
        w[2] = 1.0;
        w[6] = 1.0;
        w[10] = 1.0;
    }

    // Stiff relaxation source (hprstatevector.cpp:312-366, hprsourceterms.cpp:7-39):
    // theta1^-1 = 3 det(A)^(5/3) / (c_s^2 tau1), times (min(|sigma|, 1e8) /
    // sigma_Y)^n when plastic, |sigma| the sigma norm (tensoralgebra.cpp:276-283)
    // of the shear stress -rho A^T E_A.
    __device__ static __forceinline__ void source(const Params &m, const double (&U)[NF], double (&S)[NF])
    {
        double A[3][3], G[3][3], dev[3][3], Ad[3][3];
        load_A(U, A);
        const double rho = U[0];
#pragma unroll
        for (int k = 0; k < NF; k++) S[k] = 0.0;
        gram_dev(A, G, dev);
        mul_nn(A, dev, Ad);
        double th1 = 3.0 * (pow(det3(A), (5.0 / 3.0)) / (m.cs2 * m.tau1));
        if (m.is_plastic) {
            double EA[3][3], P[3][3], sg[3][3];
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int j = 0; j < 3; j++) EA[i][j] = m.cs2 * Ad[i][j];
            mul_tn(A, EA, P);
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int j = 0; j < 3; j++) sg[i][j] = (-rho) * P[i][j];
            const double c1 = ((sg[0][0] - sg[1][1]) * (sg[0][0] - sg[1][1])) + ((sg[1][1] - sg[2][2]) * (sg[1][1] - sg[2][2])) +
                              ((sg[2][2] - sg[0][0]) * (sg[2][2] - sg[0][0]));
            const double c2 = (sg[0][1] * sg[0][1]) + (sg[1][2] * sg[1][2]) + (sg[2][0] * sg[2][0]);
            const double norm = smin(sqrt((0.5 * c1) + (3.0 * c2)), 1e8);
            th1 = th1 * pow(norm / m.sigma_y, m.n_pow);
        }
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) S[2 + 3 * i + j] = -(m.cs2 * Ad[i][j]) * th1;
        if (m.is_thermal) {
            const double E = U[1] / rho;
            const double u = U[11] / rho, v = U[12] / rho, z = U[13] / rho;
            const double jx = U[14] / rho, jy = U[15] / rho, jz = U[16] / rho;
            const double p = pressure_of(m, rho, E, u, v, z, A, jx, jy, jz);
            const double T = mg_T(m, rho, p);
            const double th2 = 1.0 / (m.a2 * m.tau2 * ((rho / m.rho0) * (m.t0 / T)));
            S[14] = -rho * th2 * (m.a2 * jx);
            S[15] = -rho * th2 * (m.a2 * jy);
            S[16] = -rho * th2 * (m.a2 * jz);
        }
    }

    // FAST mode uses the exact arithmetic for HPR.
    __device__ static __forceinline__ void prim_to_cons_fast(const Params &m, const double (&w)[NF], double (&U)[NF])
    {
        prim_to_cons(m, w, U);
    }
    __device__ static __forceinline__ void cons_to_prim_fast(const Params &m, const double (&U)[NF], double (&w)[NF])
    {
        cons_to_prim(m, U, w);
    }
    template <int DIR>
    __device__ static __forceinline__ void flux_prim_fast(const Params &m, const double (&w)[NF], const double (&U)[NF],
                                                          double (&F)[NF])
    {
        (void)w;
        flux<DIR>(m, U, F);
    }
};

using HPRPhys = HPRPhysT<0>;

// 17 fields: 128-cell x tiles keep the LDS tile at 52 KB.
template <int EOS>
struct Tiles<HPRPhysT<EOS>> {
    // y strips of 16 rows: a 1024 x 1024 grid gives 64 strips x 8 blocks x 2
    // waves = 1024 waves, one per SIMD (32-row strips leave half idle); x
    // blocks of 3 rows give 6,156 waves, three full rounds at two per SIMD
    // (4 rows: 4,608 waves, 2.25 rounds)
    static constexpr int bsx = 128, rx = 3, bsy = 128, hy = 16;
    static constexpr bool tune = false;
};

}  // namespace vg
