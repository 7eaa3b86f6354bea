// Ideal MHD with GLM divergence cleaning: the physics of the reference's
// MHDStateVector / MHDWaveSpeeds on register arrays.  Primitive order
// (rho, u, v, w, p, Bx, By, Bz, psi); conserved order
// (rho, rho u, rho v, rho w, E, Bx, By, Bz, psi) -- MHD/mhdstatevector.cpp:49-106.
#pragma once

#include "vg_device.h"
#include "vg_quot.h"

namespace vg {

struct MHDPhys {
    static constexpr int NF = 9;
    static constexpr const char *name = "mhd";

    // MHDMaterialParameters (MHD/mhdmaterialparameters.cpp:3-34): gamma, p_inf,
    // GLM hyperbolic speed c_h and parabolic damping c_p; derived constants are
    // filled on the host (the same IEEE values the reference recomputes).
    struct Params {
        double gamma, p_inf, ch, cp;
        double gm1, gpinf, inv_gm1;
    };

    // Each step the reference sets c_h to the step's maximum wave speed and
    // c_p = sqrt(0.18 c_h) (mhdsecondordersolver.cpp:330-332,
    // mhdmaterialparameters.cpp:31-34).
    __device__ static __forceinline__ void begin_step(Params &m, double smax)
    {
        m.ch = smax;
        m.cp = sqrt(m.ch * 0.18);
    }

    // MHDStateVector::computeSourceTermVector (MHD/mhdstatevector.cpp:222-234):
    // GLM damping of psi, -(c_h^2 / c_p^2) psi; every other slot 0.
    __device__ static __forceinline__ void source(const Params &m, const double (&U)[NF], double (&S)[NF])
    {
#pragma unroll
        for (int k = 0; k < NF; k++) S[k] = 0.0;
        const double chs = m.ch * m.ch, cps = m.cp * m.cp;
        S[8] = -(chs / cps) * U[8];
    }

    __device__ static __forceinline__ double eos_e(const Params &m, double rho, double p)
    {
        return (p + m.gpinf) / (m.gm1 * rho);
    }
    __device__ static __forceinline__ double eos_p(const Params &m, double rho, double e)
    {
        return (e * m.gm1 * rho) - m.gpinf;
    }
    __device__ static __forceinline__ double eos_c(const Params &m, double rho, double p)
    {
        return sqrt((m.gamma * (p + m.p_inf)) / rho);
    }

    // VectorAlgebra::computeDotProduct: the sum starts at 0.0 (a zero result is +0).
    __device__ static __forceinline__ double dot3(double a0, double a1, double a2, double b0, double b1, double b2)
    {
        double d = 0.0;
        d = d + a0 * b0;
        d = d + a1 * b1;
        d = d + a2 * b2;
        return d;
    }

    // computeConservedVariableVector + computeTotalEnergy (mhdstatevector.cpp:89-106, 245-252).
    __device__ static __forceinline__ void prim_to_cons(const Params &m, const double (&w)[NF], double (&U)[NF])
    {
        const double rho = w[0], u = w[1], v = w[2], z = w[3], p = w[4];
        const double vsq = (u * u) + (v * v) + (z * z);
        const double hydro = rho * ((0.5 * vsq) + eos_e(m, rho, p));
        const double b2 = (w[5] * w[5]) + (w[6] * w[6]) + (w[7] * w[7]);
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

    // The quotients by rho of a conserved vector (u, v, w and the hydrodynamic
    // specific total energy (E - |B|^2 / 2) / rho), each the correctly rounded
    // IEEE quotient, from one shared reciprocal of rho.
    __device__ static __forceinline__ void by_rho(const double (&U)[NF], const SharedDivisor &d, double b2,
                                                  double (&q)[4])
    {
        const double a[4] = {U[1], U[2], U[3], U[4] - (0.5 * b2)};
        quot_group<4>(a, d, q);
    }

    // setConservedVariableVector (mhdstatevector.cpp:49-68); returns the shared
    // divisor of rho for a flux of the same density.
    __device__ static __forceinline__ SharedDivisor cons_to_prim_d(const Params &m, const double (&U)[NF],
                                                                   double (&w)[NF])
    {
        const double rho = U[0];
        const SharedDivisor d = shared_divisor(rho);
        const double bx = U[5], by = U[6], bz = U[7];
        const double b2 = (bx * bx) + (by * by) + (bz * bz);
        double q[4];
        by_rho(U, d, b2, q);
        const double u = q[0], v = q[1], z = q[2];
        const double et = q[3];
        const double vsq = (u * u) + (v * v) + (z * z);
        const double e = et - (0.5 * vsq);
        w[0] = rho;
        w[1] = u;
        w[2] = v;
        w[3] = z;
        w[4] = eos_p(m, rho, e);
        w[5] = bx;
        w[6] = by;
        w[7] = bz;
        w[8] = U[8];
        return d;
    }
    __device__ static __forceinline__ void cons_to_prim(const Params &m, const double (&U)[NF], double (&w)[NF])
    {
        (void)cons_to_prim_d(m, U, w);
    }

    // computeXFluxVector / computeYFluxVector (mhdstatevector.cpp:108-158,
    // 165-215), given the shared divisor of the density U[0].
    template <int DIR>
    __device__ static __forceinline__ void flux_d(const Params &m, const double (&U)[NF], const SharedDivisor &d,
                                                  double (&F)[NF])
    {
        const double rho = U[0];
        const double bx = U[5], by = U[6], bz = U[7], psi = U[8];
        const double b2 = (bx * bx) + (by * by) + (bz * bz);
        double q[4];
        by_rho(U, d, b2, q);
        const double u = q[0], v = q[1], z = q[2];
        const double E = U[4];
        const double het = q[3];
        const double vsq = (u * u) + (v * v) + (z * z);
        const double e = het - (0.5 * vsq);
        const double p = eos_p(m, rho, e);
        const double ch2 = m.ch * m.ch;
        const double ub = dot3(u, v, z, bx, by, bz);
        if (DIR == 0) {
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
    template <int DIR>
    __device__ static __forceinline__ void flux(const Params &m, const double (&U)[NF], double (&F)[NF])
    {
        flux_d<DIR>(m, U, shared_divisor(U[0]), F);
    }

    // MHDWaveSpeeds::computeX/YFastMagnetoAcousticSpeed with the Alfven speed
    // |B| / sqrt(rho) (mhdwavespeeds.cpp:7-15, 41-63).
    __device__ static __forceinline__ double fast_speed(const Params &m, const double (&w)[NF], int dir)
    {
        const double rho = w[0];
        const double c = eos_c(m, rho, w[4]);
        const double a = sqrt(dot3(w[5], w[6], w[7], w[5], w[6], w[7])) / sqrt(rho);
        const double c2 = (c * c), a2 = (a * a);
        const double bn = (dir == 0) ? w[5] : w[6];
        return sqrt(0.5 * (c2 + a2 + sqrt(((c2 + a2) * (c2 + a2)) - 4.0 * ((c2 * (bn * bn)) / rho))));
    }

    // MHDSolvers::computeMaximumWaveSpeed[2D] (mhdsolvers.cpp:286-303, 365-387).
    __device__ static __forceinline__ double wave_speed(const Params &m, const double (&w)[NF], int dim)
    {
        if (dim == 1) return fabs(w[1]) + fabs(fast_speed(m, w, 0));
        return smax(fabs(w[1]), fabs(w[2])) + smax(fabs(fast_speed(m, w, 0)), fabs(fast_speed(m, w, 1)));
    }

    __device__ static __forceinline__ bool admissible(const Params &m, const double (&w)[NF])
    {
        return w[0] > 0.0 && (w[4] + m.p_inf) > 0.0;
    }

    // MHDStateVector() (mhdstatevector.cpp:3-16).
    __device__ static __forceinline__ void default_state(double (&w)[NF])
    {
        w[0] = 1.0; w[1] = 0.0; w[2] = 0.0; w[3] = 0.0; w[4] = 1.0;
        w[5] = 0.0; w[6] = 0.0; w[7] = 0.0; w[8] = 0.0;
    }

    // ---- FAST mode ----------------------------------------------------------
    __device__ static __forceinline__ void prim_to_cons_fast(const Params &m, const double (&w)[NF], double (&U)[NF])
    {
        const double rho = w[0], u = w[1], v = w[2], z = w[3], p = w[4];
        U[0] = rho;
        U[1] = rho * u;
        U[2] = rho * v;
        U[3] = rho * z;
        U[4] = 0.5 * (U[1] * u + U[2] * v + U[3] * z) + (p + m.gpinf) * m.inv_gm1
               + 0.5 * (w[5] * w[5] + w[6] * w[6] + w[7] * w[7]);
        U[5] = w[5];
        U[6] = w[6];
        U[7] = w[7];
        U[8] = w[8];
    }
    __device__ static __forceinline__ void cons_to_prim_fast(const Params &m, const double (&U)[NF], double (&w)[NF])
    {
        const double rho = U[0], r = rcp_nr(rho);
        const double u = U[1] * r, v = U[2] * r, z = U[3] * r;
        const double b2 = U[5] * U[5] + U[6] * U[6] + U[7] * U[7];
        w[0] = rho;
        w[1] = u;
        w[2] = v;
        w[3] = z;
        w[4] = m.gm1 * (U[4] - 0.5 * (U[1] * u + U[2] * v + U[3] * z) - 0.5 * b2) - m.gpinf;
        w[5] = U[5];
        w[6] = U[6];
        w[7] = U[7];
        w[8] = U[8];
    }
    template <int DIR>
    __device__ static __forceinline__ void flux_prim_fast(const Params &m, const double (&w)[NF], const double (&U)[NF],
                                                          double (&F)[NF])
    {
        const double u = w[1], v = w[2], z = w[3], p = w[4];
        const double bx = w[5], by = w[6], bz = w[7];
        const double pt = p + 0.5 * (bx * bx + by * by + bz * bz);
        const double ub = u * bx + v * by + z * bz;
        const double ch2 = m.ch * m.ch;
        if (DIR == 0) {
            F[0] = U[0] * u;
            F[1] = U[1] * u + pt - bx * bx;
            F[2] = U[2] * u - bx * by;
            F[3] = U[3] * u - bx * bz;
            F[4] = (U[4] + pt) * u - ub * bx;
            F[5] = w[8];
            F[6] = by * u - bx * v;
            F[7] = bz * u - bx * z;
            F[8] = ch2 * bx;
        } else {
            F[0] = U[0] * v;
            F[1] = U[1] * v - by * bx;
            F[2] = U[2] * v + pt - by * by;
            F[3] = U[3] * v - by * bz;
            F[4] = (U[4] + pt) * v - ub * by;
            F[5] = bx * v - by * u;
            F[6] = w[8];
            F[7] = bz * v - by * z;
            F[8] = ch2 * by;
        }
    }
};

}  // namespace vg
