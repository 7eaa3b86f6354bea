// Two-material diffuse-interface Euler models: the physics of the reference's
// EulerMultiphysicsStateVector (Euler/Multiphysics/eulermultiphysicsstatevector.cpp)
// and EulerReducedStateVector (eulerreducedstatevector.cpp) on register arrays,
// in the reference's operation order, for the generic SLIC / FORCE sweeps.
//
// Multiphysics: primitives (alpha1, u, v, w, rho1, p1, rho2, p2), 8 stored per
// cell; conserved (rho, rho alpha1, rho u, rho v, rho w, alpha1 rho1, alpha1 E1,
// alpha2 rho2, alpha2 E2), 9.  Reduced: primitives (alpha1, u, v, w, rho1, rho2,
// p), 7 stored; conserved (rho, rho alpha1, rho u, rho v, rho w, alpha1 rho1,
// alpha2 rho2, rho E), 8.  The register state carries NF slots; the slot past
// the stored ones is zero.
#pragma once

#include "vg_phys_euler.h"

namespace vg {

// Two EulerMaterialParameters (material 1 and 2), and how often the
// second-order solve reinitialises the volume fraction (0: never).
struct TwoMaterialParams {
    EulerPhys::Params m1, m2;
    int reinit_every;
};

// The clamp of the volume fraction to [0.001, 0.999] in every conversion
// (eulermultiphysicsstatevector.cpp:47-58, 117-124; the reduced twin).
__device__ __forceinline__ double clamp_volume_fraction(double a)
{
    if (a < 0.001) return 0.001;
    if (a > 0.999) return 0.999;
    return a;
}

struct EulerMPPhys {
    static constexpr int NF = 9;
    static constexpr int NS = 8;
    static constexpr const char *name = "euler_multiphysics";
    using Params = TwoMaterialParams;

    __device__ static __forceinline__ void begin_step(Params &, double) {}

    // computeConservedVariableVector with computeMaterialkTotalEnergy
    // (eulermultiphysicsstatevector.cpp:113-140, 278-283, 300-305).
    __device__ static __forceinline__ void prim_to_cons(const Params &m, const double (&w)[NF], double (&U)[NF])
    {
        const double a = clamp_volume_fraction(w[0]), a2 = 1.0 - a;
        const double u = w[1], v = w[2], z = w[3], r1 = w[4], p1 = w[5], r2 = w[6], p2 = w[7];
        const double rho = (a * r1) + (a2 * r2);
        const double vsq = (u * u) + (v * v) + (z * z);
        U[0] = rho;
        U[1] = rho * a;
        U[2] = rho * u;
        U[3] = rho * v;
        U[4] = rho * z;
        U[5] = a * r1;
        U[6] = a * (r1 * ((0.5 * vsq) + EulerPhys::eos_e(m.m1, r1, p1)));
        U[7] = a2 * r2;
        U[8] = a2 * (r2 * ((0.5 * vsq) + EulerPhys::eos_e(m.m2, r2, p2)));
    }

    // setConservedVariableVector (eulermultiphysicsstatevector.cpp:43-75); its
    // velocity square adds w + w where w * w is meant (:67), kept for parity.
    __device__ static __forceinline__ void cons_to_prim(const Params &m, const double (&U)[NF], double (&w)[NF])
    {
        const double rho = U[0];
        const double a = clamp_volume_fraction(U[1] / rho), a2 = 1.0 - a;
        const double u = U[2] / rho, v = U[3] / rho, z = U[4] / rho;
        const double r1 = U[5] / a;
        const double vsq = (u * u) + (v * v) + (z + z);
        const double e1 = (U[6] / (r1 * a)) - (0.5 * vsq);
        const double r2 = U[7] / a2;
        const double e2 = (U[8] / (r2 * a2)) - (0.5 * vsq);
        w[0] = a;
        w[1] = u;
        w[2] = v;
        w[3] = z;
        w[4] = r1;
        w[5] = EulerPhys::eos_p(m.m1, r1, e1);
        w[6] = r2;
        w[7] = EulerPhys::eos_p(m.m2, r2, e2);
        w[8] = 0.0;
    }

    // computeX/YFluxVector (eulermultiphysicsstatevector.cpp:142-257).
    template <int DIR>
    __device__ static __forceinline__ void flux(const Params &m, const double (&U)[NF], double (&F)[NF])
    {
        const double rho = U[0];
        const double a = clamp_volume_fraction(U[1] / rho), a2 = 1.0 - a;
        const double u = U[2] / rho, v = U[3] / rho, z = U[4] / rho;
        const double r1 = U[5] / a;
        const double vsq = (u * u) + (v * v) + (z * z);
        const double E1 = U[6] / (a * r1);
        const double p1 = EulerPhys::eos_p(m.m1, r1, E1 - (0.5 * vsq));
        const double r2 = U[7] / a2;
        const double E2 = U[8] / (a2 * r2);
        const double p2 = EulerPhys::eos_p(m.m2, r2, E2 - (0.5 * vsq));
        const double q = (DIR == 0) ? u : v;
        const double pm = (a * p1) + (a2 * p2);
        F[0] = rho * q;
        F[1] = rho * (q * a);
        F[2] = (DIR == 0) ? (rho * (q * q)) + pm : rho * (q * u);
        F[3] = (DIR == 1) ? (rho * (q * q)) + pm : rho * (q * v);
        F[4] = rho * (q * z);
        F[5] = a * (r1 * q);
        F[6] = a * ((r1 * (q * E1)) + (q * p1));
        F[7] = a2 * (r2 * q);
        F[8] = a2 * ((r2 * (q * E2)) + (q * p2));
    }

    // computeMaximumWaveSpeed[2D] per cell (Solvers/solvers.cpp:278-296, 341-363).
    __device__ static __forceinline__ double wave_speed(const Params &m, const double (&w)[NF], int dim)
    {
        const double s = (dim == 1) ? fabs(w[1]) : smax(fabs(w[1]), fabs(w[2]));
        return s + smax(EulerPhys::eos_c(m.m1, w[4], w[5]), EulerPhys::eos_c(m.m2, w[6], w[7]));
    }

    // The partial state of a vanishing material (alpha_k near 0.001) may leave
    // the EOS's range -- the reference's Toro run reaches negative p2 -- and the
    // reference carries on (a NaN sound speed drops out of the CFL maximum), so
    // only non-finite values count as a failure.
    __device__ static __forceinline__ bool admissible(const Params &, const double (&)[NF]) { return true; }

    // EulerMultiphysicsStateVector() (eulermultiphysicsstatevector.cpp:3-15): the 2D corner ghosts.
    __device__ static __forceinline__ void default_state(double (&w)[NF])
    {
        w[0] = 0.999; w[1] = 0.0; w[2] = 0.0; w[3] = 0.0; w[4] = 1.0; w[5] = 1.0; w[6] = 1.0; w[7] = 1.0;
        w[8] = 0.0;
    }

    // One arithmetic for both modes.
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

struct EulerReducedPhys {
    static constexpr int NF = 8;
    static constexpr int NS = 7;
    static constexpr const char *name = "euler_reduced";
    using Params = TwoMaterialParams;

    __device__ static __forceinline__ void begin_step(Params &, double) {}

    // computeConservedVariableVector (eulerreducedstatevector.cpp:94-126): the
    // material energies from the total density and the interface pressure.
    __device__ static __forceinline__ void prim_to_cons(const Params &m, const double (&w)[NF], double (&U)[NF])
    {
        const double a = clamp_volume_fraction(w[0]), a2 = 1.0 - a;
        const double u = w[1], v = w[2], z = w[3], r1 = w[4], r2 = w[5], p = w[6];
        const double rho = (a * r1) + (a2 * r2);
        U[0] = rho;
        U[1] = rho * a;
        U[2] = rho * u;
        U[3] = rho * v;
        U[4] = rho * z;
        U[5] = a * r1;
        U[6] = a2 * r2;
        const double vsq = (u * u) + (v * v) + (z * z);
        const double E1 = EulerPhys::eos_e(m.m1, rho, p) + (0.5 * vsq);
        const double E2 = EulerPhys::eos_e(m.m2, rho, p) + (0.5 * vsq);
        U[7] = rho * ((a * E1) + (a2 * E2));
    }

    // The mixture pressure alpha1 p1(rho, e) + alpha2 p2(rho, e)
    // (eulerreducedstatevector.cpp:71-74, 166-168).
    __device__ static __forceinline__ double pressure(const Params &m, double a, double a2, double rho, double e)
    {
        return (a * EulerPhys::eos_p(m.m1, rho, e)) + (a2 * EulerPhys::eos_p(m.m2, rho, e));
    }

    // setConservedVariableVector (eulerreducedstatevector.cpp:43-75).
    __device__ static __forceinline__ void cons_to_prim(const Params &m, const double (&U)[NF], double (&w)[NF])
    {
        const double rho = U[0];
        const double a = clamp_volume_fraction(U[1] / rho), a2 = 1.0 - a;
        const double u = U[2] / rho, v = U[3] / rho, z = U[4] / rho;
        const double vsq = (u * u) + (v * v) + (z * z);
        const double e = (U[7] / rho) - (0.5 * vsq);
        w[0] = a;
        w[1] = u;
        w[2] = v;
        w[3] = z;
        w[4] = U[5] / a;
        w[5] = U[6] / a2;
        w[6] = pressure(m, a, a2, rho, e);
        w[7] = 0.0;
    }

    // computeX/YFluxVector (eulerreducedstatevector.cpp:128-231).
    template <int DIR>
    __device__ static __forceinline__ void flux(const Params &m, const double (&U)[NF], double (&F)[NF])
    {
        const double rho = U[0];
        const double a = clamp_volume_fraction(U[1] / rho), a2 = 1.0 - a;
        const double u = U[2] / rho, v = U[3] / rho, z = U[4] / rho;
        const double r1 = U[5] / a;
        const double r2 = U[6] / a2;
        const double E = U[7] / rho;
        const double vsq = (u * u) + (v * v) + (z * z);
        const double p = pressure(m, a, a2, rho, E - (0.5 * vsq));
        const double q = (DIR == 0) ? u : v;
        F[0] = rho * q;
        F[1] = rho * (q * a);
        F[2] = (DIR == 0) ? (rho * (q * q)) + p : rho * (q * u);
        F[3] = (DIR == 1) ? (rho * (q * q)) + p : rho * (q * v);
        F[4] = rho * (q * z);
        F[5] = a * (r1 * q);
        F[6] = a2 * (r2 * q);
        F[7] = (rho * (q * E)) + (q * p);
    }

    // computeMaximumWaveSpeed[2D] per cell (Solvers/solvers.cpp:298-316, 365-387):
    // the material sound speeds at the interface pressure.
    __device__ static __forceinline__ double wave_speed(const Params &m, const double (&w)[NF], int dim)
    {
        const double s = (dim == 1) ? fabs(w[1]) : smax(fabs(w[1]), fabs(w[2]));
        return s + smax(EulerPhys::eos_c(m.m1, w[4], w[6]), EulerPhys::eos_c(m.m2, w[5], w[6]));
    }

    // As the multiphysics state: only non-finite values count as a failure.
    __device__ static __forceinline__ bool admissible(const Params &, const double (&)[NF]) { return true; }

    // EulerReducedStateVector() (eulerreducedstatevector.cpp:3-14).
    __device__ static __forceinline__ void default_state(double (&w)[NF])
    {
        w[0] = 0.999; w[1] = 0.0; w[2] = 0.0; w[3] = 0.0; w[4] = 1.0; w[5] = 1.0; w[6] = 1.0; w[7] = 0.0;
    }

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

// Nine fields: 128-cell x tiles keep the LDS tile at 27 KB and the registers
// within budget.
template <>
struct Tiles<EulerMPPhys> {
    static constexpr int bsx = 128, rx = 4, bsy = 128, hy = 32;
    static constexpr bool tune = true;
};
template <>
struct Tiles<EulerReducedPhys> {
    static constexpr int bsx = 128, rx = 4, bsy = 128, hy = 32;
    static constexpr bool tune = true;
};

}  // namespace vg
