// Euler equations with a stiffened-gas EOS and a reaction-progress variable:
// the physics of the reference's EulerStateVector, as device functions on
// register arrays.  Primitive order (rho, u, v, w, p, lambda); conserved order
// (rho, rho u, rho v, rho w, E, rho lambda) -- Euler/eulerstatevector.h:55-61.
#pragma once

#include "vg_device.h"
#include "vg_quot.h"

namespace vg {

struct EulerPhys {
    static constexpr int NF = 6;
    static constexpr const char *name = "euler";

    // EulerMaterialParameters (Euler/eulermaterialparameters.h:30-36).
    // Derived constants (gm1 = gamma - 1, gpinf = gamma * p_inf, inv_gm1) are
    // filled on the host; each is the same IEEE value the reference recomputes
    // at every use, so EXACT mode stays bit-identical.
    struct Params {
        double gamma, p_inf, cv, e_form, t_ign, k_rate;
        double gm1, gpinf, inv_gm1;
    };

    // Euler has no parameter that follows the time step.
    __device__ static __forceinline__ void begin_step(Params &, double) {}

    // Specific internal energy, pressure and sound speed of the stiffened gas
    // (Euler/eulerequationofstate.cpp:7-29).
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

    // computeConservedVariableVector + computeTotalEnergy
    // (Euler/eulerstatevector.cpp:78-91, 192-198).
    __device__ static __forceinline__ void prim_to_cons(const Params &m, const double (&w)[NF], double (&U)[NF])
    {
        const double rho = w[0], u = w[1], v = w[2], z = w[3], p = w[4], lam = w[5];
        const double vsq = (u * u) + (v * v) + (z * z);
        U[0] = rho;
        U[1] = rho * u;
        U[2] = rho * v;
        U[3] = rho * z;
        U[4] = rho * ((0.5 * vsq) + eos_e(m, rho, p) + (m.e_form * (lam - 1.0)));
        U[5] = rho * lam;
    }

    // The quotients by rho of a conserved vector (u, v, w, lambda, E / rho), each
    // the correctly rounded IEEE quotient, from one shared reciprocal of rho.
    __device__ static __forceinline__ void by_rho(const double (&U)[NF], const SharedDivisor &d, double (&q)[5])
    {
        const double a[5] = {U[1], U[2], U[3], U[5], U[4]};
        quot_group<5>(a, d, q);
    }

    // setConservedVariableVector (Euler/eulerstatevector.cpp:47-61); returns
    // the shared divisor of rho for a flux of the same density.
    __device__ static __forceinline__ SharedDivisor cons_to_prim_d(const Params &m, const double (&U)[NF], double (&w)[NF])
    {
        const double rho = U[0];
        const SharedDivisor d = shared_divisor(rho);
        double q[5];
        by_rho(U, d, q);
        const double u = q[0], v = q[1], z = q[2];
        const double lam = q[3];
        const double vsq = (u * u) + (v * v) + (z * z);
        const double e = q[4] - (0.5 * vsq) - (m.e_form * (lam - 1.0));
        w[0] = rho;
        w[1] = u;
        w[2] = v;
        w[3] = z;
        w[4] = eos_p(m, rho, e);
        w[5] = lam;
        return d;
    }
    __device__ static __forceinline__ void cons_to_prim(const Params &m, const double (&U)[NF], double (&w)[NF])
    {
        (void)cons_to_prim_d(m, U, w);
    }

    // computeXFluxVector / computeYFluxVector of a conserved vector
    // (Euler/eulerstatevector.cpp:93-120, 127-154), given the shared divisor
    // of its density U[0].
    template <int DIR>
    __device__ static __forceinline__ void flux_d(const Params &m, const double (&U)[NF], const SharedDivisor &d,
                                                  double (&F)[NF])
    {
        const double rho = U[0];
        double q[5];
        by_rho(U, d, q);
        const double u = q[0], v = q[1], z = q[2];
        const double lam = q[3];
        const double vsq = (u * u) + (v * v) + (z * z);
        const double et = q[4];
        const double e = et - (0.5 * vsq) - (m.e_form * (lam - 1.0));
        const double p = eos_p(m, rho, e);
        if (DIR == 0) {
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
    template <int DIR>
    __device__ static __forceinline__ void flux(const Params &m, const double (&U)[NF], double (&F)[NF])
    {
        flux_d<DIR>(m, U, shared_divisor(U[0]), F);
    }

    // Wave speed of Solvers::computeMaximumWaveSpeed (1D, |u| + c) and
    // computeMaximumWaveSpeed2D (max(|u|, |v|) + c) (Solvers/solvers.cpp:267, 329).
    __device__ static __forceinline__ double wave_speed(const Params &m, const double (&w)[NF], int dim)
    {
        const double a = (dim == 1) ? fabs(w[1]) : smax(fabs(w[1]), fabs(w[2]));
        return a + eos_c(m, w[0], w[4]);
    }

    // A state whose sound speed is real: rho > 0 and p + p_inf > 0.
    __device__ static __forceinline__ bool admissible(const Params &m, const double (&w)[NF])
    {
        return w[0] > 0.0 && (w[4] + m.p_inf) > 0.0;
    }

    // EulerStateVector() (Euler/eulerstatevector.cpp:3-12): fills the 2D corner ghosts.
    __device__ static __forceinline__ void default_state(double (&w)[NF])
    {
        w[0] = 1.0; w[1] = 0.0; w[2] = 0.0; w[3] = 0.0; w[4] = 1.0; w[5] = 1.0;
    }

    // ---- FAST mode ----------------------------------------------------------
    // Same scheme, fewer VALU operations: one Newton-refined reciprocal of rho
    // per state, the EOS constants folded, fluxes taken from primitives without
    // a conserved round trip, multiply-adds fused.  Not bit-identical to the
    // reference.
    __device__ static __forceinline__ void prim_to_cons_fast(const Params &m, const double (&w)[NF], double (&U)[NF])
    {
#pragma clang fp contract(fast)
        const double rho = w[0], u = w[1], v = w[2], z = w[3], p = w[4], lam = w[5];
        U[0] = rho;
        U[1] = rho * u;
        U[2] = rho * v;
        U[3] = rho * z;
        U[4] = 0.5 * (U[1] * u + U[2] * v + U[3] * z) + (p + m.gpinf) * m.inv_gm1 + m.e_form * (rho * lam - rho);
        U[5] = rho * lam;
    }
    __device__ static __forceinline__ void cons_to_prim_fast(const Params &m, const double (&U)[NF], double (&w)[NF])
    {
#pragma clang fp contract(fast)
        const double rho = U[0], r = rcp_nr(rho);
        const double u = U[1] * r, v = U[2] * r, z = U[3] * r, lam = U[5] * r;
        const double ke = 0.5 * ((U[1] * u) + (U[2] * v) + (U[3] * z));
        w[0] = rho;
        w[1] = u;
        w[2] = v;
        w[3] = z;
        w[4] = m.gm1 * (U[4] - ke - m.e_form * (U[5] - rho)) - m.gpinf;
        w[5] = lam;
    }
    // Flux from a primitive state and its conserved vector.
    template <int DIR>
    __device__ static __forceinline__ void flux_prim_fast(const Params &m, const double (&w)[NF], const double (&U)[NF],
                                                          double (&F)[NF])
    {
#pragma clang fp contract(fast)
        (void)m;
        const double vn = (DIR == 0) ? w[1] : w[2];
        const double p = w[4];
        F[0] = U[0] * vn;
        F[1] = U[1] * vn + (DIR == 0 ? p : 0.0);
        F[2] = U[2] * vn + (DIR == 1 ? p : 0.0);
        F[3] = U[3] * vn;
        F[4] = (U[4] + p) * vn;
        F[5] = U[5] * vn;
    }
};

// The same Euler state when its out-of-plane velocity is +0 and its reaction
// progress is 1 everywhere (the 2D problems without reaction): the sweeps then
// leave rho w at +0 and rho lambda equal to rho bit for bit (every w term is a
// signed zero that the updates return to +0; rho lambda has rho's flux, slope
// and update, so lambda = rho lambda / rho = 1 exactly), and the terms they add
// to the other fields are exact zeros.  This physics carries the other four
// fields -- (rho, u, v, p), planes 0, 1, 2 and 4 of the grid -- in the same
// operation order with those zero terms dropped, so EXACT results are the full
// state's bit for bit; the two constant planes are neither read nor written.
struct EulerPlanarPhys {
    static constexpr int NF = 4;
    static constexpr const char *name = "euler_planar";
    using Params = EulerPhys::Params;

    __device__ static __forceinline__ void begin_step(Params &, double) {}

    __device__ static __forceinline__ void prim_to_cons(const Params &m, const double (&w)[NF], double (&U)[NF])
    {
        const double rho = w[0], u = w[1], v = w[2], p = w[3];
        const double vsq = (u * u) + (v * v);
        U[0] = rho;
        U[1] = rho * u;
        U[2] = rho * v;
        U[3] = rho * ((0.5 * vsq) + EulerPhys::eos_e(m, rho, p));
    }
    __device__ static __forceinline__ void by_rho(const double (&U)[NF], const SharedDivisor &d, double (&q)[3])
    {
        const double a[3] = {U[1], U[2], U[3]};
        quot_group<3>(a, d, q);
    }
    __device__ static __forceinline__ SharedDivisor cons_to_prim_d(const Params &m, const double (&U)[NF], double (&w)[NF])
    {
        const double rho = U[0];
        const SharedDivisor d = shared_divisor(rho);
        double q[3];
        by_rho(U, d, q);
        const double u = q[0], v = q[1];
        const double vsq = (u * u) + (v * v);
        const double e = q[2] - (0.5 * vsq);
        w[0] = rho;
        w[1] = u;
        w[2] = v;
        w[3] = EulerPhys::eos_p(m, rho, e);
        return d;
    }
    __device__ static __forceinline__ void cons_to_prim(const Params &m, const double (&U)[NF], double (&w)[NF])
    {
        (void)cons_to_prim_d(m, U, w);
    }
    template <int DIR>
    __device__ static __forceinline__ void flux_d(const Params &m, const double (&U)[NF], const SharedDivisor &d,
                                                  double (&F)[NF])
    {
        const double rho = U[0];
        double q[3];
        by_rho(U, d, q);
        const double u = q[0], v = q[1];
        const double vsq = (u * u) + (v * v);
        const double et = q[2];
        const double e = et - (0.5 * vsq);
        const double p = EulerPhys::eos_p(m, rho, e);
        if (DIR == 0) {
            F[0] = rho * u;
            F[1] = (rho * (u * u)) + p;
            F[2] = rho * (u * v);
            F[3] = (rho * (u * et)) + (u * p);
        } else {
            F[0] = rho * v;
            F[1] = rho * (v * u);
            F[2] = (rho * (v * v)) + p;
            F[3] = (rho * (v * et)) + (v * p);
        }
    }
    template <int DIR>
    __device__ static __forceinline__ void flux(const Params &m, const double (&U)[NF], double (&F)[NF])
    {
        flux_d<DIR>(m, U, shared_divisor(U[0]), F);
    }
    __device__ static __forceinline__ double wave_speed(const Params &m, const double (&w)[NF], int dim)
    {
        const double a = (dim == 1) ? fabs(w[1]) : smax(fabs(w[1]), fabs(w[2]));
        return a + EulerPhys::eos_c(m, w[0], w[3]);
    }
    __device__ static __forceinline__ bool admissible(const Params &m, const double (&w)[NF])
    {
        return w[0] > 0.0 && (w[3] + m.p_inf) > 0.0;
    }
    __device__ static __forceinline__ void default_state(double (&w)[NF])
    {
        w[0] = 1.0; w[1] = 0.0; w[2] = 0.0; w[3] = 1.0;
    }

    // FAST mode, as EulerPhys's.
    __device__ static __forceinline__ void prim_to_cons_fast(const Params &m, const double (&w)[NF], double (&U)[NF])
    {
#pragma clang fp contract(fast)
        const double rho = w[0], u = w[1], v = w[2], p = w[3];
        U[0] = rho;
        U[1] = rho * u;
        U[2] = rho * v;
        U[3] = 0.5 * (U[1] * u + U[2] * v) + (p + m.gpinf) * m.inv_gm1;
    }
    __device__ static __forceinline__ void cons_to_prim_fast(const Params &m, const double (&U)[NF], double (&w)[NF])
    {
#pragma clang fp contract(fast)
        const double rho = U[0], r = rcp_nr(rho);
        const double u = U[1] * r, v = U[2] * r;
        const double ke = 0.5 * ((U[1] * u) + (U[2] * v));
        w[0] = rho;
        w[1] = u;
        w[2] = v;
        w[3] = m.gm1 * (U[3] - ke) - m.gpinf;
    }
    template <int DIR>
    __device__ static __forceinline__ void flux_prim_fast(const Params &m, const double (&w)[NF], const double (&U)[NF],
                                                          double (&F)[NF])
    {
#pragma clang fp contract(fast)
        (void)m;
        const double vn = (DIR == 0) ? w[1] : w[2];
        const double p = w[3];
        F[0] = U[0] * vn;
        F[1] = U[1] * vn + (DIR == 0 ? p : 0.0);
        F[2] = U[2] * vn + (DIR == 1 ? p : 0.0);
        F[3] = (U[3] + p) * vn;
    }
};

// (rho, u, v, p) are planes 0, 1, 2 and 4 of the Euler grid.
template <>
struct PlaneMap<EulerPlanarPhys> {
    __device__ static constexpr int of(int f) { return f < 3 ? f : 4; }
};

template <>
struct ReducedPhys<EulerPhys> {
    using type = EulerPlanarPhys;
    // The planes EulerPlanarPhys leaves alone and the bits they must hold:
    // w = +0 (plane 3) and lambda = 1.0 (plane 5).
    static constexpr int kConst = 2;
    __host__ __device__ static constexpr int plane(int i) { return i == 0 ? 3 : 5; }
    __host__ __device__ static constexpr unsigned long long bits(int i) { return i == 0 ? 0ull : 0x3FF0000000000000ull; }
};

}  // namespace vg
