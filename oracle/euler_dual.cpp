// Forward-mode (dual-number) restatement of the oracle's Euler SLIC step --
// TEST INFRASTRUCTURE ONLY (oracle).
//
// The same arithmetic as scheme_oracle.c + euler_oracle.c (EulerStateVector,
// Solvers/secondordersolver.cpp:7-17, 80-98, 140-236; Solvers/solvers.cpp:454-466,
// 550-565, 601-609; Solvers/firstordersolver.cpp:79-97, 165-171, 227-230,
// 428-433; slopelimiters.cpp:7-74, 217-246), templated on the scalar type so
// that T = Dual carries one tangent direction through every operation.  With
// T = double it reproduces the oracle bit for bit (checked by the tests); with
// T = Dual it gives the exact Jacobian-vector product of the discrete map at
// fixed time steps -- the independent check of the GPU adjoint (a dot-product
// test <seed, J v> = <J^T seed, v>).  Branches (limiter, clamp) follow the
// primal values, as the adjoint's piecewise derivative does.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "oracle.h"

namespace {

struct Dual {
    double v, d;
    Dual(double a = 0.0, double b = 0.0) : v(a), d(b) {}
};
inline Dual operator+(Dual a, Dual b) { return Dual(a.v + b.v, a.d + b.d); }
inline Dual operator-(Dual a, Dual b) { return Dual(a.v - b.v, a.d - b.d); }
inline Dual operator-(Dual a) { return Dual(-a.v, -a.d); }
inline Dual operator*(Dual a, Dual b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
inline Dual operator/(Dual a, Dual b) { return Dual(a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)); }
inline Dual operator+(Dual a, double b) { return Dual(a.v + b, a.d); }
inline Dual operator+(double a, Dual b) { return Dual(a + b.v, b.d); }
inline Dual operator-(Dual a, double b) { return Dual(a.v - b, a.d); }
inline Dual operator-(double a, Dual b) { return Dual(a - b.v, -b.d); }
inline Dual operator*(Dual a, double b) { return Dual(a.v * b, a.d * b); }
inline Dual operator*(double a, Dual b) { return Dual(a * b.v, a * b.d); }
inline Dual operator/(Dual a, double b) { return Dual(a.v / b, a.d / b); }
inline Dual operator/(double a, Dual b) { return Dual(a / b.v, -a * b.d / (b.v * b.v)); }
inline double val(double x) { return x; }
inline double val(Dual x) { return x.v; }
inline double fabs_v(double x) { return std::fabs(x); }
inline double fabs_v(Dual x) { return std::fabs(x.v); }

const int NF = 6;

// Hash of every branch decision the step takes (slope clamps, limiter cases,
// min selections): two runs with equal hashes evaluated the same smooth
// pieces, so a finite difference between them involves no kink.
uint64_t g_branches = 0;
inline bool note(bool b)
{
    g_branches = (g_branches ^ (uint64_t)(b ? 0x9e37u : 0x51u)) * 0x100000001b3ull;
    return b;
}

struct P {
    double gamma, p_inf, e_form;
};

template <class T>
T eos_e(const P &m, T rho, T p) { return (p + (m.gamma * m.p_inf)) / ((m.gamma - 1) * rho); }
template <class T>
T eos_p(const P &m, T rho, T e) { return (e * (m.gamma - 1) * rho) - (m.gamma * m.p_inf); }

template <class T>
void p2c(const P &m, const T *w, T *U)
{
    T rho = w[0], u = w[1], v = w[2], z = w[3], p = w[4], lam = w[5];
    T vsq = (u * u) + (v * v) + (z * z);
    U[0] = rho;
    U[1] = rho * u;
    U[2] = rho * v;
    U[3] = rho * z;
    U[4] = rho * ((0.5 * vsq) + eos_e(m, rho, p) + (m.e_form * (lam - 1.0)));
    U[5] = rho * lam;
}
template <class T>
void c2p(const P &m, const T *U, T *w)
{
    T rho = U[0];
    T u = U[1] / rho, v = U[2] / rho, z = U[3] / rho;
    T lam = U[5] / rho;
    T vsq = (u * u) + (v * v) + (z * z);
    T e = (U[4] / rho) - (0.5 * vsq) - (m.e_form * (lam - 1.0));
    w[0] = rho;
    w[1] = u;
    w[2] = v;
    w[3] = z;
    w[4] = eos_p(m, rho, e);
    w[5] = lam;
}
template <class T>
void flux(const P &m, int dir, const T *U, T *F)
{
    T rho = U[0];
    T u = U[1] / rho, v = U[2] / rho, z = U[3] / rho;
    T lam = U[5] / rho;
    T vsq = (u * u) + (v * v) + (z * z);
    T et = U[4] / rho;
    T e = et - (0.5 * vsq) - (m.e_form * (lam - 1.0));
    T p = eos_p(m, rho, e);
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

template <class T>
T rcoef(T r, double bias) { return 2.0 / ((1.0 - bias) + ((1.0 + bias) * r)); }
template <class T>
T tmin(T a, T b) { return note(val(b) < val(a)) ? b : a; }

template <class T>
T limiter_value(T r, double bias, int limiter)
{
    if (note(val(r) < 0.0)) return T(0.0);
    if (limiter == 0) {
        if (note(val(r) < 0.5)) return 2.0 * r;
        if (note(val(r) < 1.0)) return T(1.0);
        return tmin(tmin(r, rcoef(r, bias)), T(2.0));
    }
    if (limiter == 1) return tmin((2.0 * r) / (1.0 + r), rcoef(r, bias));
    if (note(val(r) < 1.0)) return r;
    return tmin(T(1.0), rcoef(r, bias));
}

template <class T>
void evolve(const P &m, int dir, const T *UL, const T *UM, const T *UR, double bias, int limiter, double half_c,
            T *left, T *right)
{
    T xl[NF], xr[NF], Fl[NF], Fr[NF];
    for (int k = 0; k < NF; k++) {
        T dl = UM[k] - UL[k];
        T dr = UR[k] - UM[k];
        T d = ((0.5 * (1.0 + bias)) * dl) + ((0.5 * (1.0 - bias)) * dr);
        T num = dl, den = dr;
        if (note(fabs_v(num) < 1e-5)) num = T(1e-5);
        if (note(fabs_v(den) < 1e-5)) den = T(1e-5);
        d = d * limiter_value(num / den, bias, limiter);
        xl[k] = UM[k] - (0.5 * d);
        xr[k] = UM[k] + (0.5 * d);
    }
    flux(m, dir, xl, Fl);
    flux(m, dir, xr, Fr);
    for (int k = 0; k < NF; k++) {
        T evo = half_c * (Fl[k] - Fr[k]);
        left[k] = xl[k] + evo;
        right[k] = xr[k] + evo;
    }
}

// FORCE flux of two evolved conserved states, each taken through
// setConservedVariableVector and the state-overload flux (p2c(c2p(X))).
template <class T>
void force(const P &m, int dir, const T *Xa, const T *Xb, double dx, double ts, T *F)
{
    T wa[NF], wb[NF], Ua[NF], Ub[NF], Fa[NF], Fb[NF], us[NF], ri[NF];
    c2p(m, Xa, wa);
    p2c(m, wa, Ua);
    flux(m, dir, Ua, Fa);
    c2p(m, Xb, wb);
    p2c(m, wb, Ub);
    flux(m, dir, Ub, Fb);
    for (int k = 0; k < NF; k++) {
        T lf = 0.5 * ((Fa[k] + Fb[k]) + ((dx / ts) * (Ua[k] - Ub[k])));
        us[k] = 0.5 * ((Ua[k] + Ub[k]) + ((ts / dx) * (Fa[k] - Fb[k])));
        F[k] = lf;   // hold LF until the Richtmyer flux is known
    }
    flux(m, dir, us, ri);
    for (int k = 0; k < NF; k++) F[k] = 0.5 * (F[k] + ri[k]);
}

int gidx(int k, int n, int bc)
{
    if (k >= 0 && k < n) return k;
    if (bc == 1) return k < 0 ? k + n : k - n;
    return k < 0 ? -1 - k : 2 * n - 1 - k;
}

// One SLIC sweep of a line of n cells (stride `st` in the AoS arrays).
template <class T>
void sweep_line(const P &m, int dir, int n, size_t st, const T *in, T *out, double dx, double ts, double bias,
                int limiter, int bc)
{
    std::vector<T> U((size_t)(n + 4) * NF), L((size_t)(n + 4) * NF), R((size_t)(n + 4) * NF), F((size_t)(n + 1) * NF);
    for (int k = -2; k < n + 2; k++) p2c(m, in + (size_t)gidx(k, n, bc) * st * NF, &U[(size_t)(k + 2) * NF]);
    const double half_c = 0.5 * (ts / dx);
    for (int k = -1; k < n + 1; k++)
        evolve(m, dir, &U[(size_t)(k + 1) * NF], &U[(size_t)(k + 2) * NF], &U[(size_t)(k + 3) * NF], bias, limiter,
               half_c, &L[(size_t)(k + 2) * NF], &R[(size_t)(k + 2) * NF]);
    for (int f = 0; f <= n; f++)   // face f - 1/2 between cells f-1 and f
        force(m, dir, &R[(size_t)(f + 1) * NF], &L[(size_t)(f + 2) * NF], dx, ts, &F[(size_t)f * NF]);
    for (int j = 0; j < n; j++) {
        T N[NF];
        for (int k = 0; k < NF; k++)
            N[k] = U[(size_t)(j + 2) * NF + k] + ((ts / dx) * (F[(size_t)j * NF + k] - F[(size_t)(j + 1) * NF + k]));
        c2p(m, N, out + (size_t)j * st * NF);
    }
}

template <class T>
void step(const P &m, int dim, int ny, int nx, double dx, double dt, double bias, int limiter, int bc,
          std::vector<T> &a, std::vector<T> &b)
{
    if (dim == 1) {
        sweep_line(m, 0, nx, 1, a.data(), b.data(), dx, dt, bias, limiter, bc);
        a.swap(b);
        return;
    }
    const double half = 0.5 * dt;
    for (int i = 0; i < ny; i++)
        sweep_line(m, 0, nx, 1, a.data() + (size_t)i * nx * NF, b.data() + (size_t)i * nx * NF, dx, half, bias,
                   limiter, bc);
    for (int j = 0; j < nx; j++)
        sweep_line(m, 1, ny, nx, b.data() + (size_t)j * NF, a.data() + (size_t)j * NF, dx, dt, bias, limiter, bc);
    for (int i = 0; i < ny; i++)
        sweep_line(m, 0, nx, 1, a.data() + (size_t)i * nx * NF, b.data() + (size_t)i * nx * NF, dx, half, bias,
                   limiter, bc);
    a.swap(b);
}

}  // namespace

extern "C" {

/* SLIC solve with prescribed time steps: primal (prim_out) and, when tangent_in
 * is not NULL, the directional derivative (tangent_out) along tangent_in. */
void orc_euler_fixed_dt_jvp(int dim, int ny, int nx, double dx, const double *dts, int n_steps, double bias,
                            int limiter, int bc, const orc_euler_params *mp, const double *prim_in,
                            const double *tangent_in, double *prim_out, double *tangent_out)
{
    const P m = {mp->gamma, mp->p_inf, mp->e_form};
    if (dim == 1) ny = 1;
    g_branches = 0xcbf29ce484222325ull;
    const size_t cells = (size_t)ny * nx;
    if (!tangent_in) {
        std::vector<double> a(prim_in, prim_in + NF * cells), b(NF * cells);
        for (int s = 0; s < n_steps; s++) step(m, dim, ny, nx, dx, dts[s], bias, limiter, bc, a, b);
        std::memcpy(prim_out, a.data(), sizeof(double) * NF * cells);
        return;
    }
    std::vector<Dual> a(NF * cells), b(NF * cells);
    for (size_t k = 0; k < NF * cells; k++) a[k] = Dual(prim_in[k], tangent_in[k]);
    for (int s = 0; s < n_steps; s++) step(m, dim, ny, nx, dx, dts[s], bias, limiter, bc, a, b);
    for (size_t k = 0; k < NF * cells; k++) {
        prim_out[k] = a[k].v;
        tangent_out[k] = a[k].d;
    }
}

/* Hash of the branch decisions of the last orc_euler_fixed_dt_jvp call. */
uint64_t orc_euler_branch_hash(void) { return g_branches; }

}  // extern "C"
