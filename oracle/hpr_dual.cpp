// Forward-mode (dual-number) restatement of the oracle's HPR second-order
// step -- TEST INFRASTRUCTURE ONLY (oracle).
//
// The same arithmetic as hpr_oracle.c + scheme_oracle.c (HPRStateVector,
// HPREquationOfState, the stiffened-gas HPRMieGruneisen, HPRSourceTerms:
// hprstatevector.cpp:59-231, 312-366, hprequationofstate.cpp:7-81,
// hprmiegruneisen.cpp:7-263, hprsourceterms.cpp:7-39; the SLIC sweep of
// HPRSecondOrderSolver, hprsecondordersolver.cpp:7-17, 314-352, 502-546, over
// Solvers/solvers.cpp:454-466, 550-565, 601-609 and firstordersolver.cpp:79-97,
// 165-171, 227-230, 428-433; the RK4 source sub-cycles of HPRForcingSolver,
// hprforcingsolver.cpp:7-44, 134-192, 332-356, 394-417), templated on the scalar
// type.  With T = double it reproduces orc_hpr_solve bit for bit (checked by
// the tests); with T = Dual it gives the exact Jacobian-vector product of the
// step at fixed time steps, along a direction in the initial state and in the
// two relaxation times tau1 (strain dissipation) and tau2 (thermal impulse) --
// the independent check of the GPU's HPR adjoint.  Branches (slope clamps,
// limiter cases) follow the primal values and are hashed (dual.h).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "dual.h"
#include "oracle.h"

namespace {

using namespace orc_dual;

const int NF = 17;

struct HP {
    int is_thermal;
    double rho0, t_ref, t0, cv, gamma, p_inf, cs, alpha;
    double tau1, tau2, dtau1, dtau2;   // relaxation times and their tangents
};

template <class T>
using M3 = T[3][3];

template <class T>
void load_A(const T *s, M3<T> &A)
{
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) A[i][j] = s[2 + 3 * i + j];
}
template <class T>
void matmul(const M3<T> &a, const M3<T> &b, M3<T> &c)
{
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            T s = T(0.0);
            for (int k = 0; k < 3; k++) s = s + a[i][k] * b[k][j];
            c[i][j] = s;
        }
}
template <class T>
void transpose(const M3<T> &a, M3<T> &t)
{
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) t[j][i] = a[i][j];
}
template <class T>
T det3(const M3<T> &m)
{
    return (m[0][0] * ((m[1][1] * m[2][2]) - (m[2][1] * m[1][2]))) - (m[0][1] * ((m[1][0] * m[2][2]) - (m[1][2] * m[2][0]))) +
           (m[0][2] * ((m[1][0] * m[2][1]) - (m[1][1] * m[2][0])));
}
template <class T>
void gram_dev(const M3<T> &A, M3<T> &dev)
{
    M3<T> At, G;
    transpose(A, At);
    matmul(At, A, G);
    T tr = T(0.0);
    for (int i = 0; i < 3; i++) tr = tr + G[i][i];
    T s = tr / 3.0;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) dev[i][j] = G[i][j] + (-1.0 * (s * (i == j ? 1.0 : 0.0)));
}

double mg_gamma(const HP &m) { return m.gamma - 1.0; }
template <class T>
T mg_e(const HP &m, T rho, T p) { return (m.p_inf / rho) + ((p - (-m.p_inf)) / (rho * mg_gamma(m))); }
template <class T>
T mg_p(const HP &m, T rho, T e) { return ((e - (m.p_inf / rho)) * rho * mg_gamma(m)) + (-m.p_inf); }
template <class T>
T mg_T(const HP &m, T rho, T p) { return m.t_ref + ((p - (-m.p_inf)) / (rho * mg_gamma(m) * m.cv)); }

template <class T>
T e_dist(const HP &m, const M3<T> &A)
{
    M3<T> dev;
    gram_dev(A, dev);
    T total = T(0.0);
    for (int i = 0; i < 3; i++) {
        T row = T(0.0);
        for (int j = 0; j < 3; j++) row = row + dev[i][j] * dev[i][j];
        total = total + row;
    }
    return (0.25 * (m.cs * m.cs)) * total;
}
template <class T>
T e_therm(const HP &m, T jx, T jy, T jz) { return (0.5 * (m.alpha * m.alpha)) * ((jx * jx) + (jy * jy) + (jz * jz)); }
template <class T>
T e_macro(T u, T v, T w) { return 0.5 * ((u * u) + (v * v) + (w * w)); }

template <class T>
T total_energy(const HP &m, T rho, T p, T u, T v, T w, const M3<T> &A, T jx, T jy, T jz)
{
    T E = mg_e(m, rho, p) + e_dist(m, A) + e_macro(u, v, w);
    if (m.is_thermal) E = E + e_therm(m, jx, jy, jz);
    return E;
}
template <class T>
T pressure_of(const HP &m, T rho, T E, T u, T v, T w, const M3<T> &A, T jx, T jy, T jz)
{
    T e = E - (e_macro(u, v, w) + e_dist(m, A));
    if (m.is_thermal) e = e - e_therm(m, jx, jy, jz);
    return mg_p(m, rho, e);
}
template <class T>
void e_A(const HP &m, const M3<T> &A, M3<T> &out)
{
    M3<T> dev, Ad;
    gram_dev(A, dev);
    matmul(A, dev, Ad);
    const double cs2 = m.cs * m.cs;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) out[i][j] = cs2 * Ad[i][j];
}
template <class T>
void shear(const HP &m, T rho, const M3<T> &A, M3<T> &sig)
{
    M3<T> EA, At, P;
    e_A(m, A, EA);
    transpose(A, At);
    matmul(At, EA, P);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) sig[i][j] = (-rho) * P[i][j];
}

template <class T>
void p2c(const HP &m, const T *w, T *U)
{
    M3<T> A;
    load_A(w, A);
    T rho = w[0];
    U[0] = rho;
    U[1] = rho * total_energy(m, rho, w[1], w[11], w[12], w[13], A, w[14], w[15], w[16]);
    for (int k = 2; k < 11; k++) U[k] = w[k];
    U[11] = rho * w[11];
    U[12] = rho * w[12];
    U[13] = rho * w[13];
    for (int k = 14; k < 17; k++) U[k] = m.is_thermal ? rho * w[k] : w[k];
}
template <class T>
void c2p(const HP &m, const T *U, T *w)
{
    M3<T> A;
    load_A(U, A);
    T rho = U[0];
    T u = U[11] / rho, v = U[12] / rho, z = U[13] / rho;
    T jx = U[14], jy = U[15], jz = U[16];
    if (m.is_thermal) {
        jx = U[14] / rho;
        jy = U[15] / rho;
        jz = U[16] / rho;
    }
    T E = U[1] / rho;
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
template <class T>
void flux(const HP &m, int dir, const T *U, T *F)
{
    M3<T> A, sig;
    load_A(U, A);
    T rho = U[0];
    shear(m, rho, A, sig);
    T u = U[11] / rho, v = U[12] / rho, z = U[13] / rho;
    T jx = m.is_thermal ? U[14] / rho : T(0.0);
    T jy = m.is_thermal ? U[15] / rho : T(0.0);
    T jz = m.is_thermal ? U[16] / rho : T(0.0);
    T E = U[1] / rho;
    T p = pressure_of(m, rho, E, u, v, z, A, jx, jy, jz);
    for (int k = 0; k < NF; k++) F[k] = T(0.0);
    T vn = (dir == 0) ? u : v;
    F[0] = rho * vn;
    F[1] = (rho * (vn * E)) + (p * vn);
    F[11] = rho * (vn * u);
    F[12] = rho * (vn * v);
    F[13] = rho * (vn * z);
    F[11 + dir] = (rho * (vn * vn)) + p;
    T Av[3];
    for (int i = 0; i < 3; i++) {
        T s = T(0.0);
        s = s + A[i][0] * u;
        s = s + A[i][1] * v;
        s = s + A[i][2] * z;
        Av[i] = s;
    }
    T sv = T(0.0);
    sv = sv + sig[dir][0] * u;
    sv = sv + sig[dir][1] * v;
    sv = sv + sig[dir][2] * z;
    F[1] = F[1] - sv;
    for (int i = 0; i < 3; i++) F[11 + i] = F[11 + i] - sig[dir][i];
    for (int i = 0; i < 3; i++) F[2 + 3 * i + dir] = Av[i];
    if (m.is_thermal) {
        T Tt = mg_T(m, rho, p);
        const double a2 = m.alpha * m.alpha;
        T q = Tt * (a2 * (dir == 0 ? jx : jy));
        F[1] = F[1] + q;
        F[14] = rho * (vn * jx);
        F[15] = rho * (vn * jy);
        F[16] = rho * (vn * jz);
        F[14 + dir] = (rho * (vn * (dir == 0 ? jx : jy))) + Tt;
    }
}
template <class T>
void source(const HP &m, const T *U, T *S)
{
    M3<T> A, EA;
    load_A(U, A);
    T rho = U[0];
    for (int k = 0; k < NF; k++) S[k] = T(0.0);
    e_A(m, A, EA);
    const double cs2 = m.cs * m.cs;
    const T tau1 = make<T>(m.tau1, m.dtau1), tau2 = make<T>(m.tau2, m.dtau2);
    T th1 = 3.0 * (pow_t(det3(A), (5.0 / 3.0)) / (cs2 * tau1));
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) S[2 + 3 * i + j] = (-EA[i][j]) * th1;
    if (m.is_thermal) {
        T E = U[1] / rho;
        T u = U[11] / rho, v = U[12] / rho, z = U[13] / rho;
        T jx = U[14] / rho, jy = U[15] / rho, jz = U[16] / rho;
        T p = pressure_of(m, rho, E, u, v, z, A, jx, jy, jz);
        T Tt = mg_T(m, rho, p);
        const double a2 = m.alpha * m.alpha;
        T ej[3] = {a2 * jx, a2 * jy, a2 * jz};
        T th2 = 1.0 / (a2 * tau2 * ((rho / m.rho0) * (m.t0 / Tt)));
        for (int i = 0; i < 3; i++) S[14 + i] = (-rho) * th2 * ej[i];
    }
}

// ---- the generic scheme over T (scheme_oracle.c) -----------------------------------
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
void slope(const T *UL, const T *UM, const T *UR, double bias, int limiter, T *d)
{
    for (int k = 0; k < NF; k++) {
        T dl = UM[k] - UL[k];
        T dr = UR[k] - UM[k];
        d[k] = ((0.5 * (1.0 + bias)) * dl) + ((0.5 * (1.0 - bias)) * dr);
        T num = dl, den = dr;
        if (note(fabs_v(num) < 1e-5)) num = T(1e-5);
        if (note(fabs_v(den) < 1e-5)) den = T(1e-5);
        d[k] = d[k] * limiter_value(num / den, bias, limiter);
    }
}
// orc_evolve_half: one side's half-step extrapolation, as a primitive state.
template <class T>
void evolve_half(const HP &m, int dir, const T *wl, const T *wm, const T *wr, double dx, double ts, double bias,
                 int limiter, int side, T *out)
{
    T Ul[NF], Um[NF], Ur[NF], d[NF], xl[NF], xr[NF], Fl[NF], Fr[NF], U[NF];
    p2c(m, wl, Ul);
    p2c(m, wm, Um);
    p2c(m, wr, Ur);
    slope(Ul, Um, Ur, bias, limiter, d);
    for (int k = 0; k < NF; k++) {
        xl[k] = Um[k] - (0.5 * d[k]);
        xr[k] = Um[k] + (0.5 * d[k]);
    }
    flux(m, dir, xl, Fl);
    flux(m, dir, xr, Fr);
    const double c = 0.5 * (ts / dx);
    for (int k = 0; k < NF; k++) U[k] = (side == 0 ? xl[k] : xr[k]) + (c * (Fl[k] - Fr[k]));
    c2p(m, U, out);
}
// orc_force_flux between two primitive states.
template <class T>
void force_flux(const HP &m, int dir, const T *wl, const T *wr, double dx, double ts, T *F)
{
    T Ul[NF], Ur[NF], Fl[NF], Fr[NF], lf[NF], us[NF], ri[NF];
    p2c(m, wl, Ul);
    p2c(m, wr, Ur);
    flux(m, dir, Ul, Fl);
    flux(m, dir, Ur, Fr);
    const double lc = dx / ts, rc = ts / dx;
    for (int k = 0; k < NF; k++) {
        lf[k] = 0.5 * ((Fl[k] + Fr[k]) + (lc * (Ul[k] - Ur[k])));
        us[k] = 0.5 * ((Ul[k] + Ur[k]) + (rc * (Fl[k] - Fr[k])));
    }
    flux(m, dir, us, ri);
    for (int k = 0; k < NF; k++) F[k] = 0.5 * (lf[k] + ri[k]);
}
template <class T>
void slic_flux(const HP &m, int dir, const T *wa, const T *wb, const T *wc, const T *wd, double dx, double ts,
               double bias, int limiter, T *F)
{
    T er[NF], el[NF];
    evolve_half(m, dir, wa, wb, wc, dx, ts, bias, limiter, 1, er);
    evolve_half(m, dir, wb, wc, wd, dx, ts, bias, limiter, 0, el);
    force_flux(m, dir, er, el, dx, ts, F);
}

int gidx(int k, int n, int bc)
{
    if (k >= 0 && k < n) return k;
    if (bc == 1) return k < 0 ? k + n : k - n;
    return k < 0 ? -1 - k : 2 * n - 1 - k;
}

// orc_sweep (second order) over a 1D (ny = 1) or 2D grid.
template <class T>
void sweep(const HP &m, int dir, int ny, int nx, double dx, double ts, double bias, int limiter, int bc,
           const std::vector<T> &in, std::vector<T> &out)
{
    const int n = (dir == 0) ? nx : ny, lines = (dir == 0) ? ny : nx;
    for (int l = 0; l < lines; l++) {
        auto at = [&](int k) -> const T * {
            const int kk = gidx(k, n, bc);
            return &in[(size_t)NF * (dir == 0 ? (size_t)l * nx + kk : (size_t)kk * nx + l)];
        };
        for (int j = 0; j < n; j++) {
            T Fm[NF], Fp[NF], U[NF];
            slic_flux(m, dir, at(j - 2), at(j - 1), at(j), at(j + 1), dx, ts, bias, limiter, Fm);
            slic_flux(m, dir, at(j - 1), at(j), at(j + 1), at(j + 2), dx, ts, bias, limiter, Fp);
            p2c(m, at(j), U);
            const double c = ts / dx;
            T N[NF];
            for (int k = 0; k < NF; k++) N[k] = U[k] + (c * (Fm[k] - Fp[k]));
            c2p(m, N, &out[(size_t)NF * (dir == 0 ? (size_t)l * nx + j : (size_t)j * nx + l)]);
        }
    }
}

// evolve_fractional of scheme_oracle.c.
template <class T>
void evolve_fractional(const HP &m, int dir, double frac, const T *wl, const T *wm, const T *wr, double dx, double ts,
                       double bias, int limiter, T *out)
{
    T Ul[NF], Um[NF], Ur[NF], d[NF], xl[NF], xr[NF], Fl[NF], Fr[NF], N[NF];
    p2c(m, wl, Ul);
    p2c(m, wm, Um);
    p2c(m, wr, Ur);
    slope(Ul, Um, Ur, bias, limiter, d);
    for (int k = 0; k < NF; k++) {
        xl[k] = Um[k] - (0.5 * d[k]);
        xr[k] = Um[k] + (0.5 * d[k]);
    }
    flux(m, dir, xl, Fl);
    flux(m, dir, xr, Fr);
    const double c = frac * (ts / dx);
    for (int k = 0; k < NF; k++) N[k] = Um[k] + (c * (Fl[k] - Fr[k]));
    c2p(m, N, out);
}

// orc_forcing_cell: RK4 of the source with fractional predictors.
template <class T>
void forcing_cell(const HP &m, int dim, const T *Ul_, const T *Um, const T *Ur_, const T *Ut_, const T *Ub_, double dx,
                  double ts, double bias, int limiter, T *Unew)
{
    T wl[NF], wr[NF], wt[NF], wb[NF], wm[NF], e1[NF], e2[NF], e3[NF];
    T k[4][NF], S[NF], Us[NF], Ue[NF];
    c2p(m, Ul_, wl);
    c2p(m, Ur_, wr);
    if (dim == 2) {
        c2p(m, Ut_, wt);
        c2p(m, Ub_, wb);
    }
    static const double fx2[3] = {0.25, 0.25, 0.5}, fy2[3] = {0.5, 0.5, 1.0}, fx1[3] = {0.5, 0.5, 1.0};
    source(m, Um, S);
    for (int q = 0; q < NF; q++) k[0][q] = ts * S[q];
    for (int st = 0; st < 3; st++) {
        for (int q = 0; q < NF; q++) Us[q] = Um[q] + (st < 2 ? (0.5 * k[st][q]) : k[st][q]);
        c2p(m, Us, wm);
        if (dim == 1) {
            evolve_fractional(m, 0, fx1[st], wl, wm, wr, dx, ts, bias, limiter, e3);
        } else {
            evolve_fractional(m, 0, fx2[st], wl, wm, wr, dx, ts, bias, limiter, e1);
            evolve_fractional(m, 1, fy2[st], wt, e1, wb, dx, ts, bias, limiter, e2);
            evolve_fractional(m, 0, fx2[st], wl, e2, wr, dx, ts, bias, limiter, e3);
        }
        p2c(m, e3, Ue);
        source(m, Ue, S);
        for (int q = 0; q < NF; q++) k[st + 1][q] = ts * S[q];
    }
    for (int q = 0; q < NF; q++)
        Unew[q] = Um[q] + ((1.0 / 6.0) * (k[0][q] + ((2.0 * k[1][q]) + ((2.0 * k[2][q]) + k[3][q]))));
}

// orc_forcing_step: one sub-cycle over the grid.
template <class T>
void forcing_step(const HP &m, int dim, int ny, int nx, int bc, double dx, double ts, double bias, int limiter,
                  const std::vector<T> &in, std::vector<T> &out)
{
    const size_t cells = (size_t)ny * nx;
    std::vector<T> U(NF * cells);
    for (size_t c = 0; c < cells; c++) p2c(m, &in[NF * c], &U[NF * c]);
    for (int i = 0; i < ny; i++)
        for (int j = 0; j < nx; j++) {
            const size_t c = (size_t)i * nx + j;
            const T *L = &U[NF * ((size_t)i * nx + gidx(j - 1, nx, bc))];
            const T *R = &U[NF * ((size_t)i * nx + gidx(j + 1, nx, bc))];
            const T *Tp = nullptr, *Bt = nullptr;
            if (dim == 2) {
                Tp = &U[NF * ((size_t)gidx(i - 1, ny, bc) * nx + j)];
                Bt = &U[NF * ((size_t)gidx(i + 1, ny, bc) * nx + j)];
            }
            T N[NF];
            forcing_cell(m, dim, L, &U[NF * c], R, Tp, Bt, dx, ts, bias, limiter, N);
            c2p(m, N, &out[NF * c]);
            if (!m.is_thermal)
                for (int q = 14; q < 17; q++) out[NF * c + q] = in[NF * c + q];
        }
}

// One step of orc_solve_sub at a prescribed dt.
template <class T>
void step(const HP &m, int dim, int ny, int nx, double dx, double dt, double bias, int limiter, int bc, int nsub,
          std::vector<T> &a, std::vector<T> &b)
{
    for (int s = 0; s < nsub; s++) {
        forcing_step(m, dim, ny, nx, bc, dx, 0.5 * (dt / nsub), bias, limiter, a, b);
        a.swap(b);
    }
    if (dim == 1) {
        sweep(m, 0, 1, nx, dx, dt, bias, limiter, bc, a, b);
    } else {
        const double half = 0.5 * dt;
        sweep(m, 0, ny, nx, dx, half, bias, limiter, bc, a, b);
        sweep(m, 1, ny, nx, dx, dt, bias, limiter, bc, b, a);
        sweep(m, 0, ny, nx, dx, half, bias, limiter, bc, a, b);
    }
    a.swap(b);
    for (int s = 0; s < nsub; s++) {
        forcing_step(m, dim, ny, nx, bc, dx, 0.5 * (dt / nsub), bias, limiter, a, b);
        a.swap(b);
    }
}

HP to_hp(const orc_hpr_params *p, double dtau1, double dtau2)
{
    HP m;
    m.is_thermal = p->is_thermal;
    m.rho0 = p->rho0;
    m.t_ref = p->t_ref;
    m.t0 = p->t0;
    m.cv = p->cv;
    m.gamma = p->gamma;
    m.p_inf = p->p_inf;
    m.cs = p->cs;
    m.alpha = p->alpha;
    m.tau1 = p->tau1;
    m.tau2 = p->tau2;
    m.dtau1 = dtau1;
    m.dtau2 = dtau2;
    return m;
}

}  // namespace

extern "C" {

/* HPR second-order solve (stiffened gas, not plastic) with prescribed time steps
 * and `subcycles` source sub-cycles per half step: primal (prim_out) and, when
 * tangent_in is not NULL, the directional derivative (tangent_out) along
 * (tangent_in, dtau1, dtau2).  Resets the branch hash. */
void orc_hpr_fixed_dt_jvp(int dim, int ny, int nx, double dx, const double *dts, int n_steps, double bias, int limiter,
                          int bc, int subcycles, const orc_hpr_params *mp, const double *prim_in,
                          const double *tangent_in, double dtau1, double dtau2, double *prim_out, double *tangent_out)
{
    const HP m = to_hp(mp, dtau1, dtau2);
    if (dim == 1) ny = 1;
    reset_branches();
    const size_t cells = (size_t)ny * nx;
    if (!tangent_in) {
        std::vector<double> a(prim_in, prim_in + NF * cells), b(NF * cells);
        for (int s = 0; s < n_steps; s++) step(m, dim, ny, nx, dx, dts[s], bias, limiter, bc, subcycles, a, b);
        std::memcpy(prim_out, a.data(), sizeof(double) * NF * cells);
        return;
    }
    std::vector<Dual> a(NF * cells), b(NF * cells);
    for (size_t k = 0; k < NF * cells; k++) a[k] = Dual(prim_in[k], tangent_in[k]);
    for (int s = 0; s < n_steps; s++) step(m, dim, ny, nx, dx, dts[s], bias, limiter, bc, subcycles, a, b);
    for (size_t k = 0; k < NF * cells; k++) {
        prim_out[k] = a[k].v;
        tangent_out[k] = a[k].d;
    }
}

/* Jacobian-vector products of the HPR unit functions, per record of a
 * primitive state w (17) and a direction v (17): out holds 5 x 17 doubles,
 *   p2c'(w) v | c2p'(U) v | Fx'(U) v | Fy'(U) v | S'(U) v (+ dtau1, dtau2)
 * with U = p2c(w) (v taken in conserved variables for the last four). */
void orc_hpr_jvp_units(int n, const orc_hpr_params *mp, double dtau1, double dtau2, const double *w,
                       const double *v, double *out)
{
    const HP m = to_hp(mp, dtau1, dtau2);
    for (int r = 0; r < n; r++) {
        const double *wr = w + (size_t)NF * r, *vr = v + (size_t)NF * r;
        double *o = out + (size_t)5 * NF * r;
        Dual wd[NF], y[NF], Ud[NF];
        double U[NF];
        for (int k = 0; k < NF; k++) wd[k] = Dual(wr[k], vr[k]);
        p2c(m, wd, y);
        for (int k = 0; k < NF; k++) o[k] = y[k].d;
        p2c(m, wr, U);
        for (int k = 0; k < NF; k++) Ud[k] = Dual(U[k], vr[k]);
        c2p(m, Ud, y);
        for (int k = 0; k < NF; k++) o[NF + k] = y[k].d;
        flux(m, 0, Ud, y);
        for (int k = 0; k < NF; k++) o[2 * NF + k] = y[k].d;
        flux(m, 1, Ud, y);
        for (int k = 0; k < NF; k++) o[3 * NF + k] = y[k].d;
        source(m, Ud, y);
        for (int k = 0; k < NF; k++) o[4 * NF + k] = y[k].d;
    }
}

}  // extern "C"
