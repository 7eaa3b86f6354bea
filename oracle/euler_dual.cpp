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

#include "dual.h"
#include "oracle.h"

namespace {

using namespace orc_dual;

const int NF = 6;

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

// ---- RGFM (rgfm_oracle.c restated over T) ---------------------------------------
// The HLLC interface state, the ghost fill and the step of the oracle's RGFM
// driver with materials in T.  The level set stays in doubles and reads the
// value parts of the velocities: it only decides where the ghost fill writes,
// so its own derivative never reaches the materials.

struct E {
    double gamma, p_inf, e_form;
};

template <class T>
T snd_t(const T *w, const E &m) { return sqrt_t((m.gamma * (w[4] + m.p_inf)) / w[0]); }

template <class T>
T pvrs_prod_t(const T *a, const T *b, const E &m1, const E &m2)
{
    return 0.25 * (a[0] + b[0]) * (snd_t(a, m1) + snd_t(b, m2));
}
template <class T>
T pvrs_vel_t(int dir, const T *a, const T *b, const E &m1, const E &m2)
{
    T prod = pvrs_prod_t(a, b, m1, m2);
    return (0.5 * (a[1 + dir] + b[1 + dir])) + (0.5 * (a[4] - b[4]) / prod);
}
template <class T>
T pvrs_pres_t(int dir, const T *a, const T *b, const E &m1, const E &m2)
{
    T prod = pvrs_prod_t(a, b, m1, m2);
    return (0.5 * (a[4] + b[4])) + (0.5 * (a[1 + dir] - b[1 + dir]) * prod);
}
template <class T>
T two_rare_vel_t(int dir, const T *a, const T *b, const E &m1, const E &m2)
{
    T ca = snd_t(a, m1), cb = snd_t(b, m2);
    double g = 0.5 * (m1.gamma + m2.gamma);
    double ex = (g - 1.0) / (2.0 * g);
    T pr = pow_t((a[4] / b[4]), ex);
    double co = 2.0 / (g - 1.0);
    return ((pr * (a[1 + dir] / ca)) + (b[1 + dir] / cb) + (co * (pr - 1.0))) / ((pr / ca) + (1.0 / cb));
}
template <class T>
T two_rare_pres_t(int dir, const T *a, const T *b, const E &m1, const E &m2)
{
    T ca = snd_t(a, m1), cb = snd_t(b, m2);
    double g = 0.5 * (m1.gamma + m2.gamma);
    double co = 0.5 * (g - 1.0);
    T us = two_rare_vel_t(dir, a, b, m1, m2);
    double ex = 2.0 * (g / (g - 1.0));
    T ld = 1.0 + (co * ((a[1 + dir] - us) / ca));
    T rd = 1.0 + (co * ((us - b[1 + dir]) / cb));
    return 0.5 * ((a[4] * pow_t(ld, ex)) + (b[4] * pow_t(rd, ex)));
}
template <class T>
void two_shock_coeffs_t(int dir, const T *a, const T *b, const E &m1, const E &m2, T &ga, T &gb)
{
    double g = 0.5 * (m1.gamma + m2.gamma);
    T ppv = pvrs_pres_t(dir, a, b, m1, m2);
    double A = 2.0 / (g + 1.0);
    double B = (g - 1.0) / (g + 1.0);
    ga = sqrt_t((A / a[0]) / ((B * a[4]) + ppv));
    gb = sqrt_t((A / b[0]) / ((B * b[4]) + ppv));
}
template <class T>
T two_shock_pres_t(int dir, const T *a, const T *b, const E &m1, const E &m2)
{
    T ga, gb;
    two_shock_coeffs_t(dir, a, b, m1, m2, ga, gb);
    return ((ga * a[4]) + (gb * b[4]) - (b[1 + dir] - a[1 + dir])) / (ga + gb);
}
template <class T>
T two_shock_vel_t(int dir, const T *a, const T *b, const E &m1, const E &m2)
{
    T ps = two_shock_pres_t(dir, a, b, m1, m2);
    T ga, gb;
    two_shock_coeffs_t(dir, a, b, m1, m2, ga, gb);
    return (0.5 * (a[1 + dir] + b[1 + dir])) + (0.5 * ((gb * (ps - b[4])) - (ga * (ps - a[4]))));
}
bool use_pvrs_t(double pa, double pb, double ppv, const E &m1, const E &m2)
{
    double pmin = (pb < pa) ? pb : pa, pmax = (pa < pb) ? pb : pa;
    double ratio = pmax / pmin;
    return note(((ratio <= 2.0) && (pmin <= ppv) && (ppv <= pmax)) || std::fabs(m1.gamma - m2.gamma) > 1e-8 ||
                m1.p_inf > 0.0 || m2.p_inf > 0.0);
}
template <class T>
T star_vel_t(int dir, const T *a, const T *b, const E &m1, const E &m2)
{
    T ppv = (dir == 0) ? pvrs_pres_t(0, a, b, m1, m2) : pvrs_vel_t(1, a, b, m1, m2);
    if (use_pvrs_t(val(a[4]), val(b[4]), val(ppv), m1, m2)) return pvrs_vel_t(dir, a, b, m1, m2);
    if (note(val(ppv) < ((val(b[4]) < val(a[4])) ? val(b[4]) : val(a[4])))) return two_rare_vel_t(dir, a, b, m1, m2);
    return two_shock_vel_t(dir, a, b, m1, m2);
}
template <class T>
T star_pres_t(int dir, const T *a, const T *b, const E &m1, const E &m2)
{
    T ppv = pvrs_pres_t(dir, a, b, m1, m2);
    if (use_pvrs_t(val(a[4]), val(b[4]), val(ppv), m1, m2)) return ppv;
    if (note(val(ppv) < ((val(b[4]) < val(a[4])) ? val(b[4]) : val(a[4])))) return two_rare_pres_t(dir, a, b, m1, m2);
    return two_shock_pres_t(dir, a, b, m1, m2);
}
template <class T>
T weighting_t(T ps, const T *w, const E &m)
{
    T p = w[4];
    if (note(val(ps) <= val(p))) return T(1.0);
    return sqrt_t(1.0 + ((m.gamma + 1.0) / (2.0 * m.gamma)) * ((ps + m.p_inf) / (p + m.p_inf) - 1.0));
}
template <class T>
T star_rho_t(T rho, T S, T u, T Ss) { return rho * ((S - u) / (S - Ss)); }

// HLLCSolver::solveX/Y's state at x/t = 0 (pressure and normal velocity) and
// the two star densities, as interface_state of the oracle.
template <class T>
void interface_t(int dir, const T *a, const T *b, const E &m1, const E &m2, T &p, T &u, T &ra, T &rb)
{
    const P q1 = {m1.gamma, m1.p_inf, m1.e_form}, q2 = {m2.gamma, m2.p_inf, m2.e_form};
    T Ss = star_vel_t(dir, a, b, m1, m2);
    T ps = star_pres_t(dir, a, b, m1, m2);
    T Sa = a[1 + dir] - (weighting_t(ps, a, m1) * snd_t(a, m1));
    T Sb = b[1 + dir] + (weighting_t(ps, b, m2) * snd_t(b, m2));
    if (note(0.0 <= val(Sa))) {
        p = a[4];
        u = a[1 + dir];
    } else if (note(val(Sb) <= 0.0)) {
        p = b[4];
        u = b[1 + dir];
    } else {
        const bool left = note(0.0 <= val(Ss));
        const T *w = left ? a : b;
        const P &q = left ? q1 : q2;
        T S = left ? Sa : Sb;
        T rho = w[0], pw = w[4], un = w[1 + dir];
        T Uw[NF];
        p2c(q, w, Uw);
        T rs = star_rho_t(rho, S, un, Ss);
        T et = (Ss - un) * (Ss + (pw / (rho * (S - un))));
        T U[NF], out[NF];
        U[0] = rs;
        U[1] = rs * (dir == 0 ? Ss : w[1]);
        U[2] = rs * (dir == 1 ? Ss : w[2]);
        U[3] = rs * w[3];
        U[4] = rs * ((Uw[4] / rho) + et);
        U[5] = rs;
        c2p(q, U, out);
        p = out[4];
        u = out[1 + dir];
    }
    ra = star_rho_t(a[0], Sa, a[1 + dir], Ss);
    rb = star_rho_t(b[0], Sb, b[1 + dir], Ss);
}

// fill_line of the oracle over T (interface values from the ORIGINAL states).
template <class T>
void fill_line_t(int dir, int n, size_t stride, const double *phi, size_t pstride, const T *o1, const T *o2, T *n1,
                 T *n2, const E &p1, const E &p2, int in_ghost)
{
    double prev = phi[0];
    for (int k = 1; k < n; k++) {
        double cur = phi[(size_t)k * pstride];
        if (note((cur * prev) <= 0.0) && k > 2 && k < (n - 3)) {
            T p, u, ra, rb;
            const bool g = in_ghost;
            interface_t(dir, (g ? o2 : o1) + (size_t)(k - 1) * stride * NF, (g ? o1 : o2) + (size_t)k * stride * NF,
                        g ? p2 : p1, g ? p1 : p2, p, u, ra, rb);
            T *na = g ? n2 : n1, *nb = g ? n1 : n2;
            for (int j = 0; j < 4; j++) {
                T *wa = na + (size_t)(k + j - 1) * stride * NF, *wb = nb + (size_t)(k - j) * stride * NF;
                wa[0] = ra;
                wa[1 + dir] = u;
                wa[4] = p;
                wb[0] = rb;
                wb[1 + dir] = u;
                wb[4] = p;
            }
            in_ghost = !in_ghost;
        }
        prev = cur;
    }
}

template <class T>
void ghost_fill_t(int dim, int ny, int nx, const E &p1, const E &p2, std::vector<T> &m1, std::vector<T> &m2,
                  const double *phi)
{
    std::vector<T> o1 = m1, o2 = m2;
    if (dim == 1) {
        fill_line_t(0, nx, 1, phi, 1, o1.data(), o2.data(), m1.data(), m2.data(), p1, p2, 0);
        return;
    }
    for (int i = 0; i < ny; i++)
        fill_line_t(0, nx, 1, phi + (size_t)i * nx, 1, o1.data() + (size_t)i * nx * NF, o2.data() + (size_t)i * nx * NF,
                    m1.data() + (size_t)i * nx * NF, m2.data() + (size_t)i * nx * NF, p1, p2, 1);
    o1 = m1;
    o2 = m2;
    for (int j = 0; j < nx; j++)
        fill_line_t(1, ny, (size_t)nx, phi + j, (size_t)nx, o1.data() + (size_t)j * NF, o2.data() + (size_t)j * NF,
                    m1.data() + (size_t)j * NF, m2.data() + (size_t)j * NF, p1, p2, 1);
}

int mirror_i(int k, int n) { return (k < 0) ? -1 - k : (k >= n ? 2 * n - 1 - k : k); }

// level_set_line of the oracle, velocities from the value parts.
template <class T>
void level_set_line_t(int dir, int n, const double *phi, size_t ps, const T *m1, const T *m2, size_t ms, double h,
                      double dt, double *out, int in_ghost)
{
    double prev = phi[0];
    for (int k = 0; k < n; k++) {
        double c = phi[(size_t)k * ps];
        if (note((prev * c) <= 0.0)) in_ghost = !in_ghost;
        double v = val((!in_ghost ? m1 : m2)[(size_t)k * ms * NF + 1 + dir]);
        double up;
        // the upwind side is not hashed: at v = 0 both sides give the same
        // value (v * up = 0), so the level set stays continuous there
        if (v < 0.0) {
            double a = phi[(size_t)mirror_i(k + 2, n) * ps], b = phi[(size_t)mirror_i(k + 1, n) * ps];
            up = (-a + (4.0 * b) - (3.0 * c)) / (2.0 * h);
        } else {
            double a = phi[(size_t)mirror_i(k - 1, n) * ps], b = phi[(size_t)mirror_i(k - 2, n) * ps];
            up = ((3.0 * c) - (4.0 * a) + b) / (2.0 * h);
        }
        out[(size_t)k * ps] = c - (dt * (v * up));
        prev = c;
    }
}

template <class T>
void level_set_t(int dim, int dir, int ny, int nx, double h, double dt, const std::vector<T> &m1,
                 const std::vector<T> &m2, const std::vector<double> &in, std::vector<double> &out)
{
    if (dim == 1) {
        level_set_line_t(0, nx, in.data(), 1, m1.data(), m2.data(), 1, h, dt, out.data(), 0);
    } else if (dir == 0) {
        for (int i = 0; i < ny; i++)
            level_set_line_t(0, nx, in.data() + (size_t)i * nx, 1, m1.data() + (size_t)i * nx * NF,
                             m2.data() + (size_t)i * nx * NF, 1, h, dt, out.data() + (size_t)i * nx, 1);
    } else {
        for (int j = 0; j < nx; j++)
            level_set_line_t(1, ny, in.data() + j, (size_t)nx, m1.data() + (size_t)j * NF, m2.data() + (size_t)j * NF,
                             (size_t)nx, h, dt, out.data() + j, 1);
    }
}

// One SLIC step X(dt/2) Y(dt) X(dt/2) of one material with dx, dy.
template <class T>
void material_step(const P &m, int dim, int ny, int nx, double dx, double dy, double dt, double bias, int limiter,
                   std::vector<T> &a, std::vector<T> &b)
{
    if (dim == 1) {
        sweep_line(m, 0, nx, 1, a.data(), b.data(), dx, dt, bias, limiter, 0);
        a.swap(b);
        return;
    }
    const double half = 0.5 * dt;
    for (int i = 0; i < ny; i++)
        sweep_line(m, 0, nx, 1, a.data() + (size_t)i * nx * NF, b.data() + (size_t)i * nx * NF, dx, half, bias,
                   limiter, 0);
    for (int j = 0; j < nx; j++)
        sweep_line(m, 1, ny, nx, b.data() + (size_t)j * NF, a.data() + (size_t)j * NF, dy, dt, bias, limiter, 0);
    for (int i = 0; i < ny; i++)
        sweep_line(m, 0, nx, 1, a.data() + (size_t)i * nx * NF, b.data() + (size_t)i * nx * NF, dx, half, bias,
                   limiter, 0);
    a.swap(b);
}

template <class T>
void rgfm_run(int dim, int ny, int nx, double dx, double dy, const double *dts, int n_steps, double bias, int limiter,
              const E &p1, const E &p2, std::vector<T> &m1, std::vector<T> &m2, std::vector<double> &phi)
{
    const size_t cells = (size_t)ny * nx;
    const P q1 = {p1.gamma, p1.p_inf, p1.e_form}, q2 = {p2.gamma, p2.p_inf, p2.e_form};
    std::vector<T> b1(NF * cells), b2(NF * cells);
    std::vector<double> phi2(cells);
    for (int s = 0; s < n_steps; s++) {
        const double dt = dts[s];
        ghost_fill_t(dim, ny, nx, p1, p2, m1, m2, phi.data());
        if (dim == 1) {
            level_set_t(1, 0, 1, nx, dx, dt, m1, m2, phi, phi2);
            phi.swap(phi2);
        } else {
            const double half = 0.5 * dt;
            level_set_t(2, 0, ny, nx, dx, half, m1, m2, phi, phi2);
            level_set_t(2, 1, ny, nx, dy, dt, m1, m2, phi2, phi);
            level_set_t(2, 0, ny, nx, dx, half, m1, m2, phi, phi2);
            phi.swap(phi2);
        }
        material_step(q1, dim, ny, nx, dx, dy, dt, bias, limiter, m1, b1);
        material_step(q2, dim, ny, nx, dx, dy, dt, bias, limiter, m2, b2);
    }
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
    reset_branches();
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

/* Hash of the branch decisions of the last dual-number oracle call. */
uint64_t orc_euler_branch_hash(void) { return branch_state(); }

/* HLLC RGFM solve with prescribed time steps (orc_rgfm_solve's step) and, when
 * t1_in is not NULL, the forward-mode derivative of both final material
 * states along (t1_in, t2_in) by dual numbers; the level set is not
 * differentiated (it only places the ghost fill). */
void orc_rgfm_fixed_dt_jvp(int dim, int ny, int nx, double dx, double dy, const double *dts, int n_steps, double bias,
                           int limiter, const orc_euler_params *p1, const orc_euler_params *p2, const double *m1_in,
                           const double *m2_in, const double *phi_in, const double *t1_in, const double *t2_in,
                           double *m1_out, double *m2_out, double *phi_out, double *t1_out, double *t2_out)
{
    if (dim == 1) ny = 1;
    const size_t cells = (size_t)ny * nx;
    const E e1 = {p1->gamma, p1->p_inf, p1->e_form}, e2 = {p2->gamma, p2->p_inf, p2->e_form};
    std::vector<double> phi(phi_in, phi_in + cells);
    reset_branches();
    if (!t1_in) {
        std::vector<double> a1(m1_in, m1_in + NF * cells), a2(m2_in, m2_in + NF * cells);
        rgfm_run(dim, ny, nx, dx, dy, dts, n_steps, bias, limiter, e1, e2, a1, a2, phi);
        std::memcpy(m1_out, a1.data(), sizeof(double) * NF * cells);
        std::memcpy(m2_out, a2.data(), sizeof(double) * NF * cells);
    } else {
        std::vector<Dual> a1(NF * cells), a2(NF * cells);
        for (size_t k = 0; k < NF * cells; k++) {
            a1[k] = Dual(m1_in[k], t1_in[k]);
            a2[k] = Dual(m2_in[k], t2_in[k]);
        }
        rgfm_run(dim, ny, nx, dx, dy, dts, n_steps, bias, limiter, e1, e2, a1, a2, phi);
        for (size_t k = 0; k < NF * cells; k++) {
            m1_out[k] = a1[k].v;
            t1_out[k] = a1[k].d;
            m2_out[k] = a2[k].v;
            t2_out[k] = a2[k].d;
        }
    }
    std::memcpy(phi_out, phi.data(), sizeof(double) * cells);
}

}  // extern "C"
