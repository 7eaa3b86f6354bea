// Adjoint of the HPR second-order step (time steps held fixed): the SLIC
// sweeps (HPRSecondOrderSolver::computeX/YSLICFlux and the sweeps,
// hprsecondordersolver.cpp:7-17, 89-241) and the RK4 relaxation sub-cycles
// (HPRForcingSolver::computeRungeKuttaTimeStep[2D], hprforcingsolver.cpp:7-44,
// 134-192, 332-356, 394-417), for the stiffened-gas, non-plastic HPR state.
//
// Vector-Jacobian products are derived by hand from the state functions
// (hprstatevector.cpp:59-231, 312-366; hprequationofstate.cpp:7-81;
// hprsourceterms.cpp:7-39):
//   * the distortion energy c_s^2/4 |dev(A^T A)|^2 has gradient c_s^2 A dev(A^T A)
//     with respect to A (the E_A of hprderivatives.cpp:17-24);
//   * the stiffened-gas pressure p = (e - p_inf/rho) rho G - p_inf has
//     dp/drho = G e and dp/de = G rho;
//   * the thermal source -rho theta2^-1 alpha^2 J reduces to -(rho0 T / (tau2 T0)) J;
//   * theta1^-1 = 3 det(A)^(5/3) / (c_s^2 tau1) has d/dA = (5/3) theta1^-1 cof(A) / det(A).
// Branch decisions (slope clamps, limiter cases) are recomputed in the
// forward's IEEE arithmetic; linearisation points use refined reciprocals and
// fused multiply-adds, and a conserved -> primitive -> conserved round trip
// counts as the identity.
//
// The sweep adjoint runs as four launches over a grid of cells or faces (one
// thread each, SoA planes, coalesced along x in both directions):
//   1. N-bar = c2p'(N)^T w-bar of every cell's updated conserved vector;
//   2. per face: the FORCE flux's adjoint (Richtmyer state, the two evolved
//      states) -> adjoints of the right evolved state of the cell before and the
//      left one of the cell after;
//   3. per position (ghosts included for transmissive lines): the half-step
//      evolution and slope adjoints -> the cell's own term and its terms for
//      the two neighbours;
//   4. per cell: gather, then p2c'(w)^T.
// The sub-cycle adjoint recomputes each cell's RK4 forward (forcing_cell, the
// forward's own code) with its stage states recorded in HBM, reverses it one
// stage-sized launch at a time, and writes the cell's term and its four
// neighbour terms, gathered by a last launch.  The relaxation-time derivatives
// are summed per block into fixed slots.
#pragma once

#include "vg_adjoint.h"
#include "vg_forcing.h"
#include "vg_hpr_forcing.h"
#include "vg_phys_hpr.h"

namespace vg {
namespace hpradj {

using HP = HPRPhys;
using HPar = HPRPhys::Params;
constexpr int NF = 17;
constexpr int kBS = 128;
constexpr int kPeriodic = 1;   // bc value of periodic lines (ghost_index)

// ---- linearisation point of a conserved vector -----------------------------------
struct ConsPoint {
    double rho, ir, E, e, p;
    double vel[3], j[3];
    double A[3][3], dev[3][3], EA[3][3], G[3][3];
};

__device__ __forceinline__ void cons_point(const HPar &m, const double (&U)[NF], ConsPoint &c)
{
#pragma clang fp contract(fast)
    c.rho = U[0];
    c.ir = rcp_nr(c.rho);
#pragma unroll
    for (int i = 0; i < 3; i++) {
        c.vel[i] = U[11 + i] * c.ir;
        c.j[i] = m.is_thermal ? U[14 + i] * c.ir : U[14 + i];
    }
    c.E = U[1] * c.ir;
    HP::load_A(U, c.A);
    HP::gram_dev(c.A, c.G, c.dev);
    double Ad[3][3];
    HP::mul_nn(c.A, c.dev, Ad);
    double dd = 0.0;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int k = 0; k < 3; k++) {
            c.EA[i][k] = m.cs2 * Ad[i][k];
            dd += c.dev[i][k] * c.dev[i][k];
        }
    double e = c.E - (0.5 * (c.vel[0] * c.vel[0] + c.vel[1] * c.vel[1] + c.vel[2] * c.vel[2]) + 0.25 * m.cs2 * dd);
    if (m.is_thermal) e -= 0.5 * m.a2 * (c.j[0] * c.j[0] + c.j[1] * c.j[1] + c.j[2] * c.j[2]);
    c.e = e;
    c.p = (e - m.p_inf * c.ir) * c.rho * m.G - m.p_inf;
}

// Ab += A (Gb + Gb^T), Gb = devb - tr(devb)/3 I: the adjoint of dev(A^T A).
__device__ __forceinline__ void dev_gram_vjp(const double (&A)[3][3], const double (&devb)[3][3], double (&Ab)[3][3])
{
#pragma clang fp contract(fast)
    const double t = (devb[0][0] + devb[1][1] + devb[2][2]) * (1.0 / 3.0);
    double S[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int k = 0; k < 3; k++) S[i][k] = devb[i][k] + devb[k][i] - (i == k ? 2.0 * t : 0.0);
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int k = 0; k < 3; k++) Ab[i][k] += A[i][0] * S[0][k] + A[i][1] * S[1][k] + A[i][2] * S[2][k];
}

// Adjoint of E_A = c_s^2 A dev(A^T A) given EAb: adds to Ab.
__device__ __forceinline__ void ea_vjp(const HPar &m, const ConsPoint &c, const double (&EAb)[3][3], double (&Ab)[3][3])
{
#pragma clang fp contract(fast)
    double devb[3][3];
#pragma unroll
    for (int l = 0; l < 3; l++)
#pragma unroll
        for (int i = 0; i < 3; i++) {
            // Ab[k][l] += cs2 sum_i EAb[k][i] dev[l][i];  devb[l][i] = cs2 sum_k A[k][l] EAb[k][i]
            devb[l][i] = m.cs2 * (c.A[0][l] * EAb[0][i] + c.A[1][l] * EAb[1][i] + c.A[2][l] * EAb[2][i]);
        }
#pragma unroll
    for (int k = 0; k < 3; k++)
#pragma unroll
        for (int l = 0; l < 3; l++)
            Ab[k][l] += m.cs2 * (EAb[k][0] * c.dev[l][0] + EAb[k][1] * c.dev[l][1] + EAb[k][2] * c.dev[l][2]);
    dev_gram_vjp(c.A, devb, Ab);
}

// setConservedVariableVector: U-bar = c2p'(U)^T w-bar.
__device__ __forceinline__ void vjp_c2p(const HPar &m, const ConsPoint &c, const double (&wb)[NF], double (&Ub)[NF])
{
#pragma clang fp contract(fast)
    double rb = wb[0] + wb[1] * m.G * c.e;
    const double eb = wb[1] * m.G * c.rho;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int k = 0; k < 3; k++) Ub[2 + 3 * i + k] = wb[2 + 3 * i + k] - eb * c.EA[i][k];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const double vb = wb[11 + i] - eb * c.vel[i];
        Ub[11 + i] = vb * c.ir;
        rb -= vb * c.vel[i] * c.ir;
        if (m.is_thermal) {
            const double jb = wb[14 + i] - eb * m.a2 * c.j[i];
            Ub[14 + i] = jb * c.ir;
            rb -= jb * c.j[i] * c.ir;
        } else {
            Ub[14 + i] = wb[14 + i];
        }
    }
    Ub[1] = eb * c.ir;
    rb -= eb * c.E * c.ir;
    Ub[0] = rb;
}

// computeConservedVariableVector: w-bar = p2c'(w)^T U-bar.
__device__ __forceinline__ void vjp_p2c(const HPar &m, const double (&w)[NF], const double (&Ub)[NF], double (&wb)[NF])
{
#pragma clang fp contract(fast)
    const double rho = w[0], p = w[1];
    double A[3][3], G[3][3], dev[3][3], Ad[3][3];
    HP::load_A(w, A);
    HP::gram_dev(A, G, dev);
    HP::mul_nn(A, dev, Ad);
    double dd = 0.0;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int k = 0; k < 3; k++) dd += dev[i][k] * dev[i][k];
    const double ir = rcp_nr(rho), irG = ir * rcp_nr(m.G);
    double E = m.p_inf * ir + (p + m.p_inf) * irG + 0.25 * m.cs2 * dd +
               0.5 * (w[11] * w[11] + w[12] * w[12] + w[13] * w[13]);
    if (m.is_thermal) E += 0.5 * m.a2 * (w[14] * w[14] + w[15] * w[15] + w[16] * w[16]);
    double rb = Ub[0] + Ub[1] * E;
    const double Eb = Ub[1] * rho;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        rb += Ub[11 + i] * w[11 + i];
        wb[11 + i] = Ub[11 + i] * rho + Eb * w[11 + i];
        if (m.is_thermal) {
            rb += Ub[14 + i] * w[14 + i];
            wb[14 + i] = Ub[14 + i] * rho + Eb * m.a2 * w[14 + i];
        } else {
            wb[14 + i] = Ub[14 + i];
        }
    }
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int k = 0; k < 3; k++) wb[2 + 3 * i + k] = Ub[2 + 3 * i + k] + Eb * m.cs2 * Ad[i][k];
    rb += Eb * (-(m.p_inf * ir * ir) - (p + m.p_inf) * ir * irG);
    wb[1] = Eb * irG;
    wb[0] = rb;
}

// computeX/YFluxVector: U-bar = F'(U)^T F-bar.
template <int DIR>
__device__ __forceinline__ void vjp_flux(const HPar &m, const ConsPoint &c, const double (&Fb)[NF], double (&Ub)[NF])
{
#pragma clang fp contract(fast)
    const double rho = c.rho, vn = c.vel[DIR];
    double Prow[3], sig[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        Prow[i] = c.A[0][DIR] * c.EA[0][i] + c.A[1][DIR] * c.EA[1][i] + c.A[2][DIR] * c.EA[2][i];
        sig[i] = -rho * Prow[i];
    }
    double rb = 0.0, Eb = 0.0, pb = 0.0, Tb = 0.0;
    double vb[3] = {0.0, 0.0, 0.0}, jb[3] = {0.0, 0.0, 0.0}, sgb[3] = {0.0, 0.0, 0.0};
    double Ab[3][3];
    // F0 = rho vn
    rb += vn * Fb[0];
    vb[DIR] += rho * Fb[0];
    // F1 = rho vn E + p vn - sig[DIR] . vel (+ T alpha^2 j_DIR)
    rb += vn * c.E * Fb[1];
    vb[DIR] += (rho * c.E + c.p) * Fb[1];
    Eb += rho * vn * Fb[1];
    pb += vn * Fb[1];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        sgb[i] -= c.vel[i] * Fb[1];
        vb[i] -= sig[i] * Fb[1];
    }
    const double ircv = c.ir * rcp_nr(m.G * m.cv);
    const double T = m.t_ref + (c.p + m.p_inf) * ircv;
    if (m.is_thermal) {
        Tb += m.a2 * c.j[DIR] * Fb[1];
        jb[DIR] += T * m.a2 * Fb[1];
    }
    // momentum: rho vn vel_i - sig[DIR][i] + delta p
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const double f = Fb[11 + i];
        rb += vn * c.vel[i] * f;
        vb[DIR] += rho * c.vel[i] * f;
        vb[i] += rho * vn * f;
        sgb[i] -= f;
        if (i == DIR) pb += f;
    }
    // distortion: F[2 + 3i + DIR] = sum_k A[i][k] vel_k
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const double f = Fb[2 + 3 * i + DIR];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            Ab[i][k] = c.vel[k] * f;
            vb[k] += c.A[i][k] * f;
        }
    }
    if (m.is_thermal) {
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const double f = Fb[14 + i];
            rb += vn * c.j[i] * f;
            vb[DIR] += rho * c.j[i] * f;
            jb[i] += rho * vn * f;
            if (i == DIR) Tb += f;
        }
        // T = t_ref + (p + p_inf) / (rho G cv)
        pb += Tb * ircv;
        rb -= Tb * (c.p + m.p_inf) * ircv * c.ir;
    }
    // p(rho, e), e = E - |vel|^2/2 - e_dist(A) - e_therm(j)
    rb += pb * m.G * c.e;
    const double eb = pb * m.G * rho;
    Eb += eb;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        vb[i] -= eb * c.vel[i];
        if (m.is_thermal) jb[i] -= eb * m.a2 * c.j[i];
#pragma unroll
        for (int k = 0; k < 3; k++) Ab[i][k] -= eb * c.EA[i][k];
    }
    // sig[DIR][i] = -rho Prow[i], Prow[i] = sum_k A[k][DIR] EA[k][i]
    double Pb[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        rb -= sgb[i] * Prow[i];
        Pb[i] = -rho * sgb[i];
    }
    double EAb[3][3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        Ab[k][DIR] += Pb[0] * c.EA[k][0] + Pb[1] * c.EA[k][1] + Pb[2] * c.EA[k][2];
#pragma unroll
        for (int i = 0; i < 3; i++) EAb[k][i] = c.A[k][DIR] * Pb[i];
    }
    ea_vjp(m, c, EAb, Ab);
    // back to the conserved variables
    Ub[1] = Eb * c.ir;
    rb -= Eb * c.E * c.ir;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        Ub[11 + i] = vb[i] * c.ir;
        rb -= vb[i] * c.vel[i] * c.ir;
        if (m.is_thermal) {
            Ub[14 + i] = jb[i] * c.ir;
            rb -= jb[i] * c.j[i] * c.ir;
        } else {
            Ub[14 + i] = 0.0;
        }
    }
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int k = 0; k < 3; k++) Ub[2 + 3 * i + k] = Ab[i][k];
    Ub[0] = rb;
}

// computeSourceTermVector: U-bar = S'(U)^T S-bar (S-bar read on the A slots and,
// with heat conduction, the J slots), and the relaxation-time derivatives
// added to tb1, tb2.
__device__ __forceinline__ void vjp_source(const HPar &m, const ConsPoint &c, const double (&Sb)[NF],
                                           double (&Ub)[NF], double &tb1, double &tb2)
{
#pragma clang fp contract(fast)
    const double d = HP::det3(c.A);
    const double th1 = 3.0 * (pow(d, (5.0 / 3.0)) / (m.cs2 * m.tau1));
    double EAb[3][3], Ab[3][3], th1b = 0.0;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int k = 0; k < 3; k++) {
            EAb[i][k] = -th1 * Sb[2 + 3 * i + k];
            th1b -= c.EA[i][k] * Sb[2 + 3 * i + k];
        }
    tb1 -= th1b * th1 / m.tau1;
    const double detb = th1b * (5.0 / 3.0) * th1 / d;
    const double (&A)[3][3] = c.A;
    Ab[0][0] = detb * (A[1][1] * A[2][2] - A[1][2] * A[2][1]);
    Ab[0][1] = detb * (A[1][2] * A[2][0] - A[1][0] * A[2][2]);
    Ab[0][2] = detb * (A[1][0] * A[2][1] - A[1][1] * A[2][0]);
    Ab[1][0] = detb * (A[0][2] * A[2][1] - A[0][1] * A[2][2]);
    Ab[1][1] = detb * (A[0][0] * A[2][2] - A[0][2] * A[2][0]);
    Ab[1][2] = detb * (A[0][1] * A[2][0] - A[0][0] * A[2][1]);
    Ab[2][0] = detb * (A[0][1] * A[1][2] - A[0][2] * A[1][1]);
    Ab[2][1] = detb * (A[0][2] * A[1][0] - A[0][0] * A[1][2]);
    Ab[2][2] = detb * (A[0][0] * A[1][1] - A[0][1] * A[1][0]);
    ea_vjp(m, c, EAb, Ab);
#pragma unroll
    for (int k = 0; k < NF; k++) Ub[k] = 0.0;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int k = 0; k < 3; k++) Ub[2 + 3 * i + k] = Ab[i][k];
    if (m.is_thermal) {
        // S_J = -kappa T j, kappa = rho0 / (tau2 T0), T = t_ref + (p + p_inf) / (rho G cv)
        const double kap = m.rho0 / (m.tau2 * m.t0);
        const double ircv = c.ir * rcp_nr(m.G * m.cv);
        const double T = m.t_ref + (c.p + m.p_inf) * ircv;
        double Tb = 0.0, wb[NF];
#pragma unroll
        for (int k = 0; k < NF; k++) wb[k] = 0.0;
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const double s = Sb[14 + i];
            Tb -= kap * c.j[i] * s;
            wb[14 + i] = -kap * T * s;
            tb2 += s * kap * T * c.j[i] / m.tau2;
        }
        wb[1] = Tb * ircv;
        wb[0] = -Tb * (c.p + m.p_inf) * ircv * c.ir;
        double t[NF];
        vjp_c2p(m, c, wb, t);
#pragma unroll
        for (int k = 0; k < NF; k++) Ub[k] += t[k];
    }
}

// ---- slopes -------------------------------------------------------------------------
// Slopes of the three conserved vectors in the forward's arithmetic: pieces
// (3 bits per field) and the linearised extrapolations xl, xr.
__device__ __forceinline__ void slope_states(const double (&UL)[NF], const double (&UM)[NF], const double (&UR)[NF],
                                             double bias, int limiter, unsigned long long &pieces, double (&xl)[NF],
                                             double (&xr)[NF])
{
    const double cl = 0.5 * (1.0 + bias), cr = 0.5 * (1.0 - bias);
    pieces = 0;
#pragma unroll
    for (int k = 0; k < NF; k++) {
        const double dl = UM[k] - UL[k];
        const double dr = UR[k] - UM[k];
        const double num = (fabs(dl) < 1e-5) ? 1e-5 : dl;
        const double den = (fabs(dr) < 1e-5) ? 1e-5 : dr;
        const int pc = limiter_piece(num / den, bias, limiter);
        pieces |= (unsigned long long)pc << (3 * k);
        const SlopeLin sl = slope_lin(dl, dr, pc, cl, cr, bias, limiter);
        const double d = sl.base * sl.xi;
        xl[k] = UM[k] - (0.5 * d);
        xr[k] = UM[k] + (0.5 * d);
    }
}

// Given the adjoints of xl and xr: the middle state's own term and the
// neighbours' terms of the slope.
__device__ __forceinline__ void slope_vjp(const double (&UL)[NF], const double (&UM)[NF], const double (&UR)[NF],
                                          double bias, int limiter, unsigned long long pieces, const double (&xlb)[NF],
                                          const double (&xrb)[NF], double (&own)[NF], double (&toL)[NF],
                                          double (&toR)[NF])
{
#pragma clang fp contract(fast)
    const double cl = 0.5 * (1.0 + bias), cr = 0.5 * (1.0 - bias);
#pragma unroll
    for (int k = 0; k < NF; k++) {
        const double dl = UM[k] - UL[k];
        const double dr = UR[k] - UM[k];
        const SlopeLin sl = slope_lin(dl, dr, (int)((pieces >> (3 * k)) & 7), cl, cr, bias, limiter);
        const double db = 0.5 * (xrb[k] - xlb[k]);
        const double bb = db * sl.xi;
        const double rb = db * sl.base * sl.dxi;
        const double dlb = cl * bb + (sl.cln ? 0.0 : rb * sl.iden);
        const double drb = cr * bb + (sl.cld ? 0.0 : -rb * sl.num * (sl.iden * sl.iden));
        own[k] = xlb[k] + xrb[k] + dlb - drb;
        toL[k] = -dlb;
        toR[k] = drb;
    }
}

// Adjoint of one fractional step in conserved variables, N = UM + c (F(xl) - F(xr)):
// adds the neighbours' terms to ULb / URb and returns UM's in UMb.
template <int DIR>
__device__ __forceinline__ void vjp_fractional(const HPar &m, const double (&UL)[NF], const double (&UM)[NF],
                                               const double (&UR)[NF], double bias, int limiter, double c,
                                               const double (&Nb)[NF], double (&ULb)[NF], double (&UMb)[NF],
                                               double (&URb)[NF])
{
    unsigned long long pieces;
    double xl[NF], xr[NF], s[NF], xlb[NF], xrb[NF];
    slope_states(UL, UM, UR, bias, limiter, pieces, xl, xr);
#pragma unroll
    for (int k = 0; k < NF; k++) s[k] = c * Nb[k];
    {
        ConsPoint p;
        cons_point(m, xl, p);
        vjp_flux<DIR>(m, p, s, xlb);
        cons_point(m, xr, p);
        vjp_flux<DIR>(m, p, s, xrb);
    }
#pragma unroll
    for (int k = 0; k < NF; k++) xrb[k] = -xrb[k];
    double own[NF], toL[NF], toR[NF];
    slope_vjp(UL, UM, UR, bias, limiter, pieces, xlb, xrb, own, toL, toR);
#pragma unroll
    for (int k = 0; k < NF; k++) {
        UMb[k] = Nb[k] + own[k];
        ULb[k] += toL[k];
        URb[k] += toR[k];
    }
}

// ---- sweep adjoint ---------------------------------------------------------------------
struct SweepAdjArgs {
    const double *w;        // the sweep's input (primitive planes)
    const double *w_out;    // its output
    const double *wb_out;   // adjoint of the output
    double *wb_in;          // adjoint of the input
    double *nbar;           // per cell: N-bar
    double *xlb, *xrb;      // per extended position: adjoints of the evolved left / right states
    double *own, *tol, *tor;  // per extended position: slope / evolution terms
    long long plane, eplane;
    int ny, nx, bc, limiter;
    double dx, bias, ts;
    HPar m;
};

// Extended position k in [-1, n] of a line (ghost layer included).
template <int DIR>
__device__ __forceinline__ long long ext_index(int line, int k, int nx)
{
    return (DIR == 0) ? (long long)line * (nx + 2) + (k + 1) : (long long)(k + 1) * nx + line;
}
template <int DIR>
__device__ __forceinline__ long long cell_index(int line, int k, int nx)
{
    return (DIR == 0) ? (long long)line * nx + k : (long long)k * nx + line;
}

template <int NN>
__device__ __forceinline__ void load_planes(const double *p, long long plane, long long idx, double (&v)[NN])
{
#pragma unroll
    for (int k = 0; k < NN; k++) v[k] = p[k * plane + idx];
}
template <int NN>
__device__ __forceinline__ void store_planes(double *p, long long plane, long long idx, const double (&v)[NN])
{
#pragma unroll
    for (int k = 0; k < NN; k++) p[k * plane + idx] = v[k];
}

// Thread -> (line, position) of a launch over `len` positions per line, the
// positions starting at k0.  x is the fastest index in both directions.
template <int DIR>
__device__ __forceinline__ bool line_pos(const SweepAdjArgs &a, int k0, int len, int &line, int &k)
{
    const int col = blockIdx.x * kBS + threadIdx.x, row = blockIdx.y;
    if (DIR == 0) {
        line = row;
        k = k0 + col;
        return col < len;
    }
    line = col;
    k = k0 + row;
    return col < a.nx;
}

// 1. N-bar per cell, at N = p2c(w_out).
template <int DIR>
__global__ __launch_bounds__(kBS) void hpr_adj_nbar_kernel(SweepAdjArgs a)
{
    const int n = (DIR == 0) ? a.nx : a.ny;
    int line, k;
    if (!line_pos<DIR>(a, 0, n, line, k)) return;
    const long long c = cell_index<DIR>(line, k, a.nx);
    double w[NF], U[NF], wb[NF], Nb[NF];
    load_planes(a.w_out, a.plane, c, w);
    HP::prim_to_cons(a.m, w, U);
    load_planes(a.wb_out, a.plane, c, wb);
    ConsPoint p;
    cons_point(a.m, U, p);
    vjp_c2p(a.m, p, wb, Nb);
    store_planes(a.nbar, a.plane, c, Nb);
}

// 2. Per face f (between positions f-1 and f): the FORCE flux's adjoint.
template <int DIR>
__global__ __launch_bounds__(kBS) void hpr_adj_face_kernel(SweepAdjArgs a)
{
    using S = Scheme<HP, DIR, MODE_EXACT>;
    const int n = (DIR == 0) ? a.nx : a.ny;
    const bool periodic = (a.bc == kPeriodic);
    const int faces = periodic ? n : n + 1;
    int line, f;
    if (!line_pos<DIR>(a, 0, faces, line, f)) return;
    const HPar &m = a.m;
    double U[4][NF];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        double w[NF];
        load_planes(a.w, a.plane, cell_index<DIR>(line, ghost_index(f - 2 + q, n, a.bc), a.nx), w);
        HP::prim_to_cons(m, w, U[q]);
    }
    const double half_c = 0.5 * (a.ts / a.dx), lf_c = a.dx / a.ts, ri_c = a.ts / a.dx;
    ConsFlux<NF> la, ra, lb, rb;
    S::evolve(m, U[0], U[1], U[2], a.bias, a.limiter, half_c, la, ra);   // cell f-1: its right state
    S::evolve(m, U[1], U[2], U[3], a.bias, a.limiter, half_c, lb, rb);   // cell f: its left state
    // f-bar = (ts/dx) (N-bar of cell f  -  N-bar of cell f-1); ghost cells have no update
    const int ca = f - 1, cb = f;
    double fb[NF];
    {
        double na[NF], nb[NF];
        const bool ga = !periodic && ca < 0, gb = !periodic && cb >= n;
#pragma unroll
        for (int k = 0; k < NF; k++) {
            na[k] = 0.0;
            nb[k] = 0.0;
        }
        if (!ga) load_planes(a.nbar, a.plane, cell_index<DIR>(line, ghost_index(ca, n, a.bc), a.nx), na);
        if (!gb) load_planes(a.nbar, a.plane, cell_index<DIR>(line, ghost_index(cb, n, a.bc), a.nx), nb);
#pragma unroll
        for (int k = 0; k < NF; k++) fb[k] = ri_c * (nb[k] - na[k]);
    }
    double us[NF], half[NF], usb[NF];
#pragma unroll
    for (int k = 0; k < NF; k++) {
        us[k] = 0.5 * ((ra.U[k] + lb.U[k]) + (ri_c * (ra.F[k] - lb.F[k])));
        half[k] = 0.5 * fb[k];
    }
    {
        ConsPoint p;
        cons_point(m, us, p);
        vjp_flux<DIR>(m, p, half, usb);
    }
    double Fab[NF], Fbb[NF], t[NF], Rab[NF], Lbb[NF];
#pragma unroll
    for (int k = 0; k < NF; k++) {
        Fab[k] = 0.25 * fb[k] + 0.5 * ri_c * usb[k];
        Fbb[k] = 0.25 * fb[k] - 0.5 * ri_c * usb[k];
    }
    {
        ConsPoint p;
        cons_point(m, ra.U, p);
        vjp_flux<DIR>(m, p, Fab, t);
    }
#pragma unroll
    for (int k = 0; k < NF; k++) Rab[k] = 0.25 * lf_c * fb[k] + 0.5 * usb[k] + t[k];
    {
        ConsPoint p;
        cons_point(m, lb.U, p);
        vjp_flux<DIR>(m, p, Fbb, t);
    }
#pragma unroll
    for (int k = 0; k < NF; k++) Lbb[k] = -0.25 * lf_c * fb[k] + 0.5 * usb[k] + t[k];
    const int pa = periodic ? ghost_index(ca, n, a.bc) : ca;
    store_planes(a.xrb, a.eplane, ext_index<DIR>(line, pa, a.nx), Rab);
    store_planes(a.xlb, a.eplane, ext_index<DIR>(line, cb, a.nx), Lbb);
}

// 3. Per position (ghosts -1 and n for transmissive lines): evolution and
// slope adjoints.
template <int DIR>
__global__ __launch_bounds__(kBS) void hpr_adj_evolve_kernel(SweepAdjArgs a)
{
    const int n = (DIR == 0) ? a.nx : a.ny;
    const bool periodic = (a.bc == kPeriodic);
    const int k0 = periodic ? 0 : -1, len = periodic ? n : n + 2;
    int line, k;
    if (!line_pos<DIR>(a, k0, len, line, k)) return;
    const HPar &m = a.m;
    double UL[NF], UM[NF], UR[NF];
    {
        double w[NF];
        load_planes(a.w, a.plane, cell_index<DIR>(line, ghost_index(k - 1, n, a.bc), a.nx), w);
        HP::prim_to_cons(m, w, UL);
        load_planes(a.w, a.plane, cell_index<DIR>(line, ghost_index(k, n, a.bc), a.nx), w);
        HP::prim_to_cons(m, w, UM);
        load_planes(a.w, a.plane, cell_index<DIR>(line, ghost_index(k + 1, n, a.bc), a.nx), w);
        HP::prim_to_cons(m, w, UR);
    }
    const long long e = ext_index<DIR>(line, k, a.nx);
    double xlb[NF], xrb[NF];
#pragma unroll
    for (int q = 0; q < NF; q++) {
        xlb[q] = 0.0;
        xrb[q] = 0.0;
    }
    if (k >= 0) load_planes(a.xlb, a.eplane, e, xlb);       // a left face exists
    if (k <= n - 1) load_planes(a.xrb, a.eplane, e, xrb);   // a right face exists
    unsigned long long pieces;
    double xl[NF], xr[NF], s[NF], tl[NF], tr[NF];
    slope_states(UL, UM, UR, a.bias, a.limiter, pieces, xl, xr);
    const double half_c = 0.5 * (a.ts / a.dx);
#pragma unroll
    for (int q = 0; q < NF; q++) s[q] = half_c * (xlb[q] + xrb[q]);
    {
        ConsPoint p;
        cons_point(m, xl, p);
        vjp_flux<DIR>(m, p, s, tl);
        cons_point(m, xr, p);
        vjp_flux<DIR>(m, p, s, tr);
    }
#pragma unroll
    for (int q = 0; q < NF; q++) {
        xlb[q] += tl[q];
        xrb[q] -= tr[q];
    }
    double own[NF], toL[NF], toR[NF];
    slope_vjp(UL, UM, UR, a.bias, a.limiter, pieces, xlb, xrb, own, toL, toR);
    store_planes(a.own, a.eplane, e, own);
    store_planes(a.tol, a.eplane, e, toL);
    store_planes(a.tor, a.eplane, e, toR);
}

// 4. Per cell: gather the terms of every position whose stencil maps to it,
// then the adjoint of p2c.
template <int DIR>
__global__ __launch_bounds__(kBS) void hpr_adj_gather_kernel(SweepAdjArgs a)
{
    const int n = (DIR == 0) ? a.nx : a.ny;
    const bool periodic = (a.bc == kPeriodic);
    int line, j;
    if (!line_pos<DIR>(a, 0, n, line, j)) return;
    double Ub[NF];
    load_planes(a.nbar, a.plane, cell_index<DIR>(line, j, a.nx), Ub);
    // candidate positions: j-1, j, j+1 and, on transmissive lines, the ghosts
    int cand[5], nc = 0;
    auto add = [&](int q) {
        if (periodic) q = ghost_index(q, n, a.bc);
        else if (q < -1 || q > n) return;
        for (int i = 0; i < nc; i++)
            if (cand[i] == q) return;
        cand[nc++] = q;
    };
    add(j - 1);
    add(j);
    add(j + 1);
    if (!periodic) {
        add(-1);
        add(n);
    }
    for (int i = 0; i < nc; i++) {
        const int q = cand[i];
        const long long e = ext_index<DIR>(line, q, a.nx);
        if (ghost_index(q, n, a.bc) == j)
#pragma unroll
            for (int f = 0; f < NF; f++) Ub[f] += a.own[f * a.eplane + e];
        if (ghost_index(q - 1, n, a.bc) == j)
#pragma unroll
            for (int f = 0; f < NF; f++) Ub[f] += a.tol[f * a.eplane + e];
        if (ghost_index(q + 1, n, a.bc) == j)
#pragma unroll
            for (int f = 0; f < NF; f++) Ub[f] += a.tor[f * a.eplane + e];
    }
    double w[NF], wb[NF];
    const long long c = cell_index<DIR>(line, j, a.nx);
    load_planes(a.w, a.plane, c, w);
    vjp_p2c(a.m, w, Ub, wb);
    store_planes(a.wb_in, a.plane, c, wb);
}

// ---- sub-cycle adjoint -----------------------------------------------------------------
// One sub-cycle's adjoint runs as stage-sized launches with the stage states in
// HBM, so that no thread holds a whole RK4 reverse at once: a recording
// forward, then per stage (last first) the source VJP and the VJPs of its
// fractional steps (y between two x in 2D), then the source VJP at the cell's
// own state and the conversions back to primitives; the gather kernel below
// sums each cell's own and neighbour terms.
struct ForcingAdjArgs {
    const double *w;        // the sub-cycle's input (primitive planes)
    const double *wb_out;   // adjoint of its output
    double *wb_in;          // adjoint of its input
    double *own, *tl, *tr, *tt, *tb;   // per cell: its own term and its neighbours' (primitive)
    double *rt;             // per cell: the conserved state its neighbours' predictors see
    double *um;             // per cell: the middle state entering fractional step (st, ph), 9 x 17 planes
    double *ue;             // per cell: the state whose source gives increment st + 1, 3 x 17 planes
    double *nb;             // per cell: the adjoint flowing back through the current stage
    double *umb;            // per cell: adjoint of the cell's own conserved vector
    double *kb;             // per cell: adjoints of the stage increments, 4 x kMaxSlots planes
    double *ulb, *urb, *utb, *ubb;   // per cell: conserved adjoints of its neighbours' states
    double *tau_part;       // per block: (tau1, tau2) derivative sums, accumulated
    long long plane;
    int ny, nx, dim, bc, limiter;
    double dx, bias, ts;
    HPar m;
};

// The stage states, written to their planes as forcing_cell produces them.
struct GlobalStages {
    double *um, *ue;
    long long plane, c;
    template <int N>
    __device__ __forceinline__ void mid(int st, int ph, const double (&U)[N])
    {
#pragma unroll
        for (int k = 0; k < N; k++) um[((st * 3 + ph) * N + k) * plane + c] = U[k];
    }
    template <int N>
    __device__ __forceinline__ void end(int st, const double (&U)[N])
    {
#pragma unroll
        for (int k = 0; k < N; k++) ue[(st * N + k) * plane + c] = U[k];
    }
};

__device__ __forceinline__ void cell_neighbours(const ForcingAdjArgs &a, long long c, long long &cl, long long &cr,
                                                long long &ct, long long &cb)
{
    const int i = (int)(c / a.nx), j = (int)(c - (long long)i * a.nx);
    cl = (long long)i * a.nx + ghost_index(j - 1, a.nx, a.bc);
    cr = (long long)i * a.nx + ghost_index(j + 1, a.nx, a.bc);
    ct = (a.dim == 2) ? (long long)ghost_index(i - 1, a.ny, a.bc) * a.nx + j : c;
    cb = (a.dim == 2) ? (long long)ghost_index(i + 1, a.ny, a.bc) * a.nx + j : c;
}

// Adds this block's relaxation-time derivative sums to its own slots (every
// thread of the block calls it).
__device__ __forceinline__ void tau_block_add(double *tau_part, double tb1, double tb2)
{
    __shared__ double s1[kBS], s2[kBS];
    s1[threadIdx.x] = tb1;
    s2[threadIdx.x] = tb2;
    __syncthreads();
    for (int h = kBS / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
            s1[threadIdx.x] += s1[threadIdx.x + h];
            s2[threadIdx.x] += s2[threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        tau_part[2 * blockIdx.x] += s1[0];
        tau_part[2 * blockIdx.x + 1] += s2[0];
    }
}

// 0. Every cell's conserved vector as its neighbours' predictors see it
// (forcing_neighbour), into a.rt.
__global__ __launch_bounds__(kBS) void hpr_adj_force_rt_kernel(ForcingAdjArgs a)
{
    const long long cells = (long long)a.ny * a.nx;
    const long long c = (long long)blockIdx.x * kBS + threadIdx.x;
    if (c >= cells) return;
    double RT[NF];
    forcing_neighbour<HP>(a.m, a.w, a.plane, c, RT);
    store_planes(a.rt, a.plane, c, RT);
}

// 1. The forward sub-cycle of every cell with its stage states recorded, the
// adjoint of its output conversion, and the first increment adjoints.  As in
// forcing_kernel, the workgroup covers a tile of cells (forcing_blocks) whose
// neighbours' conserved vectors (from a.rt) sit in LDS, read where the
// predictors use them: two waves per SIMD instead of one with all four held in
// registers.
template <int DIM>
__global__ __launch_bounds__(kForceBS, 2) void hpr_adj_force_record_kernel(ForcingAdjArgs a)
{
    using T = ForcingTraits<HP>;
    constexpr int NS = T::kMaxSlots;
    __shared__ double s_nb[NF * ForcingTile<DIM>::slots];
    const HPar &m = a.m;
    int i, j, i0, j0;
    forcing_tile_origin<DIM>(a.nx, i0, j0, i, j);
    fill_neighbour_tile<NF, DIM>(s_nb, i0, j0, a.ny, a.nx, a.bc,
                                 [&](long long idx, double (&U)[NF]) { load_planes(a.rt, a.plane, idx, U); });
    if (i >= a.ny || j >= a.nx) return;
    const long long c = (long long)i * a.nx + j;
    double N[NF], K[4][NS];
    {
        const TileNeighbours<DIM> nb(s_nb, i, j, i0, j0);
        const LdsNeighbour &UL = nb.l, &UR = nb.r, &UT = nb.t, &UB = nb.b;
        double w0[NF], Um[NF];
        load_planes(a.w, a.plane, c, w0);
        HP::prim_to_cons(m, w0, Um);
        GlobalStages rec{a.um, a.ue, a.plane, c};
        forcing_cell<HP, DIM>(m, UL, UR, UT, UB, Um, a.ts, a.dx, a.bias, a.limiter, N, K, rec);
    }
    double wbo[NF], own[NF], Nb[NF], zero[NF];
    load_planes(a.wb_out, a.plane, c, wbo);
#pragma unroll
    for (int k = 0; k < NF; k++) {
        own[k] = 0.0;
        zero[k] = 0.0;
    }
    if (!m.is_thermal) {   // the cell keeps its own J
#pragma unroll
        for (int k = 14; k < 17; k++) {
            own[k] = wbo[k];
            wbo[k] = 0.0;
        }
    }
    {
        ConsPoint p;
        cons_point(m, N, p);
        vjp_c2p(m, p, wbo, Nb);
    }
    store_planes(a.own, a.plane, c, own);
    store_planes(a.umb, a.plane, c, Nb);
    const int nsrc = T::count(m);
#pragma unroll
    for (int q = 0; q < NS; q++) {
        const double g = (q < nsrc) ? (1.0 / 6.0) * Nb[T::slot(q)] : 0.0;
        a.kb[(0 * NS + q) * a.plane + c] = g;
        a.kb[(1 * NS + q) * a.plane + c] = 2.0 * g;
        a.kb[(2 * NS + q) * a.plane + c] = 2.0 * g;
        a.kb[(3 * NS + q) * a.plane + c] = g;
    }
    store_planes(a.ulb, a.plane, c, zero);
    store_planes(a.urb, a.plane, c, zero);
    if (DIM == 2) {
        store_planes(a.utb, a.plane, c, zero);
        store_planes(a.ubb, a.plane, c, zero);
    }
}

// 2. Stage st: the adjoint of its increment's source, S-bar = ts K-bar[st + 1].
__global__ __launch_bounds__(kBS) void hpr_adj_force_source_kernel(ForcingAdjArgs a, int st)
{
    using T = ForcingTraits<HP>;
    constexpr int NS = T::kMaxSlots;
    const HPar &m = a.m;
    const long long cells = (long long)a.ny * a.nx;
    const long long c = (long long)blockIdx.x * kBS + threadIdx.x;
    double tb1 = 0.0, tb2 = 0.0;
    if (c < cells) {
        const int nsrc = T::count(m);
        double Sb[NF], Ue[NF], Ueb[NF];
#pragma unroll
        for (int k = 0; k < NF; k++) Sb[k] = 0.0;
#pragma unroll
        for (int q = 0; q < NS; q++)
            if (q < nsrc) Sb[T::slot(q)] = a.ts * a.kb[((st + 1) * NS + q) * a.plane + c];
#pragma unroll
        for (int k = 0; k < NF; k++) Ue[k] = a.ue[(st * NF + k) * a.plane + c];
        ConsPoint p;
        cons_point(m, Ue, p);
        vjp_source(m, p, Sb, Ueb, tb1, tb2);
        store_planes(a.nb, a.plane, c, Ueb);
    }
    tau_block_add(a.tau_part, tb1, tb2);
}

// 3. Stage st, fractional step ph (DIR: its direction): N = UM + frac (F(xl) - F(xr))
// reversed.  The neighbours' terms go to their accumulators; ph > 0 passes the
// middle state's adjoint on (nb), ph = 0 ends the stage: the adjoint of
// Us = Um + coef K[st] goes to Um and to K-bar[st].
template <int DIR>
__global__ __launch_bounds__(kBS) void hpr_adj_force_frac_kernel(ForcingAdjArgs a, int st, int ph, double frac,
                                                                  double coef)
{
    using T = ForcingTraits<HP>;
    constexpr int NS = T::kMaxSlots;
    const HPar &m = a.m;
    const long long cells = (long long)a.ny * a.nx;
    const long long c = (long long)blockIdx.x * kBS + threadIdx.x;
    if (c >= cells) return;
    long long cl, cr, ct, cb;
    cell_neighbours(a, c, cl, cr, ct, cb);
    const long long nl = (DIR == 0) ? cl : ct, nr = (DIR == 0) ? cr : cb;
    const double *um = a.um + (long long)(st * 3 + ph) * NF * a.plane;
    unsigned long long pieces;
    double xlb[NF], xrb[NF];
    {
        double UL[NF], UM[NF], UR[NF], xl[NF], xr[NF], s[NF];
        load_planes(a.rt, a.plane, nl, UL);
        load_planes(um, a.plane, c, UM);
        load_planes(a.rt, a.plane, nr, UR);
        slope_states(UL, UM, UR, a.bias, a.limiter, pieces, xl, xr);
        load_planes(a.nb, a.plane, c, s);
#pragma unroll
        for (int k = 0; k < NF; k++) s[k] *= frac;
        ConsPoint p;
        cons_point(m, xl, p);
        vjp_flux<DIR>(m, p, s, xlb);
        cons_point(m, xr, p);
        vjp_flux<DIR>(m, p, s, xrb);
    }
    // the inputs again, from memory rather than kept live across the VJPs
    asm volatile("" ::: "memory");
    double UL[NF], UM[NF], UR[NF], own[NF], toL[NF], toR[NF];
    load_planes(a.rt, a.plane, nl, UL);
    load_planes(um, a.plane, c, UM);
    load_planes(a.rt, a.plane, nr, UR);
#pragma unroll
    for (int k = 0; k < NF; k++) xrb[k] = -xrb[k];
    slope_vjp(UL, UM, UR, a.bias, a.limiter, pieces, xlb, xrb, own, toL, toR);
    double *accL = (DIR == 0) ? a.ulb : a.utb, *accR = (DIR == 0) ? a.urb : a.ubb;
#pragma unroll
    for (int k = 0; k < NF; k++) {
        accL[k * a.plane + c] += toL[k];
        accR[k * a.plane + c] += toR[k];
        own[k] += a.nb[k * a.plane + c];   // UM-bar = N-bar + the extrapolations' terms
    }
    if (ph > 0) {
        store_planes(a.nb, a.plane, c, own);
    } else {
        const int nsrc = T::count(m);
#pragma unroll
        for (int k = 0; k < NF; k++) a.umb[k * a.plane + c] += own[k];
#pragma unroll
        for (int q = 0; q < NS; q++)
            if (q < nsrc) a.kb[(st * NS + q) * a.plane + c] += coef * own[T::slot(q)];
    }
}

// 4. The source at the cell's own state (K-bar[0]), then every adjoint back to
// primitives: the cell's own term and its neighbours'.
template <int DIM>
__global__ __launch_bounds__(kBS) void hpr_adj_force_final_kernel(ForcingAdjArgs a)
{
    using T = ForcingTraits<HP>;
    constexpr int NS = T::kMaxSlots;
    const HPar &m = a.m;
    const long long cells = (long long)a.ny * a.nx;
    const long long c = (long long)blockIdx.x * kBS + threadIdx.x;
    double tb1 = 0.0, tb2 = 0.0;
    if (c < cells) {
        long long cl, cr, ct, cb;
        cell_neighbours(a, c, cl, cr, ct, cb);
        const int nsrc = T::count(m);
        double w0[NF], Umb[NF], wb[NF], wn[NF], own[NF];
        load_planes(a.w, a.plane, c, w0);
        load_planes(a.umb, a.plane, c, Umb);
        {
            double Sb[NF], Um[NF], t[NF];
#pragma unroll
            for (int k = 0; k < NF; k++) Sb[k] = 0.0;
#pragma unroll
            for (int q = 0; q < NS; q++)
                if (q < nsrc) Sb[T::slot(q)] = a.ts * a.kb[q * a.plane + c];
            HP::prim_to_cons(m, w0, Um);
            ConsPoint p;
            cons_point(m, Um, p);
            vjp_source(m, p, Sb, t, tb1, tb2);
#pragma unroll
            for (int k = 0; k < NF; k++) Umb[k] += t[k];
        }
        vjp_p2c(m, w0, Umb, wb);
        load_planes(a.own, a.plane, c, own);
#pragma unroll
        for (int k = 0; k < NF; k++) own[k] += wb[k];
        store_planes(a.own, a.plane, c, own);
        double ub[NF];
        load_planes(a.w, a.plane, cl, wn);
        load_planes(a.ulb, a.plane, c, ub);
        vjp_p2c(m, wn, ub, wb);
        store_planes(a.tl, a.plane, c, wb);
        load_planes(a.w, a.plane, cr, wn);
        load_planes(a.urb, a.plane, c, ub);
        vjp_p2c(m, wn, ub, wb);
        store_planes(a.tr, a.plane, c, wb);
        if (DIM == 2) {
            load_planes(a.w, a.plane, ct, wn);
            load_planes(a.utb, a.plane, c, ub);
            vjp_p2c(m, wn, ub, wb);
            store_planes(a.tt, a.plane, c, wb);
            load_planes(a.w, a.plane, cb, wn);
            load_planes(a.ubb, a.plane, c, ub);
            vjp_p2c(m, wn, ub, wb);
            store_planes(a.tb, a.plane, c, wb);
        }
    }
    tau_block_add(a.tau_part, tb1, tb2);
}

// Per cell: its own term plus the neighbour terms of the cells whose one-layer
// stencil maps to it.
__global__ __launch_bounds__(kBS) void hpr_adj_forcing_gather_kernel(ForcingAdjArgs a)
{
    const long long cells = (long long)a.ny * a.nx;
    const long long c = (long long)blockIdx.x * kBS + threadIdx.x;
    if (c >= cells) return;
    const int i = (int)(c / a.nx), j = (int)(c - (long long)i * a.nx);
    double acc[NF];
    load_planes(a.own, a.plane, c, acc);
    // x: sources q in {j-1, j, j+1}; q's left term lands on ghost_index(q-1), its right on ghost_index(q+1)
    int seen[3], ns = 0;
    for (int d = -1; d <= 1; d++) {
        const int q = (a.bc == kPeriodic) ? ghost_index(j + d, a.nx, a.bc) : j + d;
        if (q < 0 || q >= a.nx) continue;
        bool dup = false;
        for (int s = 0; s < ns; s++) dup = dup || seen[s] == q;
        if (dup) continue;
        seen[ns++] = q;
        const long long src = (long long)i * a.nx + q;
        if (ghost_index(q - 1, a.nx, a.bc) == j)
#pragma unroll
            for (int f = 0; f < NF; f++) acc[f] += a.tl[f * a.plane + src];
        if (ghost_index(q + 1, a.nx, a.bc) == j)
#pragma unroll
            for (int f = 0; f < NF; f++) acc[f] += a.tr[f * a.plane + src];
    }
    if (a.dim == 2) {
        ns = 0;
        for (int d = -1; d <= 1; d++) {
            const int q = (a.bc == kPeriodic) ? ghost_index(i + d, a.ny, a.bc) : i + d;
            if (q < 0 || q >= a.ny) continue;
            bool dup = false;
            for (int s = 0; s < ns; s++) dup = dup || seen[s] == q;
            if (dup) continue;
            seen[ns++] = q;
            const long long src = (long long)q * a.nx + j;
            if (ghost_index(q - 1, a.ny, a.bc) == i)
#pragma unroll
                for (int f = 0; f < NF; f++) acc[f] += a.tt[f * a.plane + src];
            if (ghost_index(q + 1, a.ny, a.bc) == i)
#pragma unroll
                for (int f = 0; f < NF; f++) acc[f] += a.tb[f * a.plane + src];
        }
    }
    store_planes(a.wb_in, a.plane, c, acc);
}

}  // namespace hpradj
}  // namespace vg
