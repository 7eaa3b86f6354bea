// Hand-written adjoint (reverse mode) of the Euler SLIC sweep, time step held
// fixed.  The forward map of one sweep in direction DIR is, per cell j
// (Solvers/secondordersolver.cpp:7-17, 80-98; Solvers/solvers.cpp:454-466,
// 550-565, 601-609; Solvers/firstordersolver.cpp:79-97, 165-171, 227-230, 428-433):
//
//   U_k = p2c(w_k)
//   d_k = slope(U_{k-1}, U_k, U_{k+1})                (limiter xi(r), clamp 1e-5)
//   xl_k = U_k - d_k/2, xr_k = U_k + d_k/2
//   e_k = h (f(xl_k) - f(xr_k)),  h = dt/(2 dx)
//   XL_k = xl_k + e_k, XR_k = xr_k + e_k               (evolved conserved states)
//   a = cf(XR_j), b = cf(XL_{j+1}),  cf(X) = (U', f(U')), U' = p2c(c2p(X))
//   F_{j+1/2} = (LF + f(Us)) / 2,  LF = ((Fa + Fb) + (dx/dt)(Ua - Ub)) / 2,
//                                  Us = ((Ua + Ub) + (dt/dx)(Fa - Fb)) / 2
//   N_j = U_j + (dt/dx)(F_{j-1/2} - F_{j+1/2}),  w'_j = c2p(N_j)
//
// and the reverse pass runs the same chain backwards with analytic Jacobians of
// p2c, c2p and the flux.  The limiter is taken piecewise: its derivative is that
// of the branch the forward comparisons select (at a kink, the branch chosen by
// the same `<` tests -- a one-sided subgradient).  The 1e-5 clamp is a constant
// (zero derivative), as in the forward.
//
// Only the branch decisions need the forward's exact values, and they depend
// only on U = p2c(w) and its differences: those are recomputed with the
// forward's IEEE arithmetic, so every comparison comes out as in the forward.
// Everything else is a linearisation point, where a last-bit difference moves
// the derivative by a relative O(1e-16): it is evaluated with Newton-refined
// reciprocals, and the round trip U' = p2c(c2p(X)) is taken as the identity it
// is in exact arithmetic (U' = X, Jacobian I).
//
// One workgroup handles a tile of LW adjacent lines x BS / LW consecutive
// positions along the sweep, a halo of 4 on each side: outputs need N-bar of
// cells j-2..j+2.  Positions beyond the
// real grid carry no output adjoint (their N-bar is 0); the adjoint that lands on
// the two ghost positions on each side is written to a per-line ghost buffer and
// folded into the interior cells it copies (transmissive mirror or periodic wrap).
#pragma once

#include "vg_phys_euler.h"
#include "vg_sweep.h"

namespace vg {

using EPar = EulerPhys::Params;

// ---- analytic vector-Jacobian products of the Euler primitives ------------------
// These and the piece evaluations run at linearisation points only, so their
// multiply-adds are fused; the branch decisions (limiter_piece and the
// forward's conversions) keep the forward's separately rounded arithmetic.

// w-bar = (d p2c / d w)^T U-bar  (p2c: Euler/eulerstatevector.cpp:78-91).
__device__ __forceinline__ void vjp_p2c(const EPar &m, const double (&w)[6], const double (&Ub)[6], double (&wb)[6])
{
#pragma clang fp contract(fast)
    const double rho = w[0], u = w[1], v = w[2], z = w[3], lam = w[5];
    const double vsq = u * u + v * v + z * z;
    wb[0] = Ub[0] + u * Ub[1] + v * Ub[2] + z * Ub[3] + (0.5 * vsq + m.e_form * (lam - 1.0)) * Ub[4] + lam * Ub[5];
    wb[1] = rho * Ub[1] + rho * u * Ub[4];
    wb[2] = rho * Ub[2] + rho * v * Ub[4];
    wb[3] = rho * Ub[3] + rho * z * Ub[4];
    wb[4] = Ub[4] * m.inv_gm1;
    wb[5] = rho * m.e_form * Ub[4] + rho * Ub[5];
}

// U-bar = (d c2p / d U)^T w-bar  (c2p: eulerstatevector.cpp:47-61);
// p = (gamma-1)(U4 - |m|^2/(2 U0) - E_f (U5 - U0)) - gamma p_inf.
__device__ __forceinline__ void vjp_c2p(const EPar &m, const double (&U)[6], const double (&wb)[6], double (&Ub)[6])
{
#pragma clang fp contract(fast)
    const double q = rcp_nr(U[0]);
    const double u = U[1] * q, v = U[2] * q, z = U[3] * q, lam = U[5] * q;
    const double g1 = m.gm1;
    const double vsq = u * u + v * v + z * z;
    Ub[0] = wb[0] - (u * wb[1] + v * wb[2] + z * wb[3] + lam * wb[5]) * q + g1 * (0.5 * vsq + m.e_form) * wb[4];
    Ub[1] = wb[1] * q - g1 * u * wb[4];
    Ub[2] = wb[2] * q - g1 * v * wb[4];
    Ub[3] = wb[3] * q - g1 * z * wb[4];
    Ub[4] = g1 * wb[4];
    Ub[5] = wb[5] * q - g1 * m.e_form * wb[4];
}

// U-bar = (d f_DIR / d U)^T F-bar for the flux of a conserved vector
// (eulerstatevector.cpp:93-120, 127-154): with n = DIR, u_n = U_{1+n} / U0,
//   F0 = U_{1+n},  F_{1+c} = U_{1+n} vel_c + p [c == n],  F4 = u_n (U4 + p),  F5 = u_n U5,
//   p = (gamma-1)(U4 - |m|^2 / (2 U0) - E_f (U5 - U0)) - gamma p_inf.
// Collecting terms: U-bar = s dp/dU + (terms of u_n) + u_n F-bar on the own
// components, with s = F-bar_{1+n} + u_n F-bar_4 the weight of the pressure.
template <int DIR>
__device__ __forceinline__ void vjp_flux(const EPar &m, const double (&U)[6], const double (&Fb)[6], double (&Ub)[6])
{
#pragma clang fp contract(fast)
    const double q = rcp_nr(U[0]);
    const double vel[3] = {U[1] * q, U[2] * q, U[3] * q};
    const double lam = U[5] * q;
    const double un = vel[DIR];
    const double g1 = m.gm1;
    const double vsq = vel[0] * vel[0] + vel[1] * vel[1] + vel[2] * vel[2];
    const double p = g1 * (U[4] - 0.5 * U[0] * vsq - m.e_form * (U[5] - U[0])) - m.gpinf;
    const double H = (U[4] + p) * q;
    // d(u_n)/dU0 = -u_n / U0, d(u_n)/dU_{1+n} = 1 / U0: every row with a u_n
    // factor contributes T (its coefficient of u_n, times U0) to both
    const double T = vel[0] * Fb[1] + vel[1] * Fb[2] + vel[2] * Fb[3] + H * Fb[4] + lam * Fb[5];
    const double sp = g1 * (Fb[1 + DIR] + un * Fb[4]);   // weight of dp/dU, (gamma-1) folded in
    Ub[0] = sp * (0.5 * vsq + m.e_form) - un * T;
    Ub[1] = un * Fb[1] - sp * vel[0];
    Ub[2] = un * Fb[2] - sp * vel[1];
    Ub[3] = un * Fb[3] - sp * vel[2];
    Ub[1 + DIR] += T + Fb[0];
    Ub[4] = sp + un * Fb[4];
    Ub[5] = un * Fb[5] - sp * m.e_form;
}

// Pieces of the slope limiters (Solvers/slopelimiters.cpp:7-74): which
// expression of r the forward's comparisons select.
enum LimiterPiece : int { kXiZero = 0, kXiTwoR, kXiOne, kXiR, kXiRcoef, kXiTwo, kXiVanLeer };

// The forward's comparisons, in its IEEE arithmetic (limiter_value with
// smin(a, b) = b < a ? b : a), reduced to the piece they select.
__device__ __forceinline__ int limiter_piece(double r, double bias, int limiter)
{
    if (r < 0.0) return kXiZero;
    if (limiter == 0) {
        if (r < 0.5) return kXiTwoR;
        if (r < 1.0) return kXiOne;
        const double R = r_coefficient(r, bias);
        const double a = (R < r) ? R : r;     // smin(r, R)
        if (2.0 < a) return kXiTwo;           // smin(a, 2)
        return (R < r) ? kXiRcoef : kXiR;
    }
    if (limiter == 1) {
        const double R = r_coefficient(r, bias);
        const double vl = (2.0 * r) / (1.0 + r);
        return (R < vl) ? kXiRcoef : kXiVanLeer;   // smin(vl, R)
    }
    if (r < 1.0) return kXiR;
    return (r_coefficient(r, bias) < 1.0) ? kXiRcoef : kXiOne;   // smin(1, R)
}

// Value and r-derivative of a piece at a linearisation point, by selection
// (no divergent branches: the pieces of one limiter are all cheap).
__device__ __forceinline__ void limiter_piece_eval(int piece, double r, double bias, int limiter, double &xi,
                                                   double &dxi)
{
#pragma clang fp contract(fast)
    const double iv = rcp_nr((1.0 - bias) + (1.0 + bias) * r);
    double x = 0.0, d = 0.0;
    x = (piece == kXiTwoR) ? 2.0 * r : x;
    d = (piece == kXiTwoR) ? 2.0 : d;
    x = (piece == kXiOne) ? 1.0 : x;
    x = (piece == kXiR) ? r : x;
    d = (piece == kXiR) ? 1.0 : d;
    x = (piece == kXiTwo) ? 2.0 : x;
    x = (piece == kXiRcoef) ? 2.0 * iv : x;
    d = (piece == kXiRcoef) ? -2.0 * (1.0 + bias) * (iv * iv) : d;
    if (limiter == 1) {   // uniform over the launch
        const double ivl = rcp_nr(1.0 + r);
        x = (piece == kXiVanLeer) ? 2.0 * r * ivl : x;
        d = (piece == kXiVanLeer) ? 2.0 * (ivl * ivl) : d;
    }
    xi = x;
    dxi = d;
}

// Slope of one field at a linearisation point, given the forward's piece:
// base = cl dl + cr dr, r = num / den with |.| < 1e-5 replaced by +1e-5.
struct SlopeLin {
    double base, xi, dxi, num, iden;
    bool cln, cld;
};

__device__ __forceinline__ SlopeLin slope_lin(double dl, double dr, int piece, double cl, double cr, double bias,
                                              int limiter)
{
#pragma clang fp contract(fast)
    SlopeLin s;
    s.base = (cl * dl) + (cr * dr);
    s.cln = fabs(dl) < 1e-5;
    s.cld = fabs(dr) < 1e-5;
    s.num = s.cln ? 1e-5 : dl;
    s.iden = rcp_nr(s.cld ? 1e-5 : dr);
    limiter_piece_eval(piece, s.num * s.iden, bias, limiter, s.xi, s.dxi);
    return s;
}

// ---- forward pieces at linearisation points ---------------------------------------

// Flux of a conserved vector with one refined reciprocal of rho.
template <int DIR>
__device__ __forceinline__ void lin_flux(const EPar &m, const double (&U)[6], double (&F)[6])
{
    double w[6];
    EulerPhys::cons_to_prim_fast(m, U, w);
    EulerPhys::template flux_prim_fast<DIR>(m, w, U, F);
}

// ---- the physics of an adjoint sweep --------------------------------------------------
// AdjFull: the six Euler fields.  AdjPlanar: the four fields (rho, u, v, p) --
// planes 0, 1, 2 and 4 -- of a grid whose forward took the 4-field sweeps
// (w = +0 and lambda = 1 everywhere) with no formation energy, for seeds whose
// w and lambda parts are zero.  Then the w and lambda adjoints stay zero (their
// fluxes and the pressure carry no term that feeds them, and the terms they
// add to the other fields are products with w = 0, with the zero adjoints or
// with E_f = 0), so the 4-field VJPs are the 6-field ones with those zero
// terms left out, in the same order.

struct AdjFull {
    using P = EulerPhys;
    static constexpr int NF = 6;
    static constexpr const char *key = "euler_adjoint";
    __device__ static __forceinline__ constexpr int plane(int f) { return f; }
    __device__ static __forceinline__ void vjp_p2c(const EPar &m, const double (&w)[NF], const double (&Ub)[NF],
                                                   double (&wb)[NF])
    {
        vg::vjp_p2c(m, w, Ub, wb);
    }
    __device__ static __forceinline__ void vjp_c2p(const EPar &m, const double (&U)[NF], const double (&wb)[NF],
                                                   double (&Ub)[NF])
    {
        vg::vjp_c2p(m, U, wb, Ub);
    }
    template <int DIR>
    __device__ static __forceinline__ void vjp_flux(const EPar &m, const double (&U)[NF], const double (&Fb)[NF],
                                                    double (&Ub)[NF])
    {
        vg::vjp_flux<DIR>(m, U, Fb, Ub);
    }
};

struct AdjPlanar {
    using P = EulerPlanarPhys;
    static constexpr int NF = 4;
    static constexpr const char *key = "euler_planar_adjoint";
    __device__ static __forceinline__ constexpr int plane(int f) { return f < 3 ? f : 4; }
    // vjp_p2c with z = 0, lambda = 1, E_f = 0 and the w, lambda adjoints zero.
    __device__ static __forceinline__ void vjp_p2c(const EPar &m, const double (&w)[NF], const double (&Ub)[NF],
                                                   double (&wb)[NF])
    {
#pragma clang fp contract(fast)
        const double rho = w[0], u = w[1], v = w[2];
        const double vsq = u * u + v * v;
        wb[0] = Ub[0] + u * Ub[1] + v * Ub[2] + (0.5 * vsq) * Ub[3];
        wb[1] = rho * Ub[1] + rho * u * Ub[3];
        wb[2] = rho * Ub[2] + rho * v * Ub[3];
        wb[3] = Ub[3] * m.inv_gm1;
    }
    __device__ static __forceinline__ void vjp_c2p(const EPar &m, const double (&U)[NF], const double (&wb)[NF],
                                                   double (&Ub)[NF])
    {
#pragma clang fp contract(fast)
        const double q = rcp_nr(U[0]);
        const double u = U[1] * q, v = U[2] * q;
        const double g1 = m.gm1;
        const double vsq = u * u + v * v;
        Ub[0] = wb[0] - (u * wb[1] + v * wb[2]) * q + g1 * (0.5 * vsq) * wb[3];
        Ub[1] = wb[1] * q - g1 * u * wb[3];
        Ub[2] = wb[2] * q - g1 * v * wb[3];
        Ub[3] = g1 * wb[3];
    }
    template <int DIR>
    __device__ static __forceinline__ void vjp_flux(const EPar &m, const double (&U)[NF], const double (&Fb)[NF],
                                                    double (&Ub)[NF])
    {
#pragma clang fp contract(fast)
        const double q = rcp_nr(U[0]);
        const double vel[2] = {U[1] * q, U[2] * q};
        const double un = vel[DIR];
        const double g1 = m.gm1;
        const double vsq = vel[0] * vel[0] + vel[1] * vel[1];
        const double p = g1 * (U[3] - 0.5 * U[0] * vsq) - m.gpinf;
        const double H = (U[3] + p) * q;
        const double T = vel[0] * Fb[1] + vel[1] * Fb[2] + H * Fb[3];
        const double sp = g1 * (Fb[1 + DIR] + un * Fb[3]);
        Ub[0] = sp * (0.5 * vsq) - un * T;
        Ub[1] = un * Fb[1] - sp * vel[0];
        Ub[2] = un * Fb[2] - sp * vel[1];
        Ub[1 + DIR] += T + Fb[0];
        Ub[3] = sp + un * Fb[3];
    }
};

// Flux of a conserved vector of physics Q with one refined reciprocal of rho.
template <class Q, int DIR, int NF>
__device__ __forceinline__ void lin_flux_of(const EPar &m, const double (&U)[NF], double (&F)[NF])
{
    double w[NF];
    Q::cons_to_prim_fast(m, U, w);
    Q::template flux_prim_fast<DIR>(m, w, U, F);
}

// ---- the adjoint tile kernel ----------------------------------------------------------

struct AdjointArgs {
    const double *w;        // forward input of the sweep (SoA planes)
    const double *w35;      // where planes 3 and 5 (w, lambda) are read: w, or a grid holding the
                            // constant planes the 4-field forward sweeps leave out of their outputs
    const double *wbar_out; // adjoint of the sweep's output (SoA planes)
    double *wbar_in;        // adjoint of the sweep's input (SoA planes), interior cells
    double *ghost;          // per line: 4 positions (-2, -1, n, n+1) x 6 fields
    long long plane;
    int ny, nx, bc, limiter;
    double dx, bias, ts;    // ts: the sweep's time step (dt x its split fraction)
    EPar m;
};

constexpr int kAdjHalo = 4;

template <int DIR>
__device__ __forceinline__ long long adj_cell(int line, int k, int nx)
{
    return (DIR == 0) ? (long long)line * nx + k : (long long)k * nx + line;
}

// Tile of one block: LW adjacent lines (the fastest-varying thread index, so a
// Y sweep reads LW consecutive doubles per row) x BS / LW positions along the
// sweep, of which the middle BS / LW - 2 * kAdjHalo are outputs.
template <int BS, int LW>
struct AdjTile {
    static constexpr int lines = LW, len = BS / LW, out = BS / LW - 2 * kAdjHalo;
    static_assert(BS % LW == 0 && out > 0, "tile too short for the halo");
};

// Hides a register value's origin from the optimiser (an empty asm that may
// "modify" it), so a later recomputation from it is not merged with the
// earlier one and the earlier intermediates need not stay live.
template <int K>
__device__ __forceinline__ void opaque(double (&v)[K])
{
#pragma unroll
    for (int k = 0; k < K; k++) asm volatile("" : "+v"(v[k]));
}

// The tile runs in phases separated by barriers; three LDS planes of six
// fields carry what neighbours exchange, and only w, the slopes' differences
// and the evolved states stay in registers from the forward recomputation to
// the reverse pass (the rest is recomputed there), to keep occupancy up.
//   F1  U = p2c(w)                                   L0 <- U
//   F2  slopes, evolved XL/XR                        L1 <- XL, L2 <- f(XL)
//   F3  face q+1/2 flux (needs the next cell's XL)   L0 <- F
//   F4  N = U + c (F_{q-1/2} - F_{q+1/2}), N-bar     L1 <- N-bar
//   R1  face q+1/2 reverse                           L0 <- U'-bar, L2 <- F'-bar (next cell's left state)
//   R2  evolved states, evolution and slope reverse  L0 <- to the cell before, L2 <- to the cell after
//   R3  gather, w-bar = (d p2c / d w)^T U-bar
//
// WPE is the occupancy the register allocation must allow (waves per SIMD): 2
// keeps every value in registers (about 234 live at the peak), 3 spills some
// to scratch for half as many waves again in flight.  The arithmetic is the same.
// The state a position is linearised at.  A forward state that is not finite
// (or has rho <= 0) influences no finite result downstream -- every output that
// reads it is itself non-finite -- so its adjoint is zero in exact arithmetic;
// IEEE arithmetic would still form 0 x NaN = NaN in the products of its
// Jacobian and pass NaN to its finite neighbours.  Such a position is
// linearised at the default state (1, 0, 0, 0, 1, 1) instead, which makes its
// zero contributions exact zeros.  Finite states with rho > 0 -- every state of
// a solve whose grid stays admissible, such as the plain Euler adjoint's -- are
// left as they are.  In the two-material adjoint such states arise in a
// material's grid far from the region the level set makes real.
template <class A>
__device__ __forceinline__ void linearisation_state(double (&w)[A::NF])
{
    bool ok = w[0] > 0.0;
#pragma unroll
    for (int f = 0; f < A::NF; f++) ok = ok && isfinite(w[f]);
    if (!ok) A::P::default_state(w);
}

template <class A, int DIR, int BS, int LW, int WPE>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(WPE))) void euler_adjoint_sweep_kernel(
    AdjointArgs a)
{
    constexpr int NF = A::NF;
    using Q = typename A::P;
    using T = AdjTile<BS, LW>;
    __shared__ double L0[NF * BS], L1[NF * BS], L2[NF * BS];
    const int tid = threadIdx.x;
    const int q = tid / LW;                                  // position index within the tile
    const int n = (DIR == 0) ? a.nx : a.ny, n_lines = (DIR == 0) ? a.ny : a.nx;
    const int line_raw = (int)blockIdx.y * LW + tid % LW;
    const bool line_ok = line_raw < n_lines;
    const int line = line_ok ? line_raw : n_lines - 1;
    const int pos = -2 + (int)blockIdx.x * T::out - kAdjHalo + q;   // position along the line
    const EPar &m = a.m;
    const double half_c = 0.5 * (a.ts / a.dx), lf_c = a.dx / a.ts, ri_c = a.ts / a.dx, up_c = a.ts / a.dx;
    const double cl = 0.5 * (1.0 + a.bias), cr = 0.5 * (1.0 - a.bias);
    const bool real = (pos >= 0 && pos < n);
    const bool has_l = q > 0, has_r = q < T::len - 1;
    const int tl = has_l ? tid - LW : tid, tr = has_r ? tid + LW : tid;

    // ---- F1 ---------------------------------------------------------------------------
    double w[NF];
    {
        const int k = (pos >= -2 && pos < n + 2) ? ghost_index(pos, n, a.bc) : (pos < 0 ? 0 : n - 1);
        const long long c = adj_cell<DIR>(line, k, a.nx);
#pragma unroll
        for (int f = 0; f < NF; f++) w[f] = ((f == 3 || f == 5) && NF == 6 ? a.w35 : a.w)[A::plane(f) * a.plane + c];
        linearisation_state<A>(w);
        double U[NF];
        Q::prim_to_cons(m, w, U);
#pragma unroll
        for (int f = 0; f < NF; f++) L0[f * BS + tid] = U[f];
    }
    __syncthreads();
    // ---- F2 ---------------------------------------------------------------------------
    double dl[NF], dr[NF], XL[NF], XR[NF], URp[NF], FRp[NF];
    int pieces = 0;   // 3 bits per field: the limiter piece the forward selects
    {
        double xl[NF], xr[NF];
#pragma unroll
        for (int k = 0; k < NF; k++) {
            const double u = L0[k * BS + tid];
            dl[k] = u - L0[k * BS + tl];
            dr[k] = L0[k * BS + tr] - u;
            const double num = (fabs(dl[k]) < 1e-5) ? 1e-5 : dl[k];
            const double den = (fabs(dr[k]) < 1e-5) ? 1e-5 : dr[k];
            const int pc = limiter_piece(num / den, a.bias, a.limiter);
            pieces |= pc << (3 * k);
            const SlopeLin sl = slope_lin(dl[k], dr[k], pc, cl, cr, a.bias, a.limiter);
            const double d = sl.base * sl.xi;
            xl[k] = u - (0.5 * d);
            xr[k] = u + (0.5 * d);
        }
        double Fl[NF], Fr[NF];
        lin_flux_of<Q, DIR>(m, xl, Fl);
        lin_flux_of<Q, DIR>(m, xr, Fr);
#pragma unroll
        for (int k = 0; k < NF; k++) {
            const double e = half_c * (Fl[k] - Fr[k]);
            XL[k] = xl[k] + e;
            XR[k] = xr[k] + e;
        }
    }
    {
        double Fp[NF];
        lin_flux_of<Q, DIR>(m, XL, Fp);
#pragma unroll
        for (int f = 0; f < NF; f++) {
            L1[f * BS + tid] = XL[f];
            L2[f * BS + tid] = Fp[f];
        }
#pragma unroll
        for (int f = 0; f < NF; f++) URp[f] = XR[f];
        lin_flux_of<Q, DIR>(m, URp, FRp);
    }
    __syncthreads();
    // ---- F3: face q+1/2, a = right state of this cell, b = left state of the next ----
    double Us[NF];
    {
        double lf[NF], ri[NF];
#pragma unroll
        for (int k = 0; k < NF; k++) {
            const double ub = L1[k * BS + tr], fb = L2[k * BS + tr];
            lf[k] = 0.5 * ((FRp[k] + fb) + (lf_c * (URp[k] - ub)));
            Us[k] = 0.5 * ((URp[k] + ub) + (ri_c * (FRp[k] - fb)));
        }
        lin_flux_of<Q, DIR>(m, Us, ri);
#pragma unroll
        for (int k = 0; k < NF; k++) L0[k * BS + tid] = 0.5 * (lf[k] + ri[k]);
    }
    __syncthreads();
    // ---- F4 ---------------------------------------------------------------------------
    opaque(w);
    {
        double U[NF], N[NF], wb[NF], Nb[NF];
        Q::prim_to_cons_fast(m, w, U);
#pragma unroll
        for (int k = 0; k < NF; k++) N[k] = U[k] + (up_c * (L0[k * BS + tl] - L0[k * BS + tid]));
        if (real) {
            const long long c = adj_cell<DIR>(line, pos, a.nx);
#pragma unroll
            for (int f = 0; f < NF; f++) wb[f] = a.wbar_out[A::plane(f) * a.plane + c];
        } else {
#pragma unroll
            for (int f = 0; f < NF; f++) wb[f] = 0.0;
        }
        A::vjp_c2p(m, N, wb, Nb);
#pragma unroll
        for (int f = 0; f < NF; f++) L1[f * BS + tid] = Nb[f];
    }
    __syncthreads();
    // ---- R1: face q+1/2, F-bar = up_c (N-bar_{q+1} - N-bar_q) -------------------------
    double Uab[NF], Fab[NF];
    {
        double Fbar[NF], RIb[NF], Usb[NF];
#pragma unroll
        for (int k = 0; k < NF; k++) {
            Fbar[k] = has_r ? up_c * (L1[k * BS + tr] - L1[k * BS + tid]) : 0.0;
            RIb[k] = 0.5 * Fbar[k];
        }
        A::template vjp_flux<DIR>(m, Us, RIb, Usb);
#pragma unroll
        for (int k = 0; k < NF; k++) {
            const double LFb = 0.5 * Fbar[k];
            Uab[k] = 0.5 * lf_c * LFb + 0.5 * Usb[k];
            Fab[k] = 0.5 * LFb + 0.5 * ri_c * Usb[k];
            L0[k * BS + tid] = -0.5 * lf_c * LFb + 0.5 * Usb[k];   // next cell's left state: U'-bar
            L2[k * BS + tid] = 0.5 * LFb - 0.5 * ri_c * Usb[k];    // and F'-bar
        }
    }
    __syncthreads();
    // ---- R2 ---------------------------------------------------------------------------
    double Ut[NF], Ft[NF];
#pragma unroll
    for (int f = 0; f < NF; f++) {
        Ut[f] = has_l ? L0[f * BS + tl] : 0.0;
        Ft[f] = has_l ? L2[f * BS + tl] : 0.0;
    }
    __syncthreads();   // L0 and L2 are free for the slope exchange below
    opaque(w);
    opaque(dl);
    opaque(dr);
    double Ubar[NF];
    {
        double XLb[NF], XRb[NF];
        A::template vjp_flux<DIR>(m, XL, Ft, XLb);
        A::template vjp_flux<DIR>(m, XR, Fab, XRb);
#pragma unroll
        for (int k = 0; k < NF; k++) {
            XLb[k] += Ut[k];
            XRb[k] += Uab[k];
        }
        // evolution and extrapolation -> U-bar (own) and d-bar, then the slope
        double U[NF], xl[NF], xr[NF], eb[NF], tl1[NF], tr1[NF];
        Q::prim_to_cons_fast(m, w, U);
#pragma unroll
        for (int k = 0; k < NF; k++) {
            const SlopeLin sl = slope_lin(dl[k], dr[k], (pieces >> (3 * k)) & 7, cl, cr, a.bias, a.limiter);
            const double d = sl.base * sl.xi;
            xl[k] = U[k] - (0.5 * d);
            xr[k] = U[k] + (0.5 * d);
            eb[k] = half_c * (XLb[k] + XRb[k]);
        }
        A::template vjp_flux<DIR>(m, xl, eb, tl1);
        A::template vjp_flux<DIR>(m, xr, eb, tr1);
#pragma unroll
        for (int k = 0; k < NF; k++) {
            const double xlb = XLb[k] + tl1[k];
            const double xrb = XRb[k] - tr1[k];
            Ubar[k] = L1[k * BS + tid] + xlb + xrb;
            const double db = 0.5 * (xrb - xlb);
            const SlopeLin sl = slope_lin(dl[k], dr[k], (pieces >> (3 * k)) & 7, cl, cr, a.bias, a.limiter);
            const double bb = db * sl.xi;
            const double rb = db * sl.base * sl.dxi;
            const double dlb = cl * bb + (sl.cln ? 0.0 : rb * sl.iden);
            const double drb = cr * bb + (sl.cld ? 0.0 : -rb * sl.num * (sl.iden * sl.iden));
            Ubar[k] += dlb - drb;
            L0[k * BS + tid] = -dlb;   // for the cell before
            L2[k * BS + tid] = drb;    // for the cell after
        }
    }
    __syncthreads();
    // ---- R3 ---------------------------------------------------------------------------
#pragma unroll
    for (int f = 0; f < NF; f++) {
        if (has_r) Ubar[f] += L0[f * BS + tr];
        if (has_l) Ubar[f] += L2[f * BS + tl];
    }
    double wb[NF];
    A::vjp_p2c(m, w, Ubar, wb);
    if (line_ok && q >= kAdjHalo && q < T::len - kAdjHalo && pos >= -2 && pos < n + 2) {
        if (real) {
            const long long c = adj_cell<DIR>(line, pos, a.nx);
#pragma unroll
            for (int f = 0; f < NF; f++) a.wbar_in[A::plane(f) * a.plane + c] = wb[f];
        } else {
            const int g = (pos < 0) ? pos + 2 : 2 + (pos - n);   // -2,-1,n,n+1 -> 0..3
            double *dst = a.ghost + ((long long)line * 4 + g) * 6;
#pragma unroll
            for (int f = 0; f < NF; f++) dst[f] = wb[f];
        }
    }
}

// Marks the latest reverse step (forward step index) after which an adjoint
// field holds a non-finite value: *first = max(*first, step) when any field of
// any cell is not finite.
static __global__ __launch_bounds__(256) void adjoint_nonfinite_kernel(const double *__restrict__ a, long long plane,
                                                                       long long cells, int step, int *first)
{
    bool bad = false;
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < cells; c += (long long)gridDim.x * blockDim.x) {
#pragma unroll
        for (int f = 0; f < 6; f++) bad = bad || !isfinite(a[f * plane + c]);
    }
    if (bad) atomicMax(first, step);
}

// Clears *ok unless planes 3 and 5 (w, lambda) of the SoA buffer are zero in
// every cell.
static __global__ __launch_bounds__(256) void zero_planes_kernel(const double *__restrict__ a, long long plane,
                                                                 long long cells, int *ok)
{
    bool zero = true;
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < cells; c += (long long)gridDim.x * blockDim.x)
        zero = zero && a[3 * plane + c] == 0.0 && a[5 * plane + c] == 0.0;
    if (!zero) *ok = 0;
}

// Adds the ghost positions' adjoint to the cells they copy: transmissive
// g[-1] = c[0], g[-2] = c[1], g[n] = c[n-1], g[n+1] = c[n-2]; periodic wraps.
template <class A, int DIR>
__global__ void euler_adjoint_fold_kernel(double *wbar_in, const double *ghost, long long plane, int ny, int nx,
                                          int bc)
{
    const int line = blockIdx.x * blockDim.x + threadIdx.x;
    const int lines = (DIR == 0) ? ny : nx;
    const int n = (DIR == 0) ? nx : ny;
    if (line >= lines) return;
    const int pos[4] = {-2, -1, n, n + 1};
    // Apply in a fixed order per line, so the sums are deterministic.
    for (int g = 0; g < 4; g++) {
        const int k = ghost_index(pos[g], n, bc);
        const long long c = adj_cell<DIR>(line, k, nx);
        const double *src = ghost + ((long long)line * 4 + g) * 6;
        for (int f = 0; f < A::NF; f++) wbar_in[A::plane(f) * plane + c] += src[f];
    }
}

}  // namespace vg
