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
// One workgroup handles a tile of BS consecutive cells along the sweep, a halo
// of 4 on each side: outputs need N-bar of cells j-2..j+2.  Positions beyond the
// real grid carry no output adjoint (their N-bar is 0); the adjoint that lands on
// the two ghost positions on each side is written to a per-line ghost buffer and
// folded into the interior cells it copies (transmissive mirror or periodic wrap).
#pragma once

#include "vg_phys_euler.h"
#include "vg_sweep.h"

namespace vg {

using EPar = EulerPhys::Params;

// ---- analytic vector-Jacobian products of the Euler primitives ------------------

// w-bar = (d p2c / d w)^T U-bar  (p2c: Euler/eulerstatevector.cpp:78-91).
__device__ __forceinline__ void vjp_p2c(const EPar &m, const double (&w)[6], const double (&Ub)[6], double (&wb)[6])
{
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
    const double q = 1.0 / U[0];
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
// (eulerstatevector.cpp:93-120, 127-154): F = (m_n, m_n u + p e_n, ..., n (E + p), m_n lam).
template <int DIR>
__device__ __forceinline__ void vjp_flux(const EPar &m, const double (&U)[6], const double (&Fb)[6], double (&Ub)[6])
{
    const double q = 1.0 / U[0];
    const double vel[3] = {U[1] * q, U[2] * q, U[3] * q};
    const double lam = U[5] * q;
    const double un = vel[DIR];
    const double g1 = m.gm1;
    const double vsq = vel[0] * vel[0] + vel[1] * vel[1] + vel[2] * vel[2];
    const double p = g1 * (U[4] - 0.5 * U[0] * vsq - m.e_form * (U[5] - U[0])) - m.gpinf;
    // dp/dU
    double dp[6];
    dp[0] = g1 * (0.5 * vsq + m.e_form);
    dp[1] = -g1 * vel[0];
    dp[2] = -g1 * vel[1];
    dp[3] = -g1 * vel[2];
    dp[4] = g1;
    dp[5] = -g1 * m.e_form;
    // momentum rows: F_{1+c} = m_n vel_c + p [c == DIR]; d(m_n vel_c)/dU:
    //   dU0: -un vel_c, dU_{1+DIR}: vel_c, dU_{1+c}: un (both when c == DIR: 2 un)
    double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    // mass row F0 = U_{1+DIR}
    acc[1 + DIR] += Fb[0];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const double b = Fb[1 + c];
        acc[0] += -un * vel[c] * b;
        acc[1 + DIR] += vel[c] * b;
        acc[1 + c] += un * b;
        if (c == DIR) {
#pragma unroll
            for (int k = 0; k < 6; k++) acc[k] += dp[k] * b;
        }
    }
    // energy row F4 = un (U4 + p)
    {
        const double b = Fb[4];
        const double H = (U[4] + p) * q;
        acc[0] += (-un * H) * b;
        acc[1 + DIR] += H * b;
        acc[4] += un * b;
#pragma unroll
        for (int k = 0; k < 6; k++) acc[k] += un * dp[k] * b;
    }
    // reaction row F5 = un U5
    {
        const double b = Fb[5];
        acc[0] += (-un * lam) * b;
        acc[1 + DIR] += lam * b;
        acc[5] += un * b;
    }
#pragma unroll
    for (int k = 0; k < 6; k++) Ub[k] = acc[k];
}

// d xi / d r for the branch the forward evaluation takes (slopelimiters.cpp:7-74).
__device__ __forceinline__ double limiter_derivative(double r, double bias, int limiter)
{
    if (r < 0.0) return 0.0;
    const double den = (1.0 - bias) + (1.0 + bias) * r;
    const double dR = -2.0 * (1.0 + bias) / (den * den);   // d/dr of 2 / ((1-w) + (1+w) r)
    const double R = 2.0 / den;
    if (limiter == 0) {
        if (r < 0.5) return 2.0;
        if (r < 1.0) return 0.0;
        // smin(smin(r, R), 2): the first argument wins unless the second is smaller
        const double a = (R < r) ? R : r;
        const double da = (R < r) ? dR : 1.0;
        return (2.0 < a) ? 0.0 : da;
    }
    if (limiter == 1) {
        const double vl = (2.0 * r) / (1.0 + r);
        const double dvl = 2.0 / ((1.0 + r) * (1.0 + r));
        return (R < vl) ? dR : dvl;
    }
    if (r < 1.0) return 1.0;
    return (R < 1.0) ? dR : 0.0;
}

// ---- forward pieces (EXACT arithmetic, as vg_scheme.h) ------------------------------

template <int DIR>
__device__ __forceinline__ void fwd_flux(const EPar &m, const double (&U)[6], double (&F)[6])
{
    EulerPhys::template flux<DIR>(m, U, F);
}

// ---- the adjoint tile kernel ----------------------------------------------------------

struct AdjointArgs {
    const double *w;        // forward input of the sweep (SoA planes)
    const double *wbar_out; // adjoint of the sweep's output (SoA planes)
    double *wbar_in;        // adjoint of the sweep's input (SoA planes), interior cells
    double *ghost;          // per line: 4 positions (-2, -1, n, n+1) x 6 fields
    long long plane;
    int ny, nx, bc, limiter;
    double dx, bias, ts;    // ts: the sweep's time step (dt x its split fraction)
    EPar m;
};

constexpr int kAdjBS = 128;
constexpr int kAdjHalo = 4;
constexpr int kAdjOut = kAdjBS - 2 * kAdjHalo;

template <int DIR>
__device__ __forceinline__ long long adj_cell(int line, int k, int nx)
{
    return (DIR == 0) ? (long long)line * nx + k : (long long)k * nx + line;
}

template <int DIR>
__global__ __launch_bounds__(kAdjBS) void euler_adjoint_sweep_kernel(AdjointArgs a)
{
    constexpr int NF = 6;
    __shared__ double s1[2 * NF * kAdjBS];
    __shared__ double s2[NF * kAdjBS];
    const int tid = threadIdx.x;
    const int line = blockIdx.y;
    const int n = (DIR == 0) ? a.nx : a.ny;
    const int pos = -2 + (int)blockIdx.x * kAdjOut - kAdjHalo + tid;   // position along the line
    const EPar &m = a.m;
    const double half_c = 0.5 * (a.ts / a.dx), lf_c = a.dx / a.ts, ri_c = a.ts / a.dx, up_c = a.ts / a.dx;
    const double cl = 0.5 * (1.0 + a.bias), cr = 0.5 * (1.0 - a.bias);
    const bool real = (pos >= 0 && pos < n);

    // ---- forward recompute -------------------------------------------------------
    double w[NF], U[NF];
    {
        const int k = (pos >= -2 && pos < n + 2) ? ghost_index(pos, n, a.bc) : (pos < 0 ? 0 : n - 1);
        const long long c = adj_cell<DIR>(line, k, a.nx);
#pragma unroll
        for (int f = 0; f < NF; f++) w[f] = a.w[f * a.plane + c];
    }
    EulerPhys::prim_to_cons(m, w, U);
#pragma unroll
    for (int f = 0; f < NF; f++) s2[f * kAdjBS + tid] = U[f];
    __syncthreads();
    double UL[NF], UR[NF];
    const int tl = tid > 0 ? tid - 1 : tid, tr = tid < kAdjBS - 1 ? tid + 1 : tid;
#pragma unroll
    for (int f = 0; f < NF; f++) {
        UL[f] = s2[f * kAdjBS + tl];
        UR[f] = s2[f * kAdjBS + tr];
    }
    // slope, extrapolation, evolution (keeping what the reverse pass needs)
    double dl[NF], dr[NF], base[NF], xi[NF], rr[NF], xl[NF], xr[NF];
    bool cln[NF], cld[NF];
#pragma unroll
    for (int k = 0; k < NF; k++) {
        dl[k] = U[k] - UL[k];
        dr[k] = UR[k] - U[k];
        base[k] = (cl * dl[k]) + (cr * dr[k]);
        cln[k] = fabs(dl[k]) < 1e-5;
        cld[k] = fabs(dr[k]) < 1e-5;
        const double num = cln[k] ? 1e-5 : dl[k];
        const double den = cld[k] ? 1e-5 : dr[k];
        rr[k] = num / den;
        xi[k] = limiter_value(rr[k], a.bias, a.limiter);
        const double d = base[k] * xi[k];
        xl[k] = U[k] - (0.5 * d);
        xr[k] = U[k] + (0.5 * d);
    }
    double XL[NF], XR[NF];
    {
        double Fl[NF], Fr[NF];
        fwd_flux<DIR>(m, xl, Fl);
        fwd_flux<DIR>(m, xr, Fr);
#pragma unroll
        for (int k = 0; k < NF; k++) {
            const double e = half_c * (Fl[k] - Fr[k]);
            XL[k] = xl[k] + e;
            XR[k] = xr[k] + e;
        }
    }
    // round trips of the evolved states: W = c2p(X), U' = p2c(W), F' = f(U')
    double WL[NF], WR[NF], ULp[NF], URp[NF], FLp[NF], FRp[NF];
    EulerPhys::cons_to_prim(m, XL, WL);
    EulerPhys::prim_to_cons(m, WL, ULp);
    fwd_flux<DIR>(m, ULp, FLp);
    EulerPhys::cons_to_prim(m, XR, WR);
    EulerPhys::prim_to_cons(m, WR, URp);
    fwd_flux<DIR>(m, URp, FRp);
    __syncthreads();
#pragma unroll
    for (int f = 0; f < NF; f++) {
        s1[f * kAdjBS + tid] = ULp[f];
        s1[(NF + f) * kAdjBS + tid] = FLp[f];
    }
    __syncthreads();
    // face tid+1/2: a = right state of this cell, b = left state of the next
    double Ub[NF], Fb[NF], Us[NF], Fface[NF];
#pragma unroll
    for (int f = 0; f < NF; f++) {
        Ub[f] = s1[f * kAdjBS + tr];
        Fb[f] = s1[(NF + f) * kAdjBS + tr];
    }
    {
        double lf[NF], ri[NF];
#pragma unroll
        for (int k = 0; k < NF; k++) {
            lf[k] = 0.5 * ((FRp[k] + Fb[k]) + (lf_c * (URp[k] - Ub[k])));
            Us[k] = 0.5 * ((URp[k] + Ub[k]) + (ri_c * (FRp[k] - Fb[k])));
        }
        fwd_flux<DIR>(m, Us, ri);
#pragma unroll
        for (int k = 0; k < NF; k++) Fface[k] = 0.5 * (lf[k] + ri[k]);
    }
    __syncthreads();
#pragma unroll
    for (int f = 0; f < NF; f++) s2[f * kAdjBS + tid] = Fface[f];
    __syncthreads();
    double N[NF];
#pragma unroll
    for (int k = 0; k < NF; k++) N[k] = U[k] + (up_c * (s2[k * kAdjBS + tl] - Fface[k]));

    // ---- reverse ----------------------------------------------------------------------
    double Nb[NF];
    {
        double wb[NF];
        if (real) {
            const long long c = adj_cell<DIR>(line, pos, a.nx);
#pragma unroll
            for (int f = 0; f < NF; f++) wb[f] = a.wbar_out[f * a.plane + c];
        } else {
#pragma unroll
            for (int f = 0; f < NF; f++) wb[f] = 0.0;
        }
        vjp_c2p(m, N, wb, Nb);
    }
    __syncthreads();
#pragma unroll
    for (int f = 0; f < NF; f++) s2[f * kAdjBS + tid] = Nb[f];
    __syncthreads();
    // face tid+1/2: F-bar = up_c (N-bar_{t+1} - N-bar_t)
    double Uab[NF], Fab[NF], Ubb[NF], Fbb[NF];
    {
        double Fbar[NF], LFb[NF], RIb[NF], Usb[NF];
#pragma unroll
        for (int k = 0; k < NF; k++) {
            Fbar[k] = (tid < kAdjBS - 1) ? up_c * (s2[k * kAdjBS + tr] - Nb[k]) : 0.0;
            LFb[k] = 0.5 * Fbar[k];
            RIb[k] = 0.5 * Fbar[k];
        }
        vjp_flux<DIR>(m, Us, RIb, Usb);
#pragma unroll
        for (int k = 0; k < NF; k++) {
            Uab[k] = 0.5 * lf_c * LFb[k] + 0.5 * Usb[k];
            Ubb[k] = -0.5 * lf_c * LFb[k] + 0.5 * Usb[k];
            Fab[k] = 0.5 * LFb[k] + 0.5 * ri_c * Usb[k];
            Fbb[k] = 0.5 * LFb[k] - 0.5 * ri_c * Usb[k];
        }
    }
    __syncthreads();
#pragma unroll
    for (int f = 0; f < NF; f++) {
        s1[f * kAdjBS + tid] = Ubb[f];
        s1[(NF + f) * kAdjBS + tid] = Fbb[f];
    }
    __syncthreads();
    // round-trip chains of both evolved states -> X-bar
    double XLb[NF], XRb[NF];
    {
        double Ut[NF], Ft[NF], t1[NF], Wb[NF];
        // left state of this cell receives (U-bar, F-bar) from face tid-1/2
#pragma unroll
        for (int f = 0; f < NF; f++) {
            Ut[f] = (tid > 0) ? s1[f * kAdjBS + tl] : 0.0;
            Ft[f] = (tid > 0) ? s1[(NF + f) * kAdjBS + tl] : 0.0;
        }
        vjp_flux<DIR>(m, ULp, Ft, t1);
#pragma unroll
        for (int k = 0; k < NF; k++) t1[k] += Ut[k];
        vjp_p2c(m, WL, t1, Wb);
        vjp_c2p(m, XL, Wb, XLb);
        // right state of this cell: from this thread's face
        vjp_flux<DIR>(m, URp, Fab, t1);
#pragma unroll
        for (int k = 0; k < NF; k++) t1[k] += Uab[k];
        vjp_p2c(m, WR, t1, Wb);
        vjp_c2p(m, XR, Wb, XRb);
    }
    // evolution and extrapolation -> U-bar (own) and d-bar
    double Ubar[NF], db[NF];
    {
        double eb[NF], tl1[NF], tr1[NF];
#pragma unroll
        for (int k = 0; k < NF; k++) eb[k] = half_c * (XLb[k] + XRb[k]);
        vjp_flux<DIR>(m, xl, eb, tl1);
        vjp_flux<DIR>(m, xr, eb, tr1);
#pragma unroll
        for (int k = 0; k < NF; k++) {
            const double xlb = XLb[k] + tl1[k];
            const double xrb = XRb[k] - tr1[k];
            Ubar[k] = Nb[k] + xlb + xrb;
            db[k] = 0.5 * (xrb - xlb);
        }
    }
    // slope -> contributions to U-bar of this cell and both neighbours
    double toL[NF], toR[NF];
#pragma unroll
    for (int k = 0; k < NF; k++) {
        const double bb = db[k] * xi[k];
        const double rb = db[k] * base[k] * limiter_derivative(rr[k], a.bias, a.limiter);
        const double num = cln[k] ? 1e-5 : dl[k];
        const double den = cld[k] ? 1e-5 : dr[k];
        const double dlb = cl * bb + (cln[k] ? 0.0 : rb / den);
        const double drb = cr * bb + (cld[k] ? 0.0 : -rb * num / (den * den));
        Ubar[k] += dlb - drb;
        toL[k] = -dlb;
        toR[k] = drb;
    }
    __syncthreads();
#pragma unroll
    for (int f = 0; f < NF; f++) {
        s1[f * kAdjBS + tid] = toL[f];          // for cell tid-1
        s1[(NF + f) * kAdjBS + tid] = toR[f];   // for cell tid+1
    }
    __syncthreads();
#pragma unroll
    for (int f = 0; f < NF; f++) {
        if (tid < kAdjBS - 1) Ubar[f] += s1[f * kAdjBS + tr];
        if (tid > 0) Ubar[f] += s1[(NF + f) * kAdjBS + tl];
    }
    double wb[NF];
    vjp_p2c(m, w, Ubar, wb);
    // ---- store: interior cells, or the ghost buffer ---------------------------------
    if (tid >= kAdjHalo && tid < kAdjBS - kAdjHalo && pos >= -2 && pos < n + 2) {
        if (real) {
            const long long c = adj_cell<DIR>(line, pos, a.nx);
#pragma unroll
            for (int f = 0; f < NF; f++) a.wbar_in[f * a.plane + c] = wb[f];
        } else {
            const int g = (pos < 0) ? pos + 2 : 2 + (pos - n);   // -2,-1,n,n+1 -> 0..3
            double *dst = a.ghost + ((long long)line * 4 + g) * NF;
#pragma unroll
            for (int f = 0; f < NF; f++) dst[f] = wb[f];
        }
    }
}

// Adds the ghost positions' adjoint to the cells they copy: transmissive
// g[-1] = c[0], g[-2] = c[1], g[n] = c[n-1], g[n+1] = c[n-2]; periodic wraps.
template <int DIR>
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
        for (int f = 0; f < 6; f++) wbar_in[f * plane + c] += src[f];
    }
}

// Clock step with a prescribed time step: dts[it] instead of the CFL formula
// (the forward of a frozen-dt derivative check, and recomputation in the
// adjoint).  Same loop test and t / it bookkeeping as clock_begin_kernel.
static __global__ void clock_fixed_kernel(Clock *c, const double *dts, int n_dts)
{
    if (c->done) return;
    if (c->pending) {
        c->t += c->dt;
        c->it += 1;
        c->pending = 0;
    }
    if (c->it >= n_dts || (c->max_steps > 0 && c->it >= c->max_steps)) {
        c->done = 1;
        return;
    }
    const double dt = dts[c->it];
    c->dt = dt;
    if (c->steps_this_run == 0 || dt < c->dt_min) c->dt_min = dt;
    if (c->steps_this_run == 0 || dt > c->dt_max) c->dt_max = dt;
    c->steps_this_run += 1;
    c->pending = 1;
}

}  // namespace vg
