// Two-material Euler with the real ghost fluid method: interface Riemann
// solvers (HLLC and exact) and the line kernels of the ghost fill and the
// level-set advection (RGFM/rgfmsolver.cpp, RGFM/hllcsolver.cpp,
// RGFM/exactsolver.cpp).
//
// The reference scans each grid line sequentially, toggling a material flag at
// every sign change of the level set.  Here one workgroup owns a line and
// replaces the scan by a block-wide prefix count of the sign changes: the flag
// at position k is the initial flag xor the parity of the crossings before k.
//
// Ghost fill: every crossing solves its interface Riemann problem from the
// states as they were before the pass (phase 1, results in a per-cell record),
// then every cell takes the value of the last crossing, in line order, whose
// 4-cell window covers it (phase 2, in place) -- the same overwrite order as
// the reference's loop.
#pragma once

#include "vg_phys_euler.h"
#include "vg_sweep.h"

namespace vg {

using EP = EulerPhys::Params;
typedef double State[6];

__device__ __forceinline__ double snd(const State &w, const EP &m) { return EulerPhys::eos_c(m, w[0], w[4]); }

// ---- scalar types of the interface solver ------------------------------------------
// The HLLC interface state is written once over a scalar type T: double for the
// fill (every operation exactly as before, the EOS and conversions being
// EulerPhys's own), and a forward-mode dual number for the adjoint of the fill,
// whose branch tests read the value part (the same values as the double run).

struct Dual {
    double v, d;
};
using ::pow;
using ::sqrt;
__device__ __forceinline__ Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.d + b.d}; }
__device__ __forceinline__ Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.d - b.d}; }
__device__ __forceinline__ Dual operator*(Dual a, Dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
__device__ __forceinline__ Dual operator/(Dual a, Dual b) { return {a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)}; }
__device__ __forceinline__ Dual operator+(Dual a, double b) { return {a.v + b, a.d}; }
__device__ __forceinline__ Dual operator+(double a, Dual b) { return {a + b.v, b.d}; }
__device__ __forceinline__ Dual operator-(Dual a, double b) { return {a.v - b, a.d}; }
__device__ __forceinline__ Dual operator-(double a, Dual b) { return {a - b.v, -b.d}; }
__device__ __forceinline__ Dual operator*(Dual a, double b) { return {a.v * b, a.d * b}; }
__device__ __forceinline__ Dual operator*(double a, Dual b) { return {a * b.v, a * b.d}; }
__device__ __forceinline__ Dual operator/(Dual a, double b) { return {a.v / b, a.d / b}; }
__device__ __forceinline__ Dual operator/(double a, Dual b) { return {a / b.v, -a * b.d / (b.v * b.v)}; }
__device__ __forceinline__ Dual sqrt(Dual a)
{
    const double r = sqrt(a.v);
    return {r, a.d / (2.0 * r)};
}
__device__ __forceinline__ Dual pow(Dual a, double e)
{
    const double r = pow(a.v, e);
    return {r, e * pow(a.v, e - 1.0) * a.d};
}
__device__ __forceinline__ double val(double x) { return x; }
__device__ __forceinline__ double val(Dual x) { return x.v; }
// A constant of the scalar type of `like`.
__device__ __forceinline__ double t_const(double c, double) { return c; }
__device__ __forceinline__ Dual t_const(double c, Dual) { return {c, 0.0}; }
template <class T>
__device__ __forceinline__ T tmin(T a, T b) { return (val(b) < val(a)) ? b : a; }
template <class T>
__device__ __forceinline__ T tmax(T a, T b) { return (val(a) < val(b)) ? b : a; }

__device__ __forceinline__ double t_snd(const double (&w)[6], const EP &m) { return EulerPhys::eos_c(m, w[0], w[4]); }
__device__ __forceinline__ Dual t_snd(const Dual (&w)[6], const EP &m)
{
    return sqrt((m.gamma * (w[4] + m.p_inf)) / w[0]);
}
__device__ __forceinline__ void t_p2c(const EP &m, const double (&w)[6], double (&U)[6])
{
    EulerPhys::prim_to_cons(m, w, U);
}
__device__ __forceinline__ void t_p2c(const EP &m, const Dual (&w)[6], Dual (&U)[6])
{
    const Dual rho = w[0], u = w[1], v = w[2], z = w[3], p = w[4], lam = w[5];
    const Dual vsq = (u * u) + (v * v) + (z * z);
    U[0] = rho;
    U[1] = rho * u;
    U[2] = rho * v;
    U[3] = rho * z;
    U[4] = rho * ((0.5 * vsq) + ((p + m.gpinf) / (m.gm1 * rho)) + (m.e_form * (lam - 1.0)));
    U[5] = rho * lam;
}
__device__ __forceinline__ void t_c2p(const EP &m, const double (&U)[6], double (&w)[6])
{
    EulerPhys::cons_to_prim(m, U, w);
}
__device__ __forceinline__ void t_c2p(const EP &m, const Dual (&U)[6], Dual (&w)[6])
{
    const Dual rho = U[0];
    const Dual u = U[1] / rho, v = U[2] / rho, z = U[3] / rho;
    const Dual lam = U[5] / rho;
    const Dual vsq = (u * u) + (v * v) + (z * z);
    const Dual e = (U[4] / rho) - (0.5 * vsq) - (m.e_form * (lam - 1.0));
    w[0] = rho;
    w[1] = u;
    w[2] = v;
    w[3] = z;
    w[4] = (e * m.gm1 * rho) - m.gpinf;
    w[5] = lam;
}

// ---- HLLC (hllcsolver.cpp) ---------------------------------------------------

// computeStarRegionDensity (hllcsolver.cpp:7-10)
template <class T>
__device__ __forceinline__ T hllc_star_rho(T rho, T S, T u, T Ss)
{
    return rho * ((S - u) / (S - Ss));
}

// PVRS (:182-254); the product 0.25 (rho_a + rho_b)(c_a + c_b).
template <int DIR, class T>
__device__ __forceinline__ T pvrs_prod(const T (&a)[6], const T (&b)[6], const EP &m1, const EP &m2)
{
    return 0.25 * (a[0] + b[0]) * (t_snd(a, m1) + t_snd(b, m2));
}
template <int DIR, class T>
__device__ __forceinline__ T pvrs_vel(const T (&a)[6], const T (&b)[6], const EP &m1, const EP &m2)
{
    const T prod = pvrs_prod<DIR>(a, b, m1, m2);
    return (0.5 * (a[1 + DIR] + b[1 + DIR])) + (0.5 * (a[4] - b[4]) / prod);
}
template <int DIR, class T>
__device__ __forceinline__ T pvrs_pres(const T (&a)[6], const T (&b)[6], const EP &m1, const EP &m2)
{
    const T prod = pvrs_prod<DIR>(a, b, m1, m2);
    return (0.5 * (a[4] + b[4])) + (0.5 * (a[1 + DIR] - b[1 + DIR]) * prod);
}

// Two-rarefaction estimates (:256-348), equal-gamma materials only.
template <int DIR, class T>
__device__ __forceinline__ T two_rare_vel(const T (&a)[6], const T (&b)[6], const EP &m1, const EP &m2)
{
    const T ca = t_snd(a, m1), cb = t_snd(b, m2);
    const double g = 0.5 * (m1.gamma + m2.gamma);
    const double ex = (g - 1.0) / (2.0 * g);
    const T pr = pow((a[4] / b[4]), ex);
    const double co = 2.0 / (g - 1.0);
    return ((pr * (a[1 + DIR] / ca)) + (b[1 + DIR] / cb) + (co * (pr - 1.0))) / ((pr / ca) + (1.0 / cb));
}
template <int DIR, class T>
__device__ __forceinline__ T two_rare_pres(const T (&a)[6], const T (&b)[6], const EP &m1, const EP &m2)
{
    const T ca = t_snd(a, m1), cb = t_snd(b, m2);
    const double g = 0.5 * (m1.gamma + m2.gamma);
    const double co = 0.5 * (g - 1.0);
    const T us = two_rare_vel<DIR>(a, b, m1, m2);
    const double ex = 2.0 * (g / (g - 1.0));
    const T ld = 1.0 + (co * ((a[1 + DIR] - us) / ca));
    const T rd = 1.0 + (co * ((us - b[1 + DIR]) / cb));
    return 0.5 * ((a[4] * pow(ld, ex)) + (b[4] * pow(rd, ex)));
}

// Two-shock estimates (:350-462).
template <int DIR, class T>
__device__ __forceinline__ void two_shock_coeffs(const T (&a)[6], const T (&b)[6], const EP &m1, const EP &m2, T &ga,
                                                 T &gb)
{
    const double g = 0.5 * (m1.gamma + m2.gamma);
    const T ppv = pvrs_pres<DIR>(a, b, m1, m2);
    const double A = 2.0 / (g + 1.0);
    const double B = (g - 1.0) / (g + 1.0);
    ga = sqrt((A / a[0]) / ((B * a[4]) + ppv));
    gb = sqrt((A / b[0]) / ((B * b[4]) + ppv));
}
template <int DIR, class T>
__device__ __forceinline__ T two_shock_pres(const T (&a)[6], const T (&b)[6], const EP &m1, const EP &m2)
{
    T ga, gb;
    two_shock_coeffs<DIR>(a, b, m1, m2, ga, gb);
    return ((ga * a[4]) + (gb * b[4]) - (b[1 + DIR] - a[1 + DIR])) / (ga + gb);
}
template <int DIR, class T>
__device__ __forceinline__ T two_shock_vel(const T (&a)[6], const T (&b)[6], const EP &m1, const EP &m2)
{
    const T ps = two_shock_pres<DIR>(a, b, m1, m2);
    T ga, gb;
    two_shock_coeffs<DIR>(a, b, m1, m2, ga, gb);
    return (0.5 * (a[1 + DIR] + b[1 + DIR])) + (0.5 * ((gb * (ps - b[4])) - (ga * (ps - a[4]))));
}

// The adaptive estimate (:464-606).  For the y direction the velocity's
// branch test compares the PVRS *velocity* with the pressures, as the
// reference does (hllcsolver.cpp:552).
__device__ __forceinline__ bool use_pvrs(double pa, double pb, double ppv, const EP &m1, const EP &m2)
{
    const double pmin = smin(pa, pb), pmax = smax(pa, pb);
    const double ratio = pmax / pmin;
    return ((ratio <= 2.0) && (pmin <= ppv) && (ppv <= pmax)) || fabs(m1.gamma - m2.gamma) > 1e-8 || m1.p_inf > 0.0 ||
           m2.p_inf > 0.0;
}
template <int DIR, class T>
__device__ __forceinline__ T hllc_star_vel(const T (&a)[6], const T (&b)[6], const EP &m1, const EP &m2)
{
    const T ppv = (DIR == 0) ? pvrs_pres<0>(a, b, m1, m2) : pvrs_vel<1>(a, b, m1, m2);
    if (use_pvrs(val(a[4]), val(b[4]), val(ppv), m1, m2)) return pvrs_vel<DIR>(a, b, m1, m2);
    if (val(ppv) < smin(val(a[4]), val(b[4]))) return two_rare_vel<DIR>(a, b, m1, m2);
    return two_shock_vel<DIR>(a, b, m1, m2);
}
template <int DIR, class T>
__device__ __forceinline__ T hllc_star_pres(const T (&a)[6], const T (&b)[6], const EP &m1, const EP &m2)
{
    const T ppv = pvrs_pres<DIR>(a, b, m1, m2);
    if (use_pvrs(val(a[4]), val(b[4]), val(ppv), m1, m2)) return ppv;
    if (val(ppv) < smin(val(a[4]), val(b[4]))) return two_rare_pres<DIR>(a, b, m1, m2);
    return two_shock_pres<DIR>(a, b, m1, m2);
}

// computeWaveSpeedWeighting (:608-623).
template <class T>
__device__ __forceinline__ T hllc_weighting(T ps, const T (&w)[6], const EP &m)
{
    const T p = w[4];
    if (val(ps) <= val(p)) return t_const(1.0, p);
    return sqrt(1.0 + ((m.gamma + 1.0) / (2.0 * m.gamma)) * ((ps + m.p_inf) / (p + m.p_inf) - 1.0));
}

// The interface state of HLLCSolver::solveX/solveY (:114-180) -- only its
// pressure and normal velocity are used -- and the two star densities
// (computeLeft/Right/Top/BottomStarRegionDensity, :12-58).
template <int DIR, class T>
__device__ __forceinline__ void hllc_interface(const T (&a)[6], const T (&b)[6], const EP &m1, const EP &m2, T &p,
                                               T &u, T &ra, T &rb)
{
    const T Ss = hllc_star_vel<DIR>(a, b, m1, m2);
    const T ps = hllc_star_pres<DIR>(a, b, m1, m2);
    const T Sa = a[1 + DIR] - (hllc_weighting(ps, a, m1) * t_snd(a, m1));
    const T Sb = b[1 + DIR] + (hllc_weighting(ps, b, m2) * t_snd(b, m2));
    if (0.0 <= val(Sa)) {
        p = a[4];
        u = a[1 + DIR];
    } else if (val(Sb) <= 0.0) {
        p = b[4];
        u = b[1 + DIR];
    } else {
        // computeStarRegionX/YConservedVariableVector (:60-112), then
        // setConservedVariableVector with that side's material.
        const bool left = (0.0 <= val(Ss));
        const T(&w)[6] = left ? a : b;
        const EP &m = left ? m1 : m2;
        const T S = left ? Sa : Sb;
        const T rho = w[0], pw = w[4], un = w[1 + DIR];
        T W[6];
#pragma unroll
        for (int f = 0; f < 6; f++) W[f] = w[f];
        T Uw[6];
        t_p2c(m, W, Uw);
        const T rs = hllc_star_rho(rho, S, un, Ss);
        const T et = (Ss - un) * (Ss + (pw / (rho * (S - un))));
        T U[6];
        U[0] = rs;
        U[1] = rs * (DIR == 0 ? Ss : w[1]);
        U[2] = rs * (DIR == 1 ? Ss : w[2]);
        U[3] = rs * w[3];
        U[4] = rs * ((Uw[4] / rho) + et);
        U[5] = rs;
        T out[6];
        t_c2p(m, U, out);
        p = out[4];
        u = out[1 + DIR];
    }
    ra = hllc_star_rho(a[0], Sa, a[1 + DIR], Ss);
    rb = hllc_star_rho(b[0], Sb, b[1 + DIR], Ss);
}

// ---- exact Riemann solver (exactsolver.cpp; Toro, ch. 4, stiffened gas) --------

__device__ __forceinline__ double ex_A(double rho, double g) { return 2.0 / (rho * (g + 1.0)); }
__device__ __forceinline__ double ex_B(double p, double g, double pinf)
{
    return (p * ((g - 1.0) / (g + 1.0))) + ((2.0 * g * pinf) / (g + 1.0));
}
// computeWaveJumpFunctionComponent (:165-186) and its derivative (:202-221).
__device__ __forceinline__ double ex_f(double ps, const State &w, const EP &m)
{
    const double rho = w[0], p = w[4], c = snd(w, m), g = m.gamma, pi = m.p_inf;
    if (ps > p) return (ps - p) * sqrt(ex_A(rho, g) / (ps + ex_B(p, g, pi)));
    const double co = pow((ps + pi) / (p + pi), (g - 1.0) / (2.0 * g));
    return 2.0 * (c / (g - 1.0)) * (co - 1.0);
}
__device__ __forceinline__ double ex_df(double ps, const State &w, const EP &m)
{
    const double rho = w[0], p = w[4], c = snd(w, m), g = m.gamma, pi = m.p_inf;
    if (ps > p) {
        const double co = sqrt(ex_A(rho, g) / (ps + ex_B(p, g, pi)));
        return (1.0 - ((ps - p) / (2.0 * (ps + ex_B(p, g, pi))))) * co;
    }
    const double co = pow((ps + pi) / (p + pi), -(g + 1.0) / (2.0 * g));
    return (co * c) / (g * (p + pi));
}
__device__ __forceinline__ double ex_star_rho(double ps, const State &w, const EP &m)
{
    const double g = m.gamma, pi = m.p_inf;
    if (ps < w[4]) return w[0] * pow((ps + pi) / (w[4] + pi), 1.0 / g);
    const double q = (ps + pi) / (w[4] + pi), k = (g - 1.0) / (g + 1.0);
    return w[0] * ((q + k) / (k * q + 1.0));
}
// Newton for p* (:115-163), then the state at x/t = 0 (solveX/solveY, :262-446):
// its pressure and normal velocity, and the two star densities at that pressure.
template <int DIR>
__device__ __forceinline__ void exact_interface(const State &a, const State &b, const EP &m1, const EP &m2, double &p,
                                                double &u, double &ra, double &rb)
{
    double cur = 0.5 * (a[4] + b[4]);
    double nw = 3.0 * cur;
    for (int it = 0; (2.0 * (fabs(nw - cur) / fabs(nw + cur))) >= 1e-8 && it < 10000; it++) {
        cur = nw;
        const double f = ex_f(cur, a, m1) + ex_f(cur, b, m2) + b[1 + DIR] - a[1 + DIR];
        const double df = ex_df(cur, a, m1) + ex_df(cur, b, m2);
        nw = cur - (f / df);
        if (nw < 1e-8) nw = 1e-8;
    }
    const double ps = nw;
    const double us = 0.5 * (a[1 + DIR] + b[1 + DIR] + ex_f(ps, b, m2) - ex_f(ps, a, m1));
    const double s = 0.0;
    const bool left = (s < us);
    const State &w = left ? a : b;
    const EP &m = left ? m1 : m2;
    const double sign = left ? -1.0 : 1.0;
    const double rho = w[0], un = w[1 + DIR], pw = w[4], c = snd(w, m), g = m.gamma, pi = m.p_inf;
    p = ps;
    u = us;
    if (ps < pw) {
        const double head = un + sign * c;
        const double tail = us + sign * (c * pow((ps + pi) / (pw + pi), (g - 1.0) / (2.0 * g)));
        if ((sign < 0 && s < head) || (sign > 0 && s > head)) {
            p = pw;
            u = un;
        } else if ((sign < 0 && s < tail) || (sign > 0 && s > tail)) {
            const double cs = -sign * c;
            const double k = (2.0 / (g + 1.0)) + (g - 1.0) * ((un - s) / ((g + 1.0) * cs));
            u = 2.0 * (cs + (g - 1.0) * (un / 2.0) + s) / (g + 1.0);
            p = ((pw + pi) * pow(k, (2.0 * g) / (g - 1.0))) - pi;
        }
    } else {
        const double Q = sqrt((ps + ex_B(pw, g, pi)) / ex_A(rho, g));
        const double shock = un + sign * (Q / rho);
        if ((sign < 0 && s < shock) || (sign > 0 && s > shock)) {
            p = pw;
            u = un;
        }
    }
    ra = ex_star_rho(p, a, m1);
    rb = ex_star_rho(p, b, m2);
}

// ---- line kernels -------------------------------------------------------------

// Loop test of the step about to start (secondordersolver.cpp:295 style
// `while (currentTime < finalTime)`): folds in the previous step and stops the
// clock before anything of the next step runs.
static __global__ void clock_check_kernel(Clock *c)
{
    if (c->done) return;
    double t = c->t;
    int it = c->it;
    if (c->pending) {
        t += c->dt;
        it += 1;
    }
    if (!(t < c->t_final) || (c->max_steps > 0 && it >= c->max_steps)) {
        c->t = t;
        c->it = it;
        c->pending = 0;
        c->done = 1;
    }
}

constexpr int kLineBS = 256;

// Exclusive prefix count of `flag` over the block (kLineBS threads); `total`
// receives the block's count.  Ends with a barrier.
__device__ __forceinline__ int block_exclusive_count(bool flag, int &total, int *s_w)
{
    const unsigned long long mask = __ballot(flag);
    const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
    const unsigned long long below = (lane == 0) ? 0ull : (mask & ((~0ull) >> (kWave - lane)));
    const int in_wave = __popcll(below);
    if (lane == 0) s_w[wv] = __popcll(mask);
    __syncthreads();
    int before = 0;
    total = 0;
#pragma unroll
    for (int k = 0; k < kLineBS / kWave; k++) {
        const int c = s_w[k];
        before += (k < wv) ? c : 0;
        total += c;
    }
    __syncthreads();
    return before + in_wave;
}

struct LineArgs {
    double *m1, *m2;          // material planes (SoA, 6 fields), modified in place by the ghost fill
    const double *phi;        // level set
    double *phi_out;          // level-set passes only
    double *rec;              // ghost fill: 4 doubles per cell at crossing cells
    long long plane;          // plane stride of the material planes
    int ny, nx;
    int init_flag;            // the material flag at the start of each line (1D: 0, 2D: 1)
    int riemann;              // 0 HLLC, 1 exact
    double h, dt_scale;       // level set: spacing along the line and fraction of dt
    EP p1, p2;
    const Clock *clk;
};

// Position k of line `line` -> flat cell index.
template <int DIR>
__device__ __forceinline__ long long line_cell(int line, int k, int nx)
{
    return (DIR == 0) ? (long long)line * nx + k : (long long)k * nx + line;
}

__device__ __forceinline__ void load_state(const double *m, long long plane, long long c, State &w)
{
#pragma unroll
    for (int f = 0; f < 6; f++) w[f] = m[f * plane + c];
}

// applyRGFMBoundaryConditions{HLLC,Exact}[2D] (rgfmsolver.cpp:92-491) for one
// direction; one workgroup per line, line length at most kMaxLine.
constexpr int kMaxLine = 16384;

template <int DIR>
__global__ __launch_bounds__(kLineBS) void rgfm_fill_kernel(LineArgs a)
{
    __shared__ unsigned char s_code[kMaxLine];   // 0: none; 1: crossing with flag 0; 2: with flag 1
    __shared__ int s_w[kLineBS / kWave];
    if (a.clk->done) return;
    const int line = blockIdx.x;
    const int n = (DIR == 0) ? a.nx : a.ny;
    const int tid = threadIdx.x;
    // Phase 1: crossings, their flags, their interface states (from the states
    // before this pass).
    int running = 0;
    for (int k0 = 0; k0 < n; k0 += kLineBS) {
        const int k = k0 + tid;
        bool flag = false;
        if (k > 2 && k < n - 3) {
            const double c = a.phi[line_cell<DIR>(line, k, a.nx)], pv = a.phi[line_cell<DIR>(line, k - 1, a.nx)];
            flag = (c * pv) <= 0.0;
        }
        int total;
        const int before = block_exclusive_count(flag, total, s_w);
        if (k < n) s_code[k] = 0;
        if (flag) {
            const int f = a.init_flag ^ ((running + before) & 1);
            s_code[k] = (unsigned char)(1 + f);
            State L, R;
            const long long cl = line_cell<DIR>(line, k - 1, a.nx), cr = line_cell<DIR>(line, k, a.nx);
            // flag 0: material 1 | material 2 ; flag 1: material 2 | material 1
            load_state(f ? a.m2 : a.m1, a.plane, cl, L);
            load_state(f ? a.m1 : a.m2, a.plane, cr, R);
            const EP &pa = f ? a.p2 : a.p1;
            const EP &pb = f ? a.p1 : a.p2;
            double p, u, ra, rb;
            if (a.riemann == 0) hllc_interface<DIR>(L, R, pa, pb, p, u, ra, rb);
            else exact_interface<DIR>(L, R, pa, pb, p, u, ra, rb);
            double *r = a.rec + 4 * cr;
            r[0] = p;
            r[1] = u;
            r[2] = ra;
            r[3] = rb;
        }
        running += total;
    }
    __syncthreads();
    // Phase 2: each cell takes the last covering write, per material.  A flag-0
    // crossing at kk writes material 1 on [kk-1, kk+2] (ra) and material 2 on
    // [kk-3, kk] (rb); a flag-1 crossing the reverse.
    for (int k0 = 0; k0 < n; k0 += kLineBS) {
        const int k = k0 + tid;
        if (k >= n) continue;
        int best1 = -1, best2 = -1, which1 = 0, which2 = 0;
        for (int kk = k - 2; kk <= k + 3; kk++) {
            if (kk < 3 || kk >= n - 3) continue;
            const int code = s_code[kk];
            if (code == 0) continue;
            const bool fwd = (k >= kk - 1 && k <= kk + 2), bwd = (k >= kk - 3 && k <= kk);
            if (code == 1) {
                if (fwd) { best1 = kk; which1 = 2; }
                if (bwd) { best2 = kk; which2 = 3; }
            } else {
                if (fwd) { best2 = kk; which2 = 2; }
                if (bwd) { best1 = kk; which1 = 3; }
            }
        }
        const long long c = line_cell<DIR>(line, k, a.nx);
        if (best1 >= 0) {
            const double *r = a.rec + 4 * line_cell<DIR>(line, best1, a.nx);
            a.m1[0 * a.plane + c] = r[which1];
            a.m1[(1 + DIR) * a.plane + c] = r[1];
            a.m1[4 * a.plane + c] = r[0];
        }
        if (best2 >= 0) {
            const double *r = a.rec + 4 * line_cell<DIR>(line, best2, a.nx);
            a.m2[0 * a.plane + c] = r[which2];
            a.m2[(1 + DIR) * a.plane + c] = r[1];
            a.m2[4 * a.plane + c] = r[0];
        }
    }
}

// updateLevelSetFunction / update2DLevelSetFunctionX/Y (rgfmsolver.cpp:493-646):
// one workgroup per line; the material flag toggles at every sign change
// (including position 0 against itself), the velocity component along the line
// comes from material 1 (flag 0) or 2 (flag 1), and the one-sided second-order
// upwind difference uses mirror-ordered ghosts.
template <int DIR>
__global__ __launch_bounds__(kLineBS) void rgfm_level_set_kernel(LineArgs a)
{
    __shared__ int s_w[kLineBS / kWave];
    if (a.clk->done) return;
    const double dt = a.clk->dt * a.dt_scale;
    const int line = blockIdx.x;
    const int n = (DIR == 0) ? a.nx : a.ny;
    const int tid = threadIdx.x;
    int running = 0;
    for (int k0 = 0; k0 < n; k0 += kLineBS) {
        const int k = k0 + tid;
        bool flag = false;
        double c = 0.0;
        if (k < n) {
            c = a.phi[line_cell<DIR>(line, k, a.nx)];
            const double pv = a.phi[line_cell<DIR>(line, k > 0 ? k - 1 : 0, a.nx)];
            flag = (pv * c) <= 0.0;
        }
        int total;
        const int before = block_exclusive_count(flag, total, s_w);
        if (k < n) {
            const int f = a.init_flag ^ ((running + before + (flag ? 1 : 0)) & 1);
            const long long cell = line_cell<DIR>(line, k, a.nx);
            const double v = (f ? a.m2 : a.m1)[(1 + DIR) * a.plane + cell];
            double up;
            if (v < 0.0) {
                const double p2 = a.phi[line_cell<DIR>(line, ghost_index(k + 2, n, 0), a.nx)];
                const double p1 = a.phi[line_cell<DIR>(line, ghost_index(k + 1, n, 0), a.nx)];
                up = (-p2 + (4.0 * p1) - (3.0 * c)) / (2.0 * a.h);
            } else {
                const double m1v = a.phi[line_cell<DIR>(line, ghost_index(k - 1, n, 0), a.nx)];
                const double m2v = a.phi[line_cell<DIR>(line, ghost_index(k - 2, n, 0), a.nx)];
                up = ((3.0 * c) - (4.0 * m1v) + m2v) / (2.0 * a.h);
            }
            a.phi_out[cell] = c - (dt * (v * up));
        }
        running += total;
    }
}

// The y pass of the level-set update with coalesced rows.  A workgroup takes
// kLsCols adjacent columns, each split into kLsSegs segments of consecutive
// rows, one thread per column segment, so each row read covers kLsCols
// adjacent doubles (one workgroup per column, as in rgfm_level_set_kernel<1>,
// reads one double per row at a stride of nx).  Pass 1 counts each segment's
// sign changes; their prefix over the column's segments gives the count before
// the segment; pass 2 walks the segment again, counting on, and updates phi.
// The flag is that count's parity, so flags, velocities and the upwind
// arithmetic are those of rgfm_level_set_kernel<1>, bit for bit.
constexpr int kLsCols = 8, kLsSegs = 64;

static __global__ __launch_bounds__(kLsCols * kLsSegs) void rgfm_level_set_y_kernel(LineArgs a)
{
    __shared__ int s_cnt[kLsSegs][kLsCols];
    if (a.clk->done) return;
    const double dt = a.clk->dt * a.dt_scale;
    const int tid = threadIdx.x;
    const int cl = tid % kLsCols, seg = tid / kLsCols;
    const int col = (int)blockIdx.x * kLsCols + cl;
    const int n = a.ny, nx = a.nx;
    const bool ok = col < nx;
    const int len = (n + kLsSegs - 1) / kLsSegs;
    const int k0 = seg * len, k1 = (k0 + len < n) ? k0 + len : n;
    int cnt = 0;
    if (ok) {
#pragma unroll 4
        for (int k = k0; k < k1; k++) {
            const double c = a.phi[(long long)k * nx + col];
            const double pv = a.phi[(long long)(k > 0 ? k - 1 : 0) * nx + col];
            cnt += ((pv * c) <= 0.0) ? 1 : 0;
        }
    }
    s_cnt[seg][cl] = cnt;
    __syncthreads();
    int run = 0;
    for (int s2 = 0; s2 < seg; s2++) run += s_cnt[s2][cl];
    if (!ok) return;
#pragma unroll 4
    for (int k = k0; k < k1; k++) {
        const long long cell = (long long)k * nx + col;
        const double c = a.phi[cell];
        const double pv = a.phi[(long long)(k > 0 ? k - 1 : 0) * nx + col];
        run += ((pv * c) <= 0.0) ? 1 : 0;
        const int f = a.init_flag ^ (run & 1);
        const double v = (f ? a.m2 : a.m1)[2 * a.plane + cell];
        double up;
        if (v < 0.0) {
            const double p2 = a.phi[(long long)ghost_index(k + 2, n, 0) * nx + col];
            const double p1 = a.phi[(long long)ghost_index(k + 1, n, 0) * nx + col];
            up = (-p2 + (4.0 * p1) - (3.0 * c)) / (2.0 * a.h);
        } else {
            const double m1v = a.phi[(long long)ghost_index(k - 1, n, 0) * nx + col];
            const double m2v = a.phi[(long long)ghost_index(k - 2, n, 0) * nx + col];
            up = ((3.0 * c) - (4.0 * m1v) + m2v) / (2.0 * a.h);
        }
        a.phi_out[cell] = c - (dt * (v * up));
    }
}

// Interface circulation partial sums: (dv/dx - du/dy) dx dy over interior cells
// with phi < 0, from material 1's velocity; one partial per block.
static __global__ __launch_bounds__(256) void rgfm_circulation_kernel(const double *m1, long long plane, const double *phi,
                                                               int ny, int nx, double dx, double dy, double *partial)
{
    __shared__ double s[256];
    double acc = 0.0;
    const long long cells = (long long)ny * nx;
    for (long long c = (long long)blockIdx.x * 256 + threadIdx.x; c < cells; c += (long long)gridDim.x * 256) {
        const int i = (int)(c / nx), j = (int)(c - (long long)i * nx);
        if (i < 1 || i >= ny - 1 || j < 1 || j >= nx - 1 || !(phi[c] < 0.0)) continue;
        const double dvdx = (m1[2 * plane + c + 1] - m1[2 * plane + c - 1]) / (2.0 * dx);
        const double dudy = (m1[1 * plane + c + nx] - m1[1 * plane + c - nx]) / (2.0 * dy);
        acc += (dvdx - dudy) * (dx * dy);
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = s[0];
}

// Corner ghosts of both materials' 2D ghosted grids hold the default state
// (Euler/eulerstatevector.cpp:3-12); the larger of the two speeds.
static __global__ void rgfm_corner_kernel(EP p1, EP p2, Clock *clk)
{
    double w[6];
    EulerPhys::default_state(w);
    clk->corner = smax(EulerPhys::wave_speed(p1, w, 2), EulerPhys::wave_speed(p2, w, 2));
}

}  // namespace vg
