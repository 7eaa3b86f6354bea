// Two-material ideal MHD with the real ghost fluid method: the MHD HLLC
// interface state (RGFM/mhdhllcsolver.cpp) and the ghost-fill line kernel
// (MHDRGFMSolver::applyRGFMBoundaryConditions[2D], RGFM/mhdrgfmsolver.cpp:7-378).
//
// The fill has the structure of the Euler RGFM fill (vg_rgfm.h): one workgroup
// per grid line, the material flag of a crossing from a block-wide prefix count
// of the crossings before it, interface states from the states before the pass
// (phase 1, into a per-cell record), then every cell takes the last covering
// crossing's ghost (phase 2, in place).  An MHD ghost carries the whole
// interface state (velocities, pressure floored at 1e-8, magnetic field, and
// psi in the 2D fills only) with the side's star density.
#pragma once

#include "vg_phys_mhd.h"
#include "vg_rgfm.h"

namespace vg {

using MP = MHDPhys::Params;

// computeTotalPressure (mhdhllcsolver.cpp:268-279).
__device__ __forceinline__ double mhd_total_pressure(const double (&w)[9])
{
    const double b2 = (w[5] * w[5]) + (w[6] * w[6]) + (w[7] * w[7]);
    return w[4] + (0.5 * b2);
}

// The HLL average (S_b q_b - S_a q_a - (f_b - f_a)) / (S_b - S_a) of the
// computeX/YStarRegion* functions (mhdhllcsolver.cpp:547-922); sqa = S_a q_a,
// sqb = S_b q_b as the caller forms them.
__device__ __forceinline__ double mhd_hll(double Sa, double Sb, double sqa, double sqb, double fa, double fb)
{
    return (sqb - sqa - (fb - fa)) / (Sb - Sa);
}

// MHDHLLCSolver::solveX / solveY (mhdhllcsolver.cpp:200-266) between a (left
// or top, parameters pa) and b (right or bottom, pb) along DIR, with the star
// densities of both sides (computeStarRegionDensity and its four sides,
// :7-58) and the star speed (computeStarRegionX/YVelocity, :281-333).  Outer
// wave speeds: min / max of the normal velocities -/+ the larger fast speed
// (:1495-1537).  In the star branches every HLL field is the average (the
// selections of computeX/YHLL*, :923-1334, pick the average when
// S_a < 0 < S_b); the star state is built on the side of S* that holds x/t = 0
// and converted with that side's parameters.
template <int DIR>
__device__ __forceinline__ void mhd_hllc_interface(const double (&a)[9], const double (&b)[9], const MP &pa,
                                                   const MP &pb, double (&s)[9], double &ra, double &rb, double &Ss)
{
    constexpr int n = 1 + DIR, f = 5 + DIR;
    const double fa_ = MHDPhys::fast_speed(pa, a, DIR), fb_ = MHDPhys::fast_speed(pb, b, DIR);
    const double Sa = smin(a[n], b[n]) - smax(fa_, fb_);
    const double Sb = smax(a[n], b[n]) + smax(fa_, fb_);
    const double ca = a[0] * (Sa - a[n]);
    const double cb = b[0] * (Sb - b[n]);
    Ss = ((cb * b[n]) - (ca * a[n]) + mhd_total_pressure(a) - mhd_total_pressure(b) - (a[f] * a[f]) + (b[f] * b[f])) /
         (cb - ca);
    ra = a[0] * ((Sa - a[n]) / (Sa - Ss));
    rb = b[0] * ((Sb - b[n]) / (Sb - Ss));
    if (0.0 <= Sa) {
#pragma unroll
        for (int k = 0; k < 9; k++) s[k] = a[k];
        return;
    }
    if (Sb <= 0.0) {
#pragma unroll
        for (int k = 0; k < 9; k++) s[k] = b[k];
        return;
    }
    // HLL fields: H[0] density, H[1..3] momenta, H[5..7] field, H[8] psi
    double H[9];
    H[0] = mhd_hll(Sa, Sb, Sa * a[0], Sb * b[0], a[0] * a[n], b[0] * b[n]);
#pragma unroll
    for (int k = 1; k <= 3; k++) {
        double fa, fb;
        if (k == n) {
            const double ba2 = (a[5] * a[5]) + (a[6] * a[6]) + (a[7] * a[7]);
            const double bb2 = (b[5] * b[5]) + (b[6] * b[6]) + (b[7] * b[7]);
            fa = (a[0] * (a[n] * a[n])) + a[4] + (0.5 * ba2) - (a[f] * a[f]);
            fb = (b[0] * (b[n] * b[n])) + b[4] + (0.5 * bb2) - (b[f] * b[f]);
        } else {
            fa = (a[0] * (a[n] * a[k])) - (a[f] * a[4 + k]);
            fb = (b[0] * (b[n] * b[k])) - (b[f] * b[4 + k]);
        }
        H[k] = mhd_hll(Sa, Sb, (Sa * a[0]) * a[k], (Sb * b[0]) * b[k], fa, fb);
    }
    H[4] = 0.0;
#pragma unroll
    for (int q = 5; q <= 7; q++) {
        double fa, fb;
        if (q == f) {
            fa = a[8];
            fb = b[8];
        } else {
            // induction flux B_q u_n - B_n u_q
            fa = (a[q] * a[n]) - (a[f] * a[q - 4]);
            fb = (b[q] * b[n]) - (b[f] * b[q - 4]);
        }
        H[q] = mhd_hll(Sa, Sb, Sa * a[q], Sb * b[q], fa, fb);
    }
    H[8] = mhd_hll(Sa, Sb, Sa * a[8], Sb * b[8], (pa.ch * pa.ch) * a[f], (pb.ch * pb.ch) * b[f]);

    const bool left = 0.0 <= Ss;
    const double(&w)[9] = left ? a : b;
    const MP &m = left ? pa : pb;
    const double S = left ? Sa : Sb;
    const double rs = left ? ra : rb;
    double U[9];
    U[0] = rs;
#pragma unroll
    for (int k = 1; k <= 3; k++) {
        // tangential star momenta (:403-545): rho* vel_t - (Bn* Bt* - Bn Bt) / (S - S*)
        U[k] = (k == n) ? rs * Ss : (rs * w[k]) - (((H[f] * H[4 + k]) - (w[f] * w[4 + k])) / (S - Ss));
    }
    // star total pressure (:335-401) and total energy (:1335-1493)
    const double pt = mhd_total_pressure(w);
    const double pts = (w[0] * (S - w[n]) * (Ss - w[n])) + pt - (w[f] * w[f]) + (H[f] * H[f]);
    const double rh = H[0];
    const double ux = H[1] / rh, uy = H[2] / rh, uz = H[3] / rh;
    const double bv_s = (H[5] * ux) + (H[6] * uy) + (H[7] * uz);
    const double bv = (w[5] * w[1]) + (w[6] * w[2]) + (w[7] * w[3]);
    double Uw[9];
    MHDPhys::prim_to_cons(m, w, Uw);
    const double first = Uw[4] * ((S - w[n]) / (S - Ss));
    const double second = ((pts * Ss) - (pt * w[n]) - ((H[f] * bv_s) - (w[f] * bv))) / (S - Ss);
    U[4] = first + second;
    U[5] = H[5];
    U[6] = H[6];
    U[7] = H[7];
    U[8] = H[8];
    MHDPhys::cons_to_prim(m, U, s);
}

// Record of one crossing: the interface state (pressure floored), then the
// star densities of the a and b sides.
constexpr int kMHDRec = 11;

struct MHDLineArgs {
    double *m1, *m2;      // material planes (SoA, 9 fields), modified in place
    const double *phi;    // level set
    double *rec;          // kMHDRec doubles per cell at crossing cells
    long long plane;
    int ny, nx;
    int init_flag;        // material flag at the start of each line (1D: 0, 2D: 1)
    int with_psi;         // the 2D fills also write psi
    MP p1, p2;
    const Clock *clk;
};

__device__ __forceinline__ void load_mhd(const double *m, long long plane, long long c, double (&w)[9])
{
#pragma unroll
    for (int k = 0; k < 9; k++) w[k] = m[k * plane + c];
}

// One direction of applyRGFMBoundaryConditions[2D] (mhdrgfmsolver.cpp:7-378):
// one workgroup per line, line length at most kMaxLine.  A flag-0 crossing at
// kk (material 1 left of it) writes material 1's ghosts on [kk-1, kk+2] with
// the a-side star density and material 2's on [kk-3, kk] with the b side's; a
// flag-1 crossing the reverse.
template <int DIR>
__global__ __launch_bounds__(kLineBS) void mhd_rgfm_fill_kernel(MHDLineArgs a)
{
    __shared__ unsigned char s_code[kMaxLine];   // 0: none; 1: crossing with flag 0; 2: with flag 1
    __shared__ int s_w[kLineBS / kWave];
    if (a.clk->done) return;
    const int line = blockIdx.x;
    const int n = (DIR == 0) ? a.nx : a.ny;
    const int tid = threadIdx.x;
    const double pfloor = 1e-8;   // pow(10, -8) of the reference: the double nearest 10^-8
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
            const int fl = a.init_flag ^ ((running + before) & 1);
            s_code[k] = (unsigned char)(1 + fl);
            double L[9], R[9], sol[9];
            const long long cl = line_cell<DIR>(line, k - 1, a.nx), cr = line_cell<DIR>(line, k, a.nx);
            load_mhd(fl ? a.m2 : a.m1, a.plane, cl, L);
            load_mhd(fl ? a.m1 : a.m2, a.plane, cr, R);
            double ra, rb, ss;
            mhd_hllc_interface<DIR>(L, R, fl ? a.p2 : a.p1, fl ? a.p1 : a.p2, sol, ra, rb, ss);
            if (sol[4] < pfloor) sol[4] = pfloor;
            double *r = a.rec + (long long)kMHDRec * cr;
#pragma unroll
            for (int q = 0; q < 9; q++) r[q] = sol[q];
            r[9] = ra;
            r[10] = rb;
        }
        running += total;
    }
    __syncthreads();
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
                if (fwd) { best1 = kk; which1 = 9; }
                if (bwd) { best2 = kk; which2 = 10; }
            } else {
                if (fwd) { best2 = kk; which2 = 9; }
                if (bwd) { best1 = kk; which1 = 10; }
            }
        }
        const long long c = line_cell<DIR>(line, k, a.nx);
        const int last = a.with_psi ? 9 : 8;
        if (best1 >= 0) {
            const double *r = a.rec + (long long)kMHDRec * line_cell<DIR>(line, best1, a.nx);
            a.m1[c] = r[which1];
            for (int q = 1; q < last; q++) a.m1[q * a.plane + c] = r[q];
        }
        if (best2 >= 0) {
            const double *r = a.rec + (long long)kMHDRec * line_cell<DIR>(line, best2, a.nx);
            a.m2[c] = r[which2];
            for (int q = 1; q < last; q++) a.m2[q * a.plane + c] = r[q];
        }
    }
}

// Default-state corner ghosts of both materials' 2D grids (MHDStateVector(),
// mhdstatevector.cpp:3-16): the larger of the two wave speeds.
static __global__ void mhd_rgfm_corner_kernel(MP p1, MP p2, Clock *clk)
{
    double w[9];
    MHDPhys::default_state(w);
    clk->corner = smax(MHDPhys::wave_speed(p1, w, 2), MHDPhys::wave_speed(p2, w, 2));
}

}  // namespace vg
