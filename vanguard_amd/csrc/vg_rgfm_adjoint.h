// Adjoint of the RGFM ghost fill (applyRGFMBoundaryConditionsHLLC[2D],
// rgfmsolver.cpp:293-491) for one direction, time step held fixed.
//
// The fill maps the states before the pass to the states after it: a cell
// covered by a crossing's window takes, in the fields density, normal velocity
// and pressure, the crossing's interface values (HLLC p*, u*, and the star
// density of its side), computed from the two real states next to the
// crossing as they were before the pass; every other field and cell passes
// through.  Its adjoint therefore (1) passes the adjoint of uncovered fields
// through, (2) sums, per crossing, the adjoints of the cells whose last
// covering crossing it is into the adjoint of its four interface values, and
// (3) sends that, through the vector-Jacobian product of the HLLC interface
// state, to the two real cells the crossing read.  The level set only decides
// where the crossings are, so it carries no adjoint.
//
// The VJP of the interface state runs the same templated code as the fill on
// forward-mode dual numbers, one input direction at a time (12 directions, one
// lane each): its branches see the same values as the fill.
//
// The adjoint is updated in place and only near crossings: uncovered fields
// keep their adjoint, so one workgroup per line (1) finds the crossings, (2)
// computes every crossing's 12 VJP components in parallel into a scratch
// buffer (reading the adjoints of covered cells), and, after a barrier, (3)
// rewrites the cells that are covered or next to a crossing.
#pragma once

#include "vg_rgfm.h"

namespace vg {

struct FillAdjArgs {
    const double *m1, *m2;        // material states before the pass (SoA)
    const double *phi;            // level set of the step
    double *adj1, *adj2;          // in: adjoints of the states after the pass; out: before it
    double *vjp;                  // scratch, 12 planes: the VJP components at each crossing cell
    long long plane;
    int ny, nx, init_flag;
    EP p1, p2;
};

// Last crossing (in line order) whose window covers position k for material q
// (1 or 2), and which interface value it writes there (2: ra, 3: rb); -1 if
// none.  The same rule as phase 2 of rgfm_fill_kernel.
__device__ __forceinline__ int covering(const unsigned char *code, int k, int n, int q, int &which)
{
    int best = -1;
    which = 0;
    for (int kk = k - 2; kk <= k + 3; kk++) {
        if (kk < 3 || kk >= n - 3) continue;
        const int c = code[kk];
        if (c == 0) continue;
        const bool fwd = (k >= kk - 1 && k <= kk + 2), bwd = (k >= kk - 3 && k <= kk);
        // flag 0 (code 1): material 1 on [kk-1, kk+2] (ra), material 2 on [kk-3, kk] (rb)
        const bool mine_fwd = (c == 1) == (q == 1);
        if (mine_fwd ? fwd : bwd) {
            best = kk;
            which = mine_fwd ? 2 : 3;
        }
    }
    return best;
}

// Component i (0-5: the left real state's fields, 6-11: the right one's) of
// the adjoint of the two real states crossing kk read, given the adjoints of
// the cells it covers last.  L is the real state left of the crossing
// (material 1 if its flag is 0, else 2), R the right one.
template <int DIR>
__device__ double crossing_vjp(const FillAdjArgs &a, const unsigned char *code, int line, int kk, int n, int i)
{
    const int f = code[kk] - 1;
    // adjoint of the interface values (p, u, ra, rb), summed in line order
    double rbar[4] = {0.0, 0.0, 0.0, 0.0};
    for (int k = kk - 3; k <= kk + 2; k++) {
        if (k < 0 || k >= n) continue;
        const long long c = line_cell<DIR>(line, k, a.nx);
        for (int q = 1; q <= 2; q++) {
            int which;
            if (covering(code, k, n, q, which) != kk) continue;
            const double *bar = (q == 1) ? a.adj1 : a.adj2;
            rbar[which] += bar[0 * a.plane + c];
            rbar[1] += bar[(1 + DIR) * a.plane + c];
            rbar[0] += bar[4 * a.plane + c];
        }
    }
    double L[6], R[6];
    const long long cl = line_cell<DIR>(line, kk - 1, a.nx), cr = line_cell<DIR>(line, kk, a.nx);
    load_state(f ? a.m2 : a.m1, a.plane, cl, L);
    load_state(f ? a.m1 : a.m2, a.plane, cr, R);
    const EP &pa = f ? a.p2 : a.p1;
    const EP &pb = f ? a.p1 : a.p2;
    Dual Ld[6], Rd[6];
#pragma unroll
    for (int k = 0; k < 6; k++) {
        Ld[k] = {L[k], (i == k) ? 1.0 : 0.0};
        Rd[k] = {R[k], (i == 6 + k) ? 1.0 : 0.0};
    }
    Dual p, u, ra, rb;
    hllc_interface<DIR>(Ld, Rd, pa, pb, p, u, ra, rb);
    return rbar[0] * p.d + rbar[1] * u.d + rbar[2] * ra.d + rbar[3] * rb.d;
}

constexpr int kVjpLanes = 16;   // lanes per crossing in the VJP phase (12 used)

template <int DIR>
__global__ __launch_bounds__(kLineBS) void rgfm_fill_adjoint_kernel(FillAdjArgs a)
{
    __shared__ unsigned char s_code[kMaxLine];      // 0: none; 1: crossing with flag 0; 2: with flag 1
    __shared__ unsigned short s_cross[kMaxLine];    // positions of the crossings, in line order
    __shared__ int s_w[kLineBS / kWave];
    const int line = blockIdx.x;
    const int n = (DIR == 0) ? a.nx : a.ny;
    const int tid = threadIdx.x;
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
        if (k < n) s_code[k] = flag ? (unsigned char)(1 + (a.init_flag ^ ((running + before) & 1))) : 0;
        if (flag) s_cross[running + before] = (unsigned short)k;
        running += total;
    }
    __syncthreads();
    // (2) the VJP components of every crossing, kVjpLanes lanes per crossing
    const int ncross = running;
    for (int j0 = 0; j0 < ncross; j0 += kLineBS / kVjpLanes) {
        const int j = j0 + tid / kVjpLanes, i = tid % kVjpLanes;
        if (j < ncross && i < 12) {
            const int kk = s_cross[j];
            a.vjp[i * a.plane + line_cell<DIR>(line, kk, a.nx)] = crossing_vjp<DIR>(a, s_code, line, kk, n, i);
        }
    }
    __syncthreads();
    // (3) covered cells lose the adjoint of the fields the fill wrote; the
    // real cells next to a crossing gain its VJP
    for (int k0 = 0; k0 < n; k0 += kLineBS) {
        const int k = k0 + tid;
        if (k >= n) continue;
        int which;
        const bool cov1 = covering(s_code, k, n, 1, which) >= 0;
        const bool cov2 = covering(s_code, k, n, 2, which) >= 0;
        // this cell is the left real state of a crossing at k + 1 and/or the
        // right real state of a crossing at k
        const bool left_of = k + 1 > 2 && k + 1 < n - 3 && s_code[k + 1];
        const bool right_of = k > 2 && k < n - 3 && s_code[k];
        if (!cov1 && !cov2 && !left_of && !right_of) continue;
        const long long c = line_cell<DIR>(line, k, a.nx);
        double o1[6], o2[6];
#pragma unroll
        for (int f = 0; f < 6; f++) {
            const bool written = (f == 0 || f == 1 + DIR || f == 4);
            o1[f] = (cov1 && written) ? 0.0 : a.adj1[f * a.plane + c];
            o2[f] = (cov2 && written) ? 0.0 : a.adj2[f * a.plane + c];
        }
        if (left_of) {
            const long long cc = line_cell<DIR>(line, k + 1, a.nx);
            double *o = (s_code[k + 1] == 1) ? o1 : o2;
#pragma unroll
            for (int f = 0; f < 6; f++) o[f] += a.vjp[f * a.plane + cc];
        }
        if (right_of) {
            double *o = (s_code[k] == 1) ? o2 : o1;
#pragma unroll
            for (int f = 0; f < 6; f++) o[f] += a.vjp[(6 + f) * a.plane + c];
        }
#pragma unroll
        for (int f = 0; f < 6; f++) {
            a.adj1[f * a.plane + c] = o1[f];
            a.adj2[f * a.plane + c] = o2[f];
        }
    }
}

}  // namespace vg
