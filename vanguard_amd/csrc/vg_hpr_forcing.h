// HPR stiff-relaxation source sub-cycles: the traits of the generic RK4 source
// stage (vg_forcing.h) for the HPR state.  Only the A slots (2..10) and, with
// heat conduction, the J slots (14..16) of S (hprstatevector.cpp:312-366) are
// non-zero; without heat conduction a cell keeps its own J
// (HPRForcingSolver::computeRungeKuttaTimeStep[2D], hprforcingsolver.cpp:341-356,
// 394-417, sets A only).
#pragma once

#include "vg_forcing.h"
#include "vg_phys_hpr.h"

namespace vg {

template <int EOS>
struct ForcingTraits<HPRPhysT<EOS>> {
    static constexpr int kMaxSlots = 12;
    static constexpr const char *kTimer = "hpr_source_rk4";
    __device__ static __forceinline__ constexpr int slot(int q) { return q < 9 ? 2 + q : 14 + (q - 9); }
    __device__ static __forceinline__ int count(const HPRParams &m) { return m.is_thermal ? 12 : 9; }
    __device__ static __forceinline__ void finish(const HPRParams &m, const double (&w0)[17], double (&wn)[17])
    {
        if (!m.is_thermal) {
            wn[14] = w0[14];
            wn[15] = w0[15];
            wn[16] = w0[16];
        }
    }
};

}  // namespace vg
