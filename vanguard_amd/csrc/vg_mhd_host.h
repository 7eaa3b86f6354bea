// The MHD source stage shared by every translation unit that builds
// GridState<MHDPhys>: the RK4 GLM-damping sub-cycles of MHDForcingSolver
// (Solvers/mhdforcingsolver.cpp:7-44, 134-194, 341-356, 394-417).
#pragma once

#include "vg_forcing.h"
#include "vg_grid.h"
#include "vg_phys_mhd.h"

namespace vg {

// Only psi's slot of the MHD source is non-zero (mhdstatevector.cpp:222-234).
template <>
struct ForcingTraits<MHDPhys> {
    static constexpr int kMaxSlots = 1;
    static constexpr const char *kTimer = "mhd_source_rk4";
    __device__ static __forceinline__ constexpr int slot(int) { return 8; }
    __device__ static __forceinline__ int count(const MHDPhys::Params &) { return 1; }
    __device__ static __forceinline__ void finish(const MHDPhys::Params &, const double (&)[9], double (&)[9]) {}
};

// The GLM damping sub-cycle launch for GridState<MHDPhys>::enqueue_step.
template <>
struct SourceStage<MHDPhys> : ForcingStage<MHDPhys> {};

// MHDMaterialParameters (MHD/mhdmaterialparameters.cpp:3-34) with the derived
// constants the kernels use.
inline MHDPhys::Params mhd_params(const vg_mhd_params *p)
{
    MHDPhys::Params m;
    m.gamma = p->gamma;
    m.p_inf = p->p_inf;
    m.ch = p->ch;
    m.cp = p->cp;
    m.gm1 = p->gamma - 1;
    m.gpinf = p->gamma * p->p_inf;
    m.inv_gm1 = 1.0 / (p->gamma - 1);
    return m;
}

}  // namespace vg
