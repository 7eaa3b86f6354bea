// Host side shared by the HPR solve and the HPR adjoint: the grid's source
// stage, the material parameters in kernel form and their checks.
#pragma once

#include <cmath>

#include "vg_grid.h"
#include "vg_hpr_forcing.h"
#include "vg_phys_hpr.h"

namespace vg {

// The relaxation sub-cycle launch for GridState<HPRPhys>::enqueue_step.
template <>
struct SourceStage<HPRPhys> : ForcingStage<HPRPhys> {};

using HPRGrid = GridState<HPRPhys>;

inline HPRPhys::Params hpr_to_params(const vg_hpr_params *p)
{
    HPRPhys::Params m;
    m.is_thermal = p->is_thermal ? 1 : 0;
    m.is_plastic = p->is_plastic ? 1 : 0;
    m.rho0 = p->reference_density;
    m.t_ref = p->reference_temperature;
    m.t0 = p->initial_temperature;
    m.cv = p->specific_heat_capacity;
    m.gamma = p->adiabatic_index;
    m.p_inf = p->stiffening_parameter;
    m.cs = p->transverse_wave_speed;
    m.tau1 = p->strain_dissipation_time;
    m.tau2 = p->thermal_impulse_relaxation_time;
    m.alpha = p->heat_wave_speed;
    // The same IEEE values the reference recomputes per call.
    m.G = m.gamma - 1.0;
    m.cs2 = m.cs * m.cs;
    m.a2 = m.alpha * m.alpha;
    return m;
}

inline int hpr_check_params(const vg_hpr_params *mp, int subcycles)
{
    VG_CHECK(mp != nullptr, "null material parameters");
    if (mp->eos != VG_HPR_EOS_STIFFENED_GAS) {
        vg::set_error("HPR: only the StiffenedGas equation of state is implemented");
        return VG_ERR_UNSUPPORTED;
    }
    if (mp->is_plastic) {
        vg::set_error("HPR: the plastic (power-law) strain relaxation is not implemented");
        return VG_ERR_UNSUPPORTED;
    }
    VG_CHECK(subcycles >= 0, "subcycles must be >= 0");
    VG_CHECK(subcycles == 0 || (mp->strain_dissipation_time > 0.0 &&
                                (!mp->is_thermal || mp->thermal_impulse_relaxation_time > 0.0)),
             "relaxation times must be positive when sub-cycling the source");
    return VG_OK;
}

}  // namespace vg
