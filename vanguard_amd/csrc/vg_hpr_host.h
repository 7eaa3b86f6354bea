// Host side shared by the HPR solve and the HPR adjoint: the grid's source
// stage, the material parameters in kernel form and their checks.
#pragma once

#include <cmath>

#include "vg_grid.h"
#include "vg_hpr_forcing.h"
#include "vg_phys_hpr.h"

namespace vg {

// The relaxation sub-cycle launch for GridState<HPRPhysT<EOS>>::enqueue_step.
template <int EOS>
struct SourceStage<HPRPhysT<EOS>> : ForcingStage<HPRPhysT<EOS>> {};

template <int EOS>
using HPRGridT = GridState<HPRPhysT<EOS>>;
using HPRGrid = HPRGridT<0>;

inline HPRParams hpr_to_params(const vg_hpr_params *p)
{
    HPRParams m;
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
    m.eos = p->eos;
    m.p_ref = p->reference_pressure;
    m.c0 = p->reference_sound_speed;
    m.gamma0 = p->reference_gruneisen_coefficient;
    m.s_h = p->hugoniot_slope_coefficient;
    m.alpha_gr = p->alpha_parameter;
    m.sigma_y = p->elastic_plastic_transition_parameter;
    m.n_pow = p->power_law_index;
    // The same IEEE values the reference recomputes per call.
    m.G = m.gamma - 1.0;
    m.cs2 = m.cs * m.cs;
    m.a2 = m.alpha * m.alpha;
    m.c02 = m.c0 * m.c0;
    return m;
}

// The adjoint's hand-derived derivatives are those of the stiffened gas without
// the plastic power law.
inline int hpr_check_adjoint_params(const vg_hpr_params *mp)
{
    VG_CHECK(mp != nullptr, "null material parameters");
    if (mp->eos != VG_HPR_EOS_STIFFENED_GAS || mp->is_plastic) {
        vg::set_error("HPR adjoint: only the StiffenedGas equation of state without plastic relaxation is "
                      "differentiated");
        return VG_ERR_UNSUPPORTED;
    }
    return VG_OK;
}

inline int hpr_check_params(const vg_hpr_params *mp, int subcycles)
{
    VG_CHECK(mp != nullptr, "null material parameters");
    VG_CHECK(mp->eos == VG_HPR_EOS_STIFFENED_GAS || mp->eos == VG_HPR_EOS_SHOCK_MIE_GRUNEISEN ||
                 mp->eos == VG_HPR_EOS_GODUNOV_ROMENSKI,
             "unknown HPR equation of state");
    VG_CHECK(!mp->is_plastic || mp->elastic_plastic_transition_parameter > 0.0,
             "the plastic relaxation needs a positive elastic-plastic transition parameter");
    VG_CHECK(subcycles >= 0, "subcycles must be >= 0");
    VG_CHECK(subcycles == 0 || (mp->strain_dissipation_time > 0.0 &&
                                (!mp->is_thermal || mp->thermal_impulse_relaxation_time > 0.0)),
             "relaxation times must be positive when sub-cycling the source");
    return VG_OK;
}

}  // namespace vg
