"""ctypes bindings of libvanguard_hip.so (the C ABI declared in include/vanguard.h).

Loading fails loudly when the library is missing: there is no CPU fallback.
"""

from __future__ import annotations

import ctypes as C
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libvanguard_hip.so")

VG_OK = 0
VG_ERR_ARG = -1
VG_ERR_HIP = -2
VG_ERR_NONFINITE = -3
VG_ERR_UNSUPPORTED = -4

VG_BC_TRANSMISSIVE = 0
VG_BC_PERIODIC = 1
VG_MODE_EXACT = 0
VG_MODE_FAST = 1

dptr = C.POINTER(C.c_double)
vptr = C.c_void_p


class EulerParams(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("gamma", "p_inf", "cv", "e_form", "t_ign", "k_rate")]


class MHDParams(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("gamma", "p_inf", "ch", "cp")]


VG_RIEMANN_HLLC = 0
VG_RIEMANN_EXACT = 1

VG_HPR_EOS_STIFFENED_GAS = 0
VG_HPR_EOS_SHOCK_MIE_GRUNEISEN = 1
VG_HPR_EOS_GODUNOV_ROMENSKI = 2

HPR_DOUBLE_FIELDS = (
    "ideal_gas_constant", "reference_density", "reference_pressure", "reference_temperature", "initial_temperature",
    "specific_heat_capacity", "adiabatic_index", "stiffening_parameter", "reference_sound_speed",
    "reference_gruneisen_coefficient", "hugoniot_slope_coefficient", "reference_internal_energy",
    "transverse_wave_speed", "strain_dissipation_time", "thermal_impulse_relaxation_time",
    "dynamic_viscosity_coefficient", "prandtl_number", "elastic_plastic_transition_parameter", "power_law_index",
    "heat_wave_speed", "alpha_parameter", "beta_parameter", "gamma_parameter")


class HPRParams(C.Structure):
    _fields_ = [("eos", C.c_int32), ("is_thermal", C.c_int32), ("is_plastic", C.c_int32), ("reserved", C.c_int32)] + [
        (n, C.c_double) for n in HPR_DOUBLE_FIELDS]


class RunStats(C.Structure):
    _fields_ = [("steps", C.c_int32), ("total_steps", C.c_int32), ("t", C.c_double), ("dt_min", C.c_double),
                ("dt_max", C.c_double), ("wall_ms", C.c_double), ("device_ms", C.c_double),
                ("nonfinite_cells", C.c_int64), ("first_nonfinite", C.c_int64)]


# name -> (restype, argtypes)
SIGNATURES = {
    "vg_ctx_create": (C.c_int, [C.POINTER(vptr), C.c_int]),
    "vg_ctx_destroy": (None, [vptr]),
    "vg_last_error": (C.c_char_p, []),
    "vg_version": (C.c_char_p, []),
    "vg_ctx_set_timing": (C.c_int, [vptr, C.c_int]),
    "vg_ctx_kernel_time": (C.c_int, [vptr, C.c_char_p, C.POINTER(C.c_int64), dptr]),
    "vg_ctx_reset_timing": (C.c_int, [vptr]),
    "vg_ctx_set_graphs": (C.c_int, [vptr, C.c_int]),
    "vg_ctx_set_autotune": (C.c_int, [vptr, C.c_int]),
    "vg_ctx_set_reduced_fields": (C.c_int, [vptr, C.c_int]),
    "vg_ctx_set_batch_streams": (C.c_int, [vptr, C.c_int]),
    "vg_ctx_tuned_shape": (C.c_int, [vptr, C.c_char_p, C.POINTER(C.c_int)]),
    "vg_ctx_set_tuned_shape": (C.c_int, [vptr, C.c_char_p, C.c_int]),
    "vg_ctx_release_workspace": (C.c_int, [vptr]),
    "vg_ctx_set_adjoint_memory": (C.c_int, [vptr, C.c_int64]),
    "vg_euler_solve_1d": (C.c_int, [vptr, dptr, dptr, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int,
                                    C.c_double, C.c_int, C.c_int, C.c_int, C.POINTER(EulerParams), C.c_int,
                                    dptr, C.c_int, C.POINTER(RunStats)]),
    "vg_euler_solve_2d": (C.c_int, [vptr, dptr, dptr, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                    C.c_int, C.c_double, C.c_int, C.c_int, C.c_int, C.POINTER(EulerParams),
                                    C.c_int, dptr, C.c_int, C.POINTER(RunStats)]),
    "vg_euler_multiphysics_solve_1d": (C.c_int, [vptr, dptr, dptr, C.c_int, C.c_double, C.c_double, C.c_double,
                                                 C.c_int, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int,
                                                 C.POINTER(EulerParams), C.POINTER(EulerParams), dptr, C.c_int,
                                                 C.POINTER(RunStats)]),
    "vg_euler_multiphysics_solve_2d": (C.c_int, [vptr, dptr, dptr, C.c_int, C.c_int, C.c_double, C.c_double,
                                                 C.c_double, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int,
                                                 C.POINTER(EulerParams), C.POINTER(EulerParams), dptr, C.c_int,
                                                 C.POINTER(RunStats)]),
    "vg_euler_reduced_solve_1d": (C.c_int, [vptr, dptr, dptr, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int,
                                            C.c_double, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(EulerParams),
                                            C.POINTER(EulerParams), dptr, C.c_int, C.POINTER(RunStats)]),
    "vg_euler_reduced_solve_2d": (C.c_int, [vptr, dptr, dptr, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                            C.c_int, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int,
                                            C.POINTER(EulerParams), C.POINTER(EulerParams), dptr, C.c_int,
                                            C.POINTER(RunStats)]),
    "vg_euler_multiphysics_time_step_1d": (C.c_int, [vptr, dptr, dptr, C.c_int, C.c_double, C.c_double, C.c_double,
                                                     C.c_int, C.c_int, C.POINTER(EulerParams), C.POINTER(EulerParams)]),
    "vg_euler_multiphysics_time_step_2d": (C.c_int, [vptr, dptr, dptr, C.c_int, C.c_int, C.c_int, C.c_double,
                                                     C.c_double, C.c_double, C.c_int, C.c_int, C.POINTER(EulerParams),
                                                     C.POINTER(EulerParams)]),
    "vg_euler_reduced_time_step_1d": (C.c_int, [vptr, dptr, dptr, C.c_int, C.c_double, C.c_double, C.c_double,
                                                C.c_int, C.c_int, C.POINTER(EulerParams), C.POINTER(EulerParams)]),
    "vg_euler_reduced_time_step_2d": (C.c_int, [vptr, dptr, dptr, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                                C.c_double, C.c_int, C.c_int, C.POINTER(EulerParams), C.POINTER(EulerParams)]),
    "vg_euler_multiphysics_reinitialise": (C.c_int, [vptr, dptr, C.c_int, C.c_int, C.c_int]),
    "vg_euler_reduced_reinitialise": (C.c_int, [vptr, dptr, C.c_int, C.c_int, C.c_int]),
    "vg_euler_adjoint_1d": (C.c_int, [vptr, dptr, dptr, dptr, dptr, C.c_int, C.c_double, C.c_double, C.c_double,
                                      C.c_int, C.c_double, C.c_int, C.c_int, C.POINTER(EulerParams), C.c_int, dptr,
                                      C.c_int, C.POINTER(RunStats)]),
    "vg_euler_adjoint_2d": (C.c_int, [vptr, dptr, dptr, dptr, dptr, C.c_int, C.c_int, C.c_double, C.c_double,
                                      C.c_double, C.c_int, C.c_double, C.c_int, C.c_int, C.POINTER(EulerParams),
                                      C.c_int, dptr, C.c_int, C.POINTER(RunStats)]),
    "vg_euler_adjoint_1d_device": (C.c_int, [vptr, vptr, vptr, vptr, vptr, C.c_int, C.c_double, C.c_double,
                                             C.c_double, C.c_int, C.c_double, C.c_int, C.c_int,
                                             C.POINTER(EulerParams), C.c_int, dptr, C.c_int, C.POINTER(RunStats)]),
    "vg_euler_adjoint_2d_device": (C.c_int, [vptr, vptr, vptr, vptr, vptr, C.c_int, C.c_int, C.c_double, C.c_double,
                                             C.c_double, C.c_int, C.c_double, C.c_int, C.c_int,
                                             C.POINTER(EulerParams), C.c_int, dptr, C.c_int, C.POINTER(RunStats)]),
    "vg_euler_solve_fixed_dt_1d": (C.c_int, [vptr, dptr, dptr, C.c_int, C.c_double, dptr, C.c_int, C.c_double,
                                             C.c_int, C.c_int, C.POINTER(EulerParams)]),
    "vg_euler_solve_fixed_dt_2d": (C.c_int, [vptr, dptr, dptr, C.c_int, C.c_int, C.c_double, dptr, C.c_int,
                                             C.c_double, C.c_int, C.c_int, C.POINTER(EulerParams)]),
    "vg_euler_state_create": (C.c_int, [vptr, C.c_int, C.c_int, C.c_int, C.POINTER(vptr)]),
    "vg_euler_state_destroy": (None, [vptr]),
    "vg_euler_state_load": (C.c_int, [vptr, vptr, C.c_int]),
    "vg_euler_state_store": (C.c_int, [vptr, vptr, C.c_int]),
    "vg_euler_state_planes": (C.c_int, [vptr, C.POINTER(vptr), C.POINTER(C.c_int64)]),
    "vg_euler_state_run": (C.c_int, [vptr, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int, C.c_double,
                                     C.c_int, C.c_int, C.POINTER(EulerParams), C.c_int, dptr, C.c_int,
                                     C.POINTER(RunStats)]),
    "vg_euler_state_max_wave_speed": (C.c_int, [vptr, C.c_int, C.POINTER(EulerParams), dptr]),
    "vg_euler_state_sweep": (C.c_int, [vptr, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int,
                                       C.POINTER(EulerParams), C.c_int]),
    "vg_euler_time_step_1d": (C.c_int, [vptr, dptr, dptr, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int,
                                        C.c_int, C.POINTER(EulerParams), C.c_int]),
    "vg_euler_time_step_2d": (C.c_int, [vptr, dptr, dptr, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                        C.c_double, C.c_int, C.c_int, C.POINTER(EulerParams), C.c_int]),
    "vg_euler_time_step_2d_noalloc": (C.c_int, [vptr, vptr, vptr, C.c_int, C.c_int, C.c_int, C.c_double,
                                                C.c_double, C.c_double, C.c_int, C.c_int, C.POINTER(EulerParams),
                                                C.c_int]),
    "vg_mhd_time_step_1d": (C.c_int, [vptr, dptr, dptr, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int,
                                      C.c_int, C.POINTER(MHDParams), C.c_int]),
    "vg_mhd_time_step_2d": (C.c_int, [vptr, dptr, dptr, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                      C.c_double, C.c_int, C.c_int, C.POINTER(MHDParams), C.c_int]),
    "vg_euler_solve_batch": (C.c_int, [vptr, C.c_int, C.POINTER(C.c_int64), dptr, dptr, dptr,
                                       C.POINTER(EulerParams), dptr, C.c_double, C.c_int, C.c_double, C.c_int,
                                       C.c_int, C.c_int, C.POINTER(C.c_int32), dptr, C.POINTER(RunStats)]),
    "vg_mhd_solve_1d": (C.c_int, [vptr, dptr, dptr, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int,
                                  C.c_double, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(MHDParams), C.c_int, dptr,
                                  C.c_int, C.POINTER(RunStats)]),
    "vg_mhd_solve_2d": (C.c_int, [vptr, dptr, dptr, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int,
                                  C.c_double, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(MHDParams), C.c_int, dptr,
                                  C.c_int, C.POINTER(RunStats)]),
    "vg_mhd_solve_batch": (C.c_int, [vptr, C.c_int, C.POINTER(C.c_int64), dptr, dptr, dptr, C.POINTER(MHDParams),
                                     dptr, C.c_double, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int,
                                     C.POINTER(C.c_int32), dptr, C.POINTER(RunStats)]),
    "vg_division_selfcheck": (C.c_int, [vptr, C.c_int64, dptr, dptr, dptr, dptr, dptr]),
    "vg_mhd_rgfm_solve_1d": (C.c_int, [vptr, dptr, dptr, dptr, dptr, dptr, dptr, C.c_int, C.c_double, C.c_double,
                                       C.c_double, C.c_int, C.c_double, C.c_int, C.c_int, C.POINTER(MHDParams),
                                       C.POINTER(MHDParams), C.c_int, dptr, C.c_int, C.POINTER(RunStats)]),
    "vg_mhd_rgfm_solve_2d": (C.c_int, [vptr, dptr, dptr, dptr, dptr, dptr, dptr, C.c_int, C.c_int, C.c_double,
                                       C.c_double, C.c_double, C.c_int, C.c_double, C.c_int, C.c_int,
                                       C.POINTER(MHDParams), C.POINTER(MHDParams), C.c_int, dptr, C.c_int,
                                       C.POINTER(RunStats)]),
    "vg_mhd_rgfm_ghost_fill": (C.c_int, [vptr, C.c_int, C.c_int, C.c_int, dptr, dptr, dptr, C.POINTER(MHDParams),
                                         C.POINTER(MHDParams)]),
    "vg_rgfm_solve_1d": (C.c_int, [vptr, dptr, dptr, dptr, dptr, dptr, dptr, C.c_int, C.c_double, C.c_double,
                                   C.c_double, C.c_int, C.c_double, C.c_int, C.c_int, C.POINTER(EulerParams),
                                   C.POINTER(EulerParams), C.c_int, dptr, C.c_int, C.POINTER(RunStats)]),
    "vg_rgfm_solve_2d": (C.c_int, [vptr, dptr, dptr, dptr, dptr, dptr, dptr, C.c_int, C.c_int, C.c_double, C.c_double,
                                   C.c_double, C.c_double, C.c_int, C.c_double, C.c_int, C.c_int,
                                   C.POINTER(EulerParams), C.POINTER(EulerParams), C.c_int, dptr, C.c_int,
                                   C.POINTER(RunStats), dptr]),
    "vg_rgfm_solve_batch_2d": (C.c_int, [vptr, C.c_int, dptr, dptr, dptr, dptr, dptr, dptr, C.c_int, C.c_int,
                                         C.c_double, C.c_double, C.c_double, dptr, C.c_int, C.c_double, C.c_int,
                                         C.c_int, C.POINTER(EulerParams), C.POINTER(EulerParams), C.c_int,
                                         C.POINTER(C.c_int32), dptr, dptr, C.POINTER(RunStats)]),
    "vg_rgfm_adjoint_1d": (C.c_int, [vptr, dptr, dptr, dptr, dptr, dptr, dptr, dptr, dptr, dptr, dptr, C.c_int,
                                     C.c_double, C.c_double, C.c_double, C.c_int, C.c_double, C.c_int,
                                     C.POINTER(EulerParams), C.POINTER(EulerParams), C.c_int, dptr, C.c_int,
                                     C.POINTER(RunStats)]),
    "vg_rgfm_adjoint_2d": (C.c_int, [vptr, dptr, dptr, dptr, dptr, dptr, dptr, dptr, dptr, dptr, dptr, C.c_int,
                                     C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int, C.c_double,
                                     C.c_int, C.POINTER(EulerParams), C.POINTER(EulerParams), C.c_int, dptr,
                                     C.c_int, C.POINTER(RunStats)]),
    "vg_rgfm_adjoint_batch_2d": (C.c_int, [vptr, C.c_int, dptr, dptr, dptr, dptr, dptr, dptr, dptr, dptr, dptr,
                                           dptr, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, dptr,
                                           C.c_int, C.c_double, C.c_int, C.POINTER(EulerParams),
                                           C.POINTER(EulerParams), C.c_int, C.POINTER(C.c_int32),
                                           C.POINTER(C.c_int32), C.POINTER(RunStats)]),
    "vg_rgfm_solve_fixed_dt_1d": (C.c_int, [vptr, dptr, dptr, dptr, dptr, dptr, dptr, C.c_int, C.c_double, dptr,
                                            C.c_int, C.c_double, C.c_int, C.POINTER(EulerParams),
                                            C.POINTER(EulerParams)]),
    "vg_rgfm_solve_fixed_dt_2d": (C.c_int, [vptr, dptr, dptr, dptr, dptr, dptr, dptr, C.c_int, C.c_int, C.c_double,
                                            C.c_double, dptr, C.c_int, C.c_double, C.c_int, C.POINTER(EulerParams),
                                            C.POINTER(EulerParams)]),
    "vg_hpr_solve_1d": (C.c_int, [vptr, dptr, dptr, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int,
                                  C.c_double, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(HPRParams), C.c_int, dptr,
                                  C.c_int, C.POINTER(RunStats)]),
    "vg_hpr_solve_2d": (C.c_int, [vptr, dptr, dptr, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int,
                                  C.c_double, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(HPRParams), C.c_int, dptr,
                                  C.c_int, C.POINTER(RunStats)]),
    "vg_hpr_state_create": (C.c_int, [vptr, C.c_int, C.c_int, C.c_int, C.POINTER(vptr)]),
    "vg_hpr_state_destroy": (None, [vptr]),
    "vg_hpr_state_load": (C.c_int, [vptr, vptr, C.c_int]),
    "vg_hpr_state_store": (C.c_int, [vptr, vptr, C.c_int]),
    "vg_hpr_state_planes": (C.c_int, [vptr, C.POINTER(vptr), C.POINTER(C.c_int64)]),
    "vg_hpr_state_sweep": (C.c_int, [vptr, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int,
                                     C.POINTER(HPRParams)]),
    "vg_hpr_time_step_1d": (C.c_int, [vptr, dptr, dptr, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int,
                                      C.c_int, C.POINTER(HPRParams)]),
    "vg_hpr_time_step_2d": (C.c_int, [vptr, dptr, dptr, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                      C.c_double, C.c_int, C.c_int, C.POINTER(HPRParams)]),
    "vg_hpr_state_run": (C.c_int, [vptr, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int, C.c_double, C.c_int,
                                   C.c_int, C.c_int, C.POINTER(HPRParams), C.c_int, dptr, C.c_int,
                                   C.POINTER(RunStats)]),
    "vg_hpr_state_max_wave_speed": (C.c_int, [vptr, C.c_int, C.POINTER(HPRParams), dptr]),
    "vg_hpr_adjoint_1d": (C.c_int, [vptr, dptr, dptr, dptr, dptr, dptr, C.c_int, C.c_double, C.c_double, C.c_double,
                                    C.c_int, C.c_double, C.c_int, C.c_int, C.c_int, C.POINTER(HPRParams), C.c_int, dptr,
                                    C.c_int, C.POINTER(RunStats)]),
    "vg_hpr_adjoint_2d": (C.c_int, [vptr, dptr, dptr, dptr, dptr, dptr, C.c_int, C.c_int, C.c_double, C.c_double,
                                    C.c_double, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int, C.POINTER(HPRParams),
                                    C.c_int, dptr, C.c_int, C.POINTER(RunStats)]),
    "vg_hpr_solve_fixed_dt_1d": (C.c_int, [vptr, dptr, dptr, C.c_int, C.c_double, dptr, C.c_int, C.c_double, C.c_int,
                                           C.c_int, C.c_int, C.POINTER(HPRParams)]),
    "vg_hpr_solve_fixed_dt_2d": (C.c_int, [vptr, dptr, dptr, C.c_int, C.c_int, C.c_double, dptr, C.c_int, C.c_double,
                                           C.c_int, C.c_int, C.c_int, C.POINTER(HPRParams)]),
}

_lib = None


def load(path: str = LIB_PATH) -> C.CDLL:
    """Loads the HIP library (raises if it is absent or incomplete)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with `python -m vanguard_amd.build`")
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class VanguardError(RuntimeError):
    def __init__(self, status: int, what: str):
        super().__init__(f"{what} failed with status {status}: {load().vg_last_error().decode()}")
        self.status = status


def check(status: int, what: str) -> int:
    if status != VG_OK:
        raise VanguardError(status, what)
    return status


def as_dptr(a):
    return a.ctypes.data_as(dptr)
