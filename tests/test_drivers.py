"""The reference's test drivers end to end, pinned by the files the compiled
reference's own drivers wrote (tests/golden/driver_fixtures.npz, from
`ref_harness driver ...`: the reference's initial-condition code, solve and
`.dat` writer).

CPU: the initial conditions of `problems.py` run through the oracle's solve and
written by `dat.py` give the reference drivers' files byte for byte -- a wrong
initial state, material parameter or final time shows in the files.
GPU: `vanguard_amd.drivers` (the drivers' own names, the solves through the C
ABI) writes the same files byte for byte.
"""

import os

import numpy as np
import pytest

import oracle_lib as O
from vanguard_amd import dat, drivers, problems
from vanguard_amd import solvers as S

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = np.load(os.path.join(HERE, "golden", "driver_fixtures.npz"))

NAMES = ["euler_toro1", "euler_2d_toro1", "rgfm_toro1_hllc", "rgfm_fedkiw_hllc", "rgfm_2d_toro1_hllc",
         "rgfm_2d_fedkiw_hllc", "mhd_dumbser1", "mhd_dumbser2", "mhd_dumbser3", "mhd_2d_dumbser1", "mhd_2d_dumbser2",
         "hpr_stokes", "hpr_heat"]


def assert_files(name, directory):
    files = sorted(f for f in os.listdir(directory) if f.endswith(".dat"))
    assert files == list(FIX[f"{name}__files"]), (files, FIX[f"{name}__files"])
    for f in files:
        with open(os.path.join(directory, f), "rb") as fh:
            got = fh.read()
        want = FIX[f"{name}__file__{f}"].tobytes()
        assert got == want, f"{name}/{f} differs from the reference driver's file"


EULER = (1.4, 0.0, 0.0, 0.0, 0.0, 0.0)


def oracle_hpr_params(p):
    return dict(is_thermal=int(p.is_thermal), is_plastic=int(p.is_plastic), rho0=p.reference_density,
                t_ref=p.reference_temperature, t0=p.initial_temperature, cv=p.specific_heat_capacity,
                gamma=p.adiabatic_index, p_inf=p.stiffening_parameter, cs=p.transverse_wave_speed,
                tau1=p.strain_dissipation_time, tau2=p.thermal_impulse_relaxation_time, alpha=p.heat_wave_speed,
                eos=0, p_ref=p.reference_pressure)


def run_oracle(name, n, k, d):
    """The driver `name` with the oracle's solve in place of the reference's."""
    def euler_params(g):
        return (g,) + EULER[1:]

    if name == "euler_toro1":
        out, _ = O.euler_solve(problems.sod_1d(n), order=2, dim=1, dx=1.0 / n, cfl=0.8, tfinal=0.25, max_steps=0,
                               bias=0.0, limiter=0, params=EULER)
        dat.euler_output_solution(out, d)
    elif name == "euler_2d_toro1":
        out, _ = O.euler_solve(problems.euler_toro1_2d(n), order=2, dim=2, dx=1.0 / n, cfl=0.8, tfinal=0.125,
                               max_steps=0, bias=0.0, limiter=0, params=EULER)
        dat.euler_output_solution_2d(out, d)
    elif name.startswith("rgfm_"):
        build, t, g2, dim = {"rgfm_toro1_hllc": (problems.rgfm_toro1_1d, 0.25, 1.4, 1),
                             "rgfm_fedkiw_hllc": (problems.rgfm_fedkiw_1d, 0.0012, 1.67, 1),
                             "rgfm_2d_toro1_hllc": (problems.rgfm_toro1_2d, 0.125, 1.4, 2),
                             "rgfm_2d_fedkiw_hllc": (problems.rgfm_fedkiw_2d, 0.0006, 1.67, 2)}[name]
        m1, m2, phi = build(n)
        o1, o2, oph, _ = O.rgfm_solve(m1, m2, phi, dim=dim, dx=1.0 / n, cfl=0.8, tfinal=t, max_steps=0, bias=0.0,
                                      limiter=0, p1=euler_params(1.4), p2=euler_params(g2))
        system = S.MultimaterialSystem(o1, o2, oph)
        (dat.rgfm_output_solution if dim == 1 else dat.rgfm_output_solution_2d)(system, d)
    elif name.startswith("mhd_2d_"):
        test = int(name[-1])
        _, _, _, gamma, cfl, t = problems.MHD_DUMBSER_2D[test]
        out, _ = O.mhd_solve(problems.mhd_dumbser_2d(n, test), order=2, dim=2, dx=1.0 / n, cfl=cfl, tfinal=t,
                             max_steps=0, bias=0.0, limiter=0, gamma=gamma, subcycles=k)
        dat.mhd_output_solution_2d(out, d)
    elif name.startswith("mhd_"):
        test = int(name[-1])
        _, _, _, gamma, t = problems.MHD_DUMBSER[test]
        out, _ = O.mhd_solve(problems.mhd_dumbser_1d(n, test), order=2, dim=1, dx=1.0 / n, cfl=0.8, tfinal=t,
                             max_steps=0, bias=0.0, limiter=0, gamma=gamma, subcycles=k)
        dat.mhd_output_solution(out, d)
    else:
        p = drivers.stokes_parameters() if name == "hpr_stokes" else drivers.heat_conduction_parameters()
        w = problems.hpr_stokes_1d(n) if name == "hpr_stokes" else problems.hpr_heat_conduction_1d(n)
        out, _ = O.hpr_solve(w, order=2, dim=1, dx=1.0 / n, cfl=0.8, tfinal=1.0, max_steps=0, bias=0.0, limiter=0,
                             subcycles=k, params=oracle_hpr_params(p))
        dat.hpr_output_solution(out, p, d)


@pytest.mark.parametrize("name", NAMES)
def test_driver_files_through_the_oracle(name, tmp_path):
    n, k = (int(x) for x in FIX[f"{name}__args"])
    run_oracle(name, n, k, str(tmp_path))
    assert_files(name, str(tmp_path))


def test_stokes_parameters_follow_the_reference_formulas():
    """tau1 = 6 mu / (p_ref cs^2); with T0 = 0 the thermal-impulse time is
    infinite (hprmaterialparameters.cpp:101-119)."""
    p = drivers.stokes_parameters()
    assert p.strain_dissipation_time == 6.0 * (1e-2 / ((1.0 / 1.4) * 1.0))
    assert p.thermal_impulse_relaxation_time == np.inf
    q = drivers.heat_conduction_parameters()
    assert q.thermal_impulse_relaxation_time == 1e-2 * (1.0 / (1.0 * (2.0 * 2.0)))


GPU_DRIVERS = {
    "euler_toro1": lambda n, k, d: drivers.EulerTests.solveToroTest1(n, directory=d),
    "euler_2d_toro1": lambda n, k, d: drivers.EulerTests.solve2DToroTest1(n, directory=d),
    "rgfm_toro1_hllc": lambda n, k, d: drivers.EulerRGFMTests.solveToroTest1HLLC(n, k, directory=d),
    "rgfm_fedkiw_hllc": lambda n, k, d: drivers.EulerRGFMTests.solveFedkiwTestHLLC(n, k, directory=d),
    "rgfm_2d_toro1_hllc": lambda n, k, d: drivers.EulerRGFMTests.solve2DToroTest1HLLC(n, k, directory=d),
    "rgfm_2d_fedkiw_hllc": lambda n, k, d: drivers.EulerRGFMTests.solve2DFedkiwTestHLLC(n, k, directory=d),
    "mhd_dumbser1": lambda n, k, d: drivers.MHDTests.solveDumbserTest1(n, k, directory=d),
    "mhd_dumbser2": lambda n, k, d: drivers.MHDTests.solveDumbserTest2(n, k, directory=d),
    "mhd_dumbser3": lambda n, k, d: drivers.MHDTests.solveDumbserTest3(n, k, directory=d),
    "mhd_2d_dumbser1": lambda n, k, d: drivers.MHDTests.solve2DDumbserTest1(n, k, directory=d),
    "mhd_2d_dumbser2": lambda n, k, d: drivers.MHDTests.solve2DDumbserTest2(n, k, directory=d),
    "hpr_stokes": lambda n, k, d: drivers.HPRTests.solveStokesFirstProblem(n, k, directory=d),
    "hpr_heat": lambda n, k, d: drivers.HPRTests.solveHeatConductionProblem(n, k, directory=d),
}


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_driver_files_on_the_gpu(vg_ctx, name, tmp_path):
    n, k = (int(x) for x in FIX[f"{name}__args"])
    GPU_DRIVERS[name](n, k, str(tmp_path))
    assert_files(name, str(tmp_path))
