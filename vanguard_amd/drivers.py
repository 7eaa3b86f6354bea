"""The reference's test drivers on the GPU path.

Each driver of the reference (Euler/eulertests.cpp, RGFM/eulerrgfmtests.cpp,
MHD/mhdtests.cpp, HPR/hprtests.cpp) builds an initial condition, calls a solver
and writes the solution's `.dat` files into the working directory.  The
functions here do the same with the same names and arguments: the initial
conditions of `problems.py`, the solves through the C ABI (`solvers.py`), the
files through `dat.py`, into `directory` (the working directory by default).
The drivers that call the exact Riemann solver are absent: the reference's
exact solver does not compile (RGFM/exactsolver.cpp:295), so those drivers
cannot run there either.

Like the reference's drivers, these write whatever state the solve ends with:
a final state with inadmissible cells (the 2D Dumbser disc at 32 x 32 ends with
a negative pressure, which its files do not show) is written, not raised.  Pass
raise_on_error=True to have the solve's status raised instead.
"""

from __future__ import annotations

from . import dat, problems
from . import solvers as S


def _quiet(kw):
    return dict(kw, raise_on_error=kw.get("raise_on_error", False))


class EulerTests:
    """EulerTests (Euler/eulertests.h:8-23)."""

    @staticmethod
    def solveToroTest1(cellCount: int, directory: str = ".", **kw) -> None:
        """eulertests.cpp:7-27: Sod's tube to t = 0.25, CFL 0.8, gamma 1.4."""
        out = S.SecondOrderSolver.solve(problems.sod_1d(cellCount), 1.0 / cellCount, 0.8, 0.25, 0.0, 0, 0,
                                        S.EulerMaterialParameters(1.4), **_quiet(kw))
        dat.euler_output_solution(out, directory)

    @staticmethod
    def solve2DToroTest1(cellCount: int, directory: str = ".", **kw) -> None:
        """eulertests.cpp:29-52: the cylindrical explosion to t = 0.125."""
        out = S.SecondOrderSolver.solve2D(problems.euler_toro1_2d(cellCount), 1.0 / cellCount, 0.8, 0.125, 0.0, 0, 0,
                                          S.EulerMaterialParameters(1.4), **_quiet(kw))
        dat.euler_output_solution_2d(out, directory)


class EulerRGFMTests:
    """EulerRGFMTests (RGFM/eulerrgfmtests.h:8-27), the HLLC drivers."""

    @staticmethod
    def _run(system_arrays, cellCount, final_time, p1, p2, reinit, directory, writer, **kw):
        system = S.MultimaterialSystem(*system_arrays)
        out = S.RGFMSolver.solveHLLC(system, 1.0 / cellCount, 0.8, final_time, 0.0, 0, 0, reinit,
                                     S.EulerMaterialParameters(p1), S.EulerMaterialParameters(p2), **_quiet(kw))
        writer(out, directory)

    @staticmethod
    def solveToroTest1HLLC(cellCount: int, reinitialisationFrequency: int = 0, directory: str = ".", **kw) -> None:
        """eulerrgfmtests.cpp:28-48: Sod's tube as two materials to t = 0.25."""
        EulerRGFMTests._run(problems.rgfm_toro1_1d(cellCount), cellCount, 0.25, 1.4, 1.4, reinitialisationFrequency,
                            directory, dat.rgfm_output_solution, **kw)

    @staticmethod
    def solveFedkiwTestHLLC(cellCount: int, reinitialisationFrequency: int = 0, directory: str = ".", **kw) -> None:
        """eulerrgfmtests.cpp:84-111: air shocked into helium to t = 0.0012."""
        EulerRGFMTests._run(problems.rgfm_fedkiw_1d(cellCount), cellCount, 0.0012, 1.4, 1.67,
                            reinitialisationFrequency, directory, dat.rgfm_output_solution, **kw)

    @staticmethod
    def solve2DToroTest1HLLC(cellCount: int, reinitialisationFrequency: int = 0, directory: str = ".",
                             **kw) -> None:
        """eulerrgfmtests.cpp:136-155: the two-material explosion to t = 0.125."""
        EulerRGFMTests._run(problems.rgfm_toro1_2d(cellCount), cellCount, 0.125, 1.4, 1.4, reinitialisationFrequency,
                            directory, dat.rgfm_output_solution_2d, **kw)

    @staticmethod
    def solve2DFedkiwTestHLLC(cellCount: int, reinitialisationFrequency: int = 0, directory: str = ".",
                              **kw) -> None:
        """eulerrgfmtests.cpp:183-213: a shocked disc next to helium to t = 0.0006."""
        EulerRGFMTests._run(problems.rgfm_fedkiw_2d(cellCount), cellCount, 0.0006, 1.4, 1.67,
                            reinitialisationFrequency, directory, dat.rgfm_output_solution_2d, **kw)


class MHDTests:
    """MHDTests (MHD/mhdtests.h:8-23)."""

    @staticmethod
    def _solve1d(test, cellCount, subcyclingIterations, directory, **kw):
        _, _, _, gamma, final_time = problems.MHD_DUMBSER[test]
        out = S.MHDSecondOrderSolver.solve(problems.mhd_dumbser_1d(cellCount, test), 1.0 / cellCount, 0.8, final_time,
                                           0.0, 0, subcyclingIterations, S.MHDMaterialParameters(gamma), **_quiet(kw))
        dat.mhd_output_solution(out, directory)

    @staticmethod
    def _solve2d(test, cellCount, subcyclingIterations, directory, **kw):
        _, _, _, gamma, cfl, final_time = problems.MHD_DUMBSER_2D[test]
        out = S.MHDSecondOrderSolver.solve2D(problems.mhd_dumbser_2d(cellCount, test), 1.0 / cellCount, cfl,
                                             final_time, 0.0, 0, subcyclingIterations, S.MHDMaterialParameters(gamma),
                                             **_quiet(kw))
        dat.mhd_output_solution_2d(out, directory)

    @staticmethod
    def solveDumbserTest1(cellCount: int, subcyclingIterations: int = 0, directory: str = ".", **kw) -> None:
        """mhdtests.cpp:7-27: Brio-Wu, gamma 2, to t = 0.1."""
        MHDTests._solve1d(1, cellCount, subcyclingIterations, directory, **kw)

    @staticmethod
    def solveDumbserTest2(cellCount: int, subcyclingIterations: int = 0, directory: str = ".", **kw) -> None:
        """mhdtests.cpp:29-49: gamma 5/3, split at 0.4, to t = 0.2."""
        MHDTests._solve1d(2, cellCount, subcyclingIterations, directory, **kw)

    @staticmethod
    def solveDumbserTest3(cellCount: int, subcyclingIterations: int = 0, directory: str = ".", **kw) -> None:
        """mhdtests.cpp:51-71: gamma 5/3, to t = 0.16."""
        MHDTests._solve1d(3, cellCount, subcyclingIterations, directory, **kw)

    @staticmethod
    def solve2DDumbserTest1(cellCount: int, subcyclingIterations: int = 0, directory: str = ".", **kw) -> None:
        """mhdtests.cpp:73-96: the Brio-Wu states in a disc, CFL 0.6, to t = 0.05."""
        MHDTests._solve2d(1, cellCount, subcyclingIterations, directory, **kw)

    @staticmethod
    def solve2DDumbserTest2(cellCount: int, subcyclingIterations: int = 0, directory: str = ".", **kw) -> None:
        """mhdtests.cpp:98-121: test 2's states in a disc, to t = 0.1."""
        MHDTests._solve2d(2, cellCount, subcyclingIterations, directory, **kw)


def stokes_parameters() -> S.HPRMaterialParameters:
    """The HPRMaterialParameters of HPRTests::solveStokesFirstProblem
    (hprtests.cpp:10-14), strain-dissipation and thermal-impulse relaxation
    times configured from the viscosity and the heat-conduction coefficient."""
    p = S.HPRMaterialParameters("StiffenedGas", False, False, 8.31445985, 1.0, 1.0 / 1.4, 0.0, 0.0, 1.0, 1.4, 0.0,
                                0.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 10.0 ** -2.0, 0.75, 0.0, 1.0, 10.0 ** -16.0, 0.0,
                                0.0, 0.0)
    p.configure_strain_dissipation_time()
    p.configure_thermal_impulse_relaxation_time()
    return p


def heat_conduction_parameters() -> S.HPRMaterialParameters:
    """The HPRMaterialParameters of HPRTests::solveHeatConductionProblem
    (hprtests.cpp:36-40)."""
    p = S.HPRMaterialParameters("StiffenedGas", True, False, 8.31445985, 1.0, 1.0, 0.0, 1.0, 2.5, 1.4, 0.0, 0.0, 0.0,
                                0.0, 0.0, 1.0, 0.0, 0.0, 10.0 ** -2.0, 0.75, 0.0, 1.0, 2.0, 0.0, 0.0, 0.0)
    p.configure_strain_dissipation_time()
    p.configure_thermal_impulse_relaxation_time(10.0 ** -2.0)
    return p


class HPRTests:
    """HPRTests (HPR/hprtests.h:8-26), the stiffened-gas drivers."""

    @staticmethod
    def solveStokesFirstProblem(cellCount: int, subcyclingIterations: int = 0, directory: str = ".", **kw) -> None:
        """hprtests.cpp:7-31: Stokes' first problem to t = 1."""
        p = stokes_parameters()
        out = S.HPRSecondOrderSolver.solve(problems.hpr_stokes_1d(cellCount), 1.0 / cellCount, 0.8, 1.0, 0.0, 0,
                                           subcyclingIterations, p, **_quiet(kw))
        dat.hpr_output_solution(out, p, directory)

    @staticmethod
    def solveHeatConductionProblem(cellCount: int, subcyclingIterations: int = 0, directory: str = ".",
                                   **kw) -> None:
        """hprtests.cpp:33-57: the heat-conduction problem to t = 1."""
        p = heat_conduction_parameters()
        out = S.HPRSecondOrderSolver.solve(problems.hpr_heat_conduction_1d(cellCount), 1.0 / cellCount, 0.8, 1.0, 0.0,
                                           0, subcyclingIterations, p, **_quiet(kw))
        dat.hpr_output_solution(out, p, directory)
