"""The reference's `.dat` result files, written on the host from solver output.

The reference's test drivers end with an `outputSolution` / `outputSolution2D`
call that writes one text file per plotted quantity into the working
directory: `x value` per line in 1D, `y x value` in 2D, doubles printed the
way a C++ stream prints them by default (six significant digits), one line per
cell.  These functions write the same files, byte for byte, from the arrays
the solvers here return, so plotting scripts written for the reference keep
working:

    EulerTests::outputSolution[2D]      (Euler/eulertests.cpp:54-108)
    MHDTests::outputSolution[2D]        (MHD/mhdtests.cpp:123-160)
    HPRTests::outputSolution[2D]        (HPR/hprtests.cpp:219-277)
    EulerRGFMTests::outputSolution[2D]  (RGFM/eulerrgfmtests.cpp:215-334)

Their particulars are kept: the coordinates are `(1.0 / rows) * index` in both
directions; the 2D Euler writer creates its velocity files empty; the RGFM
writers pick each cell's material by toggling at level-set sign changes
(starting outside the ghost region in 1D, inside it at the start of every 2D
row) and write the level-set slice of the middle row without a separator
between its two numbers.  At index 0 the reference's toggle test also reads
the level set one cell before the start (outside the array).  That value is
multiplied by the level set's first value, and the test reads it only when the
first value's square is not positive, i.e. when the first value is 0 (then
the product is 0 or NaN, never positive) or so small that its square
underflows to 0 (then the reference's outcome depends on memory outside the
array).  Here the term is taken as not positive: the same files in every case
the reference defines.
"""

from __future__ import annotations

import math
import os

import numpy as np


def cxx(x: float) -> str:
    """A double as `std::ostream << double` prints it by default (%g, six
    significant digits; NaN with its sign, as glibc prints it)."""
    x = float(x)
    if x != x:
        return "-nan" if math.copysign(1.0, x) < 0.0 else "nan"
    return "%.6g" % x


def _write(directory: str, name: str, lines) -> None:
    with open(os.path.join(directory, name), "w", newline="\n") as f:
        f.writelines(lines)


def _lines_1d(h: float, values) -> list[str]:
    return [f"{cxx(h * i)} {cxx(v)}\n" for i, v in enumerate(values)]


def _lines_2d(h: float, values: np.ndarray) -> list[str]:
    rows, cols = values.shape
    return [f"{cxx(h * i)} {cxx(h * j)} {cxx(values[i, j])}\n" for i in range(rows) for j in range(cols)]


def euler_output_solution(cells: np.ndarray, directory: str = ".") -> None:
    """EulerTests::outputSolution: density, x/y/z velocity and pressure of a
    1D Euler state (n x 6 primitives)."""
    w = np.asarray(cells, dtype=np.float64)
    h = 1.0 / w.shape[0]
    for name, f in (("density", 0), ("xVelocity", 1), ("yVelocity", 2), ("zVelocity", 3), ("pressure", 4)):
        _write(directory, f"{name}.dat", _lines_1d(h, w[:, f]))


def euler_output_solution_2d(cells: np.ndarray, directory: str = ".") -> None:
    """EulerTests::outputSolution2D: density and pressure of a 2D Euler state
    (rows x columns x 6); the three velocity files are created empty, as the
    reference does."""
    w = np.asarray(cells, dtype=np.float64)
    h = 1.0 / w.shape[0]
    _write(directory, "density.dat", _lines_2d(h, w[..., 0]))
    for name in ("xVelocity", "yVelocity", "zVelocity"):
        _write(directory, f"{name}.dat", [])
    _write(directory, "pressure.dat", _lines_2d(h, w[..., 4]))


def mhd_output_solution(cells: np.ndarray, directory: str = ".") -> None:
    """MHDTests::outputSolution: density and y magnetic field (n x 9)."""
    w = np.asarray(cells, dtype=np.float64)
    h = 1.0 / w.shape[0]
    _write(directory, "density.dat", _lines_1d(h, w[:, 0]))
    _write(directory, "yMagneticField.dat", _lines_1d(h, w[:, 6]))


def mhd_output_solution_2d(cells: np.ndarray, directory: str = ".") -> None:
    """MHDTests::outputSolution2D: density and y magnetic field (rows x columns x 9)."""
    w = np.asarray(cells, dtype=np.float64)
    h = 1.0 / w.shape[0]
    _write(directory, "density.dat", _lines_2d(h, w[..., 0]))
    _write(directory, "yMagneticField.dat", _lines_2d(h, w[..., 6]))


def _hpr_temperature(w: np.ndarray, params) -> np.ndarray:
    # HPRMieGruneisen::computeTemperature for the stiffened gas, Grueneisen
    # coefficient gamma - 1: T_ref + (p - (-p_inf)) / (rho G c_v), left to right
    g = params.adiabatic_index - 1.0
    rho, p = w[..., 0], w[..., 1]
    return params.reference_temperature + ((p - (-params.stiffening_parameter)) /
                                           ((rho * g) * params.specific_heat_capacity))


def hpr_output_solution(cells: np.ndarray, params, directory: str = ".") -> None:
    """HPRTests::outputSolution: density, x/y velocity, temperature and the x
    heat flux T * (alpha^2 J_x) of a 1D HPR state (n x 17 primitives)."""
    if params.equation_of_state != "StiffenedGas":
        raise ValueError("only the StiffenedGas equation of state is implemented")
    w = np.asarray(cells, dtype=np.float64)
    h = 1.0 / w.shape[0]
    temperature = _hpr_temperature(w, params)
    a2 = params.heat_wave_speed * params.heat_wave_speed
    heat_flux = temperature * (a2 * w[:, 14])
    _write(directory, "density.dat", _lines_1d(h, w[:, 0]))
    _write(directory, "xVelocity.dat", _lines_1d(h, w[:, 11]))
    _write(directory, "yVelocity.dat", _lines_1d(h, w[:, 12]))
    _write(directory, "temperature.dat", _lines_1d(h, temperature))
    _write(directory, "heatFlux.dat", _lines_1d(h, heat_flux))


def hpr_output_solution_2d(cells: np.ndarray, params, directory: str = ".") -> None:
    """HPRTests::outputSolution2D: density and x velocity (rows x columns x 17)."""
    w = np.asarray(cells, dtype=np.float64)
    h = 1.0 / w.shape[0]
    _write(directory, "density.dat", _lines_2d(h, w[..., 0]))
    _write(directory, "xVelocity.dat", _lines_2d(h, w[..., 11]))


def _toggles(phi_row: np.ndarray, start: bool) -> np.ndarray:
    """Per cell: whether it is in the ghost region (material 2 plotted)."""
    out = np.empty(phi_row.shape[0], dtype=bool)
    ghost = start
    value = phi_row[0]
    for j in range(phi_row.shape[0]):
        prev_positive = j > 0 and (phi_row[j - 1] * value) > 0.0
        if (phi_row[j] * value) <= 0.0 and prev_positive:
            ghost = not ghost
        out[j] = ghost
        value = phi_row[j]
    return out


def rgfm_output_solution(system, directory: str = ".") -> None:
    """EulerRGFMTests::outputSolution: level set, and density and pressure of
    the material each 1D cell shows (material 1 outside the ghost region)."""
    m1 = np.asarray(system.material1, dtype=np.float64)
    m2 = np.asarray(system.material2, dtype=np.float64)
    phi = np.asarray(system.level_set, dtype=np.float64)
    n = phi.shape[0]
    h = 1.0 / n
    ghost = _toggles(phi, False)
    shown = np.where(ghost[:, None], m2, m1)
    _write(directory, "levelSetFunction.dat", _lines_1d(h, phi))
    _write(directory, "density.dat", _lines_1d(h, shown[:, 0]))
    _write(directory, "pressure.dat", _lines_1d(h, shown[:, 4]))


def rgfm_output_solution_2d(system, directory: str = ".") -> None:
    """EulerRGFMTests::outputSolution2D: level set, density and pressure of the
    material each cell shows (every row starts in the ghost region), and the
    same three along the middle row (row floor(rows / 2))."""
    m1 = np.asarray(system.material1, dtype=np.float64)
    m2 = np.asarray(system.material2, dtype=np.float64)
    phi = np.asarray(system.level_set, dtype=np.float64)
    rows, cols = phi.shape
    h = 1.0 / rows
    ghost = np.stack([_toggles(phi[i], True) for i in range(rows)])
    shown = np.where(ghost[..., None], m2, m1)
    mid = math.floor(0.5 * rows)
    _write(directory, "levelSetFunction.dat", _lines_2d(h, phi))
    _write(directory, "density.dat", _lines_2d(h, shown[..., 0]))
    _write(directory, "pressure.dat", _lines_2d(h, shown[..., 4]))
    _write(directory, "levelSetFunctionSlice.dat", [f"{cxx(h * j)}{cxx(phi[mid, j])}\n" for j in range(cols)])
    _write(directory, "densitySlice.dat", [f"{cxx(h * j)} {cxx(shown[mid, j, 0])}\n" for j in range(cols)])
    _write(directory, "pressureSlice.dat", [f"{cxx(h * j)} {cxx(shown[mid, j, 4])}\n" for j in range(cols)])
