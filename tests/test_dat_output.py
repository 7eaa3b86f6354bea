"""The `.dat` writers (vanguard_amd/dat.py) against the reference's own
outputSolution functions: tests/golden/dat_fixtures.npz holds the files the
compiled reference wrote (tests/golden/gen_dat.py) for seeded Euler, MHD, HPR
and RGFM states -- signed zeros, NaNs of both signs, values near the
six-digit rounding boundaries, level sets that cross twice, start at exactly 0,
and 2D grids with an even and an odd row count -- and every file must come out
byte for byte the same, with no file more or fewer."""

import os
import types

import numpy as np
import pytest

from vanguard_amd import dat
from vanguard_amd.solvers import HPRMaterialParameters

FIX = os.path.join(os.path.dirname(__file__), "golden", "dat_fixtures.npz")


@pytest.fixture(scope="module")
def fx():
    return np.load(FIX)


def hpr_params(a):
    return HPRMaterialParameters(is_thermal=bool(a[0]), reference_density=a[1], reference_pressure=a[2],
                                 reference_temperature=a[3], initial_temperature=a[4], specific_heat_capacity=a[5],
                                 adiabatic_index=a[6], stiffening_parameter=a[7], transverse_wave_speed=a[8],
                                 strain_dissipation_time=a[9], thermal_impulse_relaxation_time=a[10],
                                 heat_wave_speed=a[11])


def write(fx, name, kind, dim, directory):
    if kind == "rgfm":
        system = types.SimpleNamespace(material1=fx[f"{name}__m1"], material2=fx[f"{name}__m2"],
                                       level_set=fx[f"{name}__phi"])
        (dat.rgfm_output_solution if dim == "1d" else dat.rgfm_output_solution_2d)(system, directory)
        return
    w = fx[f"{name}__in"]
    if kind == "euler":
        (dat.euler_output_solution if dim == "1d" else dat.euler_output_solution_2d)(w, directory)
    elif kind == "mhd":
        (dat.mhd_output_solution if dim == "1d" else dat.mhd_output_solution_2d)(w, directory)
    else:
        p = hpr_params(fx["hpr_params"])
        (dat.hpr_output_solution if dim == "1d" else dat.hpr_output_solution_2d)(w, p, directory)


CASES = [("euler1d", "euler", "1d"), ("euler2d", "euler", "2d"), ("mhd1d", "mhd", "1d"), ("mhd2d", "mhd", "2d"),
         ("hpr1d", "hpr", "1d"), ("hpr2d", "hpr", "2d"), ("rgfm1d", "rgfm", "1d"), ("rgfm1d_phi0", "rgfm", "1d"),
         ("rgfm2d", "rgfm", "2d"), ("rgfm2d_odd", "rgfm", "2d")]


@pytest.mark.parametrize("name,kind,dim", CASES)
def test_dat_files_match_the_reference_writers(fx, name, kind, dim, tmp_path):
    write(fx, name, kind, dim, str(tmp_path))
    expected = sorted(str(f) for f in fx[f"{name}__files"])
    assert sorted(os.listdir(tmp_path)) == expected
    for f in expected:
        got = (tmp_path / f).read_bytes()
        ref = bytes(fx[f"{name}__file__{f}"])
        assert got == ref, f"{name}/{f}"


def test_cxx_formatting_closed_forms():
    """What a C++ stream prints by default: %g with six significant digits,
    no trailing zeros, two-digit exponents, NaN with its sign."""
    assert dat.cxx(0.0) == "0" and dat.cxx(-0.0) == "-0"
    assert dat.cxx(1.0 / 3.0) == "0.333333" and dat.cxx(123456.5) == "123456"
    assert dat.cxx(1234567.0) == "1.23457e+06" and dat.cxx(1e-300) == "1e-300"
    assert dat.cxx(float("inf")) == "inf" and dat.cxx(float("-inf")) == "-inf"
    assert dat.cxx(float("nan")) == "nan" and dat.cxx(np.copysign(np.nan, -1.0)) == "-nan"
