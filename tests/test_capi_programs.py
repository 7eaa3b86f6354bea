"""Compiled callers of the library: a plain C program against include/vanguard.h
and a reference-style C++ driver against include/vanguard_compat.hpp, both
linked with -lvanguard_hip.  The CPU tests compile and link them; the GPU tests
run them and compare their outputs with the compiled reference's golden runs,
bit for bit."""

import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle_lib as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "vanguard_amd")


def build(tmp, src, compiler, std):
    exe = os.path.join(tmp, os.path.splitext(os.path.basename(src))[0])
    cmd = [compiler, f"-std={std}", "-Wall", "-Wextra", "-Werror", "-O1", "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", src), "-L", LIBDIR, "-lvanguard_hip", f"-Wl,-rpath,{LIBDIR}", "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, f"{' '.join(cmd)}\n{r.stdout}\n{r.stderr}"
    return exe


@pytest.mark.parametrize("src,compiler,std", [("capi_check.c", "gcc", "c99"), ("compat_driver.cpp", "g++", "c++11")])
def test_programs_compile_and_link(tmp_path, src, compiler, std):
    if shutil.which(compiler) is None:
        pytest.skip(f"{compiler} not installed")
    build(str(tmp_path), src, compiler, std)


def load(tmp, name, shape):
    return np.fromfile(os.path.join(tmp, name), dtype=np.float64).reshape(shape)


@pytest.mark.gpu
def test_c_program_reproduces_reference(tmp_path):
    tmp = str(tmp_path)
    exe = build(tmp, "capi_check.c", "gcc", "c99")
    disc = O.golden("euler_2d_disc64_l0")
    assert disc["meta"]["dx"] == 1.0 / 64 and disc["meta"]["max_steps"] == 200
    np.ascontiguousarray(disc["prim_in"], dtype=np.float64).tofile(os.path.join(tmp, "disc_in.f64"))
    np.ascontiguousarray(disc["dt"], dtype=np.float64).tofile(os.path.join(tmp, "disc_dt.f64"))
    r = subprocess.run([exe, tmp], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    sod = O.golden("euler_c1_sod1000")
    O.assert_bitwise(np.fromfile(os.path.join(tmp, "sod_dt.f64")), sod["dt"], "C program: Sod time steps")
    O.assert_bitwise(load(tmp, "sod_solve.f64", (1000, 6)), sod["prim_out"], "C program: vg_euler_solve_1d")
    O.assert_bitwise(load(tmp, "sod_steps.f64", (1000, 6)), sod["prim_out"], "C program: vg_euler_time_step_1d")
    for name in ("disc_solve", "disc_steps", "disc_state"):
        O.assert_bitwise(load(tmp, name + ".f64", (64, 64, 6)), disc["prim_out"], f"C program: {name}")


@pytest.mark.gpu
def test_compat_driver_reproduces_reference(tmp_path):
    tmp = str(tmp_path)
    exe = build(tmp, "compat_driver.cpp", "g++", "c++11")
    r = subprocess.run([exe, tmp], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    sod = O.golden("euler_c1_sod1000")
    O.assert_bitwise(load(tmp, "compat_solve.f64", (1000, 6)), sod["prim_out"], "SecondOrderSolver::solve")
    O.assert_bitwise(load(tmp, "compat_loop.f64", (1000, 6)), sod["prim_out"], "per-sweep solve loop")
