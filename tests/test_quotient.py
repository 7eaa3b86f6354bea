"""The shared-reciprocal quotient of the EXACT Euler conversions
(vanguard_amd/csrc/vg_quot.h) equals IEEE division bit for bit: the header is
compiled on the host (with hardware FMA) and checked over random operands in
its documented range, quotients next to rounding midpoints, exact quotients,
divisor mantissas next to all-ones and powers of two, signed zeros and the
fall-back divisors (tests/quotient_check.cpp)."""

import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("compiler", ["g++", "/opt/rocm/llvm/bin/clang++"])
def test_shared_reciprocal_quotient_is_ieee_division(tmp_path, compiler):
    if shutil.which(compiler) is None and not os.path.exists(compiler):
        pytest.skip(f"{compiler} not installed")
    exe = str(tmp_path / "quotient_check")
    r = subprocess.run([compiler, "-O2", "-mfma", "-std=c++17", "-ffp-contract=off", "-I",
                        os.path.join(ROOT, "vanguard_amd", "csrc"), os.path.join(ROOT, "tests", "quotient_check.cpp"),
                        "-o", exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([exe, "3000000"], capture_output=True, text=True, timeout=600)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert "mismatches 0" in r.stdout
