"""ctypes loader for the CPU oracle (oracle/_build/liboracle.so).

Test infrastructure: the oracle is the checker, never the thing under test.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB = os.path.join(ORACLE_DIR, "_build", "liboracle.so")
GOLDEN = os.path.join(ROOT, "tests", "golden")

dptr = C.POINTER(C.c_double)


class OrcEulerParams(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("gamma", "p_inf", "cv", "e_form", "t_ign", "k_rate")]


_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        subprocess.run(["make", "-s", "-C", ORACLE_DIR, "oracle"], check=True)
    lib = C.CDLL(LIB)
    lib.orc_euler_solve.restype = C.c_int
    lib.orc_euler_solve.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int,
                                    C.c_double, C.c_int, C.c_int, C.POINTER(OrcEulerParams), dptr, dptr, dptr, C.c_int]
    lib.orc_euler_units.restype = None
    lib.orc_euler_units.argtypes = [C.c_int, C.c_double, C.c_int, C.c_double, C.c_double, C.POINTER(OrcEulerParams),
                                    dptr, dptr]
    lib.orc_limiter_units.restype = None
    lib.orc_limiter_units.argtypes = [C.c_int, C.c_double, dptr, dptr]
    lib.orc_euler_sweep.restype = None
    lib.orc_euler_sweep.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                    C.c_int, C.c_int, C.POINTER(OrcEulerParams), dptr, dptr]
    lib.orc_euler_max_wave_speed.restype = C.c_double
    lib.orc_euler_max_wave_speed.argtypes = [C.c_int, C.c_int, C.c_int, C.POINTER(OrcEulerParams), dptr]
    lib.orc_mhd_solve.restype = C.c_int
    lib.orc_mhd_solve.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int,
                                  C.c_double, C.c_int, C.c_int, C.POINTER(OrcMHDParams), dptr, dptr, dptr, C.c_int]
    lib.orc_mhd_units.restype = None
    lib.orc_mhd_units.argtypes = [C.c_int, C.c_double, C.c_int, C.c_double, C.c_double, C.POINTER(OrcMHDParams),
                                  dptr, dptr]
    _lib = lib
    return lib


class OrcMHDParams(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("gamma", "p_inf", "ch", "cp")]


def mhd_solve(prim, *, order, dim, dx, cfl, tfinal, max_steps, bias, limiter, gamma, p_inf=0.0, bc=0,
              dt_cap=200000):
    lib = load()
    w = np.ascontiguousarray(prim, dtype=np.float64)
    out = np.empty_like(w)
    dts = np.empty(dt_cap)
    ny, nx = (1, w.shape[0]) if dim == 1 else w.shape[:2]
    mp = OrcMHDParams(gamma, p_inf, 0.0, 0.0)
    n = lib.orc_mhd_solve(order, dim, ny, nx, dx, cfl, tfinal, max_steps, bias, limiter, bc, C.byref(mp), _p(w),
                          _p(out), _p(dts), dt_cap)
    return out, dts[:min(n, dt_cap)].copy()


def mhd_units(states, *, bias, limiter, dx, ts, gamma, p_inf, ch):
    lib = load()
    s = np.ascontiguousarray(states, dtype=np.float64)
    out = np.empty((s.shape[0], 85))
    mp = OrcMHDParams(gamma, p_inf, ch, 0.0)
    lib.orc_mhd_units(s.shape[0], bias, limiter, dx, ts, C.byref(mp), _p(s), _p(out))
    return out


def seq_sum(x):
    """Left-to-right floating-point sum (the solve loop's t += dt)."""
    t = 0.0
    for v in x:
        t += float(v)
    return t


def _p(a):
    return a.ctypes.data_as(dptr)


def euler_solve(prim, *, order, dim, dx, cfl, tfinal, max_steps, bias, limiter, params, bc=0, dt_cap=100000):
    lib = load()
    w = np.ascontiguousarray(prim, dtype=np.float64)
    out = np.empty_like(w)
    dts = np.empty(dt_cap)
    ny, nx = (1, w.shape[0]) if dim == 1 else w.shape[:2]
    mp = OrcEulerParams(*params)
    n = lib.orc_euler_solve(order, dim, ny, nx, dx, cfl, tfinal, max_steps, bias, limiter, bc, C.byref(mp), _p(w),
                            _p(out), _p(dts), dt_cap)
    return out, dts[:min(n, dt_cap)].copy()


def euler_units(states, *, bias, limiter, dx, ts, params):
    lib = load()
    s = np.ascontiguousarray(states, dtype=np.float64)
    n = s.shape[0]
    out = np.empty((n, 92))
    mp = OrcEulerParams(*params)
    lib.orc_euler_units(n, bias, limiter, dx, ts, C.byref(mp), _p(s), _p(out))
    return out


def limiter_units(r, bias):
    lib = load()
    r = np.ascontiguousarray(r, dtype=np.float64)
    out = np.empty((r.size, 5))
    lib.orc_limiter_units(r.size, bias, _p(r), _p(out))
    return out


def golden(name):
    """Loads a golden fixture (.npz of plain arrays; no pickles)."""
    import json
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as z:
        d = {k: z[k] for k in z.files}
    if "meta" in d:
        d["meta"] = json.loads(bytes(d["meta"]).decode())
    return d


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def assert_bitwise(actual, expected, what=""):
    a, e = bits(actual), bits(expected)
    assert a.shape == e.shape, f"{what}: shape {a.shape} != {e.shape}"
    diff = a != e
    if diff.any():
        idx = np.argwhere(diff)[0]
        raise AssertionError(f"{what}: {int(diff.sum())} of {diff.size} values differ; first at {tuple(idx)}: "
                             f"{np.asarray(actual)[tuple(idx)]!r} vs {np.asarray(expected)[tuple(idx)]!r}")
