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
    lib.orc_mhd_sweep.restype = None
    lib.orc_mhd_sweep.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                  C.c_int, C.c_int, C.POINTER(OrcMHDParams), dptr, dptr]
    lib.orc_euler_sweep.restype = None
    lib.orc_euler_sweep.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                    C.c_int, C.c_int, C.POINTER(OrcEulerParams), dptr, dptr]
    lib.orc_euler_max_wave_speed.restype = C.c_double
    lib.orc_euler_max_wave_speed.argtypes = [C.c_int, C.c_int, C.c_int, C.POINTER(OrcEulerParams), dptr]
    lib.orc_mhd_solve.restype = C.c_int
    lib.orc_mhd_solve.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int,
                                  C.c_double, C.c_int, C.c_int, C.c_int, C.POINTER(OrcMHDParams), dptr, dptr, dptr,
                                  C.c_int]
    lib.orc_mhd_units.restype = None
    lib.orc_mhd_units.argtypes = [C.c_int, C.c_double, C.c_int, C.c_double, C.c_double, C.POINTER(OrcMHDParams),
                                  dptr, dptr]
    lib.orc_hpr_solve.restype = C.c_int
    lib.orc_hpr_solve.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int,
                                  C.c_double, C.c_int, C.c_int, C.c_int, C.POINTER(OrcHPRParams), dptr, dptr, dptr,
                                  C.c_int]
    lib.orc_hpr_eos_units.restype = None
    lib.orc_hpr_eos_units.argtypes = [C.c_int, C.POINTER(OrcHPRParams), dptr, dptr]
    lib.orc_hpr_units.restype = None
    lib.orc_hpr_units.argtypes = [C.c_int, C.c_double, C.c_int, C.c_double, C.c_double, C.POINTER(OrcHPRParams),
                                  dptr, dptr]
    P = C.POINTER(OrcEulerParams)
    lib.orc_rgfm_solve.restype = C.c_int
    lib.orc_rgfm_solve.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                   C.c_int, C.c_double, C.c_int, P, P, dptr, dptr, dptr, dptr, dptr, dptr, dptr,
                                   C.c_int]
    lib.orc_rgfm_ghost_fill.restype = None
    lib.orc_rgfm_ghost_fill.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, P, P, dptr, dptr, dptr]
    lib.orc_rgfm_level_set.restype = None
    lib.orc_rgfm_level_set.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, dptr, dptr, dptr,
                                       dptr]
    lib.orc_hllc_units.restype = None
    lib.orc_hllc_units.argtypes = [C.c_int, P, P, dptr, dptr]
    lib.orc_exact_units.restype = None
    lib.orc_exact_units.argtypes = [C.c_int, C.c_int, C.c_double, P, P, dptr, dptr]
    lib.orc_euler_fixed_dt_jvp.restype = None
    lib.orc_euler_fixed_dt_jvp.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, dptr, C.c_int, C.c_double, C.c_int,
                                           C.c_int, C.POINTER(OrcEulerParams), dptr, dptr, dptr, dptr]
    lib.orc_rgfm_fixed_dt_jvp.restype = None
    lib.orc_rgfm_fixed_dt_jvp.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, dptr, C.c_int, C.c_double,
                                          C.c_int, C.POINTER(OrcEulerParams), C.POINTER(OrcEulerParams), dptr, dptr,
                                          dptr, dptr, dptr, dptr, dptr, dptr, dptr, dptr]
    HP = C.POINTER(OrcHPRParams)
    lib.orc_hpr_fixed_dt_jvp.restype = None
    lib.orc_hpr_fixed_dt_jvp.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, dptr, C.c_int, C.c_double, C.c_int,
                                         C.c_int, C.c_int, HP, dptr, dptr, C.c_double, C.c_double, dptr, dptr]
    lib.orc_hpr_jvp_units.restype = None
    lib.orc_hpr_jvp_units.argtypes = [C.c_int, HP, C.c_double, C.c_double, dptr, dptr, dptr]
    lib.orc_euler_branch_hash.restype = C.c_uint64
    lib.orc_euler_branch_hash.argtypes = []
    lib.orc_rgfm_circulation.restype = C.c_double
    lib.orc_rgfm_circulation.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, dptr, dptr]
    mq = C.POINTER(OrcMHDParams)
    lib.orc_mhd_rgfm_solve.restype = C.c_int
    lib.orc_mhd_rgfm_solve.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int,
                                       C.c_double, C.c_int, C.c_int, mq, mq, dptr, dptr, dptr, dptr, dptr, dptr, dptr,
                                       C.c_int]
    lib.orc_mhd_hllc_units.restype = None
    lib.orc_mhd_hllc_units.argtypes = [C.c_int, mq, mq, dptr, dptr]
    lib.orc_mhd_rgfm_ghost_fill.restype = None
    lib.orc_mhd_rgfm_ghost_fill.argtypes = [C.c_int, C.c_int, C.c_int, mq, mq, dptr, dptr, dptr]
    ep = C.POINTER(OrcEulerParams)
    lib.orc_mp_solve.restype = C.c_int
    lib.orc_mp_solve.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                 C.c_int, C.c_double, C.c_int, C.c_int, C.c_int, ep, ep, dptr, dptr, dptr, C.c_int]
    lib.orc_mp_sweep.restype = None
    lib.orc_mp_sweep.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                 C.c_double, C.c_int, C.c_int, ep, ep, dptr, dptr]
    lib.orc_mp_reinit.restype = None
    lib.orc_mp_reinit.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, dptr]
    lib.orc_mp_units.restype = None
    lib.orc_mp_units.argtypes = [C.c_int, C.c_int, ep, ep, C.c_double, C.c_int, C.c_double, C.c_double, dptr, dptr]
    _lib = lib
    return lib


def rgfm_solve(m1, m2, phi, *, dim, dx, cfl, tfinal, max_steps, bias, limiter, p1, p2, dy=None, riemann=0,
               dt_cap=200000):
    lib = load()
    a1 = np.ascontiguousarray(m1, dtype=np.float64)
    a2 = np.ascontiguousarray(m2, dtype=np.float64)
    ph = np.ascontiguousarray(phi, dtype=np.float64)
    o1, o2, oph = np.empty_like(a1), np.empty_like(a2), np.empty_like(ph)
    dts = np.empty(dt_cap)
    ny, nx = (1, ph.shape[0]) if dim == 1 else ph.shape
    q1, q2 = OrcEulerParams(*p1), OrcEulerParams(*p2)
    n = lib.orc_rgfm_solve(riemann, dim, ny, nx, dx, dx if dy is None else dy, cfl, tfinal, max_steps, bias, limiter,
                           C.byref(q1), C.byref(q2), _p(a1), _p(a2), _p(ph), _p(o1), _p(o2), _p(oph), _p(dts), dt_cap)
    return o1, o2, oph, dts[:min(n, dt_cap)].copy()


def hllc_units(states, p1, p2):
    lib = load()
    s = np.ascontiguousarray(states, dtype=np.float64)
    out = np.empty((s.shape[0], 24))
    q1, q2 = OrcEulerParams(*p1), OrcEulerParams(*p2)
    lib.orc_hllc_units(s.shape[0], C.byref(q1), C.byref(q2), _p(s), _p(out))
    return out


def exact_units(states, p1, p2, s=0.0, direction=0):
    lib = load()
    st = np.ascontiguousarray(states, dtype=np.float64)
    out = np.empty((st.shape[0], 10))
    q1, q2 = OrcEulerParams(*p1), OrcEulerParams(*p2)
    lib.orc_exact_units(st.shape[0], direction, s, C.byref(q1), C.byref(q2), _p(st), _p(out))
    return out


def rgfm_level_set(dim, direction, phi, m1, m2, h, dt):
    lib = load()
    ph = np.ascontiguousarray(phi, dtype=np.float64)
    out = np.empty_like(ph)
    ny, nx = (1, ph.shape[0]) if dim == 1 else ph.shape
    lib.orc_rgfm_level_set(dim, direction, ny, nx, h, dt, _p(np.ascontiguousarray(m1, dtype=np.float64)),
                           _p(np.ascontiguousarray(m2, dtype=np.float64)), _p(ph), _p(out))
    return out


def rgfm_circulation(m1, phi, dx, dy):
    lib = load()
    a = np.ascontiguousarray(m1, dtype=np.float64)
    ph = np.ascontiguousarray(phi, dtype=np.float64)
    return lib.orc_rgfm_circulation(ph.shape[0], ph.shape[1], dx, dy, _p(a), _p(ph))


class OrcHPRParams(C.Structure):
    _fields_ = [("is_thermal", C.c_int32), ("is_plastic", C.c_int32)] + [
        (n, C.c_double) for n in ("rho0", "t_ref", "t0", "cv", "gamma", "p_inf", "cs", "tau1", "tau2", "alpha")] + [
        ("eos", C.c_int32), ("reserved", C.c_int32)] + [
        (n, C.c_double) for n in ("p_ref", "c0", "gamma0", "s_h", "alpha_gr", "sigma_y", "n_pow")]


def hpr_params(thermal, *, rho0=1.0, p0=1.0 / 1.4, t_ref=0.0, t0=1.0, cv=2.5, gamma=1.4, p_inf=0.0, cs=1.0, mu=1e-2,
               alpha=2.0, kappa=1e-2):
    """HPR parameters with the relaxation times the reference's
    configureStrainDissipationTime / configureThermalImpulseRelaxationTime set
    (hprmaterialparameters.cpp:101-119): tau1 = 6 mu / (p0 cs^2),
    tau2 = kappa rho0 / (T0 alpha^2)."""
    tau1 = 6.0 * (mu / (p0 * (cs * cs)))
    tau2 = kappa * (rho0 / (t0 * (alpha * alpha))) if thermal else 0.0
    return dict(is_thermal=int(thermal), rho0=rho0, p0=p0, t_ref=t_ref, t0=t0, cv=cv, gamma=gamma, p_inf=p_inf, cs=cs,
                tau1=tau1, tau2=tau2, alpha=alpha)


def _hpr_struct(p):
    return OrcHPRParams(int(p["is_thermal"]), int(p.get("is_plastic", 0)), p["rho0"], p["t_ref"], p["t0"], p["cv"],
                        p["gamma"], p["p_inf"], p["cs"], p["tau1"], p["tau2"], p["alpha"], int(p.get("eos", 0)), 0,
                        p.get("p_ref", 0.0), p.get("c0", 0.0), p.get("gamma0", 0.0), p.get("s_h", 0.0),
                        p.get("alpha_gr", 0.0), p.get("sigma_y", 0.0), p.get("n_pow", 0.0))


# The 26 values of the reference's full HPRMaterialParameters constructor as the
# harness takes them (EOS code, is_thermal, is_plastic, then the 23 doubles in
# order, hprmaterialparameters.h:12-16) -> the oracle's parameter dictionary.
HPR_FULL_FIELDS = ("eos", "is_thermal", "is_plastic", "ideal_gas_constant", "rho0", "p_ref", "t_ref", "t0", "cv",
                   "gamma", "p_inf", "c0", "gamma0", "s_h", "e0", "cs", "tau1", "tau2", "mu", "prandtl", "sigma_y",
                   "n_pow", "alpha", "alpha_gr", "beta", "gamma_gr")


def hpr_params_from_full(v):
    d = dict(zip(HPR_FULL_FIELDS, [float(x) for x in v]))
    for k in ("eos", "is_thermal", "is_plastic"):
        d[k] = int(d[k])
    return d


def hpr_eos_units(states, params):
    lib = load()
    st = np.ascontiguousarray(states, dtype=np.float64)
    out = np.empty((st.shape[0], 8))
    q = _hpr_struct(params)
    lib.orc_hpr_eos_units(st.shape[0], C.byref(q), _p(st), _p(out))
    return out


def hpr_solve(prim, *, order, dim, dx, cfl, tfinal, max_steps, bias, limiter, subcycles, params, bc=0,
              dt_cap=100000):
    lib = load()
    w = np.ascontiguousarray(prim, dtype=np.float64)
    out = np.empty_like(w)
    dts = np.empty(dt_cap)
    ny, nx = (1, w.shape[0]) if dim == 1 else w.shape[:2]
    mp = _hpr_struct(params)
    n = lib.orc_hpr_solve(order, dim, ny, nx, dx, cfl, tfinal, max_steps, bias, limiter, bc, subcycles, C.byref(mp),
                          _p(w), _p(out), _p(dts), dt_cap)
    return out, dts[:min(n, dt_cap)].copy()


def hpr_fixed_dt(prim, dts, *, dim, dx, bias, limiter, subcycles, params, bc=0, tangent=None, dtau=(0.0, 0.0)):
    """HPR second-order solve with prescribed time steps (dual-number
    restatement); with `tangent` (and `dtau` = tangents of tau1, tau2), also the
    forward-mode derivative of the final state.  Sets the branch hash."""
    lib = load()
    w = np.ascontiguousarray(prim, dtype=np.float64)
    d = np.ascontiguousarray(dts, dtype=np.float64)
    out = np.empty_like(w)
    ny, nx = (1, w.shape[0]) if dim == 1 else w.shape[:2]
    mp = _hpr_struct(params)
    if tangent is None:
        lib.orc_hpr_fixed_dt_jvp(dim, ny, nx, dx, _p(d), d.size, bias, limiter, bc, subcycles, C.byref(mp), _p(w),
                                 None, 0.0, 0.0, _p(out), None)
        return out
    tv = np.ascontiguousarray(tangent, dtype=np.float64)
    assert tv.shape == w.shape
    tout = np.empty_like(w)
    lib.orc_hpr_fixed_dt_jvp(dim, ny, nx, dx, _p(d), d.size, bias, limiter, bc, subcycles, C.byref(mp), _p(w),
                             _p(tv), float(dtau[0]), float(dtau[1]), _p(out), _p(tout))
    return out, tout


def hpr_jvp_units(w, v, *, params, dtau=(0.0, 0.0)):
    """Per record: JVPs of p2c at w, and of c2p, Fx, Fy, S at U = p2c(w) along
    v (and dtau) -> (n, 5, 17)."""
    lib = load()
    a = np.ascontiguousarray(w, dtype=np.float64)
    b = np.ascontiguousarray(v, dtype=np.float64)
    out = np.empty((a.shape[0], 5, 17))
    mp = _hpr_struct(params)
    lib.orc_hpr_jvp_units(a.shape[0], C.byref(mp), float(dtau[0]), float(dtau[1]), _p(a), _p(b), _p(out))
    return out


def hpr_units(states, *, bias, limiter, dx, ts, params):
    lib = load()
    s = np.ascontiguousarray(states, dtype=np.float64)
    out = np.empty((s.shape[0], 156))
    mp = _hpr_struct(params)
    lib.orc_hpr_units(s.shape[0], bias, limiter, dx, ts, C.byref(mp), _p(s), _p(out))
    return out


class OrcMHDParams(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("gamma", "p_inf", "ch", "cp")]


def mhd_solve(prim, *, order, dim, dx, cfl, tfinal, max_steps, bias, limiter, gamma, p_inf=0.0, bc=0,
              dt_cap=200000, subcycles=0, ch=0.0, cp=0.0):
    lib = load()
    w = np.ascontiguousarray(prim, dtype=np.float64)
    out = np.empty_like(w)
    dts = np.empty(dt_cap)
    ny, nx = (1, w.shape[0]) if dim == 1 else w.shape[:2]
    mp = OrcMHDParams(gamma, p_inf, ch, cp)
    n = lib.orc_mhd_solve(order, dim, ny, nx, dx, cfl, tfinal, max_steps, bias, limiter, bc, subcycles, C.byref(mp),
                          _p(w), _p(out), _p(dts), dt_cap)
    return out, dts[:min(n, dt_cap)].copy()


def mhd_sweep(prim, *, order, dim, direction, dx, ts, bias, limiter, gamma, p_inf, ch, cp, bc=0):
    """One MHD sweep with c_h, c_p as given (AoS in, AoS out)."""
    lib = load()
    w = np.ascontiguousarray(prim, dtype=np.float64)
    out = np.empty_like(w)
    ny, nx = (1, w.shape[0]) if dim == 1 else w.shape[:2]
    mp = OrcMHDParams(gamma, p_inf, ch, cp)
    lib.orc_mhd_sweep(order, dim, direction, ny, nx, dx, ts, bias, limiter, bc, C.byref(mp), _p(w), _p(out))
    return out


def euler_sweep(prim, *, order, dim, direction, dx, ts, bias, limiter, params, bc=0):
    """One Euler sweep (AoS in, AoS out)."""
    lib = load()
    w = np.ascontiguousarray(prim, dtype=np.float64)
    out = np.empty_like(w)
    ny, nx = (1, w.shape[0]) if dim == 1 else w.shape[:2]
    mp = OrcEulerParams(*params)
    lib.orc_euler_sweep(order, dim, direction, ny, nx, dx, ts, bias, limiter, bc, C.byref(mp), _p(w), _p(out))
    return out


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


def euler_fixed_dt(prim, dts, *, dim, dx, bias, limiter, params, bc=0, tangent=None):
    """SLIC solve with prescribed time steps; with `tangent`, also the
    forward-mode derivative of the final state along it (dual numbers)."""
    lib = load()
    w = np.ascontiguousarray(prim, dtype=np.float64)
    d = np.ascontiguousarray(dts, dtype=np.float64)
    out = np.empty_like(w)
    ny, nx = (1, w.shape[0]) if dim == 1 else w.shape[:2]
    mp = OrcEulerParams(*params)
    if tangent is None:
        lib.orc_euler_fixed_dt_jvp(dim, ny, nx, dx, _p(d), d.size, bias, limiter, bc, C.byref(mp), _p(w), None,
                                   _p(out), None)
        return out
    tv = np.ascontiguousarray(tangent, dtype=np.float64)
    assert tv.shape == w.shape
    tout = np.empty_like(w)
    lib.orc_euler_fixed_dt_jvp(dim, ny, nx, dx, _p(d), d.size, bias, limiter, bc, C.byref(mp), _p(w), _p(tv),
                               _p(out), _p(tout))
    return out, tout


def rgfm_fixed_dt(m1, m2, phi, dts, *, dim, dx, bias, limiter, p1, p2, dy=None, tangent=None):
    """HLLC RGFM solve with prescribed time steps; with tangent = (t1, t2), also
    the forward-mode derivative of both final material states along it.
    Returns (m1, m2, phi[, t1, t2]) and sets the branch hash."""
    lib = load()
    a1 = np.ascontiguousarray(m1, dtype=np.float64)
    a2 = np.ascontiguousarray(m2, dtype=np.float64)
    ph = np.ascontiguousarray(phi, dtype=np.float64)
    d = np.ascontiguousarray(dts, dtype=np.float64)
    o1, o2, oph = np.empty_like(a1), np.empty_like(a2), np.empty_like(ph)
    ny, nx = (1, ph.shape[0]) if dim == 1 else ph.shape
    q1, q2 = OrcEulerParams(*p1), OrcEulerParams(*p2)
    dyv = dx if dy is None else dy
    if tangent is None:
        lib.orc_rgfm_fixed_dt_jvp(dim, ny, nx, dx, dyv, _p(d), d.size, bias, limiter, C.byref(q1), C.byref(q2), _p(a1),
                                  _p(a2), _p(ph), None, None, _p(o1), _p(o2), _p(oph), None, None)
        return o1, o2, oph
    t1 = np.ascontiguousarray(tangent[0], dtype=np.float64)
    t2 = np.ascontiguousarray(tangent[1], dtype=np.float64)
    u1, u2 = np.empty_like(a1), np.empty_like(a2)
    lib.orc_rgfm_fixed_dt_jvp(dim, ny, nx, dx, dyv, _p(d), d.size, bias, limiter, C.byref(q1), C.byref(q2), _p(a1),
                              _p(a2), _p(ph), _p(t1), _p(t2), _p(o1), _p(o2), _p(oph), _p(u1), _p(u2))
    return o1, o2, oph, u1, u2


def branch_hash():
    return int(load().orc_euler_branch_hash())


def euler_branch_hash(prim, dts, **kw):
    """Hash of the branch decisions the fixed-dt solve takes from `prim`."""
    euler_fixed_dt(prim, dts, **kw)
    return int(load().orc_euler_branch_hash())


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


def mhd_rgfm_solve(m1, m2, phi, *, dim, dx, cfl, tfinal, max_steps, bias, limiter, subcycles, p1, p2,
                   dt_cap=200000):
    """p1, p2: (gamma, p_inf, c_h, c_p) of each material."""
    lib = load()
    a1 = np.ascontiguousarray(m1, dtype=np.float64)
    a2 = np.ascontiguousarray(m2, dtype=np.float64)
    ph = np.ascontiguousarray(phi, dtype=np.float64)
    o1, o2, oph = np.empty_like(a1), np.empty_like(a2), np.empty_like(ph)
    dts = np.empty(dt_cap)
    ny, nx = (1, ph.shape[0]) if dim == 1 else ph.shape
    n = lib.orc_mhd_rgfm_solve(dim, ny, nx, dx, cfl, tfinal, max_steps, bias, limiter, subcycles,
                               C.byref(OrcMHDParams(*p1)), C.byref(OrcMHDParams(*p2)), _p(a1), _p(a2), _p(ph),
                               _p(o1), _p(o2), _p(oph), _p(dts), dt_cap)
    return o1, o2, oph, dts[:min(n, dt_cap)].copy()


def mhd_rgfm_ghost_fill(m1, m2, phi, p1, p2):
    """applyRGFMBoundaryConditions[2D] on copies of m1, m2 (dim from phi)."""
    lib = load()
    a1 = np.array(m1, dtype=np.float64, order="C")
    a2 = np.array(m2, dtype=np.float64, order="C")
    ph = np.ascontiguousarray(phi, dtype=np.float64)
    ny, nx = (1, ph.shape[0]) if ph.ndim == 1 else ph.shape
    lib.orc_mhd_rgfm_ghost_fill(ph.ndim, ny, nx, C.byref(OrcMHDParams(*p1)), C.byref(OrcMHDParams(*p2)), _p(a1),
                                _p(a2), _p(ph))
    return a1, a2


def mhd_hllc_units(states, p1, p2):
    lib = load()
    st = np.ascontiguousarray(states, dtype=np.float64)
    out = np.empty((st.shape[0], 24))
    lib.orc_mhd_hllc_units(st.shape[0], C.byref(OrcMHDParams(*p1)), C.byref(OrcMHDParams(*p2)), _p(st), _p(out))
    return out


# The two-material Euler models: kind 0 multiphysics (8 primitives: alpha1, u,
# v, w, rho1, p1, rho2, p2; 9 conserved), kind 1 reduced (7: alpha1, u, v, w,
# rho1, rho2, p; 8 conserved).
MP_NP = (8, 7)
MP_NF = (9, 8)


def mp_solve(kind, prim, *, order, dim, dx, cfl, tfinal, max_steps, bias, limiter, reinit, p1, p2, bc=0,
             dt_cap=100000):
    lib = load()
    w = np.ascontiguousarray(prim, dtype=np.float64)
    out = np.empty_like(w)
    dts = np.empty(dt_cap)
    ny, nx = (1, w.shape[0]) if dim == 1 else w.shape[:2]
    n = lib.orc_mp_solve(kind, order, dim, ny, nx, dx, cfl, tfinal, max_steps, bias, limiter, bc, reinit,
                         C.byref(OrcEulerParams(*p1)), C.byref(OrcEulerParams(*p2)), _p(w), _p(out), _p(dts), dt_cap)
    return out, dts[:min(n, dt_cap)].copy()


def mp_sweep(kind, prim, *, order, dim, direction, dx, ts, bias, limiter, p1, p2, bc=0):
    lib = load()
    w = np.ascontiguousarray(prim, dtype=np.float64)
    out = np.empty_like(w)
    ny, nx = (1, w.shape[0]) if dim == 1 else w.shape[:2]
    lib.orc_mp_sweep(kind, order, dim, direction, ny, nx, dx, ts, bias, limiter, bc, C.byref(OrcEulerParams(*p1)),
                     C.byref(OrcEulerParams(*p2)), _p(w), _p(out))
    return out


def mp_reinit(kind, prim, dim):
    lib = load()
    w = np.array(prim, dtype=np.float64, order="C")
    ny, nx = (1, w.shape[0]) if dim == 1 else w.shape[:2]
    lib.orc_mp_reinit(kind, dim, ny, nx, _p(w))
    return w


def mp_units(kind, states, *, bias, limiter, dx, ts, p1, p2):
    lib = load()
    st = np.ascontiguousarray(states, dtype=np.float64)
    npf, nf = MP_NP[kind], MP_NF[kind]
    n = st.shape[0]
    out = np.empty((n, 7 * nf + 3 * npf + 2))
    lib.orc_mp_units(kind, n, C.byref(OrcEulerParams(*p1)), C.byref(OrcEulerParams(*p2)), bias, limiter, dx, ts,
                     _p(st), _p(out))
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
