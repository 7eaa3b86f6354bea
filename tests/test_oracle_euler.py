"""Pins the Euler oracle (oracle/euler_oracle.c).

1. Bit for bit against golden vectors written by the compiled reference
   (tests/golden/gen_golden.py), whole solves and per-function records.
2. Against what the mathematics fixes where a golden vector cannot speak:
   the exact Riemann solution of Sod's problem, conservation, limiter closed
   forms, EOS identities, translation invariance under periodic boundaries.
"""

import math

import numpy as np
import pytest

import oracle_lib as O

RUNS = ["euler_c1_sod1000", "euler_toro100", "euler_1d_stiff_bias", "euler_force1d", "euler_2d_disc64_l0",
        "euler_2d_disc64_l1", "euler_2d_disc64_l2", "euler_2d_c2ic64", "euler_2d_rect", "euler_force2d",
        "euler_2d_api32"]


@pytest.mark.parametrize("name", RUNS)
def test_oracle_matches_reference_run(name):
    g = O.golden(name)
    m = g["meta"]
    out, dts = O.euler_solve(g["prim_in"], order=m["order"], dim=m["dim"], dx=m["dx"], cfl=m["cfl"],
                             tfinal=m["tfinal"], max_steps=m["max_steps"], bias=m["bias"], limiter=m["limiter"],
                             params=m["params"])
    O.assert_bitwise(dts, g["dt"], f"{name} dt sequence")
    O.assert_bitwise(out, g["prim_out"], f"{name} final state")


UNIT_FIELDS = [("slope", 0, 6), ("flux_x", 6, 12), ("flux_y", 12, 18), ("evolve_x", 18, 30), ("evolve_y", 30, 42),
               ("force_x", 42, 48), ("force_y", 48, 54), ("slic_x", 54, 60), ("slic_y", 60, 66),
               ("wave_speed", 66, 68), ("cons", 68, 74), ("cons_roundtrip", 74, 80), ("lax_friedrichs", 80, 86),
               ("richtmyer", 86, 92)]


@pytest.mark.parametrize("name", ["euler_units_ideal", "euler_units_vl_bias", "euler_units_minbee"])
def test_oracle_matches_reference_functions(name):
    g = O.golden(name)
    m = g["meta"]
    out = O.euler_units(g["states"], bias=m["bias"], limiter=m["limiter"], dx=m["dx"], ts=m["ts"], params=m["params"])
    for field, a, b in UNIT_FIELDS:
        O.assert_bitwise(out[:, a:b], g["out"][:, a:b], f"{name}:{field}")


@pytest.mark.parametrize("name", ["limiter_units_b0", "limiter_units_b03"])
def test_oracle_limiters_match_reference(name):
    g = O.golden(name)
    O.assert_bitwise(O.limiter_units(g["r"], float(g["bias"])), g["out"], name)


def test_golden_units_reach_every_limiter_branch():
    """The unit fixtures are only a pin if they exercise the branches: the slope
    ratio must hit r < 0, [0, 0.5), [0.5, 1), >= 1 and the 1e-5 clamp."""
    g = O.golden("euler_units_ideal")
    s = g["states"].reshape(-1, 5, 6)
    mp = g["meta"]["params"]

    def cons(w):
        rho, u, v, z, p, lam = w.T
        e = (p + mp[0] * mp[1]) / ((mp[0] - 1) * rho)
        return np.stack([rho, rho * u, rho * v, rho * z, rho * (0.5 * (u * u + v * v + z * z) + e), rho * lam], 1)

    dl = cons(s[:, 1]) - cons(s[:, 0])
    dr = cons(s[:, 2]) - cons(s[:, 1])
    num = np.where(np.abs(dl) < 1e-5, 1e-5, dl)
    den = np.where(np.abs(dr) < 1e-5, 1e-5, dr)
    r = (num / den).ravel()
    assert (np.abs(dl) < 1e-5).any() and (np.abs(dr) < 1e-5).any()
    for lo, hi in [(-np.inf, 0), (0, 0.5), (0.5, 1.0), (1.0, np.inf)]:
        assert ((r >= lo) & (r < hi)).sum() > 20, (lo, hi)


# ---- closed forms -----------------------------------------------------------

def exact_sod_star():
    """Star state of Sod's problem from the exact Riemann solver (Toro, ch. 4),
    solved here by bisection on the pressure function."""
    g = 1.4
    rl, ul, pl = 1.0, 0.0, 1.0
    rr, ur, pr = 0.125, 0.0, 0.1

    def fk(p, rk, pk):
        ck = math.sqrt(g * pk / rk)
        if p > pk:
            A = 2 / ((g + 1) * rk)
            B = (g - 1) / (g + 1) * pk
            return (p - pk) * math.sqrt(A / (p + B))
        return 2 * ck / (g - 1) * ((p / pk) ** ((g - 1) / (2 * g)) - 1)

    lo, hi = 1e-6, 10.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if fk(mid, rl, pl) + fk(mid, rr, pr) + (ur - ul) > 0:
            hi = mid
        else:
            lo = mid
    ps = 0.5 * (lo + hi)
    us = 0.5 * (ul + ur) + 0.5 * (fk(ps, rr, pr) - fk(ps, rl, pl))
    rsl = rl * (ps / pl) ** (1 / g)
    rsr = rr * ((ps / pr + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * ps / pr + 1))
    return ps, us, rsl, rsr


def test_sod_star_state_matches_exact_solution():
    g = O.golden("euler_c1_sod1000")
    out, _ = O.euler_solve(g["prim_in"], order=2, dim=1, dx=1e-3, cfl=0.9, tfinal=0.25, max_steps=0, bias=0.0,
                           limiter=2, params=[1.4, 0, 25, 0, 0.25, 250])
    ps, us, rsl, rsr = exact_sod_star()
    assert abs(ps - 0.30313) < 1e-4 and abs(us - 0.92745) < 1e-4
    # contact at x = 0.5 + u* t = 0.7319; left star plateau around x = 0.6, right star around x = 0.8
    x = (np.arange(1000) + 0.5) / 1000
    lp = (x > 0.58) & (x < 0.70)
    rp = (x > 0.76) & (x < 0.84)
    assert np.abs(out[lp, 4] - ps).max() < 2e-3
    assert np.abs(out[lp, 1] - us).max() < 5e-3
    assert np.abs(out[lp, 0] - rsl).max() < 3e-3
    assert np.abs(out[rp, 0] - rsr).max() < 5e-3


def test_conservation_1d_sod():
    """Mass and energy are conserved while the waves are inside the domain; the
    momentum changes by (p_left - p_right) t through the transmissive ends."""
    w0 = O.golden("euler_c1_sod1000")["prim_in"]
    out, dts = O.euler_solve(w0, order=2, dim=1, dx=1e-3, cfl=0.9, tfinal=0.25, max_steps=0, bias=0.0, limiter=2,
                             params=[1.4, 0, 25, 0, 0.25, 250])

    def totals(w):
        rho, u, p = w[:, 0], w[:, 1], w[:, 4]
        return np.array([rho.sum(), (rho * u).sum(), (0.5 * rho * u * u + p / 0.4).sum()]) * 1e-3

    a, b = totals(w0), totals(out)
    assert abs(b[0] - a[0]) < 1e-13
    assert abs(b[2] - a[2]) < 1e-13
    assert abs((b[1] - a[1]) - (1.0 - 0.1) * dts.sum()) < 1e-12


def test_limiter_closed_forms():
    r = np.array([-0.5, 0.0, 0.25, 0.75, 1.0, 1.5, 3.0])
    out = O.limiter_units(r, 0.0)
    # with bias 0 the R coefficient is 2 / (1 + r)
    np.testing.assert_array_equal(out[:, 3], 2.0 / (1.0 + r))
    np.testing.assert_array_equal(out[:, 0], [0, 0, 0.5, 1.0, 1.0, min(1.5, 2 / 2.5), min(3.0, 2 / 4.0)])
    np.testing.assert_array_equal(out[:, 1], [0, 0, min(0.5 / 1.25, 2 / 1.25), min(1.5 / 1.75, 2 / 1.75), 1.0,
                                              min(3 / 2.5, 2 / 2.5), min(6 / 4.0, 2 / 4.0)])
    np.testing.assert_array_equal(out[:, 2], [0, 0, 0.25, 0.75, 1.0, min(1.0, 2 / 2.5), min(1.0, 2 / 4.0)])
    np.testing.assert_array_equal(out[:, 4], out[:, 2])   # any code other than 0/1 is minbee


def test_eos_roundtrip_identity():
    """cons(prim) inverted by prim(cons) to round-off, and the stiffened-gas
    pressure law p = (gamma-1) rho e - gamma p_inf recovered."""
    rng = np.random.default_rng(3)
    s = np.zeros((200, 30))
    for k in range(5):
        s[:, 6 * k + 0] = rng.uniform(0.1, 3, 200)
        s[:, 6 * k + 1:6 * k + 4] = rng.uniform(-2, 2, (200, 3))
        s[:, 6 * k + 4] = rng.uniform(0.1, 3, 200)
        s[:, 6 * k + 5] = rng.uniform(0, 1, 200)
    params = [3.0, 0.1, 25, 0.5, 0.25, 250]
    out = O.euler_units(s, bias=0.0, limiter=0, dx=0.01, ts=0.001, params=params)
    b = s[:, 6:12]
    np.testing.assert_allclose(out[:, 74:80], b, rtol=1e-13, atol=1e-13)
    rho, u, v, z, p, lam = b.T
    E = rho * (0.5 * (u * u + v * v + z * z) + (p + 3.0 * 0.1) / (2.0 * rho) + 0.5 * (lam - 1))
    np.testing.assert_allclose(out[:, 72], E, rtol=1e-14)
    # Euler flux F = (rho u, rho u^2 + p, rho u v, rho u w, u (E + p), rho u lambda)
    F = np.stack([rho * u, rho * u * u + p, rho * u * v, rho * u * z, u * (E + p), rho * u * lam], 1)
    np.testing.assert_allclose(out[:, 6:12], F, rtol=1e-12, atol=1e-12)


def test_periodic_translation_invariance():
    """With periodic boundaries a cyclic shift of the initial grid shifts the
    solution bit for bit (the scheme is the same at every cell)."""
    rng = np.random.default_rng(5)
    ny, nx = 12, 16
    w = np.zeros((ny, nx, 6))
    w[..., 0] = rng.uniform(0.5, 1.5, (ny, nx))
    w[..., 1:3] = rng.uniform(-0.3, 0.3, (ny, nx, 2))
    w[..., 4] = rng.uniform(0.5, 1.5, (ny, nx))
    w[..., 5] = 1.0
    kw = dict(order=2, dim=2, dx=0.1, cfl=0.8, tfinal=math.inf, max_steps=10, bias=0.0, limiter=0,
              params=[1.4, 0, 25, 0, 0.25, 250], bc=1)
    a, da = O.euler_solve(w, **kw)
    b, db = O.euler_solve(np.roll(w, (3, 5), axis=(0, 1)), **kw)
    O.assert_bitwise(db, da, "dt")
    O.assert_bitwise(b, np.roll(a, (3, 5), axis=(0, 1)), "shifted state")


def test_2d_corner_ghosts_enter_the_time_step():
    """A quiet gas whose sound speed is below the default-state corner ghosts'
    (rho = p = 1, gamma 1.4: c = sqrt(1.4)) gets its dt from the corners."""
    w = np.zeros((8, 8, 6))
    w[..., 0] = 1.0
    w[..., 4] = 0.5
    w[..., 5] = 1.0
    _, dts = O.euler_solve(w, order=2, dim=2, dx=0.1, cfl=0.5, tfinal=math.inf, max_steps=7, bias=0.0, limiter=0,
                           params=[1.4, 0, 25, 0, 0.25, 250])
    s = math.sqrt(1.4 * 1.0 / 1.0)
    assert dts[0] == 0.5 * (0.1 / s) * 0.2 and dts[6] == 0.5 * (0.1 / s)
    _, d1 = O.euler_solve(w[0], order=2, dim=1, dx=0.1, cfl=0.5, tfinal=math.inf, max_steps=1, bias=0.0, limiter=0,
                          params=[1.4, 0, 25, 0, 0.25, 250])
    assert d1[0] == 0.5 * (0.1 / math.sqrt(1.4 * 0.5)) * 0.2
