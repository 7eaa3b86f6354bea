"""Pins the RGFM oracle (oracle/rgfm_oracle.c).

HLLC path: bit for bit against golden vectors of the compiled reference --
per-function HLLC records (different gammas: always PVRS; equal gammas: the
two-shock / two-rarefaction branches and the Y-velocity quirk of
hllcsolver.cpp:552; stiffened gas) and whole solveHLLC / solveHLLC2D runs (1D
Toro and Fedkiw tests checked against the reference's own solve, 2D discs, a
shock-bubble geometry).

Exact solver (the reference's does not compile): Toro's published exact star
states for his five tests (Toro, Riemann Solvers and Numerical Methods for
Fluid Dynamics, 3rd ed., Table 4.3), the Rankine-Hugoniot conditions across
computed shocks, and the isentrope and Riemann invariant across rarefactions.
Level set and circulation: exactness of the upwind stencil on linear fields,
and the closed-form circulation of a solid-body rotation.
"""

import math

import numpy as np
import pytest

import oracle_lib as O

AIR = (1.4, 0.0, 25.0, 0.0, 0.25, 250.0)


@pytest.mark.parametrize("k", [0, 1, 2])
def test_hllc_units_match_reference(k):
    g = O.golden("hllc_units")
    out = O.hllc_units(g[f"states{k}"], tuple(g[f"p1_{k}"]), tuple(g[f"p2_{k}"]))
    O.assert_bitwise(out, g[f"out{k}"], f"hllc_units set {k}")


RUNS = ["rgfm_toro1d_hllc", "rgfm_fedkiw1d_hllc", "rgfm_2d_toro32", "rgfm_2d_fedkiw32", "rgfm_2d_fedkiw24_api",
        "rgfm_2d_bubble96"]


def run_golden(g):
    m = g["meta"]
    return O.rgfm_solve(g["m1_in"], g["m2_in"], g["phi_in"], dim=m["dim"], dx=m["dx"], cfl=m["cfl"],
                        tfinal=m["tfinal"], max_steps=m["max_steps"], bias=m["bias"], limiter=m["limiter"],
                        p1=tuple(m["p1"]), p2=tuple(m["p2"]))


@pytest.mark.parametrize("name", RUNS)
def test_rgfm_run_matches_reference(name):
    g = O.golden(name)
    o1, o2, oph, dts = run_golden(g)
    O.assert_bitwise(dts, g["dt"], f"{name} dt")
    O.assert_bitwise(o1, g["m1_out"], f"{name} material 1")
    O.assert_bitwise(o2, g["m2_out"], f"{name} material 2")
    O.assert_bitwise(oph, g["phi_out"], f"{name} level set")


# Toro's tests 1-5 (gamma 1.4): left, right, exact p*, u*, rho*_L, rho*_R.
TORO = [
    ((1.0, 0.0, 1.0), (0.125, 0.0, 0.1), 0.30313, 0.92745, 0.42632, 0.26557),
    ((1.0, -2.0, 0.4), (1.0, 2.0, 0.4), 0.00189, 0.0, 0.02185, 0.02185),
    ((1.0, 0.0, 1000.0), (1.0, 0.0, 0.01), 460.894, 19.5975, 0.57506, 5.99924),
    ((1.0, 0.0, 0.01), (1.0, 0.0, 100.0), 46.0950, -6.19633, 5.99242, 0.57511),
    ((5.99924, 19.5975, 460.894), (5.99242, -6.19633, 46.0950), 1691.64, 8.68975, 14.2823, 31.0426),
]


def pair(a, b):
    return np.array([a[0], a[1], 0.0, 0.0, a[2], 1.0, b[0], b[1], 0.0, 0.0, b[2], 1.0])


def test_exact_solver_toro_tables():
    st = np.stack([pair(a, b) for a, b, *_ in TORO])
    out = O.exact_units(st, AIR, AIR)
    # the table prints 5-6 significant digits: half a unit of the last one
    for k, (_, _, ps, us, rl, rr) in enumerate(TORO):
        for got, want, what in ((out[k, 0], ps, "p*"), (out[k, 1], us, "u*"), (out[k, 2], rl, "rho*L"),
                                (out[k, 3], rr, "rho*R")):
            assert got == pytest.approx(want, rel=6e-5, abs=5e-6), f"test {k + 1} {what}"


def test_exact_solver_jump_conditions():
    """Shocks satisfy Rankine-Hugoniot (mass and momentum flux continuous in the
    shock frame); rarefactions keep p / rho^gamma and u +- 2c/(gamma-1)."""
    rng = np.random.default_rng(5)
    g = 1.4
    for _ in range(40):
        a = (10 ** rng.uniform(-1, 1), rng.uniform(-1, 1), 10 ** rng.uniform(-1, 1))
        b = (10 ** rng.uniform(-1, 1), rng.uniform(-1, 1), 10 ** rng.uniform(-1, 1))
        ps, us, rl, rr = O.exact_units(pair(a, b)[None], AIR, AIR)[0, :4]
        for (rho, u, p), rs, side in ((a, rl, -1), (b, rr, 1)):
            if ps > p:
                S = (rs * us - rho * u) / (rs - rho)
                m_star, m_k = rs * (us - S), rho * (u - S)
                assert m_star == pytest.approx(m_k, rel=1e-7, abs=1e-9)
                assert rs * us * (us - S) + ps == pytest.approx(rho * u * (u - S) + p, rel=1e-7)
            else:
                assert ps / rs ** g == pytest.approx(p / rho ** g, rel=1e-9)
                c, cs = math.sqrt(g * p / rho), math.sqrt(g * ps / rs)
                assert us - side * 2 * cs / (g - 1) == pytest.approx(u - side * 2 * c / (g - 1), rel=1e-9, abs=1e-9)


def test_hllc_pvrs_closed_form():
    """Different gammas force the PVRS estimate: p* = (pL+pR)/2 + (uL-uR)/2 * rho_bar c_bar
    with rho_bar c_bar = (rhoL+rhoR)(cL+cR)/4."""
    he = (1.648, 0.0, 25.0, 0.0, 0.25, 250.0)
    a, b = (1.0, 0.3, 2.0), (0.2, -0.1, 1.5)
    out = O.hllc_units(pair(a, b)[None], AIR, he)[0]
    cl, cr = math.sqrt(1.4 * a[2] / a[0]), math.sqrt(1.648 * b[2] / b[0])
    prod = 0.25 * (a[0] + b[0]) * (cl + cr)
    assert out[9] == pytest.approx(0.5 * (a[2] + b[2]) + 0.5 * (a[1] - b[1]) * prod, rel=1e-14)
    assert out[8] == pytest.approx(0.5 * (a[1] + b[1]) + 0.5 * (a[2] - b[2]) / prod, rel=1e-14)


def test_level_set_exact_on_linear_fields():
    """The one-sided second-order stencil differentiates linear phi exactly, so
    phi_new = phi - dt u phi_x away from the mirrored ends, for either sign of u."""
    n = 40
    phi = 0.37 * np.arange(n) - 5.0
    for u in (0.8, -0.6):
        m = np.zeros((n, 6))
        m[:, 1] = u
        out = O.rgfm_level_set(1, 0, phi, m, m, 0.37, 0.01)
        np.testing.assert_allclose(out[3:-3], (phi - 0.01 * u)[3:-3], rtol=0, atol=1e-13)


def test_circulation_of_solid_body_rotation():
    """u = -W y, v = W x has vorticity 2W, so the circulation over the phi < 0
    interior cells is 2 W dx dy times their number."""
    ny, nx, dx, dy, W = 30, 40, 0.01, 0.02, 3.0
    x = (np.arange(nx) + 0.5) * dx
    y = (np.arange(ny) + 0.5) * dy
    X, Y = np.meshgrid(x, y)
    m1 = np.zeros((ny, nx, 6))
    m1[..., 1] = -W * Y
    m1[..., 2] = W * X
    phi = np.hypot(X - 0.2, Y - 0.3) - 0.12
    inside = (phi[1:-1, 1:-1] < 0).sum()
    assert O.rgfm_circulation(m1, phi, dx, dy) == pytest.approx(2 * W * dx * dy * inside, rel=1e-12)
