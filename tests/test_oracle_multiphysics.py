"""The two-material diffuse-interface Euler oracle (oracle/multiphysics_oracle.c)
against the compiled reference's golden vectors, bit for bit: the per-function
records (conversions, fluxes, wave speeds, slopes, half-step evolution, FORCE and
SLIC fluxes), the volume-fraction reinitialisation, and whole SLIC / FORCE solves
of the reference drivers' initial conditions (1D and 2D, with and without
reinitialisation), each of which the generator also checked against the
reference's own solve / solve2D."""

import numpy as np
import pytest

import oracle_lib as O
from vanguard_amd import problems

RUNS = ["mp_toro100", "mp_toro100_r5", "mp_fedkiw100_r10", "mp_force_fedkiw100", "rd_toro100_r5",
        "rd_fedkiw100_r10", "rd_force_toro100", "mp_2d_toro32_r4", "mp_2d_fedkiw32_r3", "rd_2d_toro32_r4",
        "rd_2d_fedkiw32_r3", "mp_force_2d_toro24"]


def run_oracle(g):
    m = g["meta"]
    return O.mp_solve(m["kind"], g["prim_in"], order=m["order"], dim=m["dim"], dx=m["dx"], cfl=m["cfl"],
                      tfinal=m["tfinal"], max_steps=m["max_steps"], bias=m["bias"], limiter=m["limiter"],
                      reinit=m["reinit"], p1=m["p1"], p2=m["p2"])


@pytest.mark.parametrize("name", RUNS)
def test_two_material_solve_bit_identical_to_reference(name):
    g = O.golden(name)
    out, dts = run_oracle(g)
    O.assert_bitwise(dts, g["dt"], f"{name} dt")
    O.assert_bitwise(out, g["prim_out"], f"{name} state")


@pytest.mark.parametrize("name", ["mp_units", "mp_units_vl_bias", "rd_units", "rd_units_mb_bias"])
def test_two_material_units_bit_identical_to_reference(name):
    g = O.golden(name)
    m = g["meta"]
    out = O.mp_units(m["kind"], g["states"], bias=m["bias"], limiter=m["limiter"], dx=m["dx"], ts=m["ts"],
                     p1=m["p1"], p2=m["p2"])
    O.assert_bitwise(out, g["out"], name)


@pytest.mark.parametrize("name", ["mp_reinit_2d", "rd_reinit_2d", "mp_reinit_1d", "rd_reinit_1d"])
def test_reinitialisation_bit_identical_to_reference(name):
    g = O.golden(name)
    m = g["meta"]
    out = O.mp_reinit(m["kind"], g["prim_in"], m["dim"])
    assert (out[..., 0] != g["prim_in"][..., 0]).any()   # the field had crossings to reset
    O.assert_bitwise(out, g["prim_out"], name)


def test_driver_initial_conditions_match_the_fixtures():
    """problems.py's two-material ICs are the ones the fixtures were run from."""
    for name, ic, n in (("mp_toro100", problems.two_material_toro_1d, 100),
                        ("mp_fedkiw100_r10", problems.two_material_fedkiw_1d, 100),
                        ("rd_toro100_r5", problems.two_material_toro_1d, 100),
                        ("rd_fedkiw100_r10", problems.two_material_fedkiw_1d, 100),
                        ("mp_2d_toro32_r4", problems.two_material_toro_2d, 32),
                        ("rd_2d_fedkiw32_r3", problems.two_material_fedkiw_2d, 32)):
        g = O.golden(name)
        O.assert_bitwise(ic(n, g["meta"]["kind"]), g["prim_in"], name)


@pytest.mark.parametrize("kind", [0, 1])
def test_uniform_state_is_a_fixed_point(kind):
    """A uniform two-material state (any volume fraction inside the clamps) is
    a fixed point of SLIC: every flux difference is exactly zero, so only the
    primitive/conserved round trip of each step moves the state, by round-off.
    The reduced model's round trip keeps the pressure only when both materials
    share an EOS (its mixture energy is not inverted by its mixture pressure
    otherwise), so it runs with one material law twice."""
    n = 24
    w = np.zeros((n, O.MP_NP[kind]))
    w[:, 0] = 0.37
    w[:, 4:] = [1.3, 0.9, 0.6, 1.1][:O.MP_NP[kind] - 4]
    p1 = (1.4, 0.0, 25.0, 0.0, 0.25, 250.0)
    p2 = (3.0, 0.2, 25.0, 0.0, 0.25, 250.0) if kind == 0 else p1
    out, dts = O.mp_solve(kind, w, order=2, dim=1, dx=1 / n, cfl=0.8, tfinal=np.inf, max_steps=3, bias=0.0,
                          limiter=0, reinit=0, p1=p1, p2=p2)
    assert len(dts) == 3
    np.testing.assert_allclose(out, w, rtol=1e-14, atol=0)


@pytest.mark.parametrize("kind", [0, 1])
def test_periodic_mass_is_conserved(kind):
    """On a periodic grid each material's mass alpha_k rho_k sums to its start
    value to round-off (the conserved slots 5 and 6 / 7 are fluxes' divergences)."""
    n = 64
    x = (np.arange(n) + 0.5) / n
    w = np.zeros((n, O.MP_NP[kind]))
    w[:, 0] = 0.5 + 0.45 * np.sin(2 * np.pi * x)
    w[:, 1] = 0.3
    w[:, 4] = 1.0 + 0.2 * np.cos(2 * np.pi * x)
    w[:, 5] = 0.8 if kind == 1 else 1.0
    w[:, 6:] = 1.0
    p = (1.4, 0.0, 25.0, 0.0, 0.25, 250.0)
    out, _ = O.mp_solve(kind, w, order=2, dim=1, dx=1 / n, cfl=0.5, tfinal=np.inf, max_steps=20, bias=0.0,
                        limiter=1, reinit=0, p1=p, p2=p, bc=1)
    rho2 = 6 if kind == 0 else 5
    for a, r in ((lambda v: v[:, 0], 4), (lambda v: 1.0 - v[:, 0], rho2)):
        m0 = float(np.sum(a(w) * w[:, r]))
        m1 = float(np.sum(a(out) * out[:, r]))
        assert abs(m1 - m0) <= 1e-12 * abs(m0)
