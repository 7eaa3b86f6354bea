"""Pins the MHD oracle (oracle/mhd_oracle.c).

Bit for bit against golden vectors of the compiled reference (whole solves with
the per-step GLM speed c_h, per-function records), plus what the mathematics
fixes: the fast magnetosonic speed reduces to the sound speed without a field
and to max(c, c_A) when B is along the normal; the flux of a state at rest with
a normal field only carries magnetic pressure terms.
"""

import math

import numpy as np
import pytest

import oracle_lib as O


MHD_RUNS = ["mhd_briowu256", "mhd_force_briowu128", "mhd_2d_ot32", "mhd_force_2d_ot16", "mhd_psi128_sub2",
            "mhd_2d_ot16_sub1"]


@pytest.mark.parametrize("name", MHD_RUNS)
def test_oracle_matches_reference_mhd_run(name):
    g = O.golden(name)
    m = g["meta"]
    out, dts = O.mhd_solve(g["prim_in"], order=m["order"], dim=m["dim"], dx=m["dx"], cfl=m["cfl"],
                           tfinal=m["tfinal"], max_steps=m["max_steps"], bias=m["bias"], limiter=m["limiter"],
                           gamma=m["gamma"], p_inf=m["p_inf"], subcycles=m.get("subcycles", 0))
    O.assert_bitwise(dts, g["dt"], f"{name} dt")
    O.assert_bitwise(out, g["prim_out"], f"{name} state")


def test_sub_cycles_and_first_order_c_h_are_seen():
    """The fixtures pin what they claim: the sub-cycled runs differ from runs
    without sub-cycles, and the first-order 2D run differs from one whose c_h
    follows the step's maximum wave speed (psi would move)."""
    for name in ("mhd_psi128_sub2", "mhd_2d_ot16_sub1"):
        g = O.golden(name)
        m = g["meta"]
        plain, _ = O.mhd_solve(g["prim_in"], order=2, dim=m["dim"], dx=m["dx"], cfl=m["cfl"], tfinal=m["tfinal"],
                               max_steps=0, bias=0.0, limiter=m["limiter"], gamma=m["gamma"])
        assert np.abs(plain[..., 8] - g["prim_out"][..., 8]).max() > 1e-4, name
    g = O.golden("mhd_force_2d_ot16")
    assert np.all(g["prim_out"][..., 8] == 0.0)   # c_h = 0: psi never moves
    m = g["meta"]
    moved, _ = O.mhd_solve(g["prim_in"], order=1, dim=2, dx=m["dx"], cfl=m["cfl"], tfinal=m["tfinal"], max_steps=0,
                           bias=0.0, limiter=0, gamma=m["gamma"], ch=1.0, cp=1.0)
    assert np.abs(moved[..., 8]).max() > 1e-6


def test_glm_source_damps_psi_exponentially():
    """A uniform state with psi != 0: the fluxes cancel, so each sub-cycle is RK4
    on psi' = -(c_h^2 / c_p^2) psi with c_p^2 = 0.18 c_h; after one step with n
    sub-cycles per half step psi is multiplied by the RK4 amplification factor
    R(z)^(2n), z = -(c_h / 0.18) * dt / (2n)."""
    w = np.zeros((32, 9))
    w[:, 0] = 1.0
    w[:, 4] = 1.0
    w[:, 5] = 0.3
    w[:, 8] = 0.25
    for n in (1, 3):
        out, dts = O.mhd_solve(w, order=2, dim=1, dx=1.0 / 32, cfl=0.5, tfinal=np.inf, max_steps=1, bias=0.0,
                               limiter=0, gamma=1.4, bc=1, subcycles=n)
        # c_h = the step's max |u| + c_f, with u = 0 and B = (0.3, 0, 0)
        cf = np.sqrt(0.5 * ((1.4 + 0.09) + np.sqrt((1.4 + 0.09) ** 2 - 4 * 1.4 * 0.09)))
        z = -(cf / 0.18) * (0.5 * dts[0] / n)
        R = 1 + z + z * z / 2 + z ** 3 / 6 + z ** 4 / 24
        np.testing.assert_allclose(out[:, 8], 0.25 * R ** (2 * n), rtol=1e-12)
        np.testing.assert_allclose(out[:, :8], w[:, :8], rtol=0, atol=1e-15)


def test_oracle_matches_reference_mhd_batch_instances():
    g = O.golden("mhd_batch12")
    m = g["meta"]
    off = g["offsets"]
    for q in range(m["n_inst"]):
        a, b = off[q], off[q + 1]
        out, dts = O.mhd_solve(g["prim_in"][a:b], order=2, dim=1, dx=float(g["dx"][q]), cfl=m["cfl"],
                               tfinal=m["tfinal"], max_steps=0, bias=m["bias"], limiter=m["limiter"],
                               gamma=float(g["gamma"][q]))
        assert len(dts) == g["steps"][q]
        assert O.seq_sum(dts) == g["t"][q]
        O.assert_bitwise(out, g["prim_out"][a:b], f"instance {q}")


FIELDS = [("slope", 0, 9), ("flux_x", 9, 18), ("flux_y", 18, 27), ("evolve_x", 27, 45), ("force_x", 45, 54),
          ("slic_x", 54, 63), ("wave_speed", 63, 65), ("cons", 65, 74), ("cons_roundtrip", 74, 83),
          ("fast_speed", 83, 84), ("alfven", 84, 85)]


def test_oracle_matches_reference_mhd_functions():
    g = O.golden("mhd_units")
    m = g["meta"]
    out = O.mhd_units(g["states"], bias=m["bias"], limiter=m["limiter"], dx=m["dx"], ts=m["ts"], gamma=m["gamma"],
                      p_inf=m["p_inf"], ch=m["ch"])
    for field, a, b in FIELDS:
        O.assert_bitwise(out[:, a:b], g["out"][:, a:b], f"mhd_units:{field}")


def test_fast_speed_limits():
    """c_f = c without a field; c_f = max(c, c_A) for a field along x."""
    rng = np.random.default_rng(1)
    n = 50
    st = np.zeros((n, 45))
    for k in range(5):
        st[:, 9 * k + 0] = rng.uniform(0.2, 3, n)
        st[:, 9 * k + 4] = rng.uniform(0.2, 3, n)
    st[:, 9 + 5] = rng.uniform(-2, 2, n)      # state b: Bx only
    out = O.mhd_units(st, bias=0.0, limiter=0, dx=0.01, ts=0.001, gamma=5 / 3, p_inf=0.0, ch=1.0)
    rho, p, bx = st[:, 9], st[:, 13], st[:, 14]
    c = np.sqrt(5 / 3 * p / rho)
    ca = np.abs(bx) / np.sqrt(rho)
    np.testing.assert_allclose(out[:, 83], np.maximum(c, ca), rtol=1e-12)
    np.testing.assert_allclose(out[:, 84], ca, rtol=1e-14)
    st2 = st.copy()
    st2[:, 14] = 0.0
    out2 = O.mhd_units(st2, bias=0.0, limiter=0, dx=0.01, ts=0.001, gamma=5 / 3, p_inf=0.0, ch=1.0)
    np.testing.assert_allclose(out2[:, 83], c, rtol=1e-14)


def test_flux_of_static_state_with_normal_field():
    """u = 0, B = (Bx, By, 0): F = (0, p + B^2/2 - Bx^2, -Bx By, 0, 0, psi, 0, 0, c_h^2 Bx)."""
    st = np.zeros((1, 45))
    st[0, 9:18] = [1.3, 0, 0, 0, 0.7, 0.4, -0.9, 0.0, 0.05]
    out = O.mhd_units(st, bias=0.0, limiter=0, dx=0.01, ts=0.001, gamma=2.0, p_inf=0.0, ch=1.7)
    bx, by = 0.4, -0.9
    b2 = bx * bx + by * by
    exp = [0, 0.7 + 0.5 * b2 - bx * bx, -bx * by, 0, 0, 0.05, 0, 0, 1.7 * 1.7 * bx]
    np.testing.assert_allclose(out[0, 9:18], exp, rtol=1e-14, atol=1e-15)


def test_brio_wu_conservation():
    """Total mass and energy are conserved before waves reach the ends; By's total
    changes only through the boundary fluxes, which vanish for a state at rest."""
    g = O.golden("mhd_briowu256")
    w0, w1 = g["prim_in"], g["prim_out"]

    def mass(w):
        return w[:, 0].sum()

    assert abs(mass(w1) - mass(w0)) < 1e-12 * mass(w0)
    # Bx is constant in 1D (its flux is psi, and the cleaning keeps it near constant)
    assert np.abs(w1[:, 5] - 0.75).max() < 0.05
