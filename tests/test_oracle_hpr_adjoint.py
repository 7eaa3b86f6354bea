"""The oracle's dual-number HPR step (oracle/hpr_dual.cpp), the independent
check of the GPU's HPR adjoint.

1. With T = double it is the oracle's HPR solve bit for bit (orc_hpr_solve at
   the solve's own time steps): 1D and 2D, with and without heat conduction,
   0-2 relaxation sub-cycles, transmissive and periodic lines.
2. Its tangent, along a direction in the initial state and in the two
   relaxation times, matches central differences of the fixed-step map where
   the branch hash (slope clamps, limiter cases) is the same at both ends.
3. Its unit JVPs (p2c, c2p, the fluxes, the source) match central differences
   of the oracle's own functions and the closed forms the mathematics fixes:
   c2p'(p2c(w)) p2c'(w) = I and the source's relaxation-time derivative -S / tau.
"""

import numpy as np
import pytest

import oracle_lib as O
from vanguard_amd import problems


def hpr_adj_params(thermal, mu=1e-3):
    """The viscous-flow problem's material: stiffened gas gamma 1.4, c_s 1,
    alpha 2, tau1 = 6 mu / (p0 c_s^2), tau2 from kappa = mu gamma c_v / Pr."""
    return O.hpr_params(thermal, mu=mu, alpha=2.0, kappa=mu * 1.4 * 2.5 / 0.75)


def hpr_adj_case(dim, thermal, n=12, seed=4):
    """The Fourier viscous-flow initial state at n^2 (or its middle row in 1D),
    with a small random distortion (and heat flux) so that every term of the
    flux and source is exercised from the first step."""
    w = problems.hpr_fourier_2d(n, seed=seed)
    rng = np.random.default_rng(seed + 100)
    w[..., 2:11] += 2e-3 * rng.standard_normal(w[..., 2:11].shape)
    if thermal:
        w[..., 14:17] += 1e-3 * rng.standard_normal(w[..., 14:17].shape)
    w[..., 0] += 0.02 * rng.standard_normal(w[..., 0].shape)
    if dim == 1:
        w = np.ascontiguousarray(w[n // 2])
    return w


CASES = [(2, True, 1, 1), (2, False, 2, 0), (1, True, 1, 0), (2, True, 0, 1), (1, False, 1, 1)]


def run_case(dim, thermal, sub, bc, steps=4, n=12):
    w = hpr_adj_case(dim, thermal, n)
    p = hpr_adj_params(thermal)
    _, dts = O.hpr_solve(w, order=2, dim=dim, dx=1.0 / n, cfl=0.8, tfinal=1.0, max_steps=steps, bias=0.0, limiter=0,
                         subcycles=sub, params=p, bc=bc)
    return w, p, dts, dict(dim=dim, dx=1.0 / n, bias=0.0, limiter=0, subcycles=sub, params=p, bc=bc)


@pytest.mark.parametrize("dim,thermal,sub,bc", CASES)
def test_hpr_dual_primal_is_the_oracle_solve(dim, thermal, sub, bc):
    w, p, dts, kw = run_case(dim, thermal, sub, bc)
    ref, _ = O.hpr_solve(w, order=2, dim=dim, dx=kw["dx"], cfl=0.8, tfinal=1.0, max_steps=dts.size, bias=0.0,
                         limiter=0, subcycles=sub, params=p, bc=bc)
    O.assert_bitwise(O.hpr_fixed_dt(w, dts, **kw), ref, "dual-number primal")
    out, _ = O.hpr_fixed_dt(w, dts, tangent=np.zeros_like(w), **kw)
    O.assert_bitwise(out, ref, "dual-number primal with a tangent")


@pytest.mark.parametrize("dim,thermal,sub,bc", CASES)
def test_hpr_dual_tangent_matches_central_differences(dim, thermal, sub, bc):
    w, p, dts, kw = run_case(dim, thermal, sub, bc)
    rng = np.random.default_rng(7)
    v = rng.standard_normal(w.shape) * 1e-2
    dtau = (0.3 * p["tau1"], 0.2 * p["tau2"] if thermal else 0.0)
    _, t = O.hpr_fixed_dt(w, dts, tangent=v, dtau=dtau, **kw)
    h0 = O.branch_hash()
    for h in (1e-7, 1e-8, 1e-6):
        pa, pb = dict(p), dict(p)
        pa["tau1"] += h * dtau[0]
        pa["tau2"] += h * dtau[1]
        pb["tau1"] -= h * dtau[0]
        pb["tau2"] -= h * dtau[1]
        a = O.hpr_fixed_dt(w + h * v, dts, **dict(kw, params=pa))
        ha = O.branch_hash()
        b = O.hpr_fixed_dt(w - h * v, dts, **dict(kw, params=pb))
        if ha == h0 == O.branch_hash():
            break
    else:
        pytest.fail("no step keeps every branch")
    fd = (a - b) / (2 * h)
    scale = np.abs(t).reshape(-1, 17).max(axis=0) + 1e-300
    err = (np.abs(fd - t).reshape(-1, 17).max(axis=0) / scale).max()
    assert err < 2e-5, (h, err)


def test_hpr_branch_hash_sees_a_clamp_crossing():
    """A perturbation that moves a distortion difference across the 1e-5 slope
    clamp changes the hash (and the map's derivative), one that does not, not."""
    w, p, dts, kw = run_case(2, True, 1, 1, steps=2)
    O.hpr_fixed_dt(w, dts, **kw)
    h0 = O.branch_hash()
    O.hpr_fixed_dt(w * (1 + 1e-13), dts, **kw)
    assert O.branch_hash() == h0
    flat = w.copy()
    flat[..., 2:11] = np.eye(3).reshape(9)   # every distortion difference 0: clamped
    O.hpr_fixed_dt(flat, dts, **kw)
    assert O.branch_hash() != h0


def unit_states(thermal, count=6, seed=3):
    rng = np.random.default_rng(seed)
    w = np.zeros((count, 17))
    w[:, 0] = 1.0 + 0.2 * rng.random(count)
    w[:, 1] = 0.7 + 0.2 * rng.random(count)
    w[:, 2:11] = np.eye(3).reshape(9) + 0.05 * rng.standard_normal((count, 9))
    w[:, 11:14] = 0.3 * rng.standard_normal((count, 3))
    w[:, 14:17] = 0.1 * rng.standard_normal((count, 3)) if thermal else 0.0
    return w


@pytest.mark.parametrize("thermal", [True, False])
def test_hpr_unit_jvps(thermal):
    """Along a primitive direction u: p2c'(w) u, and F'(U) v, S'(U) v at
    U = p2c(w), v = p2c'(w) u, against central differences of the oracle's own
    p2c, fluxes and source at p2c(w +- h u); c2p'(U) v = u exactly in exact
    arithmetic (the conversions are inverse)."""
    p = hpr_adj_params(thermal)
    w = unit_states(thermal)
    rng = np.random.default_rng(11)
    u = rng.standard_normal(w.shape)
    if not thermal:
        u[:, 14:17] = 0.0
    v = O.hpr_jvp_units(w, u, params=p)[:, 0]
    J = O.hpr_jvp_units(w, v, params=p)
    rec = lambda x: O.hpr_units(np.concatenate([x] * 5, axis=1), bias=0.0, limiter=0, dx=0.1, ts=1e-3, params=p)
    h = 1e-6
    up, um = rec(w + h * u), rec(w - h * u)
    for name, k, sl in (("p2c", None, slice(0, 17)), ("Fx", 2, slice(34, 51)), ("Fy", 3, slice(51, 68)),
                        ("S", 4, slice(68, 85))):
        fd = (up[:, sl] - um[:, sl]) / (2 * h)
        got = v if k is None else J[:, k]
        np.testing.assert_allclose(got, fd, rtol=1e-6, atol=1e-8 * (np.abs(fd).max() + 1), err_msg=name)
    np.testing.assert_allclose(J[:, 1], u, rtol=1e-9, atol=1e-11)
    # the source is proportional to 1 / tau: d S / d tau = -S / tau
    S = rec(w)[:, 68:85]
    Jt = O.hpr_jvp_units(w, np.zeros_like(u), params=p, dtau=(1.0, 0.0))
    np.testing.assert_allclose(Jt[:, 4, 2:11], -S[:, 2:11] / p["tau1"], rtol=1e-12)
    if thermal:
        Jt2 = O.hpr_jvp_units(w, np.zeros_like(u), params=p, dtau=(0.0, 1.0))
        np.testing.assert_allclose(Jt2[:, 4, 14:17], -S[:, 14:17] / p["tau2"], rtol=1e-12)
