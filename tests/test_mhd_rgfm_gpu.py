"""Two-material ideal MHD with the real ghost fluid method on the GPU through the
C ABI: MHDRGFMSolver::solve / solve2D bit-identical to the compiled
reference's golden vectors (time steps, both materials, the level set), the
standalone ghost fills against the oracle's on states that reach every branch
of the MHD HLLC interface solver and the pressure floor, and ragged 2D grids
with GLM sub-cycles against the oracle."""

import math

import numpy as np
import pytest

import oracle_lib as O
from vanguard_amd import _lib
from vanguard_amd import solvers as S

pytestmark = pytest.mark.gpu

RUNS = ["mhd_rgfm_dumbser200", "mhd_rgfm_dumbser_multi200", "mhd_rgfm_vl_glm100", "mhd_rgfm_2d_dumbser32",
        "mhd_rgfm_2d_multi32_sub1"]


@pytest.mark.parametrize("name", RUNS)
def test_mhd_rgfm_solve_bit_identical_to_reference(vg_ctx, name):
    g = O.golden(name)
    m = g["meta"]
    system = S.MHDMultimaterialSystem(g["m1_in"], g["m2_in"], g["phi_in"])
    r = S.mhd_rgfm_solve(system, m["dx"], m["cfl"], m["tfinal"], tuple(m["p1"]), tuple(m["p2"]), bias=m["bias"],
                         limiter=m["limiter"], subcycles=m["subcycles"], max_steps=m["max_steps"])
    print(f"{name}: {r.stats.steps} steps", flush=True)
    O.assert_bitwise(r.dt, g["dt"], f"{name} dt")
    O.assert_bitwise(r.system.material1, g["m1_out"], f"{name} material 1")
    O.assert_bitwise(r.system.material2, g["m2_out"], f"{name} material 2")
    O.assert_bitwise(r.system.level_set, g["phi_out"], f"{name} level set")


def test_reference_style_entry_points(vg_ctx):
    """MHDRGFMSolver.solve / solve2D with the reference's argument list."""
    g = O.golden("mhd_rgfm_vl_glm100")
    m = g["meta"]
    out = S.MHDRGFMSolver.solve(S.MHDMultimaterialSystem(g["m1_in"], g["m2_in"], g["phi_in"]), m["dx"], m["cfl"],
                                m["tfinal"], m["bias"], m["limiter"], 3, 0, S.MHDMaterialParameters(*m["p1"]),
                                S.MHDMaterialParameters(*m["p2"]))
    O.assert_bitwise(out.material1, g["m1_out"], "solve")
    g = O.golden("mhd_rgfm_2d_multi32_sub1")
    m = g["meta"]
    out = S.MHDRGFMSolver.solve2D(S.MHDMultimaterialSystem(g["m1_in"], g["m2_in"], g["phi_in"]), m["dx"], m["cfl"],
                                  m["tfinal"], m["bias"], m["limiter"], m["subcycles"], 0,
                                  S.MHDMaterialParameters(*m["p1"]), S.MHDMaterialParameters(*m["p2"]))
    O.assert_bitwise(out.material2, g["m2_out"], "solve2D")
    O.assert_bitwise(out.level_set, g["phi_out"], "solve2D level set")


def random_states(rng, shape):
    """MHD states with O(1) jumps, fast opposing flows (star pressures below
    the 1e-8 floor) and fast same-direction flows (supersonic interfaces)."""
    w = np.zeros(shape + (9,))
    w[..., 0] = rng.uniform(0.2, 3.0, shape)
    w[..., 1:4] = rng.uniform(-1.0, 1.0, shape + (3,))
    w[..., 4] = rng.uniform(0.05, 3.0, shape)
    w[..., 5:8] = rng.uniform(-1.5, 1.5, shape + (3,))
    w[..., 8] = rng.uniform(-0.3, 0.3, shape)
    pick = rng.uniform(size=shape)
    w[..., 1] = np.where(pick < 0.1, rng.choice([-12.0, 12.0], shape), w[..., 1])
    w[..., 2] = np.where(pick > 0.9, rng.choice([-12.0, 12.0], shape), w[..., 2])
    return w


def level_set_with_many_crossings(rng, shape):
    phi = rng.uniform(-1.0, 1.0, shape)
    phi[rng.uniform(size=shape) < 0.05] = 0.0   # zero values count as crossings on both sides
    return phi


@pytest.mark.parametrize("dim", [1, 2])
def test_ghost_fill_matches_oracle(vg_ctx, dim):
    rng = np.random.default_rng(40 + dim)
    shape = (300,) if dim == 1 else (23, 41)
    m1, m2 = random_states(rng, shape), random_states(rng, shape)
    phi = level_set_with_many_crossings(rng, shape)
    p1, p2 = (1.4, 0.0, 1.3, 0.4), (1.67, 0.3, 0.7, 0.9)
    out = S.mhd_rgfm_ghost_fill(S.MHDMultimaterialSystem(m1, m2, phi), p1, p2)
    r1, r2 = O.mhd_rgfm_ghost_fill(m1, m2, phi, p1, p2)
    O.assert_bitwise(out.material1, r1, "material 1")
    O.assert_bitwise(out.material2, r2, "material 2")
    # the fill reached the pressure floor and both kinds of ghost
    assert (r1[..., 4] == 1e-8).any() or (r2[..., 4] == 1e-8).any()
    changed1 = (r1 != m1).any(axis=-1)
    assert changed1.sum() > shape[-1] // 4
    if dim == 1:
        O.assert_bitwise(r1[..., 8], m1[..., 8], "1D fills leave psi alone")


def test_ghost_fill_reaches_every_interface_branch():
    """The oracle's MHD HLLC over the ghost-fill test's state distribution
    takes the left, right and both star branches, so the fill comparison above
    covers all four."""
    rng = np.random.default_rng(7)
    a, b = random_states(rng, (4000,)), random_states(rng, (4000,))
    out = O.mhd_hllc_units(np.concatenate([a, b], axis=1), (1.4, 0.0, 1.3, 0.4), (1.67, 0.3, 0.7, 0.9))
    sol = out[:, :9]
    left = (sol == a).all(axis=1)
    right = (sol == b).all(axis=1) & ~left
    star = ~left & ~right
    assert left.any() and right.any() and (star & (out[:, 22] >= 0)).any() and (star & (out[:, 22] < 0)).any()


@pytest.mark.parametrize("shape,sub,bias,limiter", [((20, 37), 2, 0.3, 1), ((33, 18), 0, -0.2, 2)])
def test_ragged_2d_matches_oracle(vg_ctx, shape, sub, bias, limiter):
    ny, nx = shape
    i = np.arange(ny)[:, None] + 0.5
    j = np.arange(nx)[None, :] + 0.5
    phi = np.sqrt((i - 0.45 * ny) ** 2 + (j - 0.55 * nx) ** 2) - 0.3 * min(ny, nx)
    rng = np.random.default_rng(ny * nx)
    m1 = np.zeros(shape + (9,))
    m1[..., 0] = 1.0 + 0.1 * rng.uniform(size=shape)
    m1[..., 1] = 0.2
    m1[..., 4] = 1.0 + 0.1 * rng.uniform(size=shape)
    m1[..., 5] = 0.75
    m1[..., 6] = 1.0
    m2 = m1.copy()
    m2[..., 0] = 0.125
    m2[..., 4] = 0.1
    m2[..., 6] = -1.0
    p1, p2 = (1.4, 0.0, 0.5, 0.3), (2.0, 0.1, 0.5, 0.3)
    dx = 1.0 / nx
    o1, o2, oph, dts = O.mhd_rgfm_solve(m1, m2, phi, dim=2, dx=dx, cfl=0.7, tfinal=math.inf, max_steps=12, bias=bias,
                                        limiter=limiter, subcycles=sub, p1=p1, p2=p2)
    r = S.mhd_rgfm_solve(S.MHDMultimaterialSystem(m1, m2, phi), dx, 0.7, math.inf, p1, p2, bias=bias,
                         limiter=limiter, subcycles=sub, max_steps=12)
    assert r.stats.steps == 12
    O.assert_bitwise(r.dt, dts, "dt")
    O.assert_bitwise(r.system.material1, o1, "material 1")
    O.assert_bitwise(r.system.material2, o2, "material 2")
    O.assert_bitwise(r.system.level_set, oph, "level set")


def test_rejected_arguments(vg_ctx):
    g = O.golden("mhd_rgfm_dumbser200")
    system = S.MHDMultimaterialSystem(g["m1_in"], g["m2_in"], g["phi_in"])
    r = S.mhd_rgfm_solve(system, 0.005, 0.8, math.inf, 2.0, 2.0, raise_on_error=False)
    assert r.status == _lib.VG_ERR_ARG
    with pytest.raises(ValueError):
        S.MHDRGFMSolver.solve2D(system, 0.005, 0.8, 0.1, 0.0, 0, 0, 0, S.MHDMaterialParameters(2.0),
                                S.MHDMaterialParameters(2.0))
    long = np.zeros((16385, 9))
    long[:, 0] = 1.0
    long[:, 4] = 1.0
    r = S.mhd_rgfm_solve(S.MHDMultimaterialSystem(long, long, np.linspace(-1, 1, 16385)), 1e-3, 0.5, 0.1, 2.0, 2.0,
                         raise_on_error=False)
    assert r.status == _lib.VG_ERR_ARG
