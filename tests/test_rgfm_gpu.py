"""Two-material RGFM on the GPU through the C ABI.

HLLC ghost fill between materials of different gamma (always the PVRS branch:
no pow) is bit-identical to the compiled reference's golden vectors and to the
oracle, 1D and 2D.  Equal-gamma problems can take the two-rarefaction branch,
which calls pow (hllcsolver.cpp:272, 298), and the exact Riemann solver's
Newton iteration calls pow throughout: those runs are held to a relative L1
error of 1e-10 per field against the oracle.
"""

import math

import numpy as np
import pytest

import oracle_lib as O
from vanguard_amd import _lib, problems
from vanguard_amd import solvers as S
from vanguard_amd._lib import VanguardError

pytestmark = pytest.mark.gpu

REL_L1 = 1e-10


def rel_l1(a, b):
    a = np.asarray(a).reshape(-1, np.asarray(a).shape[-1]) if np.asarray(a).ndim > 1 else np.asarray(a)[:, None]
    b = np.asarray(b).reshape(a.shape)
    return float((np.abs(a - b).sum(axis=0) / np.maximum(np.abs(b).sum(axis=0), 1e-300)).max())


def gpu_golden(g, **kw):
    m = g["meta"]
    return S.rgfm_solve(S.MultimaterialSystem(g["m1_in"], g["m2_in"], g["phi_in"]), m["dx"], m["cfl"], m["tfinal"],
                        tuple(m["p1"]), tuple(m["p2"]), bias=m["bias"], limiter=m["limiter"],
                        max_steps=m["max_steps"], **kw)


@pytest.mark.parametrize("name", ["rgfm_fedkiw1d_hllc", "rgfm_2d_fedkiw32", "rgfm_2d_fedkiw24_api",
                                  "rgfm_2d_bubble96"])
def test_rgfm_hllc_bit_identical_to_reference(vg_ctx, name):
    g = O.golden(name)
    r = gpu_golden(g)
    print(f"{name}: steps={r.stats.steps}", flush=True)
    O.assert_bitwise(r.dt, g["dt"], f"{name} dt")
    O.assert_bitwise(r.system.material1, g["m1_out"], f"{name} material 1")
    O.assert_bitwise(r.system.material2, g["m2_out"], f"{name} material 2")
    O.assert_bitwise(r.system.level_set, g["phi_out"], f"{name} level set")


@pytest.mark.parametrize("name", ["rgfm_toro1d_hllc", "rgfm_2d_toro32"])
def test_rgfm_hllc_equal_gamma_matches_reference(vg_ctx, name):
    g = O.golden(name)
    r = gpu_golden(g)
    assert len(r.dt) == len(g["dt"])
    np.testing.assert_allclose(r.dt, g["dt"], rtol=1e-12)
    for got, want in ((r.system.material1, g["m1_out"]), (r.system.material2, g["m2_out"])):
        assert rel_l1(got, want) <= REL_L1
    assert rel_l1(r.system.level_set, g["phi_out"]) <= REL_L1


def bubble(nx, ny, ly_over_lx=0.3, radius=0.03):
    lx = 0.325
    return problems.shock_bubble_2d(nx, ny, length=(lx, lx * ly_over_lx), center=(0.175, 0.5 * lx * ly_over_lx),
                                    radius=radius)


def test_rgfm_anisotropic_spacing_matches_oracle(vg_ctx):
    """dy != dx (not in the reference): the same arithmetic as the oracle."""
    m1, m2, phi, dx, dy = bubble(64, 30, 0.35)
    assert dx != dy
    o1, o2, oph, dts = O.rgfm_solve(m1, m2, phi, dim=2, dx=dx, dy=dy, cfl=0.8, tfinal=math.inf, max_steps=10,
                                    bias=0.0, limiter=0, p1=problems.HELIUM, p2=problems.AIR)
    r = S.rgfm_solve(S.MultimaterialSystem(m1, m2, phi), dx, 0.8, math.inf, problems.HELIUM, problems.AIR, dy=dy,
                     max_steps=10)
    O.assert_bitwise(r.dt, dts, "dt")
    O.assert_bitwise(r.system.material1, o1, "material 1")
    O.assert_bitwise(r.system.material2, o2, "material 2")
    O.assert_bitwise(r.system.level_set, oph, "level set")


@pytest.mark.parametrize("dim", [1, 2])
def test_rgfm_exact_riemann_matches_oracle(vg_ctx, dim):
    if dim == 1:
        n = 160
        x = (1.0 / n) * np.arange(n)
        m1 = np.tile([1.0, 0.0, 0.0, 0.0, 1e5, 1.0], (n, 1))
        m1[x <= 0.05] = [1.3333, 0.3535 * np.sqrt(1e5), 0.0, 0.0, 1.5e5, 1.0]
        m2 = np.tile([0.1379, 0.0, 0.0, 0.0, 1e5, 1.0], (n, 1))
        phi, dx, dy, p1, p2, steps = x - 0.5, 1.0 / n, None, problems.AIR, (1.67, 0, 25, 0, 0.25, 250), 40
    else:
        m1, m2, phi, dx, dy = bubble(48, 16)
        p1, p2, steps = problems.HELIUM, problems.AIR, 8
    o1, o2, oph, dts = O.rgfm_solve(m1, m2, phi, dim=dim, dx=dx, dy=dy, cfl=0.8, tfinal=math.inf, max_steps=steps,
                                    bias=0.0, limiter=1, p1=p1, p2=p2, riemann=1)
    r = S.rgfm_solve(S.MultimaterialSystem(m1, m2, phi), dx, 0.8, math.inf, p1, p2, dy=dy, limiter=1,
                     riemann=S.EXACT_RIEMANN, max_steps=steps)
    np.testing.assert_allclose(r.dt, dts, rtol=1e-12)
    assert rel_l1(r.system.material1, o1) <= REL_L1
    assert rel_l1(r.system.material2, o2) <= REL_L1
    assert rel_l1(r.system.level_set, oph) <= REL_L1


def test_rgfm_batch_matches_single_solves_and_circulation(vg_ctx):
    """Three shock-bubble instances (radius, Mach from the batch recipe) in one
    call: each equals its own oracle solve; the circulation equals the oracle's
    sum (a different summation order, hence a tolerance)."""
    radii, machs = problems.shock_bubble_batch(3, seed=6)
    ins = [problems.shock_bubble_2d(64, 24, length=(0.325, 0.325 * 24 / 64), center=(0.175, 0.325 * 12 / 64),
                                    radius=float(r) * 1.5, mach=float(mc)) for r, mc in zip(radii, machs)]
    dx, dy = ins[0][3], ins[0][4]
    m1 = np.stack([a[0] for a in ins])
    m2 = np.stack([a[1] for a in ins])
    phi = np.stack([a[2] for a in ins])
    tf = np.array([2e-5, 3e-5, 2.5e-5])
    r = S.rgfm_solve_batch_2d(m1, m2, phi, dx, dy, 0.8, tf, problems.HELIUM, problems.AIR)
    for q in range(3):
        o1, o2, oph, dts = O.rgfm_solve(m1[q], m2[q], phi[q], dim=2, dx=dx, dy=dy, cfl=0.8, tfinal=tf[q],
                                        max_steps=0, bias=0.0, limiter=0, p1=problems.HELIUM, p2=problems.AIR)
        assert r.steps[q] == len(dts)
        assert r.t[q] == O.seq_sum(dts)
        O.assert_bitwise(r.material1[q], o1, f"instance {q} material 1")
        O.assert_bitwise(r.material2[q], o2, f"instance {q} material 2")
        O.assert_bitwise(r.level_set[q], oph, f"instance {q} level set")
        ref = O.rgfm_circulation(o1, oph, dx, dy)
        assert r.circulation[q] == pytest.approx(ref, rel=1e-11, abs=1e-14)


def test_rgfm_batch_streams_change_nothing():
    """Five instances on 1, 2 and 4 worker streams (instances taken longest
    predicted run first from a shared queue): every output is the same bits,
    and the stats' step total matches."""
    radii, machs = problems.shock_bubble_batch(5, seed=9)
    ins = [problems.shock_bubble_2d(48, 20, length=(0.325, 0.325 * 20 / 48), center=(0.175, 0.325 * 10 / 48),
                                    radius=float(r) * 1.5, mach=float(mc)) for r, mc in zip(radii, machs)]
    dx, dy = ins[0][3], ins[0][4]
    m1 = np.stack([a[0] for a in ins])
    m2 = np.stack([a[1] for a in ins])
    phi = np.stack([a[2] for a in ins])
    tf = np.array([2e-5, 3e-5, 1e-5, 2.5e-5, 1.5e-5])
    base = None
    for streams in (1, 2, 4):
        ctx = S.Context(0)
        try:
            ctx.set_batch_streams(streams)
            print(f"batch of 5 on {streams} streams", flush=True)
            r = S.rgfm_solve_batch_2d(m1, m2, phi, dx, dy, 0.8, tf, problems.HELIUM, problems.AIR, ctx=ctx)
        finally:
            ctx.close()
        assert r.stats.steps == int(r.steps.sum()) and r.stats.nonfinite_cells == 0
        if base is None:
            base = r
            continue
        for name in ("material1", "material2", "level_set", "t", "circulation"):
            O.assert_bitwise(getattr(r, name), getattr(base, name), f"{name} on {streams} streams")
        assert np.array_equal(r.steps, base.steps)
    ctx = S.Context(0)
    try:
        for bad in (0, 17):
            with pytest.raises(VanguardError):
                ctx.set_batch_streams(bad)
    finally:
        ctx.close()


def test_rgfm_reduced_fields_give_the_same_bits():
    """Shock-bubble states hold w = +0 and lambda = 1 in both materials, so a
    run takes the 4-field sweeps; with the switch off it takes the 6-field
    ones.  Same bits in 1D and 2D (materials, level set, time steps), and the
    two constant planes come back unchanged."""
    m1, m2, phi, dx, dy = problems.shock_bubble_2d(64, 24, length=(0.325, 0.325 * 24 / 64),
                                                   center=(0.175, 0.325 * 12 / 64), radius=0.03, mach=1.25)
    for m in (m1, m2):
        assert np.all(m[..., 3].view(np.uint64) == 0) and np.all(m[..., 5] == 1.0)
    cases = {"2d": (m1, m2, phi, dy), "1d": (m1[12].copy(), m2[12].copy(), phi[12].copy(), None)}
    out = {}
    for on in (True, False):
        ctx = S.Context(0)
        try:
            ctx.set_reduced_fields(on)
            for name, (a, b, ph, ddy) in cases.items():
                print(f"rgfm {name}, reduced fields {on}", flush=True)
                out[name, on] = S.rgfm_solve(S.MultimaterialSystem(a, b, ph), dx, 0.8, 3e-5, problems.HELIUM,
                                             problems.AIR, dy=ddy, ctx=ctx)
        finally:
            ctx.close()
    for name in cases:
        a, b = out[name, True], out[name, False]
        assert a.stats.steps == b.stats.steps > 5
        O.assert_bitwise(a.dt, b.dt, f"{name} time steps")
        O.assert_bitwise(a.system.material1, b.system.material1, f"{name} material 1")
        O.assert_bitwise(a.system.material2, b.system.material2, f"{name} material 2")
        O.assert_bitwise(a.system.level_set, b.system.level_set, f"{name} level set")
        for m in (a.system.material1, a.system.material2):
            assert np.all(m[..., 3].view(np.uint64) == 0) and np.all(m[..., 5] == 1.0)


def test_rgfm_fast_mode_within_tolerance(vg_ctx):
    g = O.golden("rgfm_2d_bubble96")
    a = gpu_golden(g)
    b = gpu_golden(g, mode=S.FAST)
    assert rel_l1(b.system.material1, a.system.material1) <= REL_L1
    assert rel_l1(b.system.material2, a.system.material2) <= REL_L1


def test_rgfm_rejects_bad_arguments(vg_ctx):
    m1, m2, phi, dx, dy = bubble(16, 8)
    sysm = S.MultimaterialSystem(m1, m2, phi)
    r = S.rgfm_solve(sysm, dx, 0.8, math.inf, problems.HELIUM, problems.AIR, max_steps=0, raise_on_error=False)
    assert r.status == _lib.VG_ERR_ARG   # infinite final time without a step cap
    r = S.rgfm_solve(sysm, dx, 0.8, 1e-5, problems.HELIUM, problems.AIR, riemann=7, raise_on_error=False)
    assert r.status == _lib.VG_ERR_ARG
