"""Euler SLIC/FORCE on the GPU through the C ABI: parity with the reference.

EXACT mode must be bit-identical to the golden vectors of the compiled
reference and to the oracle (time steps and final states); FAST mode must agree
within 1e-10 relative L1 (the parity bar for conserved fields in FAST mode).
"""

import math

import numpy as np
import pytest

import oracle_lib as O
from vanguard_amd import _lib, problems
from vanguard_amd import solvers as S

pytestmark = pytest.mark.gpu

RUNS = ["euler_c1_sod1000", "euler_toro100", "euler_1d_stiff_bias", "euler_force1d", "euler_2d_disc64_l0",
        "euler_2d_disc64_l1", "euler_2d_disc64_l2", "euler_2d_c2ic64", "euler_2d_rect", "euler_force2d",
        "euler_2d_api32"]
IDEAL = S.EulerMaterialParameters(1.4)


def gpu_run(g, mode=S.EXACT, **over):
    m = dict(g["meta"])
    m.update(over)
    return S.euler_solve(g["prim_in"], m["dim"], m["dx"], m["cfl"], m["tfinal"], order=m["order"], bias=m["bias"],
                         limiter=m["limiter"], params=S.EulerMaterialParameters(*m["params"]),
                         max_steps=m["max_steps"], mode=mode)


def cons(w, gamma=1.4, p_inf=0.0, e_form=0.0):
    rho, u, v, z, p, lam = np.moveaxis(w, -1, 0)
    e = (p + gamma * p_inf) / ((gamma - 1) * rho)
    E = rho * (0.5 * (u * u + v * v + z * z) + e + e_form * (lam - 1))
    return np.stack([rho, rho * u, rho * v, rho * z, E, rho * lam], -1)


def rel_l1(a, b):
    return np.abs(a - b).sum() / np.abs(b).sum()


@pytest.mark.parametrize("name", RUNS)
def test_exact_mode_is_bit_identical_to_reference(vg_ctx, name):
    g = O.golden(name)
    r = gpu_run(g)
    print(f"{name}: steps={r.stats.steps}", flush=True)
    O.assert_bitwise(r.dt, g["dt"], f"{name} dt")
    O.assert_bitwise(r.cells, g["prim_out"], f"{name} state")


@pytest.mark.parametrize("name", ["euler_c1_sod1000", "euler_1d_stiff_bias", "euler_2d_c2ic64", "euler_2d_rect",
                                  "euler_force2d"])
def test_fast_mode_within_tolerance(vg_ctx, name):
    g = O.golden(name)
    m = g["meta"]
    gam, pinf, _, ef, _, _ = m["params"]
    r = gpu_run(g, mode=S.FAST)
    assert r.stats.steps == len(g["dt"])
    assert np.abs(r.dt - g["dt"]).max() <= 1e-10 * np.abs(g["dt"]).max()
    a = cons(r.cells, gam, pinf, ef)
    b = cons(g["prim_out"], gam, pinf, ef)
    for f in range(6):
        if np.abs(b[..., f]).sum() > 0:
            assert rel_l1(a[..., f], b[..., f]) <= 1e-10, (name, f, rel_l1(a[..., f], b[..., f]))


@pytest.mark.parametrize("shape,bc,order,limiter,bias", [
    ((37, 253), 0, 2, 0, 0.0),     # ragged: partial X tiles and Y segments
    ((300, 7), 0, 2, 1, 0.25),
    ((5, 520), 1, 2, 2, -0.3),     # periodic, X tile wrap
    ((70, 66), 1, 1, 0, 0.0),      # periodic FORCE
    ((2, 2), 0, 2, 0, 0.0),        # smallest grid
    ((33, 257), 0, 1, 0, 0.0),
])
def test_ragged_grids_match_oracle(vg_ctx, shape, bc, order, limiter, bias):
    ny, nx = shape
    rng = np.random.default_rng(ny * 1000 + nx)
    w = np.zeros((ny, nx, 6))
    w[..., 0] = rng.uniform(0.3, 1.5, shape)
    w[..., 1:4] = rng.uniform(-0.4, 0.4, shape + (3,))
    w[..., 4] = rng.uniform(0.3, 1.5, shape)
    w[..., 5] = rng.uniform(0.5, 1.0, shape)
    params = [1.6, 0.05, 25, 0.3, 0.25, 250]
    kw = dict(order=order, dim=2, dx=0.01, cfl=0.7, tfinal=math.inf, max_steps=8, bias=bias, limiter=limiter,
              params=params, bc=bc)
    ref, dref = O.euler_solve(w, **kw)
    r = S.euler_solve(w, 2, 0.01, 0.7, math.inf, order=order, bias=bias, limiter=limiter,
                      params=S.EulerMaterialParameters(*params), max_steps=8, bc=bc)
    O.assert_bitwise(r.dt, dref, "dt")
    O.assert_bitwise(r.cells, ref, "state")


@pytest.mark.parametrize("n,bc,order", [(2, 0, 2), (3, 1, 2), (255, 0, 2), (1000, 1, 1), (4097, 0, 2)])
def test_1d_sizes_match_oracle(vg_ctx, n, bc, order):
    rng = np.random.default_rng(n)
    w = np.zeros((n, 6))
    w[:, 0] = rng.uniform(0.3, 1.5, n)
    w[:, 1:4] = rng.uniform(-0.4, 0.4, (n, 3))
    w[:, 4] = rng.uniform(0.3, 1.5, n)
    w[:, 5] = 1.0
    kw = dict(order=order, dim=1, dx=1.0 / n, cfl=0.9, tfinal=math.inf, max_steps=12, bias=0.0, limiter=2,
              params=[1.4, 0, 25, 0, 0.25, 250], bc=bc)
    ref, dref = O.euler_solve(w, **kw)
    r = S.euler_solve(w, 1, 1.0 / n, 0.9, math.inf, order=order, limiter=2, max_steps=12, bc=bc)
    O.assert_bitwise(r.dt, dref, "dt")
    O.assert_bitwise(r.cells, ref, "state")


def test_final_time_and_step_cap_semantics(vg_ctx):
    w = problems.sod_1d(100)
    r0 = S.euler_solve(w, 1, 0.01, 0.8, 0.0)             # t_final <= 0: no step at all
    assert r0.stats.steps == 0
    O.assert_bitwise(r0.cells, w, "untouched")
    r = S.euler_solve(w, 1, 0.01, 0.8, 0.25, max_steps=10)
    g = O.golden("euler_toro100")
    O.assert_bitwise(r.dt, g["dt"][:10], "capped dt prefix")
    full = S.euler_solve(w, 1, 0.01, 0.8, 0.25)
    assert full.stats.t == 0.25 and full.stats.steps == len(g["dt"])


def test_device_state_runs_compose(vg_ctx):
    """Two runs of 5 steps on a device-resident grid equal one run of 10."""
    w = problems.disc_2d(48, gaussians=4, length=1.0)
    a = S.EulerGrid(2, 48, 48)
    a.load(w)
    st1, d1 = a.run(1 / 48, 0.8, max_steps=5, dt_cap=5)
    st2, d2 = a.run(1 / 48, 0.8, max_steps=5, dt_cap=5)
    assert st2.total_steps == 10
    one = S.euler_solve(w, 2, 1 / 48, 0.8, math.inf, max_steps=10)
    O.assert_bitwise(np.concatenate([d1, d2]), one.dt, "dt")
    O.assert_bitwise(a.store(), one.cells, "state")
    s = a.max_wave_speed()
    lib = O.load()
    ref = lib.orc_euler_max_wave_speed(2, 48, 48, O.OrcEulerParams(1.4, 0, 25, 0, 0.25, 250), O._p(one.cells))
    assert s == max(ref, math.sqrt(1.4))
    a.close()


def test_nonfinite_state_is_reported(vg_ctx):
    w = problems.sod_1d(64)
    w[40, 4] = -5.0          # negative pressure -> NaN sound speed and state
    r = S.euler_solve(w, 1, 1 / 64, 0.5, math.inf, max_steps=3, raise_on_error=False)
    assert r.status == _lib.VG_ERR_NONFINITE
    assert r.stats.nonfinite_cells > 0 and 30 <= r.stats.first_nonfinite <= 50


def test_invalid_arguments(vg_ctx):
    w = problems.sod_1d(16)
    for kw in [dict(order=3), dict(mode=7), dict(bc=5)]:
        r = S.euler_solve(w, 1, 0.1, 0.5, 0.1, raise_on_error=False, **kw)
        assert r.status == _lib.VG_ERR_ARG
    r = S.euler_solve(w, 1, 0.1, 0.5, math.inf, raise_on_error=False)   # endless run refused
    assert r.status == _lib.VG_ERR_ARG
    r = S.euler_solve(w[:1], 1, 0.1, 0.5, 0.1, raise_on_error=False)    # one cell: no mirror ghosts
    assert r.status == _lib.VG_ERR_ARG


@pytest.mark.parametrize("dim,order,mode", [(2, 2, S.EXACT), (2, 1, S.EXACT), (1, 2, S.EXACT), (2, 2, S.FAST)])
def test_reduced_field_sweeps_give_the_full_state_bits(vg_ctx, dim, order, mode):
    """A grid at rest out of plane (w = +0) with lambda = 1 everywhere is swept
    by the 4-field kernels; the 6-field kernels (the context switch off) give
    the same bits in EXACT mode, and both keep w = +0 and lambda = 1 exactly
    (FAST: within 1e-13).
    One cell off the constants sends the run back to the 6-field kernels."""
    w = problems.disc_2d(96)
    if dim == 1:
        w = np.ascontiguousarray(w[48])
    dx = 2.0 / 96
    full = S.Context()
    try:
        full.set_reduced_fields(False)
        kw = dict(order=order, max_steps=30, mode=mode, bc=0)
        a = S.euler_solve(w, dim, dx, 0.8, math.inf, ctx=full, **kw)
        b = S.euler_solve(w, dim, dx, 0.8, math.inf, **kw)
        if mode == S.EXACT:
            O.assert_bitwise(b.dt, a.dt, "dt")
            O.assert_bitwise(b.cells, a.cells, "state")
        else:
            assert rel_l1(cons(b.cells), cons(a.cells)) <= 1e-13
        # FAST mode's 6-field lambda = (rho lambda) rcp(rho) is 1 to within an ulp
        for r in ((a, b) if mode == S.EXACT else (b,)):
            assert not np.any(r.cells[..., 3].view(np.uint64)), "w must stay +0"
            assert np.all(r.cells[..., 5] == 1.0), "lambda must stay 1"
        w2 = w.copy()
        w2[(0,) * (w2.ndim - 1) + (5,)] = 0.999
        c = S.euler_solve(w2, dim, dx, 0.8, math.inf, **kw)
        d = S.euler_solve(w2, dim, dx, 0.8, math.inf, ctx=full, **kw)
        if mode == S.EXACT:
            O.assert_bitwise(c.cells, d.cells, "state with a reacting cell")
            ref, dref = O.euler_solve(w2, order=order, dim=dim, dx=dx, cfl=0.8, tfinal=math.inf, max_steps=30,
                                      bias=0.0, limiter=0, params=[1.4, 0.0, 25.0, 0.0, 0.25, 250.0], bc=0)
            O.assert_bitwise(c.cells, ref, "state with a reacting cell vs the oracle")
    finally:
        full.close()


def _write_plane(ptr, plane, f, values):
    """Host values into plane f of a grid's device planes (synchronous copy)."""
    import ctypes
    try:
        hip = ctypes.CDLL("libamdhip64.so")
    except OSError:
        hip = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so")
    src = np.ascontiguousarray(values, dtype=np.float64).ravel()
    dst = ctypes.c_void_p(ptr + 8 * f * plane)
    rc = hip.hipMemcpy(dst, src.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(src.nbytes), 1)
    assert rc == 0, rc


def test_planar_check_is_redone_after_a_load_or_a_planes_handout():
    """A grid remembers that its constant planes hold w = +0 and lambda = 1 and
    skips the check on later runs.  Loading a state with w != 0, or writing w
    through the planes pointer between runs, must send the next run back to
    the 6-field sweeps: the same bits as a context that never takes the 4-field
    ones."""
    n, dx = 64, 2.0 / 64
    w = problems.disc_2d(n)
    w2 = w.copy()
    w2[..., 3] = 0.05 * np.sin(2 * np.pi * (np.arange(n) + 0.5) / n)[None, :]
    full = S.Context()
    try:
        full.set_reduced_fields(False)
        grids = [S.EulerGrid(2, n, n), S.EulerGrid(2, n, n, ctx=full)]
        out = []
        for g in grids:
            g.load(w)
            g.run(dx, 0.8, max_steps=3)
            g.load(w2)
            g.run(dx, 0.8, max_steps=3)
            first = g.store()
            g.load(w)
            g.run(dx, 0.8, max_steps=3)
            ptr, plane = g.planes()
            _write_plane(ptr, plane, 3, w2[..., 3])
            g.run(dx, 0.8, max_steps=3)
            out.append((first, g.store()))
            g.close()
        O.assert_bitwise(out[0][0], out[1][0], "after loading a state with w != 0")
        O.assert_bitwise(out[0][1], out[1][1], "after writing w through the planes pointer")
        assert np.any(out[0][1][..., 3] != 0.0)
    finally:
        full.close()


def test_headline_grid_exact_mode_matches_oracle(vg_ctx):
    """The 4096 x 4096 cylindrical-explosion workload (64 Gaussian bumps, superbee,
    CFL 0.8 on [0, 2]^2), two steps in EXACT mode through the whole-solve entry
    point, against the oracle on the host: time steps and state bit for bit."""
    n = 4096
    w = problems.disc_2d(n)
    kw = dict(order=2, dim=2, dx=2.0 / n, cfl=0.8, tfinal=math.inf, max_steps=2, bias=0.0, limiter=0,
              params=[1.4, 0.0, 25.0, 0.0, 0.25, 250.0])   # EulerMaterialParameters(1.4)
    r = S.euler_solve(w, 2, 2.0 / n, 0.8, math.inf, limiter=0, max_steps=2, mode=S.EXACT)
    print(f"GPU done: dt {r.dt}", flush=True)
    ref, dref = O.euler_solve(w, **kw)
    O.assert_bitwise(r.dt, dref, "dt")
    O.assert_bitwise(r.cells, ref, "state")
