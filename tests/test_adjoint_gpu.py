"""Hand-written adjoint of the Euler SLIC solve on the GPU (vg_euler_adjoint_*),
through the C ABI.

The derivative is that of the discrete map with every time step held at the
value the forward run chose.  Checks:

1. the adjoint's forward is the solve (final state and time steps bit for bit),
   and the fixed-time-step forward is the oracle's map bit for bit;
2. dot-product test against the oracle's forward-mode derivative (dual numbers,
   oracle/euler_dual.cpp): <grad, v> = <seed, J v> for random v, on smooth and
   discontinuous data, 1D and 2D, both boundary conditions, every limiter;
3. central finite differences of the GPU's own fixed-time-step forward, on
   smooth data, through a thin torch.autograd.Function (torch.autograd.gradcheck)
   and along random directions;
4. the checkpoint interval changes nothing (bit for bit).
"""

import math
import zlib

import numpy as np
import pytest
import torch

import oracle_lib as O
from test_oracle_adjoint import IDEAL, REACTIVE, local_direction, smooth_1d, smooth_2d, sod_1d, sod_2d
from vanguard_amd import solvers as S
from vanguard_amd._lib import VanguardError

pytestmark = pytest.mark.gpu


def mat(p):
    return S.EulerMaterialParameters(*p)


def run_adjoint(w, seed, *, dim, steps, limiter=0, bias=0.0, bc=0, params=IDEAL, cfl=0.6, k=0):
    return S.euler_adjoint(w, seed, dim, 1.0 / w.shape[-2], cfl, math.inf, bias=bias, limiter=limiter,
                           params=mat(params), max_steps=steps, bc=bc, checkpoint_every=k)


CASES = [
    dict(name="sod1d_superbee", w=lambda: sod_1d(96), dim=1, limiter=0, bias=0.0, bc=0, params=IDEAL, steps=30),
    dict(name="sod1d_vanleer_bias", w=lambda: sod_1d(80), dim=1, limiter=1, bias=0.4, bc=0, params=REACTIVE,
         steps=25),
    dict(name="smooth1d_minbee_periodic", w=lambda: smooth_1d(300), dim=1, limiter=2, bias=-0.3, bc=1,
         params=IDEAL, steps=40),
    dict(name="sod2d_superbee", w=lambda: sod_2d(24, 40), dim=2, limiter=0, bias=0.0, bc=0, params=IDEAL, steps=12),
    dict(name="smooth2d_vanleer_periodic", w=lambda: smooth_2d(36, 20), dim=2, limiter=1, bias=0.0, bc=1,
         params=REACTIVE, steps=10),
    dict(name="sod2d_minbee_tall", w=lambda: sod_2d(150, 9), dim=2, limiter=2, bias=0.2, bc=1, params=IDEAL,
         steps=8),
]


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_adjoint_forward_is_the_solve(case):
    w = case["w"]()
    dx = 1.0 / w.shape[-2]
    kw = dict(bias=case["bias"], limiter=case["limiter"], params=mat(case["params"]), bc=case["bc"])
    ref = S.euler_solve(w, case["dim"], dx, 0.6, math.inf, max_steps=case["steps"], mode=S.EXACT, **kw)
    res = run_adjoint(w, np.zeros_like(w), dim=case["dim"], steps=case["steps"], limiter=case["limiter"],
                      bias=case["bias"], bc=case["bc"], params=case["params"])
    O.assert_bitwise(res.dt, ref.dt, "time steps")
    O.assert_bitwise(res.cells, ref.cells, "final state")
    assert not np.any(res.grad)   # zero seed, zero gradient
    fixed = S.euler_solve_fixed_dt(w, case["dim"], dx, ref.dt, **kw)
    O.assert_bitwise(fixed, ref.cells, "fixed-dt forward")
    orc = O.euler_fixed_dt(w, ref.dt, dim=case["dim"], dx=dx, bias=case["bias"], limiter=case["limiter"],
                           params=case["params"], bc=case["bc"])
    O.assert_bitwise(fixed, orc, "fixed-dt forward vs oracle")


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_dot_product_against_dual_numbers(case):
    w = case["w"]()
    dx = 1.0 / w.shape[-2]
    rng = np.random.default_rng(zlib.crc32(case["name"].encode()))
    seed = rng.standard_normal(w.shape)
    res = run_adjoint(w, seed, dim=case["dim"], steps=case["steps"], limiter=case["limiter"], bias=case["bias"],
                      bc=case["bc"], params=case["params"])
    assert res.dt.size == case["steps"]
    for _ in range(3):
        v = rng.standard_normal(w.shape)
        _, jv = O.euler_fixed_dt(w, res.dt, dim=case["dim"], dx=dx, bias=case["bias"], limiter=case["limiter"],
                                 params=case["params"], bc=case["bc"], tangent=v)
        lhs = float(np.sum(res.grad * v))
        rhs = float(np.sum(seed * jv))
        scale = float(np.sum(np.abs(res.grad * v)))
        assert abs(lhs - rhs) <= 1e-10 * scale, (lhs, rhs, scale)
    # the planted-mistake counterpart: the gradient of a different seed fails
    bad = float(np.sum(run_adjoint(w, seed * (1 + 1e-6 * rng.standard_normal(w.shape)), dim=case["dim"],
                                   steps=case["steps"], limiter=case["limiter"], bias=case["bias"], bc=case["bc"],
                                   params=case["params"]).grad * v))
    assert abs(bad - rhs) > 1e-10 * scale


def test_single_step_adjoint_every_cell():
    """One 1D step on 6 cells: the adjoint of every unit seed against the full
    Jacobian from the dual numbers (every ghost fold and tile position)."""
    w = sod_1d(6)
    w[:, 1] = [0.3, -0.2, 0.1, 0.5, -0.4, 0.2]
    for bc in (0, 1):
        dts = S.euler_adjoint(w, np.zeros_like(w), 1, 1.0 / 6, 0.8, math.inf, params=IDEAL_M, max_steps=1, bc=bc).dt
        J = np.zeros((36, 36))
        for i in range(36):
            v = np.zeros(36)
            v[i] = 1.0
            J[:, i] = O.euler_fixed_dt(w, dts, dim=1, dx=1.0 / 6, bias=0.0, limiter=0, params=IDEAL, bc=bc,
                                       tangent=v.reshape(6, 6))[1].ravel()
        for o in range(36):
            seed = np.zeros(36)
            seed[o] = 1.0
            res = S.euler_adjoint(w, seed.reshape(6, 6), 1, 1.0 / 6, 0.8, math.inf, params=IDEAL_M, max_steps=1,
                                  bc=bc)
            O.assert_bitwise(res.dt, dts, "time step")
            np.testing.assert_allclose(res.grad.ravel(), J[o], rtol=1e-11, atol=1e-11 * np.abs(J).max(),
                                       err_msg=f"bc {bc} output {o}")


IDEAL_M = mat(IDEAL)


class SlicSolveFrozenDt(torch.autograd.Function):
    """Thin test harness: forward = the GPU fixed-time-step solve, backward =
    the GPU adjoint (which re-derives the time steps by CFL at the saved input
    and must reproduce the frozen ones)."""

    @staticmethod
    def forward(ctx, prim, dts, dim, dx, cfl, limiter, bias, bc):
        w = prim.detach().cpu().numpy()
        out = S.euler_solve_fixed_dt(w, dim, dx, dts.cpu().numpy(), bias=bias, limiter=limiter, params=IDEAL_M,
                                     bc=bc)
        ctx.save_for_backward(prim, dts)
        ctx.args = (dim, dx, cfl, limiter, bias, bc)
        return torch.from_numpy(out)

    @staticmethod
    def backward(ctx, gout):
        prim, dts = ctx.saved_tensors
        dim, dx, cfl, limiter, bias, bc = ctx.args
        res = S.euler_adjoint(prim.detach().cpu().numpy(), gout.detach().cpu().numpy(), dim, dx, cfl, math.inf,
                              bias=bias, limiter=limiter, params=IDEAL_M, max_steps=dts.numel(), bc=bc)
        O.assert_bitwise(res.dt, dts.cpu().numpy(), "adjoint time steps")
        return torch.from_numpy(res.grad), None, None, None, None, None, None, None


@pytest.mark.parametrize("dim,limiter,bc", [(1, 0, 1), (1, 2, 0), (2, 1, 1)])
def test_gradcheck_central_differences(dim, limiter, bc):
    w = smooth_1d(10) if dim == 1 else smooth_2d(5, 4)
    dx, cfl, steps = 1.0 / w.shape[-2], 0.5, 4
    dts = S.euler_solve(w, dim, dx, cfl, math.inf, max_steps=steps, limiter=limiter, params=IDEAL_M, bc=bc).dt
    x = torch.from_numpy(w.copy()).requires_grad_(True)
    fn = lambda p: SlicSolveFrozenDt.apply(p, torch.from_numpy(dts), dim, dx, cfl, limiter, 0.0, bc)
    # eps keeps every perturbation below the 1e-5 slope clamp and away from limiter kinks
    assert torch.autograd.gradcheck(fn, (x,), eps=1e-8, atol=1e-6, rtol=1e-6, nondet_tol=0.0)


@pytest.mark.parametrize("make,dim,bc,steps", [(lambda: smooth_1d(200), 1, 1, 15), (lambda: sod_1d(100), 1, 0, 20),
                                                (lambda: smooth_2d(40, 32), 2, 1, 15),
                                                (lambda: sod_2d(24, 20), 2, 0, 8)])
def test_directional_central_differences(make, dim, bc, steps):
    """<grad, v> against central differences of <seed, .> through the GPU's own
    fixed-time-step forward, for local directions v whose perturbed runs take
    the same limiter and clamp branches as the base run (the oracle's branch
    hash certifies it), to 1e-6 of the magnitude of the sum's terms."""
    w = make()
    dx = 1.0 / w.shape[-2]
    rng = np.random.default_rng(steps + 10 * dim)
    seed = rng.standard_normal(w.shape)
    res = run_adjoint(w, seed, dim=dim, steps=steps, limiter=0, bc=bc)
    okw = dict(dim=dim, dx=dx, bias=0.0, limiter=0, params=IDEAL, bc=bc)
    h, h0, used = 1e-8, O.euler_branch_hash(w, res.dt, **okw), 0
    for _ in range(6):
        v = local_direction(rng, w.shape, dim)
        if not (O.euler_branch_hash(w + h * v, res.dt, **okw) == h0 == O.euler_branch_hash(w - h * v, res.dt, **okw)):
            continue
        fp = S.euler_solve_fixed_dt(w + h * v, dim, dx, res.dt, params=IDEAL_M, bc=bc)
        fm = S.euler_solve_fixed_dt(w - h * v, dim, dx, res.dt, params=IDEAL_M, bc=bc)
        fd = float(np.sum(seed * (fp - fm))) / (2 * h)
        ad = float(np.sum(res.grad * v))
        _, jv = O.euler_fixed_dt(w, res.dt, tangent=v, **okw)   # the sum's terms, for the scale
        assert abs(fd - ad) <= 1e-6 * float(np.sum(np.abs(seed * jv))), (fd, ad)
        used += 1
    assert used >= 3


@pytest.mark.parametrize("dim", [1, 2])
def test_checkpoint_interval_changes_nothing(dim):
    w = sod_1d(120) if dim == 1 else sod_2d(20, 28)
    seed = np.random.default_rng(4).standard_normal(w.shape)
    steps = 23 if dim == 1 else 11
    base = run_adjoint(w, seed, dim=dim, steps=steps, k=0)
    for k in (1, 2, 5, steps, steps + 7):
        other = run_adjoint(w, seed, dim=dim, steps=steps, k=k)
        O.assert_bitwise(other.grad, base.grad, f"checkpoint every {k}")
        O.assert_bitwise(other.cells, base.cells, f"final state, checkpoint every {k}")


@pytest.mark.parametrize("dim", [1, 2])
def test_recorded_forward_stops_at_the_final_time(dim):
    """With a step cap the forward writes its sweeps into recording buffers:
    every sweep input (checkpoint_every 0) or every step's start (1).  A final
    time reached before the cap ends that forward early, and both results are
    the checkpointed adjoint's bit for bit."""
    w = sod_1d(120) if dim == 1 else sod_2d(20, 28)
    seed = np.random.default_rng(5).standard_normal(w.shape)
    dx = 1.0 / w.shape[-2]
    t_final = 0.02
    kw = dict(params=mat(IDEAL), max_steps=500)
    ref = S.euler_adjoint(w, seed, dim, dx, 0.6, t_final, checkpoint_every=4, **kw)
    for k in (0, 1):
        r = S.euler_adjoint(w, seed, dim, dx, 0.6, t_final, checkpoint_every=k, **kw)
        assert 0 < r.stats.steps < 500 and abs(float(np.sum(r.dt)) - t_final) < 1e-15
        O.assert_bitwise(r.dt, ref.dt, f"time steps, layout {k}")
        O.assert_bitwise(r.cells, ref.cells, f"final state, layout {k}")
        O.assert_bitwise(r.grad, ref.grad, f"gradient, layout {k}")


@pytest.mark.parametrize("make,full", [(lambda: smooth_2d(36, 20), 5), (lambda: planar(smooth_2d(36, 20)), 3),
                                       (lambda: sod_2d(20, 28), -1)])
def test_partly_recorded_forward_gives_the_same_bits(make, full):
    """Under a device memory limit the recorded forward keeps every step's start
    and the inner sweep inputs of only its last `full` steps (the reverse pass
    recomputes the others'); a limit below the step starts (full = -1) falls
    back to checkpoints.  Time steps, final state and gradient are the
    unlimited run's bit for bit, for 6-field and 4-field grids (whose recording
    buffers hold planes 0-4 only)."""
    w = make()
    ny, nx = w.shape[:2]
    steps = 12
    seed = np.random.default_rng(7).standard_normal(w.shape)
    seed[..., 3] = 0.0
    seed[..., 5] = 0.0
    dx = 1.0 / nx
    kw = dict(params=mat(IDEAL), max_steps=steps, checkpoint_every=0)
    ref = S.euler_adjoint(w, seed, 2, dx, 0.6, math.inf, **kw)
    plane = (ny * nx + 63) // 64 * 64
    reduced = not np.any(w[..., 3]) and np.all(w[..., 5] == 1.0)
    rb, gb = (5 if reduced else 6) * plane * 8, 6 * plane * 8
    # record_layout: limit - 2 grids - 64 MiB left for (steps + 2) starts and scratch + 2 per full step
    limit = 64 * 2 ** 20 + 2 * gb + (steps + 2) * rb + (2 * full + 1) * rb if full >= 0 else 64 * 2 ** 20
    ctx = S.Context(0)
    ctx.set_adjoint_memory(limit)
    r = S.euler_adjoint(w, seed, 2, dx, 0.6, math.inf, ctx=ctx, **kw)
    assert r.stats.steps == steps
    O.assert_bitwise(r.dt, ref.dt, "time steps")
    O.assert_bitwise(r.cells, ref.cells, "final state")
    O.assert_bitwise(r.grad, ref.grad, "gradient")


def planar(w):
    w = np.array(w)
    w[..., 3] = 0.0
    w[..., 5] = 1.0
    return w


@pytest.mark.parametrize("make,dim,limiter,bias,bc,k", [
    (lambda: sod_1d(150), 1, 0, 0.0, 0, 0),
    (lambda: planar(smooth_1d(200)), 1, 1, 0.3, 1, 3),
    (lambda: sod_2d(24, 40), 2, 0, 0.0, 0, 0),
    (lambda: planar(smooth_2d(36, 20)), 2, 2, -0.2, 1, 1),
    (lambda: planar(smooth_2d(30, 30)), 2, 1, 0.0, 0, 4),
])
def test_four_field_adjoint_equals_six_field(make, dim, limiter, bias, bc, k):
    """With w = +0, lambda = 1, E_f = 0 and a seed whose w and lambda parts are
    zero, the adjoint sweeps run on the four fields (rho, u, v, p); with the
    context's 4-field switch off the six-field sweeps run.  The gradients are
    the same numbers (zeros may differ in sign), for every recording layout."""
    w = make()
    seed = np.random.default_rng(6).standard_normal(w.shape)
    seed[..., 3] = 0.0
    seed[..., 5] = 0.0
    dx = 1.0 / w.shape[-2]
    kw = dict(bias=bias, limiter=limiter, params=mat(IDEAL), max_steps=9, bc=bc, checkpoint_every=k)
    ctx6 = S.Context(0)
    ctx6.set_reduced_fields(False)
    six = S.euler_adjoint(w, seed, dim, dx, 0.6, math.inf, ctx=ctx6, **kw)
    four = S.euler_adjoint(w, seed, dim, dx, 0.6, math.inf, **kw)
    O.assert_bitwise(four.dt, six.dt, "time steps")
    O.assert_bitwise(four.cells, six.cells, "final state")
    assert np.array_equal(four.grad, six.grad), float(np.max(np.abs(four.grad - six.grad)))
    nz = four.grad != 0.0
    O.assert_bitwise(four.grad[nz], six.grad[nz], "nonzero gradient entries")
    assert not np.any(four.grad[..., 3]) and not np.any(four.grad[..., 5])
    assert np.count_nonzero(four.grad[..., 0]) > 0


def test_adjoint_tile_variants_change_nothing():
    """Every tile variant of the adjoint sweeps (shape and occupancy target)
    gives the same gradient bit for bit (each output is computed with the same
    arithmetic), and the autotuner records its picks under the grid's keys."""
    w = sod_2d(70, 45)
    seed = np.random.default_rng(12).standard_normal(w.shape)
    w1 = sod_1d(300)
    seed1 = np.random.default_rng(13).standard_normal(w1.shape)
    ctx = S.Context(0)
    try:
        kx, ky, k1 = "euler_adjoint_x/70/45", "euler_adjoint_y/70/45", "euler_adjoint_x/1/300"
        base = base1 = None
        for vx in range(6):
            for vy in range(6):
                ctx.set_tuned_key(kx, vx)
                ctx.set_tuned_key(ky, vy)
                r = S.euler_adjoint(w, seed, 2, 1.0 / 45, 0.6, math.inf, params=IDEAL_M, max_steps=6, bc=1, ctx=ctx)
                if base is None:
                    base = r.grad
                O.assert_bitwise(r.grad, base, f"tile variants x {vx} y {vy}")
        # the same for the 4-field adjoint sweeps (seed without w and lambda parts)
        seedp = seed.copy()
        seedp[..., 3] = 0.0
        seedp[..., 5] = 0.0
        px, py = "euler_planar_adjoint_x/70/45", "euler_planar_adjoint_y/70/45"
        basep = None
        for v in range(6):
            for vy in (v, 5 - v):
                ctx.set_tuned_key(px, v)
                ctx.set_tuned_key(py, vy)
                r = S.euler_adjoint(w, seedp, 2, 1.0 / 45, 0.6, math.inf, params=IDEAL_M, max_steps=6, bc=1, ctx=ctx)
                if basep is None:
                    basep = r.grad
                O.assert_bitwise(r.grad, basep, f"4-field tile variants x {v} y {vy}")
        for vx in range(6):
            ctx.set_tuned_key(k1, vx)
            r1 = S.euler_adjoint(w1, seed1, 1, 1.0 / 300, 0.6, math.inf, params=IDEAL_M, max_steps=9, ctx=ctx)
            if base1 is None:
                base1 = r1.grad
            O.assert_bitwise(r1.grad, base1, f"1D tile variant {vx}")
        for k in (kx, ky):
            ctx.set_tuned_key(k, -1)
            assert ctx.tuned_key(k) == -1
        S.euler_adjoint(w, seed, 2, 1.0 / 45, 0.6, math.inf, params=IDEAL_M, max_steps=2, bc=1, ctx=ctx)
        assert 0 <= ctx.tuned_key(kx) <= 5 and 0 <= ctx.tuned_key(ky) <= 5
    finally:
        ctx.close()


@pytest.mark.parametrize("dim", [1, 2])
def test_device_entry_points_match_host(dim):
    """vg_euler_adjoint_*_device on device arrays gives the host entry point's
    results bit for bit, and a second call reuses the context's workspace."""
    w = sod_1d(150) if dim == 1 else sod_2d(30, 26)
    seed = np.random.default_rng(31).standard_normal(w.shape)
    shape = w.shape[:-1]
    host = run_adjoint(w, seed, dim=dim, steps=13, bc=1)
    wd, sd = torch.from_numpy(w).cuda(), torch.from_numpy(seed).cuda()
    od, gd = torch.empty_like(wd), torch.empty_like(wd)
    for _ in range(2):
        dts, st = S.euler_adjoint_device(wd.data_ptr(), od.data_ptr(), sd.data_ptr(), gd.data_ptr(), shape,
                                         1.0 / w.shape[-2], 0.6, math.inf, params=IDEAL_M, max_steps=13, bc=1)
        O.assert_bitwise(dts, host.dt, "time steps")
        O.assert_bitwise(od.cpu().numpy(), host.cells, "final state")
        O.assert_bitwise(gd.cpu().numpy(), host.grad, "gradient")
    S.default_context().release_workspace()


def test_adjoint_to_final_time():
    """Stop at t_final (the last step is cut to land on it): the recorded steps
    sum to t_final and the gradient still passes the dot-product test."""
    w = sod_1d(64)
    seed = np.random.default_rng(8).standard_normal(w.shape)
    res = S.euler_adjoint(w, seed, 1, 1.0 / 64, 0.8, 0.05, params=IDEAL_M)
    assert res.dt.size > 5
    assert abs(float(np.sum(res.dt)) - 0.05) < 1e-15
    v = np.random.default_rng(9).standard_normal(w.shape)
    _, jv = O.euler_fixed_dt(w, res.dt, dim=1, dx=1.0 / 64, bias=0.0, limiter=0, params=IDEAL, bc=0, tangent=v)
    lhs, rhs = float(np.sum(res.grad * v)), float(np.sum(seed * jv))
    assert abs(lhs - rhs) <= 1e-10 * float(np.sum(np.abs(res.grad * v)))


def test_adjoint_rejects_bad_arguments():
    w = sod_1d(16)
    with pytest.raises(VanguardError):
        S.euler_adjoint(w, np.zeros_like(w), 1, 1.0 / 16, 0.8, math.inf, params=IDEAL_M, max_steps=0)
    with pytest.raises(ValueError):
        S.euler_adjoint(w, np.zeros((15, 6)), 1, 1.0 / 16, 0.8, 0.1, params=IDEAL_M)
