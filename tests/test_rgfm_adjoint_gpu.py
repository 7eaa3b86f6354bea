"""Adjoint of the HLLC RGFM solve on the GPU (vg_rgfm_adjoint_*), through the
C ABI, time steps frozen.

1. The adjoint's forward is the solve, and the fixed-time-step forward is the
   oracle's map, bit for bit.
2. Dot-product test against the oracle's forward-mode derivative
   (orc_rgfm_fixed_dt_jvp, dual numbers): <grad, v> = <seed, J v> for random v.
3. The circulation-versus-Mach derivative of the shock-bubble problem: the
   adjoint of Gamma contracted with d(initial state)/dM, against the dual
   numbers and against central differences of Gamma(M) through the GPU's own
   fixed-time-step forward.
4. The checkpoint interval changes nothing (bit for bit).
5. With w = +0 and lambda = 1 in both materials and seeds without w and lambda
   parts the adjoint runs on four fields; its gradients are the 6-field ones.
"""

import math

import numpy as np
import pytest

import oracle_lib as O
from test_oracle_adjoint import rgfm_case, rgfm_parameter_directions
from vanguard_amd import problems
from vanguard_amd import solvers as S
from vanguard_amd.problems import AIR, HELIUM

pytestmark = pytest.mark.gpu

SCALE = np.array([0.1, 10.0, 10.0, 10.0, 1e4, 0.1])   # field sizes of the states


def gpu_solve(case, steps, dim):
    m1, m2, phi, dx, dy = case
    return S.rgfm_solve(S.MultimaterialSystem(m1, m2, phi), dx, 0.8, math.inf, HELIUM, AIR, dy=dy, max_steps=steps)


def run_adjoint(case, s1, s2, steps, k=0):
    m1, m2, phi, dx, dy = case
    return S.rgfm_adjoint(S.MultimaterialSystem(m1, m2, phi), s1, s2, dx, 0.8, math.inf, HELIUM, AIR, dy=dy,
                          max_steps=steps, checkpoint_every=k)


def circulation_seed(m1, phi, dx, dy):
    """d Gamma / d (final material 1), Gamma = sum over interior cells with
    phi < 0 of (dv/dx - du/dy) dx dy by central differences (the circulation
    kernel's definition)."""
    ny, nx = phi.shape
    mask = np.zeros((ny, nx))
    mask[1:-1, 1:-1] = (phi[1:-1, 1:-1] < 0.0)
    s = np.zeros_like(m1)
    # dv/dx: v_{j+1} - v_{j-1} over 2 dx, times dx dy
    s[:, :-2, 2] -= 0.5 * dy * mask[:, 1:-1]
    s[:, 2:, 2] += 0.5 * dy * mask[:, 1:-1]
    # - du/dy
    s[:-2, :, 1] += 0.5 * dx * mask[1:-1, :]
    s[2:, :, 1] -= 0.5 * dx * mask[1:-1, :]
    return s


@pytest.mark.parametrize("dim,tie", [(1, True), (1, False), (2, True), (2, False)])
def test_forward_and_fixed_dt_are_the_solve(dim, tie):
    case = rgfm_case(dim, tie=tie)
    ref = gpu_solve(case, 25, dim)
    m1, m2, phi, dx, dy = case
    fixed = S.rgfm_solve_fixed_dt(S.MultimaterialSystem(m1, m2, phi), dx, ref.dt, HELIUM, AIR, dy=dy)
    O.assert_bitwise(fixed.material1, ref.system.material1, "fixed-dt material 1")
    O.assert_bitwise(fixed.material2, ref.system.material2, "fixed-dt material 2")
    O.assert_bitwise(fixed.level_set, ref.system.level_set, "fixed-dt level set")
    o1, o2, oph = O.rgfm_fixed_dt(m1, m2, phi, ref.dt, dim=dim, dx=dx, dy=dy, bias=0.0, limiter=0, p1=HELIUM,
                                  p2=AIR)
    O.assert_bitwise(fixed.material1, o1, "fixed-dt vs oracle 1")
    O.assert_bitwise(fixed.material2, o2, "fixed-dt vs oracle 2")
    adj = run_adjoint(case, np.zeros_like(m1), np.zeros_like(m2), 25)
    O.assert_bitwise(adj.dt, ref.dt, "adjoint time steps")
    O.assert_bitwise(adj.system.material1, ref.system.material1, "adjoint forward 1")
    O.assert_bitwise(adj.system.material2, ref.system.material2, "adjoint forward 2")
    assert not np.any(adj.grad1) and not np.any(adj.grad2)


@pytest.mark.parametrize("dim,tie,steps", [(1, True, 30), (1, False, 30), (2, True, 20), (2, False, 20)])
def test_dot_product_against_dual_numbers(dim, tie, steps):
    case = rgfm_case(dim, tie=tie)
    m1, m2, phi, dx, dy = case
    rng = np.random.default_rng(60 + 2 * dim + int(tie))
    s1, s2 = rng.standard_normal(m1.shape) / SCALE, rng.standard_normal(m2.shape) / SCALE
    adj = run_adjoint(case, s1, s2, steps)
    assert adj.dt.size == steps
    for _ in range(3):
        v1, v2 = rng.standard_normal(m1.shape) * SCALE, rng.standard_normal(m2.shape) * SCALE
        _, _, _, t1, t2 = O.rgfm_fixed_dt(m1, m2, phi, adj.dt, dim=dim, dx=dx, dy=dy, bias=0.0, limiter=0,
                                          p1=HELIUM, p2=AIR, tangent=(v1, v2))
        lhs = float(np.sum(adj.grad1 * v1) + np.sum(adj.grad2 * v2))
        rhs = float(np.sum(s1 * t1) + np.sum(s2 * t2))
        scale = float(np.sum(np.abs(adj.grad1 * v1)) + np.sum(np.abs(adj.grad2 * v2)))
        assert abs(lhs - rhs) <= 1e-9 * scale, (lhs, rhs, scale)


@pytest.mark.parametrize("tie", [True, False])
def test_circulation_mach_derivative(tie):
    """dGamma/dM of the shock-bubble problem (time steps frozen): adjoint vs
    dual numbers (1e-9) and vs central differences of the GPU's fixed-step
    forward, whose branches the oracle's hash certifies unchanged (1e-6)."""
    (m1, m2, phi, dx, dy), dirs = rgfm_parameter_directions(2, tie=tie)
    _, v1, v2 = [d for d in dirs if d[0] == "mach"][0]
    ref = gpu_solve((m1, m2, phi, dx, dy), 30, 2)
    seed1 = circulation_seed(ref.system.material1, ref.system.level_set, dx, dy)
    adj = run_adjoint((m1, m2, phi, dx, dy), seed1, np.zeros_like(m2), 30)
    dgamma = float(np.sum(adj.grad1 * v1) + np.sum(adj.grad2 * v2))
    _, _, _, t1, _ = O.rgfm_fixed_dt(m1, m2, phi, adj.dt, dim=2, dx=dx, dy=dy, bias=0.0, limiter=0, p1=HELIUM,
                                     p2=AIR, tangent=(v1, v2))
    dual = float(np.sum(seed1 * t1))
    assert abs(dgamma - dual) <= 1e-9 * float(np.sum(np.abs(seed1 * t1))), (dgamma, dual)
    okw = dict(dim=2, dx=dx, dy=dy, bias=0.0, limiter=0, p1=HELIUM, p2=AIR)
    O.rgfm_fixed_dt(m1, m2, phi, adj.dt, **okw)
    h0 = O.branch_hash()
    M = 1.3
    for h in (1e-6, 1e-7):
        a, b = rgfm_case(2, M + h, tie=tie), rgfm_case(2, M - h, tie=tie)
        O.rgfm_fixed_dt(a[0], a[1], a[2], adj.dt, **okw)
        ha = O.branch_hash()
        O.rgfm_fixed_dt(b[0], b[1], b[2], adj.dt, **okw)
        if ha == h0 == O.branch_hash():
            break
    else:
        pytest.fail("no step keeps every branch")
    fa = S.rgfm_solve_fixed_dt(S.MultimaterialSystem(a[0], a[1], a[2]), dx, adj.dt, HELIUM, AIR, dy=dy)
    fb = S.rgfm_solve_fixed_dt(S.MultimaterialSystem(b[0], b[1], b[2]), dx, adj.dt, HELIUM, AIR, dy=dy)
    np.testing.assert_array_equal(fa.level_set < 0, ref.system.level_set < 0)   # same circulation mask
    ga = O.rgfm_circulation(fa.material1, ref.system.level_set, dx, dy)
    gb = O.rgfm_circulation(fb.material1, ref.system.level_set, dx, dy)
    fd = (ga - gb) / (2 * h)
    assert abs(fd - dgamma) <= 1e-6 * abs(dgamma) + 1e-9 * float(np.sum(np.abs(seed1 * t1))), (fd, dgamma)


def test_checkpoint_interval_changes_nothing():
    case = rgfm_case(2, tie=False)
    m1, m2 = case[0], case[1]
    rng = np.random.default_rng(70)
    s1, s2 = rng.standard_normal(m1.shape) / SCALE, rng.standard_normal(m2.shape) / SCALE
    base = run_adjoint(case, s1, s2, 17)
    for k in (1, 3, 17, 40):
        other = run_adjoint(case, s1, s2, 17, k=k)
        O.assert_bitwise(other.grad1, base.grad1, f"material 1, checkpoint every {k}")
        O.assert_bitwise(other.grad2, base.grad2, f"material 2, checkpoint every {k}")


def planar_seeds(shape, seed):
    rng = np.random.default_rng(seed)
    s1, s2 = rng.standard_normal(shape) / SCALE, rng.standard_normal(shape) / SCALE
    for s in (s1, s2):
        s[..., 3] = 0.0
        s[..., 5] = 0.0
    return s1, s2


@pytest.mark.parametrize("dim,tie,steps,k", [(1, True, 30, 0), (1, False, 25, 4), (2, True, 20, 0),
                                             (2, False, 17, 3), (2, False, 9, 1)])
def test_four_field_adjoint_equals_six_field(dim, tie, steps, k):
    """The 4-field adjoint (default context) against the 6-field one (the
    context's 4-field switch off, which also gives the forward six fields):
    same time steps, same forward bits, the same gradient numbers (zeros may
    differ in sign), zero w and lambda parts."""
    case = rgfm_case(dim, tie=tie)
    m1, m2, phi, dx, dy = case
    assert not np.any(m1[..., 3]) and np.all(m1[..., 5] == 1.0) and np.all(m2[..., 5] == 1.0)
    s1, s2 = planar_seeds(m1.shape, 80 + 2 * dim + int(tie))
    ctx6 = S.Context(0)
    ctx6.set_reduced_fields(False)
    kw = dict(dy=dy, max_steps=steps, checkpoint_every=k)
    six = S.rgfm_adjoint(S.MultimaterialSystem(m1, m2, phi), s1, s2, dx, 0.8, math.inf, HELIUM, AIR, ctx=ctx6, **kw)
    four = S.rgfm_adjoint(S.MultimaterialSystem(m1, m2, phi), s1, s2, dx, 0.8, math.inf, HELIUM, AIR, **kw)
    O.assert_bitwise(four.dt, six.dt, "time steps")
    O.assert_bitwise(four.system.material1, six.system.material1, "forward 1")
    O.assert_bitwise(four.system.material2, six.system.material2, "forward 2")
    for g4, g6, name in ((four.grad1, six.grad1, "material 1"), (four.grad2, six.grad2, "material 2")):
        assert np.array_equal(g4, g6), (name, float(np.max(np.abs(g4 - g6))))
        nz = g4 != 0.0
        O.assert_bitwise(g4[nz], g6[nz], f"{name}: nonzero gradient entries")
        assert not np.any(g4[..., 3]) and not np.any(g4[..., 5])
        assert np.count_nonzero(g4[..., 0]) > 0


@pytest.mark.parametrize("dim,tie", [(1, False), (2, True)])
def test_four_field_dot_product_against_dual_numbers(dim, tie):
    """The 4-field adjoint in the dot-product test against the oracle's
    forward-mode derivative, with tangents in all six fields."""
    case = rgfm_case(dim, tie=tie)
    m1, m2, phi, dx, dy = case
    s1, s2 = planar_seeds(m1.shape, 90 + 2 * dim + int(tie))
    adj = run_adjoint(case, s1, s2, 20)
    rng = np.random.default_rng(91)
    v1, v2 = rng.standard_normal(m1.shape) * SCALE, rng.standard_normal(m2.shape) * SCALE
    _, _, _, t1, t2 = O.rgfm_fixed_dt(m1, m2, phi, adj.dt, dim=dim, dx=dx, dy=dy, bias=0.0, limiter=0,
                                      p1=HELIUM, p2=AIR, tangent=(v1, v2))
    lhs = float(np.sum(adj.grad1 * v1) + np.sum(adj.grad2 * v2))
    rhs = float(np.sum(s1 * t1) + np.sum(s2 * t2))
    scale = float(np.sum(np.abs(adj.grad1 * v1)) + np.sum(np.abs(adj.grad2 * v2)))
    assert abs(lhs - rhs) <= 1e-9 * scale, (lhs, rhs, scale)


@pytest.mark.parametrize("streams", [1, 3])
def test_batch_equals_one_call_per_instance(streams):
    """vg_rgfm_adjoint_batch_2d on worker streams against vg_rgfm_adjoint_2d per
    instance, bit for bit: four shock-bubble instances of different Mach
    numbers and radii to their own final times, planar seeds (4-field
    adjoint) for three and a full seed (6-field) for one, plus an instance
    whose seed overflows, reported in its status alone."""
    machs, radii, tfs = [1.22, 1.3, 1.15, 1.25, 1.2], [0.025, 0.02, 0.03, 0.022, 0.025], [2e-5, 3e-5, 1.5e-5,
                                                                                      2.5e-5, 1e-5]
    cases = [problems.shock_bubble_2d(48, 24, radius=r, mach=M) for r, M in zip(radii, machs)]
    dx, dy = cases[0][3], cases[0][4]
    m1 = np.stack([c[0] for c in cases])
    m2 = np.stack([c[1] for c in cases])
    phi = np.stack([c[2] for c in cases])
    rng = np.random.default_rng(95)
    s1, s2 = rng.standard_normal(m1.shape) / SCALE, rng.standard_normal(m2.shape) / SCALE
    s1[:3, ..., 3] = s1[:3, ..., 5] = s2[:3, ..., 3] = s2[:3, ..., 5] = 0.0
    s1[4] = 1e308
    ctx = S.Context(0)
    ctx.set_batch_streams(streams)
    b = S.rgfm_adjoint_batch_2d(m1, m2, phi, s1, s2, dx, dy, 0.8, np.array(tfs), HELIUM, AIR, ctx=ctx,
                                raise_on_error=False)
    assert b.rc == S._lib.VG_ERR_NONFINITE
    assert list(b.status) == [0, 0, 0, 0, S._lib.VG_ERR_NONFINITE]
    for q in range(4):
        one = S.rgfm_adjoint(S.MultimaterialSystem(m1[q], m2[q], phi[q]), s1[q], s2[q], dx, 0.8, tfs[q], HELIUM, AIR,
                             dy=dy)
        assert b.steps[q] == one.dt.size
        O.assert_bitwise(b.material1[q], one.system.material1, f"instance {q}: forward 1")
        O.assert_bitwise(b.material2[q], one.system.material2, f"instance {q}: forward 2")
        O.assert_bitwise(b.level_set[q], one.system.level_set, f"instance {q}: level set")
        O.assert_bitwise(b.grad1[q], one.grad1, f"instance {q}: gradient 1")
        O.assert_bitwise(b.grad2[q], one.grad2, f"instance {q}: gradient 2")


def test_adjoint_to_final_time_on_the_benchmark_geometry():
    """A short run to t_final (last step cut) of the shock-bubble problem at a
    small size: finite gradients and the dot-product test."""
    m1, m2, phi, dx, dy = problems.shock_bubble_2d(64, 32, radius=0.025, mach=1.22)
    rng = np.random.default_rng(71)
    s1, s2 = rng.standard_normal(m1.shape) / SCALE, rng.standard_normal(m2.shape) / SCALE
    adj = S.rgfm_adjoint(S.MultimaterialSystem(m1, m2, phi), s1, s2, dx, 0.8, 2e-5, HELIUM, AIR, dy=dy)
    assert abs(float(np.sum(adj.dt)) - 2e-5) < 1e-18
    assert np.isfinite(adj.grad1).all() and np.isfinite(adj.grad2).all()
    v1, v2 = rng.standard_normal(m1.shape) * SCALE, rng.standard_normal(m2.shape) * SCALE
    _, _, _, t1, t2 = O.rgfm_fixed_dt(m1, m2, phi, adj.dt, dim=2, dx=dx, dy=dy, bias=0.0, limiter=0, p1=HELIUM,
                                      p2=AIR, tangent=(v1, v2))
    lhs = float(np.sum(adj.grad1 * v1) + np.sum(adj.grad2 * v2))
    rhs = float(np.sum(s1 * t1) + np.sum(s2 * t2))
    assert abs(lhs - rhs) <= 1e-9 * float(np.sum(np.abs(adj.grad1 * v1)) + np.sum(np.abs(adj.grad2 * v2)))


def test_overflowing_adjoint_names_the_step():
    """An adjoint that overflows is reported, not returned: a seed of 1e308
    overflows in the reversal of the last steps, and the entry point returns
    VG_ERR_NONFINITE with the forward step (0-based) whose reversal first
    produced a non-finite value (the per-step scan of the reverse sweep)."""
    case = rgfm_case(2, tie=False)
    m1, m2, phi, dx, dy = case
    s1 = np.full(m1.shape, 1e308)
    adj = S.rgfm_adjoint(S.MultimaterialSystem(m1, m2, phi), s1, np.zeros_like(m2), dx, 0.8, math.inf, HELIUM, AIR,
                         dy=dy, max_steps=9, raise_on_error=False)
    assert adj.status == S._lib.VG_ERR_NONFINITE
    msg = S.default_context().lib.vg_last_error().decode()
    print(msg, flush=True)
    step = int(msg.split("reverse of step ")[1].split(" of ")[0])
    assert msg.endswith(" of 9") and 5 <= step <= 8
    assert adj.stats.nonfinite_cells > 0
