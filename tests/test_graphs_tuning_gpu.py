"""Captured-graph replay and the y-sweep autotuner on the GPU.

Every solver path runs once eagerly (graphs off) and once replaying captured
chunks of steps; the replays must reproduce the eager results bit for bit,
including across several runs of a device-resident grid with host work and
synchronisation between them.  The autotuner only picks among shapes of the
same kernel, so results with every tuned shape equal the untuned ones.
"""

import math

import numpy as np
import pytest

import oracle_lib as O
from vanguard_amd import problems
from vanguard_amd import solvers as S

pytestmark = pytest.mark.gpu


@pytest.fixture
def fresh_ctx():
    ctx = S.Context(0)
    yield ctx
    ctx.close()


def test_euler_graph_replay_matches_eager(fresh_ctx):
    w = problems.disc_2d(96, ny=80)
    runs = {}
    for graphs in (False, True):
        fresh_ctx.set_graphs(graphs)
        r = S.euler_solve(w, 2, 2.0 / 96, 0.8, math.inf, max_steps=37, limiter=1, ctx=fresh_ctx)
        runs[graphs] = r
    O.assert_bitwise(runs[True].dt, runs[False].dt, "dt")
    O.assert_bitwise(runs[True].cells, runs[False].cells, "state")


def test_device_state_graph_replays_between_host_work(fresh_ctx):
    """Three runs on one device-resident grid, host reads and numpy work between
    them: every intermediate state equals the eager one."""
    w = problems.disc_2d(64)
    states = {}
    for graphs in (False, True):
        fresh_ctx.set_graphs(graphs)
        g = S.EulerGrid(2, 64, 64, ctx=fresh_ctx)
        g.load(w)
        seq = []
        for k in range(3):
            g.run(2.0 / 64, 0.8, max_steps=11 + k, limiter=0)
            out = g.store()
            seq.append(out)
            _ = float(np.abs(out).sum())          # host work between replays
        g.close()
        states[graphs] = seq
    for k in range(3):
        O.assert_bitwise(states[True][k], states[False][k], f"after run {k}")


def test_hpr_graph_replay_matches_eager(fresh_ctx):
    # mu = 1e-2: at 32^2 the stiffer mu = 1e-3 relaxation is unstable for the
    # reference's RK4 sub-cycle too (the oracle blows up the same way).
    p = S.HPRMaterialParameters(is_thermal=True, dynamic_viscosity_coefficient=1e-2)
    p.configure_strain_dissipation_time()
    p.configure_thermal_impulse_relaxation_time(1e-2)
    w = problems.hpr_fourier_2d(32, seed=4)
    runs = {}
    for graphs in (False, True):
        fresh_ctx.set_graphs(graphs)
        runs[graphs] = S.hpr_solve(w, 2, 1 / 32, 0.8, math.inf, subcycles=1, params=p, max_steps=19, bc=S.PERIODIC,
                                   ctx=fresh_ctx)
    O.assert_bitwise(runs[True].dt, runs[False].dt, "dt")
    O.assert_bitwise(runs[True].cells, runs[False].cells, "state")


def test_rgfm_graph_replay_matches_eager(fresh_ctx):
    m1, m2, phi, dx, dy = problems.shock_bubble_2d(80, 32, length=(0.325, 0.13), center=(0.175, 0.065),
                                                   radius=0.03)
    runs = {}
    for graphs in (False, True):
        fresh_ctx.set_graphs(graphs)
        runs[graphs] = S.rgfm_solve(S.MultimaterialSystem(m1, m2, phi), dx, 0.8, 4e-5, problems.HELIUM,
                                    problems.AIR, dy=dy, ctx=fresh_ctx)
    a, b = runs[True], runs[False]
    O.assert_bitwise(a.dt, b.dt, "dt")
    O.assert_bitwise(a.system.material1, b.system.material1, "material 1")
    O.assert_bitwise(a.system.material2, b.system.material2, "material 2")
    O.assert_bitwise(a.system.level_set, b.system.level_set, "level set")


def test_autotuned_shape_gives_identical_results(fresh_ctx):
    w = problems.disc_2d(300, ny=70)
    fresh_ctx.set_autotune(False)
    ref = S.euler_solve(w, 2, 2.0 / 300, 0.8, math.inf, max_steps=9, ctx=fresh_ctx)
    # the disc is at rest out of plane with lambda = 1: its sweeps are the 4-field ones
    assert fresh_ctx.tuned_shape("euler_planar", 2, S.EXACT, 70, 300) == -1
    fresh_ctx.set_autotune(True)
    r = S.euler_solve(w, 2, 2.0 / 300, 0.8, math.inf, max_steps=9, ctx=fresh_ctx)
    v = fresh_ctx.tuned_shape("euler_planar", 2, S.EXACT, 70, 300)
    print(f"tuned y-sweep shape for 70 x 300: variant {v}", flush=True)
    assert 0 <= v < 8
    O.assert_bitwise(r.cells, ref.cells, "state")
    O.assert_bitwise(r.dt, ref.dt, "dt")


@pytest.mark.parametrize("mode", [S.EXACT, S.FAST])
def test_every_y_shape_gives_identical_results(fresh_ctx, mode):
    """Each of the eight y-sweep shapes, pinned, on a grid whose rows and
    columns none of them divides evenly (strips and blocks end inside the
    grid): the same bits as the default shape."""
    w = problems.disc_2d(301, ny=173)
    key = f"euler_planar/2/{mode}/173/301"
    runs = []
    for v in range(8):
        fresh_ctx.set_tuned_key(key, v)
        runs.append(S.euler_solve(w, 2, 2.0 / 301, 0.8, math.inf, max_steps=7, mode=mode, ctx=fresh_ctx))
        assert fresh_ctx.tuned_key(key) == v
    for v in range(1, 8):
        O.assert_bitwise(runs[v].cells, runs[0].cells, f"state, y shape {v}")
        O.assert_bitwise(runs[v].dt, runs[0].dt, f"dt, y shape {v}")
