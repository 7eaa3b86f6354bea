"""MHD on the GPU through the C ABI, and the batched persistent solves.

EXACT mode is bit-identical to the compiled reference's golden vectors (1D with
the per-step GLM speed, FORCE, 2D) and to the oracle; the batched launch gives
every instance exactly what its own solve call gives.
"""

import math

import numpy as np
import pytest

import oracle_lib as O
from vanguard_amd import _lib, problems
from vanguard_amd import solvers as S

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["mhd_briowu256", "mhd_force_briowu128", "mhd_2d_ot32", "mhd_force_2d_ot16",
                                  "mhd_psi128_sub2", "mhd_2d_ot16_sub1"])
def test_mhd_exact_bit_identical_to_reference(vg_ctx, name):
    g = O.golden(name)
    m = g["meta"]
    r = S.mhd_solve(g["prim_in"], m["dim"], m["dx"], m["cfl"], m["tfinal"], order=m["order"], bias=m["bias"],
                    limiter=m["limiter"], params=S.MHDMaterialParameters(m["gamma"], m["p_inf"]),
                    max_steps=m["max_steps"], subcycles=m.get("subcycles", 0))
    print(f"{name}: steps={r.stats.steps}", flush=True)
    O.assert_bitwise(r.dt, g["dt"], f"{name} dt")
    O.assert_bitwise(r.cells, g["prim_out"], f"{name} state")


def test_mhd_batch_bit_identical_to_reference(vg_ctx):
    g = O.golden("mhd_batch12")
    m = g["meta"]
    r = S.mhd_solve_batch(g["prim_in"], g["offsets"], g["dx"], g["gamma"], m["tfinal"], m["cfl"],
                          limiter=m["limiter"])
    np.testing.assert_array_equal(r.steps, g["steps"])
    O.assert_bitwise(r.t, g["t"], "final times")
    O.assert_bitwise(r.cells, g["prim_out"], "states")


def test_mhd_batch_uneven_sizes_match_oracle(vg_ctx):
    """Sizes from 2 cells to a few thousand, mixed in one launch, in the
    batched-MHD benchmark recipe; more instances than one wave of workgroups."""
    prim, off, gamma, dx = problems.mhd_batch(600, seed=11, n_min=2, n_max=3000)
    tf = np.full(600, 2e-3)
    r = S.mhd_solve_batch(prim, off, dx, gamma, tf, 0.8, limiter=1)
    rng = np.random.default_rng(0)
    for q in rng.choice(600, 12, replace=False):
        a, b = off[q], off[q + 1]
        ref, dts = O.mhd_solve(prim[a:b], order=2, dim=1, dx=float(dx[q]), cfl=0.8, tfinal=2e-3, max_steps=0,
                               bias=0.0, limiter=1, gamma=float(gamma[q]))
        assert r.steps[q] == len(dts)
        assert r.t[q] == O.seq_sum(dts)
        O.assert_bitwise(r.cells[a:b], ref, f"instance {q} (n={b - a})")


@pytest.mark.parametrize("dim,bc,sub,limiter", [(1, 1, 3, 0), (1, 0, 1, 2), (2, 1, 2, 1), (2, 0, 1, 0)])
def test_mhd_sub_cycles_match_oracle(vg_ctx, dim, bc, sub, limiter):
    """RK4 GLM sub-cycles on sizes and boundaries the fixtures do not cover,
    bit for bit against the oracle (including the first-order path, which
    ignores them and keeps the given c_h)."""
    rng = np.random.default_rng(40 + dim + 3 * sub)
    shape = (300, 9) if dim == 1 else (22, 31, 9)
    w = np.zeros(shape)
    w[..., 0] = rng.uniform(0.5, 1.5, shape[:-1])
    w[..., 1:4] = rng.uniform(-0.3, 0.3, shape[:-1] + (3,))
    w[..., 4] = rng.uniform(0.5, 1.5, shape[:-1])
    w[..., 5:8] = rng.uniform(-0.8, 0.8, shape[:-1] + (3,))
    w[..., 8] = rng.uniform(-0.1, 0.1, shape[:-1])
    dx = 1.0 / shape[-2]
    for order in (2, 1):
        p = S.MHDMaterialParameters(5.0 / 3.0, 0.0, 0.7, 0.4)
        r = S.mhd_solve(w, dim, dx, 0.6, np.inf, order=order, limiter=limiter, params=p, max_steps=12, bc=bc,
                        subcycles=sub)
        ref, dts = O.mhd_solve(w, order=order, dim=dim, dx=dx, cfl=0.6, tfinal=np.inf, max_steps=12, bias=0.0,
                               limiter=limiter, gamma=5.0 / 3.0, bc=bc, subcycles=sub, ch=0.7, cp=0.4)
        O.assert_bitwise(r.dt, dts, f"order {order} dt")
        O.assert_bitwise(r.cells, ref, f"order {order} state")


def test_mhd_sub_cycled_graph_replay_matches_eager():
    """Captured step chunks with the source stage give the eager result bit for bit."""
    g = O.golden("mhd_2d_ot16_sub1")
    m = g["meta"]
    out = []
    for graphs in (True, False):
        ctx = S.Context(0)
        try:
            ctx.set_graphs(graphs)
            r = S.mhd_solve(g["prim_in"], 2, m["dx"], m["cfl"], m["tfinal"], limiter=m["limiter"],
                            params=S.MHDMaterialParameters(m["gamma"]), subcycles=1, ctx=ctx)
            out.append(r.cells)
        finally:
            ctx.close()
    O.assert_bitwise(out[0], out[1], "graphs vs eager")
    O.assert_bitwise(out[0], g["prim_out"], "graphs vs reference")


def test_mhd_fast_mode_within_tolerance(vg_ctx):
    g = O.golden("mhd_briowu256")
    m = g["meta"]
    r = S.mhd_solve(g["prim_in"], 1, m["dx"], m["cfl"], m["tfinal"], limiter=m["limiter"],
                    params=S.MHDMaterialParameters(m["gamma"]), mode=S.FAST)
    assert r.stats.steps == len(g["dt"])
    for f in range(9):
        ref = g["prim_out"][:, f]
        if np.abs(ref).sum() > 0:
            assert np.abs(r.cells[:, f] - ref).sum() / np.abs(ref).sum() <= 1e-10, f


def test_euler_batch_matches_single_solves(vg_ctx):
    rng = np.random.default_rng(4)
    sizes = rng.integers(2, 2000, 40)
    off = np.concatenate([[0], np.cumsum(sizes)])
    prim = np.concatenate([problems.sod_1d(int(n)) for n in sizes])
    dx = 1.0 / sizes
    tf = rng.uniform(0.01, 0.05, 40)
    params = [S.EulerMaterialParameters(float(g)) for g in rng.uniform(1.2, 1.8, 40)]
    r = S.euler_solve_batch(prim, off, dx, params, tf, 0.9, limiter=2)
    for q in range(0, 40, 7):
        a, b = off[q], off[q + 1]
        ref, dts = O.euler_solve(prim[a:b], order=2, dim=1, dx=float(dx[q]), cfl=0.9, tfinal=float(tf[q]),
                                 max_steps=0, bias=0.0, limiter=2, params=[params[q].gamma, 0, 25, 0, 0.25, 250])
        assert r.steps[q] == len(dts)
        O.assert_bitwise(r.cells[a:b], ref, f"euler instance {q}")


def test_euler_1d_device_state_path(vg_ctx):
    """The multi-workgroup grid path for 1D (device-resident state) agrees with the
    persistent single-workgroup path used by vg_euler_solve_1d."""
    w = problems.sod_1d(3000)
    grid = S.EulerGrid(1, 1, 3000)
    grid.load(w)
    st, d = grid.run(1 / 3000, 0.9, 0.05, limiter=2, dt_cap=10000)
    one = S.euler_solve(w, 1, 1 / 3000, 0.9, 0.05, limiter=2)
    O.assert_bitwise(d, one.dt, "dt")
    O.assert_bitwise(grid.store(), one.cells, "state")
    grid.close()


def test_batch_rejects_bad_offsets(vg_ctx):
    prim = np.zeros((10, 9))
    prim[:, 0] = 1
    prim[:, 4] = 1
    r = S.mhd_solve_batch(prim, [0, 1, 10], [0.1, 0.1], [2.0, 2.0], [0.1, 0.1], 0.8, raise_on_error=False)
    assert r.status == _lib.VG_ERR_ARG


def test_mhd_batch_work_budget_gives_the_same_bits():
    """With more instances than resident workgroups (the MHD batch kernel's
    27 KB of LDS allows five per CU, 1280 on 256 CUs), each instance also stops
    at the launch's work budget (predicted cells x steps over the resident
    workgroups, at least 1024 steps).  One 4096-cell instance of about 1200
    steps among 3000 light ones stops at 1024 and finishes on the grid path:
    the bits and step count of solving it alone, on 1 or 4 continuation
    streams; the light instances match the oracle."""
    n_light = 3000
    prim, off, gamma, dx = problems.mhd_batch(n_light, seed=7, n_min=32, n_max=32)
    n = 4096
    heavy = np.zeros((n, 9))
    left = np.arange(n) <= 0.5 * n
    heavy[left] = (1.0, 0.0, 0.0, 0.0, 10.0, 0.5, 0.2, 0.0, 0.0)
    heavy[~left] = (0.5, 0.0, 0.0, 0.0, 1.0, 0.5, -0.2, 0.0, 0.0)
    prim = np.concatenate([heavy, prim])
    sizes = np.concatenate([[n], np.diff(off)])
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    gamma = np.concatenate([[1.4], gamma])
    dx = 1.0 / sizes.astype(np.float64)
    tf = np.full(n_light + 1, 0.01)
    tf[0] = 0.06
    alone = S.mhd_solve(heavy, 1, float(dx[0]), 0.8, 0.06, limiter=1, params=S.MHDMaterialParameters(1.4))
    assert alone.stats.steps > 1024
    base = None
    for streams in (1, 4):
        ctx = S.Context(0)
        try:
            ctx.set_batch_streams(streams)
            print(f"{n_light + 1} instances, {streams} continuation streams", flush=True)
            r = S.mhd_solve_batch(prim, off, dx, gamma, tf, 0.8, limiter=1, ctx=ctx)
        finally:
            ctx.close()
        assert r.steps[0] == alone.stats.steps and r.t[0] == alone.stats.t
        O.assert_bitwise(r.cells[:n], alone.cells, f"heavy instance vs alone, {streams} streams")
        if base is None:
            base = r
        else:
            O.assert_bitwise(r.cells, base.cells, "1 vs 4 continuation streams")
            assert np.array_equal(r.steps, base.steps)
    for q in (1, 1500, n_light):
        a, b = off[q], off[q + 1]
        ref, dts = O.mhd_solve(prim[a:b], order=2, dim=1, dx=float(dx[q]), cfl=0.8, tfinal=0.01, max_steps=0,
                               bias=0.0, limiter=1, gamma=float(gamma[q]))
        assert base.steps[q] == len(dts)
        O.assert_bitwise(base.cells[a:b], ref, f"instance {q}")


def test_mhd_batch_continuation_past_the_step_cap_gives_the_same_bits(vg_ctx):
    """The launch stops every instance at 1.25x the largest step count its
    initial wave speeds predict.  A strong blast (pressure ratio 1e5, 2048
    cells) speeds up by more than that: 957 steps predicted, 1371 taken.  It
    finishes on the grid path from where the launch stopped it, with the bits
    and step count of solving it alone; the light instances match the oracle."""
    prim, off, gamma, dx = problems.mhd_batch(200, seed=5, n_min=256, n_max=256)
    n = 2048
    blast = np.zeros((n, 9))
    left = np.arange(n) <= 0.5 * n
    blast[left] = (1.0, 0.0, 0.0, 0.0, 1000.0, 0.5, 0.0, 0.0, 0.0)
    blast[~left] = (1.0, 0.0, 0.0, 0.0, 0.01, 0.5, 0.0, 0.0, 0.0)
    prim = np.concatenate([prim[:off[100]], blast, prim[off[100]:]])
    sizes = np.concatenate([np.diff(off)[:100], [n], np.diff(off)[100:]])
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    gamma = np.concatenate([gamma[:100], [1.4], gamma[100:]])
    dx = 1.0 / sizes.astype(np.float64)
    tf = np.full(201, 0.01)
    r = S.mhd_solve_batch(prim, off, dx, gamma, tf, 0.8, limiter=1)
    a, b = off[100], off[101]
    alone = S.mhd_solve(prim[a:b], 1, float(dx[100]), 0.8, 0.01, limiter=1, params=S.MHDMaterialParameters(1.4))
    print(f"blast: {r.steps[100]} steps", flush=True)
    assert r.steps[100] == alone.stats.steps == 1371
    assert r.t[100] == alone.stats.t
    O.assert_bitwise(r.cells[a:b], alone.cells, "blast vs alone")
    for q in (0, 57, 150, 200):
        a, b = off[q], off[q + 1]
        ref, dts = O.mhd_solve(prim[a:b], order=2, dim=1, dx=float(dx[q]), cfl=0.8, tfinal=0.01, max_steps=0,
                               bias=0.0, limiter=1, gamma=float(gamma[q]))
        assert r.steps[q] == len(dts)
        O.assert_bitwise(r.cells[a:b], ref, f"instance {q}")
