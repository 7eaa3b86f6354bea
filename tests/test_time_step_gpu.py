"""The per-sweep entry points on the GPU through the C ABI.

vg_euler_time_step_*, vg_mhd_time_step_*, vg_hpr_time_step_* take the
reference's computeSLICTimeStep / computeX|YSLICTimeStep2D arguments (the
ghosted snapshot the caller built, the caller's time step) and
vg_euler_state_sweep / vg_hpr_state_sweep do the same on a device-resident
grid.  Driving them in the reference's own solve loop -- ghosts by
insertBoundaryCells[2D], the time steps the compiled reference chose -- must
reproduce its golden runs bit for bit (secondordersolver.cpp:288-314, 392-424;
firstordersolver.cpp:435-461, 521-553; hprsecondordersolver.cpp:314-352,
502-546); MHD second-order sweeps with the caller's c_h match the oracle's
sweep bit for bit, and caller-supplied ghost values are read as given.
"""

import numpy as np
import pytest

import oracle_lib as O
from vanguard_amd import solvers as S
from vanguard_amd.solvers import Solvers

from test_hpr_gpu import mat_params

pytestmark = pytest.mark.gpu


def euler_params(m):
    return S.EulerMaterialParameters(*m["params"])


@pytest.mark.parametrize("name", ["euler_c1_sod1000", "euler_toro100", "euler_1d_stiff_bias", "euler_force1d"])
def test_euler_1d_time_steps_reproduce_reference(vg_ctx, name):
    g = O.golden(name)
    m = g["meta"]
    order, mp = m["order"], euler_params(m)
    cells = np.ascontiguousarray(g["prim_in"], dtype=np.float64).copy()
    for dt in g["dt"]:
        wb = Solvers.insertBoundaryCells(cells, order)
        if order == 2:
            S.SecondOrderSolver.computeSLICTimeStep(cells, wb, m["dx"], dt, m["bias"], m["limiter"], mp)
        else:
            S.FirstOrderSolver.computeFORCETimeStep(cells, wb, m["dx"], dt, mp)
    print(f"{name}: {len(g['dt'])} per-sweep steps", flush=True)
    O.assert_bitwise(cells, g["prim_out"], name)


@pytest.mark.parametrize("name", ["euler_2d_disc64_l0", "euler_2d_c2ic64", "euler_force2d"])
def test_euler_2d_time_steps_reproduce_reference(vg_ctx, name):
    g = O.golden(name)
    m = g["meta"]
    order, mp, dx = m["order"], euler_params(m), m["dx"]
    cells = np.ascontiguousarray(g["prim_in"], dtype=np.float64).copy()
    for dt in g["dt"]:
        for direction, ts in ((0, 0.5 * dt), (1, dt), (0, 0.5 * dt)):
            wb = Solvers.insertBoundaryCells2D(cells, order)
            if order == 2:
                fn = S.SecondOrderSolver.computeXSLICTimeStep2D if direction == 0 else \
                    S.SecondOrderSolver.computeYSLICTimeStep2D
                fn(cells, wb, dx, ts, m["bias"], m["limiter"], mp)
            else:
                fn = S.FirstOrderSolver.computeXFORCETimeStep2D if direction == 0 else \
                    S.FirstOrderSolver.computeYFORCETimeStep2D
                fn(cells, wb, dx, ts, mp)
    print(f"{name}: {len(g['dt'])} per-sweep steps", flush=True)
    O.assert_bitwise(cells, g["prim_out"], name)


def test_euler_state_sweep_reproduces_reference(vg_ctx):
    """Device-resident: load once, three sweeps per step with the reference's dt."""
    g = O.golden("euler_2d_disc64_l0")
    m = g["meta"]
    ny, nx = g["prim_in"].shape[:2]
    grid = S.EulerGrid(2, ny, nx)
    grid.load(g["prim_in"])
    for dt in g["dt"]:
        for direction, ts in ((0, 0.5 * dt), (1, dt), (0, 0.5 * dt)):
            grid.sweep(direction, m["dx"], ts, bias=m["bias"], limiter=m["limiter"], params=euler_params(m))
    out = grid.store()
    grid.close()
    O.assert_bitwise(out, g["prim_out"], "device-resident sweeps")


def test_euler_time_step_reads_caller_ghosts(vg_ctx):
    """Ghost values that are not the transmissive mirror are read as given: the
    interior of a sweep over the ghosted row is the oracle's."""
    rng = np.random.default_rng(7)
    n = 50
    w = np.zeros((n + 4, 6))
    w[:, 0] = rng.uniform(0.5, 2.0, n + 4)
    w[:, 1] = rng.uniform(-0.5, 0.5, n + 4)
    w[:, 4] = rng.uniform(0.5, 2.0, n + 4)
    w[:, 5] = 1.0
    mp = S.EulerMaterialParameters(1.4)
    cells = np.zeros((n, 6))
    S.SecondOrderSolver.computeSLICTimeStep(cells, w, 0.02, 0.004, 0.0, 0, mp)
    ref = O.euler_sweep(w, order=2, dim=1, direction=0, dx=0.02, ts=0.004, bias=0.0, limiter=0,
                        params=(1.4, 0, 25, 0, 0.25, 250))
    O.assert_bitwise(cells, ref[2:-2], "ghosted sweep")
    # the boundary cells did read the supplied ghosts: the mirror gives another answer
    mirror = np.zeros((n, 6))
    S.SecondOrderSolver.computeSLICTimeStep(mirror, Solvers.insertBoundaryCells(w[2:-2], 2), 0.02, 0.004, 0.0, 0, mp)
    assert not np.array_equal(mirror[:2], cells[:2]) and np.array_equal(mirror[3:-3], cells[3:-3])


@pytest.mark.parametrize("name", ["mhd_force_briowu128", "mhd_force_2d_ot16"])
def test_mhd_force_time_steps_reproduce_reference(vg_ctx, name):
    """The first-order MHD solvers keep the given c_h (0), so the reference's
    loop over the per-sweep calls reproduces them exactly."""
    g = O.golden(name)
    m = g["meta"]
    mp = S.MHDMaterialParameters(m["gamma"], m["p_inf"])
    cells = np.ascontiguousarray(g["prim_in"], dtype=np.float64).copy()
    for dt in g["dt"]:
        if m["dim"] == 1:
            S.MHDFirstOrderSolver.computeFORCETimeStep(cells, Solvers.insertBoundaryCells(cells, 1, S.MHD_DEFAULT_STATE),
                                                       m["dx"], dt, mp)
        else:
            for direction, ts in ((0, 0.5 * dt), (1, dt), (0, 0.5 * dt)):
                wb = Solvers.insertBoundaryCells2D(cells, 1, S.MHD_DEFAULT_STATE)
                fn = S.MHDFirstOrderSolver.computeXFORCETimeStep2D if direction == 0 else \
                    S.MHDFirstOrderSolver.computeYFORCETimeStep2D
                fn(cells, wb, m["dx"], ts, mp)
    O.assert_bitwise(cells, g["prim_out"], name)


@pytest.mark.parametrize("dim,direction", [(1, 0), (2, 0), (2, 1)])
def test_mhd_slic_time_step_uses_given_glm_speeds(vg_ctx, dim, direction):
    g = O.golden("mhd_2d_ot32" if dim == 2 else "mhd_briowu256")
    m = g["meta"]
    w = np.ascontiguousarray(g["prim_out"], dtype=np.float64)
    ch, cp = 2.7, 0.61
    mp = S.MHDMaterialParameters(m["gamma"], m["p_inf"])
    mp.ch, mp.cp = ch, cp
    cells = np.zeros_like(w)
    ts = 0.3 * m["dx"]
    if dim == 1:
        S.MHDSecondOrderSolver.computeSLICTimeStep(cells, Solvers.insertBoundaryCells(w, 2, S.MHD_DEFAULT_STATE),
                                                   m["dx"], ts, 0.0, 1, mp)
    else:
        fn = S.MHDSecondOrderSolver.computeXSLICTimeStep2D if direction == 0 else \
            S.MHDSecondOrderSolver.computeYSLICTimeStep2D
        fn(cells, Solvers.insertBoundaryCells2D(w, 2, S.MHD_DEFAULT_STATE), m["dx"], ts, 0.0, 1, mp)
    ref = O.mhd_sweep(w, order=2, dim=dim, direction=direction, dx=m["dx"], ts=ts, bias=0.0, limiter=1,
                      gamma=m["gamma"], p_inf=m["p_inf"], ch=ch, cp=cp)
    O.assert_bitwise(cells, ref, f"MHD SLIC sweep dim {dim} dir {direction}")


@pytest.mark.parametrize("name", ["hpr_stokes100", "hpr_heat100", "hpr_force_stokes64", "hpr_2d16"])
def test_hpr_time_steps_reproduce_reference(vg_ctx, name):
    g = O.golden(name)
    m = g["meta"]
    assert m["subcycles"] == 0
    mp, order, dx = mat_params(m["params"]), m["order"], m["dx"]
    cells = np.ascontiguousarray(g["prim_in"], dtype=np.float64).copy()
    for dt in g["dt"]:
        if m["dim"] == 1:
            wb = Solvers.insertBoundaryCells(cells, order, S.HPR_DEFAULT_STATE)
            if order == 2:
                S.HPRSecondOrderSolver.computeSLICTimeStep(cells, wb, dx, dt, m["bias"], m["limiter"], mp)
            else:
                S.HPRFirstOrderSolver.computeFORCETimeStep(cells, wb, dx, dt, mp)
        else:
            for direction, ts in ((0, 0.5 * dt), (1, dt), (0, 0.5 * dt)):
                wb = Solvers.insertBoundaryCells2D(cells, order, S.HPR_DEFAULT_STATE)
                fn = S.HPRSecondOrderSolver.computeXSLICTimeStep2D if direction == 0 else \
                    S.HPRSecondOrderSolver.computeYSLICTimeStep2D
                fn(cells, wb, dx, ts, m["bias"], m["limiter"], mp)
    O.assert_bitwise(cells, g["prim_out"], name)


def test_hpr_state_sweep_reproduces_reference(vg_ctx):
    g = O.golden("hpr_2d16")
    m = g["meta"]
    ny, nx = g["prim_in"].shape[:2]
    grid = S.HPRGrid(2, ny, nx)
    grid.load(g["prim_in"])
    for dt in g["dt"]:
        for direction, ts in ((0, 0.5 * dt), (1, dt), (0, 0.5 * dt)):
            grid.sweep(direction, m["dx"], ts, bias=m["bias"], limiter=m["limiter"], params=mat_params(m["params"]))
    out = grid.store()
    grid.close()
    O.assert_bitwise(out, g["prim_out"], "device-resident HPR sweeps")
