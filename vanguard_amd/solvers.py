"""Python mirror of the reference's solver interface, over the C ABI.

Same names and argument meaning as the reference's static C++ methods
(Solvers/secondordersolver.h:55-65, Solvers/firstordersolver.h:87-95), with
states held as NumPy arrays of primitives in the reference's layout (cells x 6,
or rows x columns x 6).  The arithmetic runs in the HIP kernels of
libvanguard_hip.so; this module only moves arguments.
"""

from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import check

EXACT = _lib.VG_MODE_EXACT
FAST = _lib.VG_MODE_FAST
TRANSMISSIVE = _lib.VG_BC_TRANSMISSIVE
PERIODIC = _lib.VG_BC_PERIODIC


class Context:
    """A vg_ctx on one HIP device (a stream plus solver scratch)."""

    def __init__(self, device: int = 0):
        self.lib = _lib.load()
        self.handle = C.c_void_p()
        check(self.lib.vg_ctx_create(C.byref(self.handle), device), "vg_ctx_create")

    def close(self):
        if self.handle:
            self.lib.vg_ctx_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_timing(self, on: bool):
        check(self.lib.vg_ctx_set_timing(self.handle, int(on)), "vg_ctx_set_timing")

    def kernel_time(self, name: str):
        n = C.c_int64()
        ms = C.c_double()
        check(self.lib.vg_ctx_kernel_time(self.handle, name.encode(), C.byref(n), C.byref(ms)), "vg_ctx_kernel_time")
        return n.value, ms.value

    def reset_timing(self):
        check(self.lib.vg_ctx_reset_timing(self.handle), "vg_ctx_reset_timing")

    def set_graphs(self, on: bool):
        """Replay chunks of time steps as captured HIP graphs (default on)."""
        check(self.lib.vg_ctx_set_graphs(self.handle, int(on)), "vg_ctx_set_graphs")

    def set_autotune(self, on: bool):
        """Time the candidate y-sweep shapes on first use of a grid shape (default on)."""
        check(self.lib.vg_ctx_set_autotune(self.handle, int(on)), "vg_ctx_set_autotune")

    def set_reduced_fields(self, on: bool):
        """Sweep Euler grids whose out-of-plane velocity and reaction progress are
        constant with the 4-field kernels (default on; same bits either way)."""
        check(self.lib.vg_ctx_set_reduced_fields(self.handle, int(on)), "vg_ctx_set_reduced_fields")

    def set_batch_streams(self, streams: int):
        """Streams a batch call may run its instances on at once (1..16);
        results are the same bits for every count."""
        check(self.lib.vg_ctx_set_batch_streams(self.handle, int(streams)), "vg_ctx_set_batch_streams")

    def tuned_shape(self, physics: str, order: int, mode: int, ny: int, nx: int) -> int:
        v = C.c_int()
        key = f"{physics}/{order}/{mode}/{ny}/{nx}".encode()
        check(self.lib.vg_ctx_tuned_shape(self.handle, key, C.byref(v)), "vg_ctx_tuned_shape")
        return v.value

    def tuned_key(self, key: str) -> int:
        """Variant recorded for a raw key (e.g. "euler_adjoint_y/ny/nx"), -1 if none."""
        v = C.c_int()
        check(self.lib.vg_ctx_tuned_shape(self.handle, key.encode(), C.byref(v)), "vg_ctx_tuned_shape")
        return v.value

    def release_workspace(self):
        """Free the device workspace kept between calls (adjoint checkpoints)."""
        check(self.lib.vg_ctx_release_workspace(self.handle), "vg_ctx_release_workspace")

    def set_adjoint_memory(self, nbytes: int):
        """Limit the Euler adjoint's recorded forward to nbytes of device memory (0: 85 % of free)."""
        check(self.lib.vg_ctx_set_adjoint_memory(self.handle, int(nbytes)), "vg_ctx_set_adjoint_memory")

    def set_tuned_key(self, key: str, variant: int):
        """Pin a tile-shape variant for a key; -1 forgets it."""
        check(self.lib.vg_ctx_set_tuned_shape(self.handle, key.encode(), int(variant)), "vg_ctx_set_tuned_shape")


_default_ctx = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


@dataclass
class EulerMaterialParameters:
    """EulerMaterialParameters (Euler/eulermaterialparameters.h:5-37); the
    one-argument constructor's defaults (eulermaterialparameters.cpp:14-23)."""

    gamma: float = 1.0
    p_inf: float = 0.0
    cv: float = 25.0
    e_form: float = 0.0
    t_ign: float = 0.25
    k_rate: float = 250.0

    def c_struct(self) -> _lib.EulerParams:
        return _lib.EulerParams(self.gamma, self.p_inf, self.cv, self.e_form, self.t_ign, self.k_rate)


@dataclass
class SolveResult:
    cells: np.ndarray
    dt: np.ndarray
    stats: _lib.RunStats
    status: int


def euler_solve(prim: np.ndarray, dim: int, dx: float, cfl: float, t_final: float, *, order: int = 2,
                bias: float = 0.0, limiter: int = 0, params: EulerMaterialParameters | None = None,
                max_steps: int = 0, mode: int = EXACT, bc: int = TRANSMISSIVE, dt_cap: int = 100000,
                ctx: Context | None = None, raise_on_error: bool = True) -> SolveResult:
    """Whole solve through vg_euler_solve_1d / vg_euler_solve_2d."""
    ctx = ctx or default_context()
    params = params or EulerMaterialParameters(1.4)
    w = np.ascontiguousarray(prim, dtype=np.float64)
    out = np.empty_like(w)
    dts = np.empty(dt_cap)
    st = _lib.RunStats()
    mp = params.c_struct()
    if dim == 1:
        if w.ndim != 2 or w.shape[1] != 6:
            raise ValueError("1D state must be (cells, 6)")
        rc = ctx.lib.vg_euler_solve_1d(ctx.handle, _lib.as_dptr(w), _lib.as_dptr(out), w.shape[0], dx, cfl, t_final,
                                       max_steps, bias, limiter, order, bc, C.byref(mp), mode, _lib.as_dptr(dts),
                                       dt_cap, C.byref(st))
    else:
        if w.ndim != 3 or w.shape[2] != 6:
            raise ValueError("2D state must be (rows, columns, 6)")
        rc = ctx.lib.vg_euler_solve_2d(ctx.handle, _lib.as_dptr(w), _lib.as_dptr(out), w.shape[0], w.shape[1], dx,
                                       cfl, t_final, max_steps, bias, limiter, order, bc, C.byref(mp), mode,
                                       _lib.as_dptr(dts), dt_cap, C.byref(st))
    if raise_on_error and rc != _lib.VG_OK:
        check(rc, "vg_euler_solve")
    return SolveResult(out, dts[:min(st.steps, dt_cap)].copy(), st, rc)


@dataclass
class AdjointResult:
    cells: np.ndarray      # final primitive state of the forward solve
    grad: np.ndarray       # d loss / d initial primitive state
    dt: np.ndarray         # the time steps the forward chose (held fixed)
    stats: _lib.RunStats


def euler_adjoint(prim: np.ndarray, seed: np.ndarray, dim: int, dx: float, cfl: float, t_final: float, *,
                  bias: float = 0.0, limiter: int = 0, params: EulerMaterialParameters | None = None,
                  max_steps: int = 0, bc: int = TRANSMISSIVE, checkpoint_every: int = 0, dt_cap: int = 100000,
                  ctx: Context | None = None) -> AdjointResult:
    """Reverse-mode derivative of the SLIC solve (vg_euler_adjoint_1d / _2d).

    `seed` is d loss / d final primitive state (same shape as `prim`); the
    result's `grad` is d loss / d initial primitive state, with every time step
    held at the value the forward run chose."""
    ctx = ctx or default_context()
    params = params or EulerMaterialParameters(1.4)
    w = np.ascontiguousarray(prim, dtype=np.float64)
    sd = np.ascontiguousarray(seed, dtype=np.float64)
    if sd.shape != w.shape:
        raise ValueError("seed must have the shape of the state")
    out = np.empty_like(w)
    grad = np.empty_like(w)
    dts = np.empty(dt_cap)
    st = _lib.RunStats()
    mp = params.c_struct()
    if dim == 1:
        if w.ndim != 2 or w.shape[1] != 6:
            raise ValueError("1D state must be (cells, 6)")
        rc = ctx.lib.vg_euler_adjoint_1d(ctx.handle, _lib.as_dptr(w), _lib.as_dptr(out), _lib.as_dptr(sd),
                                         _lib.as_dptr(grad), w.shape[0], dx, cfl, t_final, max_steps, bias, limiter,
                                         bc, C.byref(mp), checkpoint_every, _lib.as_dptr(dts), dt_cap, C.byref(st))
    else:
        if w.ndim != 3 or w.shape[2] != 6:
            raise ValueError("2D state must be (rows, columns, 6)")
        rc = ctx.lib.vg_euler_adjoint_2d(ctx.handle, _lib.as_dptr(w), _lib.as_dptr(out), _lib.as_dptr(sd),
                                         _lib.as_dptr(grad), w.shape[0], w.shape[1], dx, cfl, t_final, max_steps,
                                         bias, limiter, bc, C.byref(mp), checkpoint_every, _lib.as_dptr(dts), dt_cap,
                                         C.byref(st))
    check(rc, "vg_euler_adjoint")
    return AdjointResult(out, grad, dts[:min(st.steps, dt_cap)].copy(), st)


def euler_adjoint_device(prim_ptr: int, prim_out_ptr: int, seed_ptr: int, grad_ptr: int, shape, dx: float,
                         cfl: float, t_final: float, *, bias: float = 0.0, limiter: int = 0,
                         params: EulerMaterialParameters | None = None, max_steps: int = 0, bc: int = TRANSMISSIVE,
                         checkpoint_every: int = 0, dt_cap: int = 100000, ctx: Context | None = None):
    """vg_euler_adjoint_1d_device / _2d_device: the four state arrays are device
    pointers (array of structures, 6 doubles per cell) of a grid of `shape`
    ((n,) or (ny, nx)).  Returns (time steps, stats)."""
    ctx = ctx or default_context()
    params = params or EulerMaterialParameters(1.4)
    dts = np.empty(dt_cap)
    st = _lib.RunStats()
    mp = params.c_struct()
    V = C.c_void_p
    if len(shape) == 1:
        rc = ctx.lib.vg_euler_adjoint_1d_device(ctx.handle, V(prim_ptr), V(prim_out_ptr), V(seed_ptr), V(grad_ptr),
                                                shape[0], dx, cfl, t_final, max_steps, bias, limiter, bc, C.byref(mp),
                                                checkpoint_every, _lib.as_dptr(dts), dt_cap, C.byref(st))
    else:
        rc = ctx.lib.vg_euler_adjoint_2d_device(ctx.handle, V(prim_ptr), V(prim_out_ptr), V(seed_ptr), V(grad_ptr),
                                                shape[0], shape[1], dx, cfl, t_final, max_steps, bias, limiter, bc,
                                                C.byref(mp), checkpoint_every, _lib.as_dptr(dts), dt_cap,
                                                C.byref(st))
    check(rc, "vg_euler_adjoint_device")
    return dts[:min(st.steps, dt_cap)].copy(), st


def euler_solve_fixed_dt(prim: np.ndarray, dim: int, dx: float, dts, *, bias: float = 0.0, limiter: int = 0,
                         params: EulerMaterialParameters | None = None, bc: int = TRANSMISSIVE,
                         ctx: Context | None = None) -> np.ndarray:
    """SLIC solve with prescribed time steps (vg_euler_solve_fixed_dt_1d / _2d)."""
    ctx = ctx or default_context()
    params = params or EulerMaterialParameters(1.4)
    w = np.ascontiguousarray(prim, dtype=np.float64)
    d = np.ascontiguousarray(dts, dtype=np.float64).ravel()
    out = np.empty_like(w)
    mp = params.c_struct()
    if dim == 1:
        if w.ndim != 2 or w.shape[1] != 6:
            raise ValueError("1D state must be (cells, 6)")
        rc = ctx.lib.vg_euler_solve_fixed_dt_1d(ctx.handle, _lib.as_dptr(w), _lib.as_dptr(out), w.shape[0], dx,
                                                _lib.as_dptr(d), d.size, bias, limiter, bc, C.byref(mp))
    else:
        if w.ndim != 3 or w.shape[2] != 6:
            raise ValueError("2D state must be (rows, columns, 6)")
        rc = ctx.lib.vg_euler_solve_fixed_dt_2d(ctx.handle, _lib.as_dptr(w), _lib.as_dptr(out), w.shape[0],
                                                w.shape[1], dx, _lib.as_dptr(d), d.size, bias, limiter, bc,
                                                C.byref(mp))
    check(rc, "vg_euler_solve_fixed_dt")
    return out


# ---- ghost cells and per-sweep calls ---------------------------------------------

EULER_DEFAULT_STATE = (1.0, 0.0, 0.0, 0.0, 1.0, 1.0)                 # Euler/eulerstatevector.cpp:3-12
# EulerMultiphysicsStateVector() / EulerReducedStateVector()
# (eulermultiphysicsstatevector.cpp:3-15, eulerreducedstatevector.cpp:3-14).
MULTIPHYSICS_DEFAULT_STATE = np.array([0.999, 0.0, 0.0, 0.0, 1.0, 1.0, 1.0, 1.0])
REDUCED_DEFAULT_STATE = np.array([0.999, 0.0, 0.0, 0.0, 1.0, 1.0, 1.0])


def _default_state(nf: int) -> np.ndarray:
    table = {6: EULER_DEFAULT_STATE, 8: MULTIPHYSICS_DEFAULT_STATE, 7: REDUCED_DEFAULT_STATE}
    if nf not in table:
        raise ValueError(f"no default state for {nf} fields per cell; pass default_state")
    return table[nf]
MHD_DEFAULT_STATE = (1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 0.0)    # MHD/mhdstatevector.cpp:3-16
HPR_DEFAULT_STATE = (1.0, 1.0 / 1.4, 1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0,
                     0.0)                                             # HPR/hprstatevector.cpp:3-16


class Solvers:
    """Host-side helpers of Solvers/solvers.cpp that build the ghosted grids the
    per-sweep calls read."""

    @staticmethod
    def insertBoundaryCells(current_cells, boundary_size, default_state=None):
        """Solvers::insertBoundaryCells (solvers.cpp:7-32, and the overloads of
        the two-material states, :34-86): transmissive ghosts, one layer copied,
        two layers mirror-ordered (g[-1] = c[0], g[-2] = c[1])."""
        c = np.asarray(current_cells, dtype=np.float64)
        default_state = _default_state(c.shape[-1]) if default_state is None else default_state
        n, b = c.shape[0], boundary_size
        out = np.empty((n + 2 * b, c.shape[1]))
        out[:] = default_state
        if b == 1:
            out[0], out[n + 1] = c[0], c[n - 1]
        elif b == 2:
            out[0], out[1], out[n + 2], out[n + 3] = c[1], c[0], c[n - 1], c[n - 2]
        out[b:b + n] = c
        return out

    @staticmethod
    def insertBoundaryCells2D(current_cells, boundary_size, default_state=None):
        """Solvers::insertBoundaryCells2D (solvers.cpp:88-258): the same per row
        and per column; the corner ghosts keep the state's default (by default
        the one of the state with this many fields)."""
        c = np.asarray(current_cells, dtype=np.float64)
        default_state = _default_state(c.shape[-1]) if default_state is None else default_state
        ny, nx, b = c.shape[0], c.shape[1], boundary_size
        out = np.empty((ny + 2 * b, nx + 2 * b, c.shape[2]))
        out[:] = default_state
        if b == 1:
            out[1:ny + 1, 0], out[1:ny + 1, nx + 1] = c[:, 0], c[:, nx - 1]
            out[0, 1:nx + 1], out[ny + 1, 1:nx + 1] = c[0], c[ny - 1]
        elif b == 2:
            out[2:ny + 2, 0], out[2:ny + 2, 1] = c[:, 1], c[:, 0]
            out[2:ny + 2, nx + 2], out[2:ny + 2, nx + 3] = c[:, nx - 1], c[:, nx - 2]
            out[0, 2:nx + 2], out[1, 2:nx + 2] = c[1], c[0]
            out[ny + 2, 2:nx + 2], out[ny + 3, 2:nx + 2] = c[ny - 1], c[ny - 2]
        out[b:b + ny, b:b + nx] = c
        return out


def _time_step(fn_name, params_struct, current_cells, with_boundary, dim, direction, cell_spacing, time_step, bias,
               limiter, order, nf, mode=None, ctx: Context | None = None):
    """One per-sweep call: current_cells (an array of the reference's layout) is
    overwritten in place with the updated states, as the reference's
    computeSLICTimeStep mutates currentCells.  params_struct: one parameter
    struct, or a tuple of them (the two-material models take two)."""
    ctx = ctx or default_context()
    cells = current_cells
    if not (isinstance(cells, np.ndarray) and cells.dtype == np.float64 and cells.flags.c_contiguous):
        raise TypeError("current_cells must be a C-contiguous float64 array (updated in place)")
    g = 2 if order == 2 else 1
    wb = np.ascontiguousarray(with_boundary, dtype=np.float64)
    if dim == 1:
        n = cells.shape[0]
        if cells.shape != (n, nf) or wb.shape != (n + 2 * g, nf):
            raise ValueError(f"expected cells ({n}, {nf}) and a ghosted grid ({n + 2 * g}, {nf})")
        args = [n]
    else:
        ny, nx = cells.shape[:2]
        if cells.shape != (ny, nx, nf) or wb.shape != (ny + 2 * g, nx + 2 * g, nf):
            raise ValueError(f"expected cells ({ny}, {nx}, {nf}) and a ghosted grid "
                             f"({ny + 2 * g}, {nx + 2 * g}, {nf})")
        args = [ny, nx, direction]
    extra = [] if mode is None else [mode]
    structs = params_struct if isinstance(params_struct, tuple) else (params_struct,)
    rc = getattr(ctx.lib, fn_name)(ctx.handle, _lib.as_dptr(cells), _lib.as_dptr(wb), *args, cell_spacing, time_step,
                                   bias, limiter, order, *[C.byref(p) for p in structs], *extra)
    check(rc, fn_name)
    return cells




# ---- two-material diffuse-interface Euler -------------------------------------
# "multiphysics": EulerMultiphysicsStateVector, 8 primitives per cell (alpha1, u,
# v, w, rho1, p1, rho2, p2); "reduced": EulerReducedStateVector, 7 (alpha1, u, v,
# w, rho1, rho2, p).
TWO_MATERIAL_FIELDS = {"multiphysics": 8, "reduced": 7}


def _two_params(params) -> bool:
    """Whether an overload's trailing arguments are two EulerMaterialParameters
    (the two-material states) rather than one (and an optional mode)."""
    return len(params) >= 2 and all(isinstance(p, EulerMaterialParameters) for p in params[-2:])


def _two_material_model(cells) -> str:
    nf = np.shape(cells)[-1]
    for model, n in TWO_MATERIAL_FIELDS.items():
        if n == nf:
            return model
    raise ValueError(f"a two-material state has 8 (multiphysics) or 7 (reduced) fields per cell, not {nf}")


def two_material_solve(model: str, prim: np.ndarray, dim: int, dx: float, cfl: float, t_final: float, *,
                       order: int = 2, bias: float = 0.0, limiter: int = 0, reinit_every: int = 0,
                       params1: EulerMaterialParameters | None = None, params2: EulerMaterialParameters | None = None,
                       max_steps: int = 0, bc: int = TRANSMISSIVE, dt_cap: int = 100000, ctx: Context | None = None,
                       raise_on_error: bool = True) -> SolveResult:
    """Whole solve through vg_euler_{multiphysics,reduced}_solve_{1d,2d}
    (SecondOrderSolver / FirstOrderSolver::solve[2D] for these states)."""
    ctx = ctx or default_context()
    nf = TWO_MATERIAL_FIELDS[model]
    m1 = (params1 or EulerMaterialParameters(1.4)).c_struct()
    m2 = (params2 or EulerMaterialParameters(1.4)).c_struct()
    w = np.ascontiguousarray(prim, dtype=np.float64)
    if w.ndim != dim + 1 or w.shape[-1] != nf:
        raise ValueError(f"{dim}D {model} state must have shape {'(cells' if dim == 1 else '(rows, columns'}, {nf})")
    out = np.empty_like(w)
    dts = np.empty(dt_cap)
    st = _lib.RunStats()
    fn = getattr(ctx.lib, f"vg_euler_{model}_solve_{dim}d")
    grid = [w.shape[0]] if dim == 1 else [w.shape[0], w.shape[1]]
    rc = fn(ctx.handle, _lib.as_dptr(w), _lib.as_dptr(out), *grid, dx, cfl, t_final, max_steps, bias, limiter, order,
            bc, reinit_every, C.byref(m1), C.byref(m2), _lib.as_dptr(dts), dt_cap, C.byref(st))
    if raise_on_error and rc != _lib.VG_OK:
        check(rc, f"vg_euler_{model}_solve_{dim}d")
    return SolveResult(out, dts[:min(st.steps, dt_cap)].copy(), st, rc)


class MultiphysicsSolvers:
    """Solvers/multiphysicssolvers.h:8-39, the Euler overloads."""

    @staticmethod
    def reinitialiseVolumeFraction(current_cells, material1_parameters=None, material2_parameters=None,
                                   ctx: Context | None = None):
        """In place, as the reference mutates currentCells; the material
        parameters do not enter (the reference's unused mixture energy)."""
        ctx = ctx or default_context()
        cells = current_cells
        if not (isinstance(cells, np.ndarray) and cells.dtype == np.float64 and cells.flags.c_contiguous):
            raise TypeError("current_cells must be a C-contiguous float64 array (updated in place)")
        model = _two_material_model(cells)
        dim = cells.ndim - 1
        ny, nx = (1, cells.shape[0]) if dim == 1 else cells.shape[:2]
        fn = f"vg_euler_{model}_reinitialise"
        check(getattr(ctx.lib, fn)(ctx.handle, _lib.as_dptr(cells), dim, ny, nx), fn)
        return cells


class SecondOrderSolver:
    """SLIC (MUSCL-Hancock + FORCE), Solvers/secondordersolver.h:7-72."""

    # The reference overloads these on the state type: EulerStateVector takes one
    # EulerMaterialParameters; EulerMultiphysicsStateVector and
    # EulerReducedStateVector (8 or 7 fields per cell here) take a
    # reinitialisation frequency (solve only) and two
    # (secondordersolver.h:34-70).

    @staticmethod
    def solve(initial_cells, cell_spacing, cfl_coefficient, final_time, bias, slope_limiter,
              subcycling_iterations, *params, **kw):
        if _two_params(params):
            reinit, m1, m2 = params
            return two_material_solve(_two_material_model(initial_cells), initial_cells, 1, cell_spacing,
                                      cfl_coefficient, final_time, order=2, bias=bias, limiter=slope_limiter,
                                      reinit_every=reinit, params1=m1, params2=m2, **kw).cells
        # subcyclingIterations is an empty loop for Euler (secondordersolver.cpp:302-305).
        (material_parameters,) = params
        return euler_solve(initial_cells, 1, cell_spacing, cfl_coefficient, final_time, order=2, bias=bias,
                           limiter=slope_limiter, params=material_parameters, **kw).cells

    @staticmethod
    def solve2D(initial_cells, cell_spacing, cfl_coefficient, final_time, bias, slope_limiter,
                subcycling_iterations, *params, **kw):
        if _two_params(params):
            reinit, m1, m2 = params
            return two_material_solve(_two_material_model(initial_cells), initial_cells, 2, cell_spacing,
                                      cfl_coefficient, final_time, order=2, bias=bias, limiter=slope_limiter,
                                      reinit_every=reinit, params1=m1, params2=m2, **kw).cells
        (material_parameters,) = params
        return euler_solve(initial_cells, 2, cell_spacing, cfl_coefficient, final_time, order=2, bias=bias,
                           limiter=slope_limiter, params=material_parameters, **kw).cells

    @staticmethod
    def _sweep(dim, direction, current_cells, current_cells_with_boundary, cell_spacing, time_step, bias,
               slope_limiter, params, mode):
        if _two_params(params):
            model = _two_material_model(current_cells)
            return _time_step(f"vg_euler_{model}_time_step_{dim}d", tuple(p.c_struct() for p in params),
                              current_cells, current_cells_with_boundary, dim, direction, cell_spacing, time_step,
                              bias, slope_limiter, 2, TWO_MATERIAL_FIELDS[model])
        if len(params) == 2:
            mode = params[1]
        return _time_step(f"vg_euler_time_step_{dim}d", params[0].c_struct(), current_cells,
                          current_cells_with_boundary, dim, direction, cell_spacing, time_step, bias, slope_limiter,
                          2, 6, mode)

    @staticmethod
    def computeSLICTimeStep(current_cells, current_cells_with_boundary, cell_spacing, time_step, bias, slope_limiter,
                            *params, mode=EXACT):
        return SecondOrderSolver._sweep(1, 0, current_cells, current_cells_with_boundary, cell_spacing, time_step,
                                        bias, slope_limiter, params, mode)

    @staticmethod
    def computeXSLICTimeStep2D(current_cells, current_cells_with_boundary, cell_spacing, time_step, bias,
                               slope_limiter, *params, mode=EXACT):
        return SecondOrderSolver._sweep(2, 0, current_cells, current_cells_with_boundary, cell_spacing, time_step,
                                        bias, slope_limiter, params, mode)

    @staticmethod
    def computeYSLICTimeStep2D(current_cells, current_cells_with_boundary, cell_spacing, time_step, bias,
                               slope_limiter, *params, mode=EXACT):
        return SecondOrderSolver._sweep(2, 1, current_cells, current_cells_with_boundary, cell_spacing, time_step,
                                        bias, slope_limiter, params, mode)


class FirstOrderSolver:
    """FORCE, Solvers/firstordersolver.h:7-101."""

    # Overloaded on the state type as SecondOrderSolver (firstordersolver.h:65-100).

    @staticmethod
    def solve(initial_cells, cell_spacing, cfl_coefficient, final_time, subcycling_iterations, *params, **kw):
        if _two_params(params):
            return two_material_solve(_two_material_model(initial_cells), initial_cells, 1, cell_spacing,
                                      cfl_coefficient, final_time, order=1, params1=params[0], params2=params[1],
                                      **kw).cells
        (material_parameters,) = params
        return euler_solve(initial_cells, 1, cell_spacing, cfl_coefficient, final_time, order=1,
                           params=material_parameters, **kw).cells

    @staticmethod
    def solve2D(initial_cells, cell_spacing, cfl_coefficient, final_time, subcycling_iterations, *params, **kw):
        if _two_params(params):
            return two_material_solve(_two_material_model(initial_cells), initial_cells, 2, cell_spacing,
                                      cfl_coefficient, final_time, order=1, params1=params[0], params2=params[1],
                                      **kw).cells
        (material_parameters,) = params
        return euler_solve(initial_cells, 2, cell_spacing, cfl_coefficient, final_time, order=1,
                           params=material_parameters, **kw).cells

    @staticmethod
    def _sweep(dim, direction, current_cells, current_cells_with_boundary, cell_spacing, time_step, params, mode):
        if _two_params(params):
            model = _two_material_model(current_cells)
            return _time_step(f"vg_euler_{model}_time_step_{dim}d", tuple(p.c_struct() for p in params),
                              current_cells, current_cells_with_boundary, dim, direction, cell_spacing, time_step,
                              0.0, 0, 1, TWO_MATERIAL_FIELDS[model])
        if len(params) == 2:
            mode = params[1]
        return _time_step(f"vg_euler_time_step_{dim}d", params[0].c_struct(), current_cells,
                          current_cells_with_boundary, dim, direction, cell_spacing, time_step, 0.0, 0, 1, 6, mode)

    @staticmethod
    def computeFORCETimeStep(current_cells, current_cells_with_boundary, cell_spacing, time_step, *params,
                             mode=EXACT):
        return FirstOrderSolver._sweep(1, 0, current_cells, current_cells_with_boundary, cell_spacing, time_step,
                                       params, mode)

    @staticmethod
    def computeXFORCETimeStep2D(current_cells, current_cells_with_boundary, cell_spacing, time_step, *params,
                                mode=EXACT):
        return FirstOrderSolver._sweep(2, 0, current_cells, current_cells_with_boundary, cell_spacing, time_step,
                                       params, mode)

    @staticmethod
    def computeYFORCETimeStep2D(current_cells, current_cells_with_boundary, cell_spacing, time_step, *params,
                                mode=EXACT):
        return FirstOrderSolver._sweep(2, 1, current_cells, current_cells_with_boundary, cell_spacing, time_step,
                                       params, mode)


class EulerGrid:
    """Device-resident grid (vg_euler_state): load once, advance many times."""

    def __init__(self, dim: int, ny: int, nx: int, ctx: Context | None = None):
        self.ctx = ctx or default_context()
        self.lib = self.ctx.lib
        self.dim, self.ny, self.nx = dim, ny, nx
        self.handle = C.c_void_p()
        check(self.lib.vg_euler_state_create(self.ctx.handle, dim, ny, nx, C.byref(self.handle)),
              "vg_euler_state_create")

    def close(self):
        if self.handle:
            self.lib.vg_euler_state_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def load(self, prim: np.ndarray):
        w = np.ascontiguousarray(prim, dtype=np.float64)
        check(self.lib.vg_euler_state_load(self.handle, w.ctypes.data_as(C.c_void_p), 0), "vg_euler_state_load")

    def load_device(self, ptr: int):
        check(self.lib.vg_euler_state_load(self.handle, C.c_void_p(ptr), 1), "vg_euler_state_load")

    def store(self) -> np.ndarray:
        shape = (self.nx, 6) if self.dim == 1 else (self.ny, self.nx, 6)
        out = np.empty(shape)
        check(self.lib.vg_euler_state_store(self.handle, out.ctypes.data_as(C.c_void_p), 0), "vg_euler_state_store")
        return out

    def run(self, dx, cfl, t_final=math.inf, max_steps=0, *, order=2, bias=0.0, limiter=0, bc=TRANSMISSIVE,
            params: EulerMaterialParameters | None = None, mode=EXACT, dt_cap=0):
        params = params or EulerMaterialParameters(1.4)
        mp = params.c_struct()
        st = _lib.RunStats()
        dts = np.empty(max(dt_cap, 1))
        check(self.lib.vg_euler_state_run(self.handle, order, dx, cfl, t_final, max_steps, bias, limiter, bc,
                                          C.byref(mp), mode, _lib.as_dptr(dts) if dt_cap else None, dt_cap,
                                          C.byref(st)), "vg_euler_state_run")
        return st, dts[:min(st.steps, dt_cap)].copy()

    def sweep(self, direction, dx, dt, *, order=2, bias=0.0, limiter=0, bc=TRANSMISSIVE,
              params: EulerMaterialParameters | None = None, mode=EXACT):
        """vg_euler_state_sweep: one sweep with the caller's time step."""
        params = params or EulerMaterialParameters(1.4)
        mp = params.c_struct()
        check(self.lib.vg_euler_state_sweep(self.handle, direction, order, dx, dt, bias, limiter, bc, C.byref(mp),
                                            mode), "vg_euler_state_sweep")

    def max_wave_speed(self, params: EulerMaterialParameters | None = None, bc=TRANSMISSIVE) -> float:
        params = params or EulerMaterialParameters(1.4)
        mp = params.c_struct()
        s = C.c_double()
        check(self.lib.vg_euler_state_max_wave_speed(self.handle, bc, C.byref(mp), C.byref(s)),
              "vg_euler_state_max_wave_speed")
        return s.value

    def planes(self) -> tuple[int, int]:
        """vg_euler_state_planes: the device address of the current
        structure-of-arrays planes (field f of cell c at element f * plane + c)
        and the plane stride in doubles."""
        ptr = C.c_void_p()
        plane = C.c_int64()
        check(self.lib.vg_euler_state_planes(self.handle, C.byref(ptr), C.byref(plane)), "vg_euler_state_planes")
        return ptr.value, plane.value


# ---- batched 1D solves ---------------------------------------------------------

@dataclass
class BatchResult:
    cells: np.ndarray
    steps: np.ndarray
    t: np.ndarray
    stats: _lib.RunStats
    status: int


def _batch_arrays(prim, offsets, dx, t_final, nf):
    w = np.ascontiguousarray(prim, dtype=np.float64)
    if w.ndim != 2 or w.shape[1] != nf:
        raise ValueError(f"batched state must be (cells, {nf})")
    off = np.ascontiguousarray(offsets, dtype=np.int64)
    n_inst = off.size - 1
    dx = np.ascontiguousarray(np.broadcast_to(np.asarray(dx, dtype=np.float64), (n_inst,)))
    tf = np.ascontiguousarray(np.broadcast_to(np.asarray(t_final, dtype=np.float64), (n_inst,)))
    return w, off, n_inst, dx, tf


def euler_solve_batch(prim, offsets, dx, params, t_final, cfl, *, max_steps=0, bias=0.0, limiter=0, order=2,
                      mode=EXACT, ctx: Context | None = None, raise_on_error=True) -> BatchResult:
    """n_inst independent 1D Euler solves in one persistent launch (vg_euler_solve_batch)."""
    ctx = ctx or default_context()
    w, off, n_inst, dx, tf = _batch_arrays(prim, offsets, dx, t_final, 6)
    if isinstance(params, EulerMaterialParameters):
        params = [params] * n_inst
    mp = (_lib.EulerParams * n_inst)(*[p.c_struct() for p in params])
    out = np.empty_like(w)
    steps = np.empty(n_inst, dtype=np.int32)
    tend = np.empty(n_inst)
    st = _lib.RunStats()
    rc = ctx.lib.vg_euler_solve_batch(ctx.handle, n_inst, off.ctypes.data_as(C.POINTER(C.c_int64)), _lib.as_dptr(w),
                                      _lib.as_dptr(out), _lib.as_dptr(dx), mp, _lib.as_dptr(tf), cfl, max_steps, bias,
                                      limiter, order, mode, steps.ctypes.data_as(C.POINTER(C.c_int32)),
                                      _lib.as_dptr(tend), C.byref(st))
    if raise_on_error and rc != _lib.VG_OK:
        check(rc, "vg_euler_solve_batch")
    return BatchResult(out, steps, tend, st, rc)


# ---- ideal MHD ---------------------------------------------------------------------

@dataclass
class MHDMaterialParameters:
    """MHDMaterialParameters (MHD/mhdmaterialparameters.h); the one-argument
    constructor sets p_inf, c_h and c_p to 0 (mhdmaterialparameters.cpp:12-19)."""

    gamma: float = 1.0
    p_inf: float = 0.0
    ch: float = 0.0
    cp: float = 0.0

    def c_struct(self) -> _lib.MHDParams:
        return _lib.MHDParams(self.gamma, self.p_inf, self.ch, self.cp)


def mhd_solve(prim: np.ndarray, dim: int, dx: float, cfl: float, t_final: float, *, order: int = 2,
              bias: float = 0.0, limiter: int = 0, params: MHDMaterialParameters | None = None, max_steps: int = 0,
              mode: int = EXACT, bc: int = TRANSMISSIVE, subcycles: int = 0, dt_cap: int = 200000,
              ctx: Context | None = None, raise_on_error: bool = True) -> SolveResult:
    """Whole MHD solve through vg_mhd_solve_1d / vg_mhd_solve_2d; `subcycles`
    RK4 GLM-damping sub-cycles around every second-order step."""
    ctx = ctx or default_context()
    params = params or MHDMaterialParameters(2.0)
    w = np.ascontiguousarray(prim, dtype=np.float64)
    out = np.empty_like(w)
    dts = np.empty(dt_cap)
    st = _lib.RunStats()
    mp = params.c_struct()
    if dim == 1:
        if w.ndim != 2 or w.shape[1] != 9:
            raise ValueError("1D MHD state must be (cells, 9)")
        rc = ctx.lib.vg_mhd_solve_1d(ctx.handle, _lib.as_dptr(w), _lib.as_dptr(out), w.shape[0], dx, cfl, t_final,
                                     max_steps, bias, limiter, order, bc, subcycles, C.byref(mp), mode,
                                     _lib.as_dptr(dts), dt_cap, C.byref(st))
    else:
        if w.ndim != 3 or w.shape[2] != 9:
            raise ValueError("2D MHD state must be (rows, columns, 9)")
        rc = ctx.lib.vg_mhd_solve_2d(ctx.handle, _lib.as_dptr(w), _lib.as_dptr(out), w.shape[0], w.shape[1], dx, cfl,
                                     t_final, max_steps, bias, limiter, order, bc, subcycles, C.byref(mp), mode,
                                     _lib.as_dptr(dts), dt_cap, C.byref(st))
    if raise_on_error and rc != _lib.VG_OK:
        check(rc, "vg_mhd_solve")
    return SolveResult(out, dts[:min(st.steps, dt_cap)].copy(), st, rc)


def mhd_solve_batch(prim, offsets, dx, gamma, t_final, cfl, *, max_steps=0, bias=0.0, limiter=1, order=2,
                    mode=EXACT, p_inf=0.0, ctx: Context | None = None, raise_on_error=True) -> BatchResult:
    """n_inst independent MHDSecondOrderSolver::solve calls in one persistent launch."""
    ctx = ctx or default_context()
    w, off, n_inst, dx, tf = _batch_arrays(prim, offsets, dx, t_final, 9)
    g = np.broadcast_to(np.asarray(gamma, dtype=np.float64), (n_inst,))
    mp = (_lib.MHDParams * n_inst)(*[_lib.MHDParams(float(x), p_inf, 0.0, 0.0) for x in g])
    out = np.empty_like(w)
    steps = np.empty(n_inst, dtype=np.int32)
    tend = np.empty(n_inst)
    st = _lib.RunStats()
    rc = ctx.lib.vg_mhd_solve_batch(ctx.handle, n_inst, off.ctypes.data_as(C.POINTER(C.c_int64)), _lib.as_dptr(w),
                                    _lib.as_dptr(out), _lib.as_dptr(dx), mp, _lib.as_dptr(tf), cfl, max_steps, bias,
                                    limiter, order, mode, steps.ctypes.data_as(C.POINTER(C.c_int32)),
                                    _lib.as_dptr(tend), C.byref(st))
    if raise_on_error and rc != _lib.VG_OK:
        check(rc, "vg_mhd_solve_batch")
    return BatchResult(out, steps, tend, st, rc)


class MHDSecondOrderSolver:
    """MHDSecondOrderSolver (Solvers/mhdsecondordersolver.h:58, 70), RK4 GLM
    sub-cycles (MHDForcingSolver) included."""

    @staticmethod
    def solve(initial_cells, cell_spacing, cfl_coefficient, final_time, bias, slope_limiter, subcycling_iterations,
              material_parameters, **kw):
        return mhd_solve(initial_cells, 1, cell_spacing, cfl_coefficient, final_time, order=2, bias=bias,
                         limiter=slope_limiter, params=material_parameters, subcycles=subcycling_iterations,
                         **kw).cells

    @staticmethod
    def solve2D(initial_cells, cell_spacing, cfl_coefficient, final_time, bias, slope_limiter, subcycling_iterations,
                material_parameters, **kw):
        return mhd_solve(initial_cells, 2, cell_spacing, cfl_coefficient, final_time, order=2, bias=bias,
                         limiter=slope_limiter, params=material_parameters, subcycles=subcycling_iterations,
                         **kw).cells

    @staticmethod
    def computeSLICTimeStep(current_cells, current_cells_with_boundary, cell_spacing, time_step, bias, slope_limiter,
                            material_parameters, mode=EXACT):
        return _time_step("vg_mhd_time_step_1d", material_parameters.c_struct(), current_cells,
                          current_cells_with_boundary, 1, 0, cell_spacing, time_step, bias, slope_limiter, 2, 9, mode)

    @staticmethod
    def computeXSLICTimeStep2D(current_cells, current_cells_with_boundary, cell_spacing, time_step, bias,
                               slope_limiter, material_parameters, mode=EXACT):
        return _time_step("vg_mhd_time_step_2d", material_parameters.c_struct(), current_cells,
                          current_cells_with_boundary, 2, 0, cell_spacing, time_step, bias, slope_limiter, 2, 9, mode)

    @staticmethod
    def computeYSLICTimeStep2D(current_cells, current_cells_with_boundary, cell_spacing, time_step, bias,
                               slope_limiter, material_parameters, mode=EXACT):
        return _time_step("vg_mhd_time_step_2d", material_parameters.c_struct(), current_cells,
                          current_cells_with_boundary, 2, 1, cell_spacing, time_step, bias, slope_limiter, 2, 9, mode)


class MHDFirstOrderSolver:
    """MHDFirstOrderSolver (Solvers/mhdfirstordersolver.h): FORCE; never
    sub-cycles and never sets c_h (the parameters' c_h is used)."""

    @staticmethod
    def solve(initial_cells, cell_spacing, cfl_coefficient, final_time, subcycling_iterations, material_parameters,
              **kw):
        return mhd_solve(initial_cells, 1, cell_spacing, cfl_coefficient, final_time, order=1,
                         params=material_parameters, **kw).cells

    @staticmethod
    def solve2D(initial_cells, cell_spacing, cfl_coefficient, final_time, subcycling_iterations, material_parameters,
                **kw):
        return mhd_solve(initial_cells, 2, cell_spacing, cfl_coefficient, final_time, order=1,
                         params=material_parameters, **kw).cells

    @staticmethod
    def computeFORCETimeStep(current_cells, current_cells_with_boundary, cell_spacing, time_step,
                             material_parameters, mode=EXACT):
        return _time_step("vg_mhd_time_step_1d", material_parameters.c_struct(), current_cells,
                          current_cells_with_boundary, 1, 0, cell_spacing, time_step, 0.0, 0, 1, 9, mode)

    @staticmethod
    def computeXFORCETimeStep2D(current_cells, current_cells_with_boundary, cell_spacing, time_step,
                                material_parameters, mode=EXACT):
        return _time_step("vg_mhd_time_step_2d", material_parameters.c_struct(), current_cells,
                          current_cells_with_boundary, 2, 0, cell_spacing, time_step, 0.0, 0, 1, 9, mode)

    @staticmethod
    def computeYFORCETimeStep2D(current_cells, current_cells_with_boundary, cell_spacing, time_step,
                                material_parameters, mode=EXACT):
        return _time_step("vg_mhd_time_step_2d", material_parameters.c_struct(), current_cells,
                          current_cells_with_boundary, 2, 1, cell_spacing, time_step, 0.0, 0, 1, 9, mode)


# ---- Godunov-Peshkov-Romenski (HPR) ----------------------------------------------

@dataclass


class HPRMaterialParameters:
    """HPRMaterialParameters (HPR/hprmaterialparameters.h:10-15): the
    constructor's arguments in order, the EOS name as a string."""

    equation_of_state: str = "StiffenedGas"
    is_thermal: bool = False
    is_plastic: bool = False
    ideal_gas_constant: float = 8.31445985
    reference_density: float = 1.0
    reference_pressure: float = 1.0 / 1.4
    reference_temperature: float = 0.0
    initial_temperature: float = 1.0
    specific_heat_capacity: float = 2.5
    adiabatic_index: float = 1.4
    stiffening_parameter: float = 0.0
    reference_sound_speed: float = 0.0
    reference_gruneisen_coefficient: float = 0.0
    hugoniot_slope_coefficient: float = 0.0
    reference_internal_energy: float = 0.0
    transverse_wave_speed: float = 1.0
    strain_dissipation_time: float = 0.0
    thermal_impulse_relaxation_time: float = 0.0
    dynamic_viscosity_coefficient: float = 1e-2
    prandtl_number: float = 0.75
    elastic_plastic_transition_parameter: float = 0.0
    power_law_index: float = 1.0
    heat_wave_speed: float = 2.0
    alpha_parameter: float = 0.0
    beta_parameter: float = 0.0
    gamma_parameter: float = 0.0

    _EOS = {"StiffenedGas": _lib.VG_HPR_EOS_STIFFENED_GAS, "ShockMieGruneisen": _lib.VG_HPR_EOS_SHOCK_MIE_GRUNEISEN,
            "GodunovRomenski": _lib.VG_HPR_EOS_GODUNOV_ROMENSKI}

    def configure_strain_dissipation_time(self):
        """tau1 = 6 mu / (p_ref c_s^2) (hprmaterialparameters.cpp:101-104)."""
        self.strain_dissipation_time = 6.0 * (self.dynamic_viscosity_coefficient /
                                              (self.reference_pressure * (self.transverse_wave_speed ** 2)))
        return self

    def configure_thermal_impulse_relaxation_time(self, heat_conduction_coefficient: float | None = None):
        """tau2 = kappa rho0 / (T0 alpha^2) (hprmaterialparameters.cpp:112-119);
        without an argument kappa = mu gamma (c_v / Pr) (:106-114).  T0 = 0 gives
        an infinite tau2, as in the reference."""
        if heat_conduction_coefficient is None:
            heat_conduction_coefficient = (self.dynamic_viscosity_coefficient * self.adiabatic_index) * (
                self.specific_heat_capacity / self.prandtl_number)
        den = np.float64(self.initial_temperature) * (np.float64(self.heat_wave_speed) * self.heat_wave_speed)
        with np.errstate(divide="ignore", invalid="ignore"):   # IEEE division, as in C++
            ratio = np.float64(self.reference_density) / den
        self.thermal_impulse_relaxation_time = float(heat_conduction_coefficient * ratio)
        return self

    def c_struct(self) -> _lib.HPRParams:
        if self.equation_of_state not in self._EOS:
            raise ValueError(f"unknown equation of state {self.equation_of_state!r}")
        return _lib.HPRParams(self._EOS[self.equation_of_state], int(self.is_thermal), int(self.is_plastic), 0,
                              *[float(getattr(self, n)) for n in _lib.HPR_DOUBLE_FIELDS])


def hpr_solve(prim: np.ndarray, dim: int, dx: float, cfl: float, t_final: float, *, order: int = 2,
              bias: float = 0.0, limiter: int = 0, subcycles: int = 0, params: HPRMaterialParameters | None = None,
              max_steps: int = 0, mode: int = EXACT, bc: int = TRANSMISSIVE, dt_cap: int = 100000,
              ctx: Context | None = None, raise_on_error: bool = True) -> SolveResult:
    """Whole HPR solve through vg_hpr_solve_1d / vg_hpr_solve_2d."""
    ctx = ctx or default_context()
    params = params or HPRMaterialParameters()
    w = np.ascontiguousarray(prim, dtype=np.float64)
    out = np.empty_like(w)
    dts = np.empty(dt_cap)
    st = _lib.RunStats()
    mp = params.c_struct()
    if dim == 1:
        if w.ndim != 2 or w.shape[1] != 17:
            raise ValueError("1D HPR state must be (cells, 17)")
        rc = ctx.lib.vg_hpr_solve_1d(ctx.handle, _lib.as_dptr(w), _lib.as_dptr(out), w.shape[0], dx, cfl, t_final,
                                     max_steps, bias, limiter, order, bc, subcycles, C.byref(mp), mode,
                                     _lib.as_dptr(dts), dt_cap, C.byref(st))
    else:
        if w.ndim != 3 or w.shape[2] != 17:
            raise ValueError("2D HPR state must be (rows, columns, 17)")
        rc = ctx.lib.vg_hpr_solve_2d(ctx.handle, _lib.as_dptr(w), _lib.as_dptr(out), w.shape[0], w.shape[1], dx, cfl,
                                     t_final, max_steps, bias, limiter, order, bc, subcycles, C.byref(mp), mode,
                                     _lib.as_dptr(dts), dt_cap, C.byref(st))
    if raise_on_error and rc != _lib.VG_OK:
        check(rc, "vg_hpr_solve")
    return SolveResult(out, dts[:min(st.steps, dt_cap)].copy(), st, rc)


@dataclass
class HPRAdjointResult:
    cells: np.ndarray      # final primitive state of the forward solve
    grad: np.ndarray       # d loss / d initial primitive state
    grad_tau: np.ndarray   # d loss / d (strain_dissipation_time, thermal_impulse_relaxation_time)
    dt: np.ndarray         # the time steps the forward chose (held fixed)
    stats: _lib.RunStats


def hpr_adjoint(prim: np.ndarray, seed: np.ndarray, dim: int, dx: float, cfl: float, t_final: float, *,
                bias: float = 0.0, limiter: int = 0, subcycles: int = 0, params: HPRMaterialParameters | None = None,
                max_steps: int = 0, bc: int = TRANSMISSIVE, checkpoint_every: int = 0, dt_cap: int = 100000,
                ctx: Context | None = None) -> HPRAdjointResult:
    """Reverse-mode derivative of the second-order HPR solve (vg_hpr_adjoint_1d /
    _2d), the time steps held at the values the forward chose: `grad` is
    d loss / d initial primitive state given `seed` = d loss / d final state,
    `grad_tau` the derivative with respect to the two relaxation times."""
    ctx = ctx or default_context()
    params = params or HPRMaterialParameters()
    w = np.ascontiguousarray(prim, dtype=np.float64)
    sd = np.ascontiguousarray(seed, dtype=np.float64)
    if sd.shape != w.shape:
        raise ValueError("seed must have the shape of the state")
    out, grad, gtau = np.empty_like(w), np.empty_like(w), np.zeros(2)
    dts = np.empty(dt_cap)
    st = _lib.RunStats()
    mp = params.c_struct()
    if dim == 1:
        if w.ndim != 2 or w.shape[1] != 17:
            raise ValueError("1D HPR state must be (cells, 17)")
        rc = ctx.lib.vg_hpr_adjoint_1d(ctx.handle, _lib.as_dptr(w), _lib.as_dptr(out), _lib.as_dptr(sd),
                                       _lib.as_dptr(grad), _lib.as_dptr(gtau), w.shape[0], dx, cfl, t_final,
                                       max_steps, bias, limiter, bc, subcycles, C.byref(mp), checkpoint_every,
                                       _lib.as_dptr(dts), dt_cap, C.byref(st))
    else:
        if w.ndim != 3 or w.shape[2] != 17:
            raise ValueError("2D HPR state must be (rows, columns, 17)")
        rc = ctx.lib.vg_hpr_adjoint_2d(ctx.handle, _lib.as_dptr(w), _lib.as_dptr(out), _lib.as_dptr(sd),
                                       _lib.as_dptr(grad), _lib.as_dptr(gtau), w.shape[0], w.shape[1], dx, cfl,
                                       t_final, max_steps, bias, limiter, bc, subcycles, C.byref(mp),
                                       checkpoint_every, _lib.as_dptr(dts), dt_cap, C.byref(st))
    check(rc, "vg_hpr_adjoint")
    return HPRAdjointResult(out, grad, gtau, dts[:min(st.steps, dt_cap)].copy(), st)


def hpr_solve_fixed_dt(prim: np.ndarray, dim: int, dx: float, dts, *, bias: float = 0.0, limiter: int = 0,
                       subcycles: int = 0, params: HPRMaterialParameters | None = None, bc: int = TRANSMISSIVE,
                       ctx: Context | None = None) -> np.ndarray:
    """The second-order HPR solve with prescribed time steps (the map
    hpr_adjoint differentiates)."""
    ctx = ctx or default_context()
    params = params or HPRMaterialParameters()
    w = np.ascontiguousarray(prim, dtype=np.float64)
    d = np.ascontiguousarray(dts, dtype=np.float64)
    out = np.empty_like(w)
    mp = params.c_struct()
    if dim == 1:
        rc = ctx.lib.vg_hpr_solve_fixed_dt_1d(ctx.handle, _lib.as_dptr(w), _lib.as_dptr(out), w.shape[0], dx,
                                              _lib.as_dptr(d), d.size, bias, limiter, bc, subcycles, C.byref(mp))
    else:
        rc = ctx.lib.vg_hpr_solve_fixed_dt_2d(ctx.handle, _lib.as_dptr(w), _lib.as_dptr(out), w.shape[0], w.shape[1],
                                              dx, _lib.as_dptr(d), d.size, bias, limiter, bc, subcycles, C.byref(mp))
    check(rc, "vg_hpr_solve_fixed_dt")
    return out


class HPRSecondOrderSolver:
    """HPRSecondOrderSolver (Solvers/hprsecondordersolver.h:56, 67)."""

    @staticmethod
    def solve(initial_cells, cell_spacing, cfl_coefficient, final_time, bias, slope_limiter, subcycling_iterations,
              material_parameters, **kw):
        return hpr_solve(initial_cells, 1, cell_spacing, cfl_coefficient, final_time, order=2, bias=bias,
                         limiter=slope_limiter, subcycles=subcycling_iterations, params=material_parameters,
                         **kw).cells

    @staticmethod
    def solve2D(initial_cells, cell_spacing, cfl_coefficient, final_time, bias, slope_limiter, subcycling_iterations,
                material_parameters, **kw):
        return hpr_solve(initial_cells, 2, cell_spacing, cfl_coefficient, final_time, order=2, bias=bias,
                         limiter=slope_limiter, subcycles=subcycling_iterations, params=material_parameters,
                         **kw).cells

    @staticmethod
    def computeSLICTimeStep(current_cells, current_cells_with_boundary, cell_spacing, time_step, bias, slope_limiter,
                            material_parameters):
        return _time_step("vg_hpr_time_step_1d", material_parameters.c_struct(), current_cells,
                          current_cells_with_boundary, 1, 0, cell_spacing, time_step, bias, slope_limiter, 2, 17)

    @staticmethod
    def computeXSLICTimeStep2D(current_cells, current_cells_with_boundary, cell_spacing, time_step, bias,
                               slope_limiter, material_parameters):
        return _time_step("vg_hpr_time_step_2d", material_parameters.c_struct(), current_cells,
                          current_cells_with_boundary, 2, 0, cell_spacing, time_step, bias, slope_limiter, 2, 17)

    @staticmethod
    def computeYSLICTimeStep2D(current_cells, current_cells_with_boundary, cell_spacing, time_step, bias,
                               slope_limiter, material_parameters):
        return _time_step("vg_hpr_time_step_2d", material_parameters.c_struct(), current_cells,
                          current_cells_with_boundary, 2, 1, cell_spacing, time_step, bias, slope_limiter, 2, 17)


class HPRFirstOrderSolver:
    """HPRFirstOrderSolver (Solvers/hprfirstordersolver.h): FORCE."""

    @staticmethod
    def solve(initial_cells, cell_spacing, cfl_coefficient, final_time, subcycling_iterations, material_parameters,
              **kw):
        return hpr_solve(initial_cells, 1, cell_spacing, cfl_coefficient, final_time, order=1,
                         subcycles=subcycling_iterations, params=material_parameters, **kw).cells

    @staticmethod
    def solve2D(initial_cells, cell_spacing, cfl_coefficient, final_time, subcycling_iterations, material_parameters,
                **kw):
        return hpr_solve(initial_cells, 2, cell_spacing, cfl_coefficient, final_time, order=1,
                         subcycles=subcycling_iterations, params=material_parameters, **kw).cells

    @staticmethod
    def computeFORCETimeStep(current_cells, current_cells_with_boundary, cell_spacing, time_step,
                             material_parameters):
        return _time_step("vg_hpr_time_step_1d", material_parameters.c_struct(), current_cells,
                          current_cells_with_boundary, 1, 0, cell_spacing, time_step, 0.0, 0, 1, 17)

    @staticmethod
    def computeXFORCETimeStep2D(current_cells, current_cells_with_boundary, cell_spacing, time_step,
                                material_parameters):
        return _time_step("vg_hpr_time_step_2d", material_parameters.c_struct(), current_cells,
                          current_cells_with_boundary, 2, 0, cell_spacing, time_step, 0.0, 0, 1, 17)

    @staticmethod
    def computeYFORCETimeStep2D(current_cells, current_cells_with_boundary, cell_spacing, time_step,
                                material_parameters):
        return _time_step("vg_hpr_time_step_2d", material_parameters.c_struct(), current_cells,
                          current_cells_with_boundary, 2, 1, cell_spacing, time_step, 0.0, 0, 1, 17)




class HPRGrid:
    """Device-resident HPR grid (vg_hpr_state)."""

    def __init__(self, dim: int, ny: int, nx: int, ctx: Context | None = None):
        self.ctx = ctx or default_context()
        self.lib = self.ctx.lib
        self.dim, self.ny, self.nx = dim, ny, nx
        self.handle = C.c_void_p()
        check(self.lib.vg_hpr_state_create(self.ctx.handle, dim, ny, nx, C.byref(self.handle)), "vg_hpr_state_create")

    def close(self):
        if self.handle:
            self.lib.vg_hpr_state_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def load(self, prim: np.ndarray):
        w = np.ascontiguousarray(prim, dtype=np.float64)
        check(self.lib.vg_hpr_state_load(self.handle, w.ctypes.data_as(C.c_void_p), 0), "vg_hpr_state_load")

    def load_device(self, ptr: int):
        check(self.lib.vg_hpr_state_load(self.handle, C.c_void_p(ptr), 1), "vg_hpr_state_load")

    def store(self) -> np.ndarray:
        shape = (self.nx, 17) if self.dim == 1 else (self.ny, self.nx, 17)
        out = np.empty(shape)
        check(self.lib.vg_hpr_state_store(self.handle, out.ctypes.data_as(C.c_void_p), 0), "vg_hpr_state_store")
        return out

    def store_device(self, ptr: int):
        check(self.lib.vg_hpr_state_store(self.handle, C.c_void_p(ptr), 1), "vg_hpr_state_store")

    def run(self, dx, cfl, t_final=math.inf, max_steps=0, *, order=2, bias=0.0, limiter=0, bc=TRANSMISSIVE,
            subcycles=0, params: HPRMaterialParameters | None = None, dt_cap=0):
        params = params or HPRMaterialParameters()
        mp = params.c_struct()
        st = _lib.RunStats()
        dts = np.empty(max(dt_cap, 1))
        check(self.lib.vg_hpr_state_run(self.handle, order, dx, cfl, t_final, max_steps, bias, limiter, bc, subcycles,
                                        C.byref(mp), EXACT, _lib.as_dptr(dts) if dt_cap else None, dt_cap,
                                        C.byref(st)), "vg_hpr_state_run")
        return st, dts[:min(st.steps, dt_cap)].copy()

    def sweep(self, direction, dx, dt, *, order=2, bias=0.0, limiter=0, bc=TRANSMISSIVE,
              params: HPRMaterialParameters | None = None):
        """vg_hpr_state_sweep: one sweep with the caller's time step (no source)."""
        params = params or HPRMaterialParameters()
        mp = params.c_struct()
        check(self.lib.vg_hpr_state_sweep(self.handle, direction, order, dx, dt, bias, limiter, bc, C.byref(mp)),
              "vg_hpr_state_sweep")

    def max_wave_speed(self, params: HPRMaterialParameters | None = None, bc=TRANSMISSIVE) -> float:
        params = params or HPRMaterialParameters()
        mp = params.c_struct()
        s = C.c_double()
        check(self.lib.vg_hpr_state_max_wave_speed(self.handle, bc, C.byref(mp), C.byref(s)),
              "vg_hpr_state_max_wave_speed")
        return s.value


# ---- two-material Euler: RGFM ------------------------------------------------------

HLLC = _lib.VG_RIEMANN_HLLC
EXACT_RIEMANN = _lib.VG_RIEMANN_EXACT


@dataclass
class MultimaterialSystem:
    """MultimaterialSystem (RGFM/multimaterialsystem.h:8-39): both materials'
    primitive states over the whole domain and the level set."""

    material1: np.ndarray
    material2: np.ndarray
    level_set: np.ndarray


@dataclass
class RGFMResult:
    system: MultimaterialSystem
    dt: np.ndarray
    stats: _lib.RunStats
    status: int
    circulation: float | None = None


def _params(p):
    if isinstance(p, EulerMaterialParameters):
        return p
    return EulerMaterialParameters(*p) if isinstance(p, (tuple, list)) else EulerMaterialParameters(float(p))


def rgfm_solve(system: MultimaterialSystem, dx: float, cfl: float, t_final: float, params1, params2, *,
               dy: float | None = None, bias: float = 0.0, limiter: int = 0, riemann: int = HLLC, max_steps: int = 0,
               mode: int = EXACT, dt_cap: int = 200000, ctx: Context | None = None,
               raise_on_error: bool = True) -> RGFMResult:
    """RGFMSolver::solveHLLC[2D] / solveExact[2D] through vg_rgfm_solve_1d / _2d."""
    ctx = ctx or default_context()
    m1 = np.ascontiguousarray(system.material1, dtype=np.float64)
    m2 = np.ascontiguousarray(system.material2, dtype=np.float64)
    phi = np.ascontiguousarray(system.level_set, dtype=np.float64)
    if m1.shape != m2.shape or m1.shape[:-1] != phi.shape or m1.shape[-1] != 6:
        raise ValueError("materials must be (..., 6) over the level set's shape")
    o1, o2, oph = np.empty_like(m1), np.empty_like(m2), np.empty_like(phi)
    dts = np.empty(dt_cap)
    st = _lib.RunStats()
    q1, q2 = _params(params1).c_struct(), _params(params2).c_struct()
    circ = C.c_double(math.nan)
    P = _lib.as_dptr
    if phi.ndim == 1:
        rc = ctx.lib.vg_rgfm_solve_1d(ctx.handle, P(m1), P(m2), P(phi), P(o1), P(o2), P(oph), phi.shape[0], dx, cfl,
                                      t_final, max_steps, bias, limiter, riemann, C.byref(q1), C.byref(q2), mode,
                                      P(dts), dt_cap, C.byref(st))
    else:
        rc = ctx.lib.vg_rgfm_solve_2d(ctx.handle, P(m1), P(m2), P(phi), P(o1), P(o2), P(oph), phi.shape[0],
                                      phi.shape[1], dx, dx if dy is None else dy, cfl, t_final, max_steps, bias,
                                      limiter, riemann, C.byref(q1), C.byref(q2), mode, P(dts), dt_cap, C.byref(st),
                                      C.byref(circ))
    if raise_on_error and rc != _lib.VG_OK:
        check(rc, "vg_rgfm_solve")
    return RGFMResult(MultimaterialSystem(o1, o2, oph), dts[:min(st.steps, dt_cap)].copy(), st, rc,
                      None if phi.ndim == 1 else circ.value)


@dataclass
class RGFMBatchResult:
    material1: np.ndarray
    material2: np.ndarray
    level_set: np.ndarray
    steps: np.ndarray
    t: np.ndarray
    circulation: np.ndarray
    stats: _lib.RunStats
    status: int


def rgfm_solve_batch_2d(m1, m2, phi, dx, dy, cfl, t_final, params1, params2, *, bias=0.0, limiter=0, riemann=HLLC,
                        max_steps=0, mode=EXACT, ctx: Context | None = None, raise_on_error=True) -> RGFMBatchResult:
    """n_inst instances (leading axis) through vg_rgfm_solve_batch_2d."""
    ctx = ctx or default_context()
    a1 = np.ascontiguousarray(m1, dtype=np.float64)
    a2 = np.ascontiguousarray(m2, dtype=np.float64)
    ph = np.ascontiguousarray(phi, dtype=np.float64)
    n_inst, ny, nx = ph.shape
    tf = np.ascontiguousarray(np.broadcast_to(np.asarray(t_final, dtype=np.float64), (n_inst,)))
    o1, o2, oph = np.empty_like(a1), np.empty_like(a2), np.empty_like(ph)
    steps = np.zeros(n_inst, dtype=np.int32)
    tend = np.zeros(n_inst)
    circ = np.zeros(n_inst)
    st = _lib.RunStats()
    q1, q2 = _params(params1).c_struct(), _params(params2).c_struct()
    P = _lib.as_dptr
    rc = ctx.lib.vg_rgfm_solve_batch_2d(ctx.handle, n_inst, P(a1), P(a2), P(ph), P(o1), P(o2), P(oph), ny, nx, dx, dy,
                                        cfl, P(tf), max_steps, bias, limiter, riemann, C.byref(q1), C.byref(q2), mode,
                                        steps.ctypes.data_as(C.POINTER(C.c_int32)), P(tend), P(circ), C.byref(st))
    if raise_on_error and rc != _lib.VG_OK:
        check(rc, "vg_rgfm_solve_batch_2d")
    return RGFMBatchResult(o1, o2, oph, steps, tend, circ, st, rc)


@dataclass
class RGFMAdjointResult:
    system: MultimaterialSystem   # final states of the forward
    grad1: np.ndarray             # d loss / d initial material 1
    grad2: np.ndarray             # d loss / d initial material 2
    dt: np.ndarray                # the time steps the forward chose (held fixed)
    stats: _lib.RunStats
    status: int = 0               # VG_OK, or VG_ERR_NONFINITE (see rgfm_adjoint)


def rgfm_adjoint(system: MultimaterialSystem, seed1: np.ndarray, seed2: np.ndarray, dx: float, cfl: float,
                 t_final: float, params1, params2, *, dy: float | None = None, bias: float = 0.0, limiter: int = 0,
                 max_steps: int = 0, checkpoint_every: int = 0, dt_cap: int = 200000,
                 ctx: Context | None = None, raise_on_error: bool = True) -> RGFMAdjointResult:
    """Reverse-mode derivative of the HLLC RGFM solve (vg_rgfm_adjoint_1d / _2d):
    seeds are d loss / d final material states, the result holds d loss / d
    initial material states, time steps frozen.  A forward that ends with
    inadmissible real cells, or a gradient that is not finite, gives
    VG_ERR_NONFINITE (raised unless raise_on_error is False; the status is in
    the result's `status`)."""
    ctx = ctx or default_context()
    m1 = np.ascontiguousarray(system.material1, dtype=np.float64)
    m2 = np.ascontiguousarray(system.material2, dtype=np.float64)
    phi = np.ascontiguousarray(system.level_set, dtype=np.float64)
    s1 = np.ascontiguousarray(seed1, dtype=np.float64)
    s2 = np.ascontiguousarray(seed2, dtype=np.float64)
    if m1.shape != m2.shape or m1.shape[:-1] != phi.shape or m1.shape[-1] != 6 or s1.shape != m1.shape \
            or s2.shape != m1.shape:
        raise ValueError("materials and seeds must be (..., 6) over the level set's shape")
    o1, o2, oph = np.empty_like(m1), np.empty_like(m2), np.empty_like(phi)
    g1, g2 = np.empty_like(m1), np.empty_like(m2)
    dts = np.empty(dt_cap)
    st = _lib.RunStats()
    q1, q2 = _params(params1).c_struct(), _params(params2).c_struct()
    P = _lib.as_dptr
    if phi.ndim == 1:
        rc = ctx.lib.vg_rgfm_adjoint_1d(ctx.handle, P(m1), P(m2), P(phi), P(o1), P(o2), P(oph), P(s1), P(s2), P(g1),
                                        P(g2), phi.shape[0], dx, cfl, t_final, max_steps, bias, limiter,
                                        C.byref(q1), C.byref(q2), checkpoint_every, P(dts), dt_cap, C.byref(st))
    else:
        rc = ctx.lib.vg_rgfm_adjoint_2d(ctx.handle, P(m1), P(m2), P(phi), P(o1), P(o2), P(oph), P(s1), P(s2), P(g1),
                                        P(g2), phi.shape[0], phi.shape[1], dx, dx if dy is None else dy, cfl,
                                        t_final, max_steps, bias, limiter, C.byref(q1), C.byref(q2),
                                        checkpoint_every, P(dts), dt_cap, C.byref(st))
    if raise_on_error and rc != _lib.VG_OK:
        check(rc, "vg_rgfm_adjoint")
    return RGFMAdjointResult(MultimaterialSystem(o1, o2, oph), g1, g2, dts[:min(st.steps, dt_cap)].copy(), st, rc)


@dataclass
class RGFMAdjointBatchResult:
    material1: np.ndarray         # final states of the forwards, (n_inst, ny, nx, 6)
    material2: np.ndarray
    level_set: np.ndarray         # (n_inst, ny, nx)
    grad1: np.ndarray             # d loss / d initial material 1, per instance
    grad2: np.ndarray
    steps: np.ndarray             # step count per instance
    status: np.ndarray            # per instance: VG_OK or VG_ERR_NONFINITE
    stats: _lib.RunStats
    rc: int


def rgfm_adjoint_batch_2d(m1, m2, phi, seed1, seed2, dx, dy, cfl, t_final, params1, params2, *, bias=0.0, limiter=0,
                          max_steps=0, checkpoint_every=0, ctx: Context | None = None,
                          raise_on_error=True) -> RGFMAdjointBatchResult:
    """The adjoints of n_inst instances (leading axis) through
    vg_rgfm_adjoint_batch_2d: each instance as rgfm_adjoint, on the context's
    worker streams.  An instance without a derivative has VG_ERR_NONFINITE in
    `status`; the call's status is then VG_ERR_NONFINITE (raised unless
    raise_on_error is False)."""
    ctx = ctx or default_context()
    a1 = np.ascontiguousarray(m1, dtype=np.float64)
    a2 = np.ascontiguousarray(m2, dtype=np.float64)
    ph = np.ascontiguousarray(phi, dtype=np.float64)
    s1 = np.ascontiguousarray(seed1, dtype=np.float64)
    s2 = np.ascontiguousarray(seed2, dtype=np.float64)
    n_inst, ny, nx = ph.shape
    if a1.shape != (n_inst, ny, nx, 6) or a2.shape != a1.shape or s1.shape != a1.shape or s2.shape != a1.shape:
        raise ValueError("materials and seeds must be (n_inst, ny, nx, 6) over the level sets' shape")
    tf = np.ascontiguousarray(np.broadcast_to(np.asarray(t_final, dtype=np.float64), (n_inst,)))
    o1, o2, oph = np.empty_like(a1), np.empty_like(a2), np.empty_like(ph)
    g1, g2 = np.empty_like(a1), np.empty_like(a2)
    steps = np.zeros(n_inst, dtype=np.int32)
    status = np.zeros(n_inst, dtype=np.int32)
    st = _lib.RunStats()
    q1, q2 = _params(params1).c_struct(), _params(params2).c_struct()
    P = _lib.as_dptr
    i32 = C.POINTER(C.c_int32)
    rc = ctx.lib.vg_rgfm_adjoint_batch_2d(ctx.handle, n_inst, P(a1), P(a2), P(ph), P(o1), P(o2), P(oph), P(s1),
                                          P(s2), P(g1), P(g2), ny, nx, dx, dy, cfl, P(tf), max_steps, bias, limiter,
                                          C.byref(q1), C.byref(q2), checkpoint_every, steps.ctypes.data_as(i32),
                                          status.ctypes.data_as(i32), C.byref(st))
    if raise_on_error and rc != _lib.VG_OK:
        check(rc, "vg_rgfm_adjoint_batch_2d")
    return RGFMAdjointBatchResult(o1, o2, oph, g1, g2, steps, status, st, rc)


def rgfm_solve_fixed_dt(system: MultimaterialSystem, dx: float, dts, params1, params2, *, dy: float | None = None,
                        bias: float = 0.0, limiter: int = 0, ctx: Context | None = None) -> MultimaterialSystem:
    """HLLC RGFM solve with prescribed time steps (vg_rgfm_solve_fixed_dt_1d / _2d)."""
    ctx = ctx or default_context()
    m1 = np.ascontiguousarray(system.material1, dtype=np.float64)
    m2 = np.ascontiguousarray(system.material2, dtype=np.float64)
    phi = np.ascontiguousarray(system.level_set, dtype=np.float64)
    d = np.ascontiguousarray(dts, dtype=np.float64).ravel()
    o1, o2, oph = np.empty_like(m1), np.empty_like(m2), np.empty_like(phi)
    q1, q2 = _params(params1).c_struct(), _params(params2).c_struct()
    P = _lib.as_dptr
    if phi.ndim == 1:
        rc = ctx.lib.vg_rgfm_solve_fixed_dt_1d(ctx.handle, P(m1), P(m2), P(phi), P(o1), P(o2), P(oph), phi.shape[0],
                                               dx, P(d), d.size, bias, limiter, C.byref(q1), C.byref(q2))
    else:
        rc = ctx.lib.vg_rgfm_solve_fixed_dt_2d(ctx.handle, P(m1), P(m2), P(phi), P(o1), P(o2), P(oph), phi.shape[0],
                                               phi.shape[1], dx, dx if dy is None else dy, P(d), d.size, bias,
                                               limiter, C.byref(q1), C.byref(q2))
    check(rc, "vg_rgfm_solve_fixed_dt")
    return MultimaterialSystem(o1, o2, oph)


class RGFMSolver:
    """RGFMSolver (RGFM/rgfmsolver.h:8-41): the whole-solve entry points."""

    @staticmethod
    def _solve(system, cell_spacing, cfl, final_time, bias, limiter, p1, p2, riemann, **kw):
        return rgfm_solve(system, cell_spacing, cfl, final_time, p1, p2, bias=bias, limiter=limiter, riemann=riemann,
                          **kw).system

    @staticmethod
    def solveHLLC(system, cell_spacing, cfl_coefficient, final_time, bias, slope_limiter, subcycling_iterations,
                  reinitialisation_frequency, material1_parameters, material2_parameters, **kw):
        return RGFMSolver._solve(system, cell_spacing, cfl_coefficient, final_time, bias, slope_limiter,
                                 material1_parameters, material2_parameters, HLLC, **kw)

    solveHLLC2D = solveHLLC

    @staticmethod
    def solveExact(system, cell_spacing, cfl_coefficient, final_time, bias, slope_limiter, subcycling_iterations,
                   reinitialisation_frequency, material1_parameters, material2_parameters, **kw):
        return RGFMSolver._solve(system, cell_spacing, cfl_coefficient, final_time, bias, slope_limiter,
                                 material1_parameters, material2_parameters, EXACT_RIEMANN, **kw)

    solveExact2D = solveExact


# ---- two-material ideal MHD: real ghost fluid method --------------------------------

@dataclass
class MHDMultimaterialSystem(MultimaterialSystem):
    """MHDMultimaterialSystem (RGFM/mhdmultimaterialsystem.h): both materials'
    MHD primitive states (rho, u, v, w, p, Bx, By, Bz, psi) and the level set."""


def _mhd_params(p):
    if isinstance(p, MHDMaterialParameters):
        return p
    return MHDMaterialParameters(*p) if isinstance(p, (tuple, list)) else MHDMaterialParameters(float(p))


def _mhd_system_arrays(system):
    m1 = np.ascontiguousarray(system.material1, dtype=np.float64)
    m2 = np.ascontiguousarray(system.material2, dtype=np.float64)
    phi = np.ascontiguousarray(system.level_set, dtype=np.float64)
    if m1.shape != m2.shape or m1.shape[:-1] != phi.shape or m1.shape[-1] != 9 or phi.ndim not in (1, 2):
        raise ValueError("materials must be (..., 9) over the level set's 1D or 2D shape")
    return m1, m2, phi


def mhd_rgfm_solve(system: MultimaterialSystem, dx: float, cfl: float, t_final: float, params1, params2, *,
                   bias: float = 0.0, limiter: int = 0, subcycles: int = 0, max_steps: int = 0, mode: int = EXACT,
                   dt_cap: int = 200000, ctx: Context | None = None, raise_on_error: bool = True) -> RGFMResult:
    """MHDRGFMSolver::solve / solve2D through vg_mhd_rgfm_solve_1d / _2d (the
    1D solve ignores `subcycles`, as the reference's does)."""
    ctx = ctx or default_context()
    m1, m2, phi = _mhd_system_arrays(system)
    o1, o2, oph = np.empty_like(m1), np.empty_like(m2), np.empty_like(phi)
    dts = np.empty(dt_cap)
    st = _lib.RunStats()
    q1, q2 = _mhd_params(params1).c_struct(), _mhd_params(params2).c_struct()
    P = _lib.as_dptr
    if phi.ndim == 1:
        rc = ctx.lib.vg_mhd_rgfm_solve_1d(ctx.handle, P(m1), P(m2), P(phi), P(o1), P(o2), P(oph), phi.shape[0], dx,
                                          cfl, t_final, max_steps, bias, limiter, subcycles, C.byref(q1), C.byref(q2),
                                          mode, P(dts), dt_cap, C.byref(st))
    else:
        rc = ctx.lib.vg_mhd_rgfm_solve_2d(ctx.handle, P(m1), P(m2), P(phi), P(o1), P(o2), P(oph), phi.shape[0],
                                          phi.shape[1], dx, cfl, t_final, max_steps, bias, limiter, subcycles,
                                          C.byref(q1), C.byref(q2), mode, P(dts), dt_cap, C.byref(st))
    if raise_on_error and rc != _lib.VG_OK:
        check(rc, "vg_mhd_rgfm_solve")
    return RGFMResult(MHDMultimaterialSystem(o1, o2, oph), dts[:min(st.steps, dt_cap)].copy(), st, rc)


def mhd_rgfm_ghost_fill(system: MultimaterialSystem, params1, params2, *,
                        ctx: Context | None = None) -> MHDMultimaterialSystem:
    """applyRGFMBoundaryConditions[2D] through vg_mhd_rgfm_ghost_fill; returns
    a new system (the level set unchanged)."""
    ctx = ctx or default_context()
    m1, m2, phi = _mhd_system_arrays(system)
    m1, m2 = m1.copy(), m2.copy()
    q1, q2 = _mhd_params(params1).c_struct(), _mhd_params(params2).c_struct()
    ny, nx = (1, phi.shape[0]) if phi.ndim == 1 else phi.shape
    rc = ctx.lib.vg_mhd_rgfm_ghost_fill(ctx.handle, phi.ndim, ny, nx, _lib.as_dptr(m1), _lib.as_dptr(m2),
                                        _lib.as_dptr(phi), C.byref(q1), C.byref(q2))
    check(rc, "vg_mhd_rgfm_ghost_fill")
    return MHDMultimaterialSystem(m1, m2, phi.copy())


class MHDRGFMSolver:
    """MHDRGFMSolver (RGFM/mhdrgfmsolver.h:8-31): ghost fills and the whole
    solves (reinitialisation_frequency is unused by the reference's solves)."""

    @staticmethod
    def applyRGFMBoundaryConditions(system, material1_parameters, material2_parameters, **kw):
        if np.ndim(system.level_set) != 1:
            raise ValueError("applyRGFMBoundaryConditions takes a 1D system")
        return mhd_rgfm_ghost_fill(system, material1_parameters, material2_parameters, **kw)

    @staticmethod
    def applyRGFMBoundaryConditions2D(system, material1_parameters, material2_parameters, **kw):
        if np.ndim(system.level_set) != 2:
            raise ValueError("applyRGFMBoundaryConditions2D takes a 2D system")
        return mhd_rgfm_ghost_fill(system, material1_parameters, material2_parameters, **kw)

    @staticmethod
    def solve(system, cell_spacing, cfl_coefficient, final_time, bias, slope_limiter, subcycling_iterations,
              reinitialisation_frequency, material1_parameters, material2_parameters, **kw):
        if np.ndim(system.level_set) != 1:
            raise ValueError("solve takes a 1D system; use solve2D")
        return mhd_rgfm_solve(system, cell_spacing, cfl_coefficient, final_time, material1_parameters,
                              material2_parameters, bias=bias, limiter=slope_limiter,
                              subcycles=subcycling_iterations, **kw).system

    @staticmethod
    def solve2D(system, cell_spacing, cfl_coefficient, final_time, bias, slope_limiter, subcycling_iterations,
                reinitialisation_frequency, material1_parameters, material2_parameters, **kw):
        if np.ndim(system.level_set) != 2:
            raise ValueError("solve2D takes a 2D system")
        return mhd_rgfm_solve(system, cell_spacing, cfl_coefficient, final_time, material1_parameters,
                              material2_parameters, bias=bias, limiter=slope_limiter,
                              subcycles=subcycling_iterations, **kw).system
