"""Adjoint of the second-order HPR solve on the GPU (vg_hpr_adjoint_*), through
the C ABI, time steps frozen.

1. The fixed-time-step forward is the solve (bit for bit on the GPU), and the
   oracle's map (bit for bit without relaxation sub-cycles; with them the stiff
   source's pow(det A, 5/3) may differ in the last bit, so 1e-10 relative L1).
2. Dot-product test against the oracle's forward-mode derivative (dual
   numbers, oracle/hpr_dual.cpp): <grad, v> + grad_tau . dtau = <seed, J (v, dtau)>
   for random directions in the initial state and the two relaxation times.
3. The viscous-flow problem's loss sum (rho - rho_target)^2 dx^2: its
   derivative with respect to the viscosity mu (through tau1(mu), tau2(mu))
   against the dual numbers and against central differences of the GPU's own
   fixed-step forward, at a step the oracle's branch hash certifies.
4. The checkpoint interval changes nothing (bit for bit).
"""

import math

import numpy as np
import pytest

import oracle_lib as O
from test_hpr_gpu import mat_params, rel_l1
from test_oracle_hpr_adjoint import CASES, hpr_adj_case, hpr_adj_params
from vanguard_amd import _lib, problems
from vanguard_amd import solvers as S
from vanguard_amd._lib import VanguardError

pytestmark = pytest.mark.gpu

SCALE = np.array([0.1, 0.1] + [0.01] * 9 + [0.1] * 3 + [0.01] * 3)   # field sizes of the primitive state


def gpu_adjoint(w, seed, dim, sub, bc, p, steps, k=0, limiter=0):
    n = w.shape[0] if dim == 1 else w.shape[1]
    return S.hpr_adjoint(w, seed, dim, 1.0 / n, 0.8, math.inf, limiter=limiter, subcycles=sub, params=mat_params(p),
                         max_steps=steps, bc=bc, checkpoint_every=k)


@pytest.mark.parametrize("dim,thermal,sub,bc", CASES)
def test_fixed_dt_forward_is_the_solve(dim, thermal, sub, bc):
    w = hpr_adj_case(dim, thermal)
    p = hpr_adj_params(thermal)
    n = w.shape[0] if dim == 1 else w.shape[1]
    print(f"case dim={dim} thermal={thermal} sub={sub} bc={bc}", flush=True)
    ref = S.hpr_solve(w, dim, 1.0 / n, 0.8, math.inf, subcycles=sub, params=mat_params(p), max_steps=5, bc=bc)
    fixed = S.hpr_solve_fixed_dt(w, dim, 1.0 / n, ref.dt, subcycles=sub, params=mat_params(p), bc=bc)
    O.assert_bitwise(fixed, ref.cells, "fixed-dt forward vs the solve")
    orc = O.hpr_fixed_dt(w, ref.dt, dim=dim, dx=1.0 / n, bias=0.0, limiter=0, subcycles=sub, params=p, bc=bc)
    if sub == 0:
        O.assert_bitwise(fixed, orc, "fixed-dt forward vs the oracle")
    else:
        assert rel_l1(fixed, orc) <= 1e-10
    adj = gpu_adjoint(w, np.zeros_like(w), dim, sub, bc, p, 5)
    O.assert_bitwise(adj.dt, ref.dt, "adjoint time steps")
    O.assert_bitwise(adj.cells, ref.cells, "adjoint forward")
    assert not np.any(adj.grad) and not np.any(adj.grad_tau)


@pytest.mark.parametrize("dim,thermal,sub,bc", CASES)
def test_dot_product_against_dual_numbers(dim, thermal, sub, bc):
    w = hpr_adj_case(dim, thermal)
    p = hpr_adj_params(thermal)
    n = w.shape[0] if dim == 1 else w.shape[1]
    print(f"case dim={dim} thermal={thermal} sub={sub} bc={bc}", flush=True)
    rng = np.random.default_rng(80 + 10 * dim + 2 * sub + bc)
    seed = rng.standard_normal(w.shape) / SCALE
    adj = gpu_adjoint(w, seed, dim, sub, bc, p, 4)
    assert adj.dt.size == 4
    for _ in range(3):
        v = rng.standard_normal(w.shape) * SCALE
        if not thermal:
            v[..., 14:17] = 0.0
        dtau = (p["tau1"] * rng.standard_normal(), p["tau2"] * rng.standard_normal() if thermal else 0.0)
        _, t = O.hpr_fixed_dt(w, adj.dt, dim=dim, dx=1.0 / n, bias=0.0, limiter=0, subcycles=sub, params=p, bc=bc,
                              tangent=v, dtau=dtau)
        lhs = float(np.sum(adj.grad * v) + adj.grad_tau[0] * dtau[0] + adj.grad_tau[1] * dtau[1])
        rhs = float(np.sum(seed * t))
        scale = float(np.sum(np.abs(adj.grad * v)) + abs(adj.grad_tau[0] * dtau[0]) + abs(adj.grad_tau[1] * dtau[1]))
        assert abs(lhs - rhs) <= 1e-9 * scale, (lhs, rhs, scale)


@pytest.mark.parametrize("limiter", [1, 2])
def test_dot_product_other_limiters(limiter):
    w = hpr_adj_case(2, True)
    p = hpr_adj_params(True)
    rng = np.random.default_rng(90 + limiter)
    seed = rng.standard_normal(w.shape) / SCALE
    adj = gpu_adjoint(w, seed, 2, 1, 1, p, 3, limiter=limiter)
    v = rng.standard_normal(w.shape) * SCALE
    _, t = O.hpr_fixed_dt(w, adj.dt, dim=2, dx=1.0 / 12, bias=0.0, limiter=limiter, subcycles=1, params=p, bc=1,
                          tangent=v)
    lhs, rhs = float(np.sum(adj.grad * v)), float(np.sum(seed * t))
    assert abs(lhs - rhs) <= 1e-9 * float(np.sum(np.abs(adj.grad * v))), (lhs, rhs)


def viscous_loss(rho, target, dx):
    return float(np.sum((rho - target) ** 2) * dx * dx)


def test_viscosity_derivative():
    """dL/dmu of the viscous-flow problem (16^2, periodic, heat conduction, one
    sub-cycle, 6 steps): adjoint vs dual numbers (1e-9) and vs central
    differences of the GPU's fixed-step forward (1e-6)."""
    n, mu = 16, 1e-3
    w = problems.hpr_fourier_2d(n, seed=4)
    target = problems.hpr_fourier_2d(n, seed=5)[..., 0] + 0.01 * np.sin(
        2 * np.pi * (np.arange(n) + 0.5) / n)[None, :]
    p = hpr_adj_params(True, mu)
    dx = 1.0 / n
    ref = S.hpr_solve(w, 2, dx, 0.8, math.inf, subcycles=1, params=mat_params(p), max_steps=6, bc=1)
    seed = np.zeros_like(w)
    seed[..., 0] = 2.0 * (ref.cells[..., 0] - target) * dx * dx
    adj = gpu_adjoint(w, seed, 2, 1, 1, p, 6)
    dtau = (p["tau1"] / mu, p["tau2"] / mu)   # both relaxation times are proportional to mu
    dl_dmu = float(adj.grad_tau[0] * dtau[0] + adj.grad_tau[1] * dtau[1])
    _, t = O.hpr_fixed_dt(w, adj.dt, dim=2, dx=dx, bias=0.0, limiter=0, subcycles=1, params=p, bc=1,
                          tangent=np.zeros_like(w), dtau=dtau)
    dual = float(np.sum(seed * t))
    assert abs(dl_dmu - dual) <= 1e-9 * abs(dual), (dl_dmu, dual)
    h0 = O.branch_hash()
    for h in (1e-7, 1e-8):
        pa, pb = hpr_adj_params(True, mu * (1 + h)), hpr_adj_params(True, mu * (1 - h))
        O.hpr_fixed_dt(w, adj.dt, dim=2, dx=dx, bias=0.0, limiter=0, subcycles=1, params=pa, bc=1)
        ha = O.branch_hash()
        O.hpr_fixed_dt(w, adj.dt, dim=2, dx=dx, bias=0.0, limiter=0, subcycles=1, params=pb, bc=1)
        if ha == h0 == O.branch_hash():
            break
    else:
        pytest.fail("no step keeps every branch")
    fa = S.hpr_solve_fixed_dt(w, 2, dx, adj.dt, subcycles=1, params=mat_params(pa), bc=1)
    fb = S.hpr_solve_fixed_dt(w, 2, dx, adj.dt, subcycles=1, params=mat_params(pb), bc=1)
    fd = (viscous_loss(fa[..., 0], target, dx) - viscous_loss(fb[..., 0], target, dx)) / (2 * h * mu)
    assert abs(fd - dl_dmu) <= 1e-6 * abs(dl_dmu), (fd, dl_dmu)
    assert np.isfinite(adj.grad).all() and np.any(adj.grad[..., 11:13])


def test_checkpoint_interval_changes_nothing():
    w = hpr_adj_case(2, True)
    p = hpr_adj_params(True)
    seed = np.random.default_rng(95).standard_normal(w.shape) / SCALE
    base = gpu_adjoint(w, seed, 2, 1, 1, p, 5)
    for k in (1, 2, 5, 9):
        other = gpu_adjoint(w, seed, 2, 1, 1, p, 5, k=k)
        O.assert_bitwise(other.grad, base.grad, f"checkpoint every {k}")
        O.assert_bitwise(other.grad_tau, base.grad_tau, f"relaxation times, checkpoint every {k}")


def test_adjoint_rejects_bad_arguments():
    w = hpr_adj_case(2, True)
    p = hpr_adj_params(True)
    with pytest.raises(VanguardError):   # fewer than 3 cells along y
        S.hpr_adjoint(w[:2], np.zeros_like(w[:2]), 2, 1 / 12, 0.8, math.inf, subcycles=1, params=mat_params(p),
                      max_steps=2)
    plastic = mat_params(p)
    plastic.is_plastic = True
    plastic.elastic_plastic_transition_parameter = 1.0
    romenski = mat_params(p)
    romenski.equation_of_state = "GodunovRomenski"
    for q in (plastic, romenski):   # the forward solver runs these; the adjoint differentiates neither
        with pytest.raises(VanguardError) as e:
            S.hpr_adjoint(w, np.zeros_like(w), 2, 1 / 12, 0.8, math.inf, subcycles=1, params=q, max_steps=2)
        assert e.value.status == _lib.VG_ERR_UNSUPPORTED
    with pytest.raises(ValueError):
        S.hpr_adjoint(w, np.zeros((12, 12, 6)), 2, 1 / 12, 0.8, math.inf, params=mat_params(p), max_steps=2)
