"""Initial-condition builders for the benchmark and test problems.

Host-side NumPy code that produces primitive-state arrays in the reference's
layout: arrays of structures, fields last, row-major with x (column j)
contiguous, as ``std::vector<std::vector<EulerStateVector>>`` holds them
(Euler field order rho, u, v, w, p, lambda: Euler/eulerstatevector.h:55-61).

The random draws use splitmix64 so that every problem is reproducible from its
seed on any machine.
"""

from __future__ import annotations

import math

import numpy as np

_MASK = (1 << 64) - 1


def splitmix64(seed: int, count: int) -> np.ndarray:
    """``count`` successive splitmix64 outputs (uint64) from ``seed``."""
    out = np.empty(count, dtype=np.uint64)
    state = seed & _MASK
    for k in range(count):
        state = (state + 0x9E3779B97F4A7C15) & _MASK
        z = state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK
        out[k] = z ^ (z >> 31)
    return out


def uniform01(seed: int, count: int) -> np.ndarray:
    """Doubles in [0, 1) from the top 53 bits of splitmix64 draws."""
    u = splitmix64(seed, count)
    return (u >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def sod_1d(n: int) -> np.ndarray:
    """Toro test 1 / Sod shock tube on n cells, discontinuity at i <= n/2.

    Same cell classification as EulerTests::solveToroTest1
    (Euler/eulertests.cpp:14-24): left (1, 0, 0, 0, 1, 1), right
    (0.125, 0, 0, 0, 0.1, 1).
    """
    w = np.empty((n, 6))
    left = np.arange(n) <= 0.5 * n
    w[left] = (1.0, 0.0, 0.0, 0.0, 1.0, 1.0)
    w[~left] = (0.125, 0.0, 0.0, 0.0, 0.1, 1.0)
    return w


def disc_2d(n: int, gaussians: int = 64, seed: int = 20261202, amplitude: float = 1e-3,
            width: float = 0.02, length: float = 2.0, ny: int | None = None) -> np.ndarray:
    """Cylindrical explosion on an n x n grid over [0, length]^2, with Gaussian bumps.

    The disc follows EulerTests::solve2DToroTest1 (Euler/eulertests.cpp:36-49):
    cell (i, j) is inside when sqrt((i - n/2)^2 + (j - n/2)^2) <= 0.2 n, giving
    (rho, p) = (1, 1) inside and (0.125, 0.1) outside, at rest, lambda = 1.

    ``gaussians`` density bumps rho += amplitude * exp(-|x - c_k|^2 / (2 width^2))
    are added at centres c_k drawn from splitmix64(seed): draws 2k and 2k+1 give
    (x_k, y_k) = length * (u >> 11) * 2^-53.  Cell centres are
    ((j + 1/2) dx, (i + 1/2) dx).  Each bump is evaluated as the separable product
    exp(-(x-cx)^2/(2w^2)) * exp(-(y-cy)^2/(2w^2)) inside a window of 12 widths
    (outside it the bump is below 1e-31 and cannot change a double of the sum).
    """
    ny = n if ny is None else ny
    nx = n
    dx = length / nx
    w = np.empty((ny, nx, 6))
    i = np.arange(ny, dtype=np.float64)[:, None]
    j = np.arange(nx, dtype=np.float64)[None, :]
    r = np.sqrt((i - 0.5 * ny) * (i - 0.5 * ny) + (j - 0.5 * nx) * (j - 0.5 * nx))
    inside = r <= 0.2 * n
    w[..., 1:4] = 0.0
    w[..., 5] = 1.0
    w[..., 0] = np.where(inside, 1.0, 0.125)
    w[..., 4] = np.where(inside, 1.0, 0.1)
    if gaussians:
        c = uniform01(seed, 2 * gaussians) * length
        bump = np.zeros((ny, nx))
        xc = (np.arange(nx) + 0.5) * dx
        yc = (np.arange(ny) + 0.5) * dx
        half = 12.0 * width
        for k in range(gaussians):
            cx, cy = c[2 * k], c[2 * k + 1]
            jx = np.nonzero(np.abs(xc - cx) <= half)[0]
            iy = np.nonzero(np.abs(yc - cy) <= half)[0]
            if jx.size == 0 or iy.size == 0:
                continue
            gx = np.exp(-((xc[jx] - cx) ** 2) / (2.0 * width * width))
            gy = np.exp(-((yc[iy] - cy) ** 2) / (2.0 * width * width))
            bump[iy[0]:iy[-1] + 1, jx[0]:jx[-1] + 1] += np.outer(gy, gx)
        w[..., 0] += amplitude * bump
    return w


def brio_wu(n: int) -> np.ndarray:
    """Brio-Wu shock tube as MHDTests::solveDumbserTest1 sets it up
    (MHD/mhdtests.cpp:12-24), gamma 2: left (1, 0, 0, 0, 1, 0.75, 1, 0, 0),
    right (0.125, 0, 0, 0, 0.1, 0.75, -1, 0, 0), split at i <= n/2.
    Primitive order (rho, u, v, w, p, Bx, By, Bz, psi)."""
    w = np.empty((n, 9))
    left = np.arange(n) <= 0.5 * n
    w[left] = (1.0, 0.0, 0.0, 0.0, 1.0, 0.75, 1.0, 0.0, 0.0)
    w[~left] = (0.125, 0.0, 0.0, 0.0, 0.1, 0.75, -1.0, 0.0, 0.0)
    return w


def mhd_batch(n_inst: int, seed: int = 3, n_min: int = 256, n_max: int = 16384):
    """Independent 1D ideal-MHD Riemann problems (the batched-MHD benchmark recipe).

    Per instance, 17 draws u_k in [0, 1) from splitmix64(seed) in this order:
    N = n_min + floor(u (n_max - n_min + 1)); gamma = 1.3 + 0.7 u; Bx = 2u - 1
    (shared by both sides); then for the left and then the right state:
    rho = 10^(2u - 1), p = 10^(4u - 2), u, v, w = 4u - 2 each, By, Bz = 4u - 2 each.
    psi = 0.  Cells i <= N/2 take the left state.  dx = 1/N.

    Returns (prim, offsets, gamma, dx): prim is the concatenation of all
    instances (total cells x 9), offsets the N+1 cell offsets.
    """
    u = uniform01(seed, 17 * n_inst).reshape(n_inst, 17)
    sizes = (n_min + np.floor(u[:, 0] * (n_max - n_min + 1))).astype(np.int64)
    sizes = np.minimum(sizes, n_max)
    gamma = 1.3 + 0.7 * u[:, 1]
    bx = 2.0 * u[:, 2] - 1.0
    offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    prim = np.empty((int(offsets[-1]), 9))
    for q in range(n_inst):
        side = []
        for s in range(2):
            d = u[q, 3 + 7 * s: 10 + 7 * s]
            side.append((10.0 ** (2 * d[0] - 1), 4 * d[2] - 2, 4 * d[3] - 2, 4 * d[4] - 2, 10.0 ** (4 * d[1] - 2),
                         bx[q], 4 * d[5] - 2, 4 * d[6] - 2, 0.0))
        n = int(sizes[q])
        blk = prim[offsets[q]:offsets[q + 1]]
        left = np.arange(n) <= 0.5 * n
        blk[left] = side[0]
        blk[~left] = side[1]
    return prim, offsets, gamma, 1.0 / sizes.astype(np.float64)


def hpr_state(rho: float, p: float, u: float = 0.0, v: float = 0.0, w: float = 0.0, A=None,
              J=(0.0, 0.0, 0.0)) -> np.ndarray:
    """One HPR primitive state: rho, p, A (3x3 row-major), u, v, w, J
    (HPRStateVector field order, HPR/hprstatevector.h)."""
    s = np.zeros(17)
    s[0], s[1] = rho, p
    s[2:11] = (np.eye(3) if A is None else np.asarray(A, dtype=np.float64)).ravel()
    s[11:14] = (u, v, w)
    s[14:17] = J
    return s


def hpr_stokes_1d(n: int, v0: float = 0.1) -> np.ndarray:
    """HPRTests::solveStokesFirstProblem (HPR/hprtests.cpp:7-31): rho = 1,
    p = 1/1.4, A = I, v = -v0 for i <= n/2 and +v0 beyond."""
    return np.stack([hpr_state(1.0, 1.0 / 1.4, 0.0, -v0 if i <= 0.5 * n else v0) for i in range(n)])


def hpr_fourier_2d(n: int, seed: int = 4, modes: int = 16, amplitude: float = 0.1, gamma: float = 1.4,
                   ny: int | None = None) -> np.ndarray:
    """Viscous-flow IC on [0,1]^2, periodic: rho = 1, p = 1/gamma, A = I, J = 0,
    and each velocity component a sum of ``modes`` Fourier modes
    amplitude * a_m * sin(2 pi (k_m . x) + phi_m) with k_m in {1..4}^2 and
    a_m in [-1, 1), phi_m in [0, 2 pi) from splitmix64(seed) (u and v draw
    their own modes: 4 draws per mode (k_x, k_y, a, phi), u first)."""
    ny = n if ny is None else ny
    r = uniform01(seed, 2 * modes * 4).reshape(2, modes, 4)
    x = (np.arange(n) + 0.5) / n
    y = (np.arange(ny) + 0.5) / ny
    X, Y = np.meshgrid(x, y)
    vel = []
    for c in range(2):
        f = np.zeros((ny, n))
        for m in range(modes):
            kx = 1 + int(r[c, m, 0] * 4)
            ky = 1 + int(r[c, m, 1] * 4)
            a = 2.0 * r[c, m, 2] - 1.0
            phi = 2.0 * np.pi * r[c, m, 3]
            f += amplitude * a * np.sin(2.0 * np.pi * (kx * X + ky * Y) + phi)
        vel.append(f)
    w = np.zeros((ny, n, 17))
    w[..., 0] = 1.0
    w[..., 1] = 1.0 / gamma
    w[..., 2] = w[..., 6] = w[..., 10] = 1.0
    w[..., 11] = vel[0]
    w[..., 12] = vel[1]
    return w


# ---- two-material (RGFM) problems ---------------------------------------------------

AIR = (1.4, 0.0, 25.0, 0.0, 0.25, 250.0)        # EulerMaterialParameters(1.4)
HELIUM = (1.648, 0.0, 25.0, 0.0, 0.25, 250.0)   # EulerMaterialParameters(1.648)


def post_shock_state(rho0: float, p0: float, gamma: float, mach: float):
    """Rankine-Hugoniot state behind a normal shock of Mach ``mach`` running
    into gas at rest (rho0, p0): returns (rho1, |u1|, p1) with
    p1/p0 = 1 + 2 gamma (M^2 - 1)/(gamma + 1),
    rho1/rho0 = (gamma + 1) M^2 / ((gamma - 1) M^2 + 2),
    |u1| = 2 c0 (M - 1/M) / (gamma + 1), the gas following the shock."""
    m2 = mach * mach
    c0 = np.sqrt(gamma * p0 / rho0)
    p1 = p0 * (1.0 + 2.0 * gamma * (m2 - 1.0) / (gamma + 1.0))
    rho1 = rho0 * (gamma + 1.0) * m2 / ((gamma - 1.0) * m2 + 2.0)
    u1 = 2.0 * c0 * (mach - 1.0 / mach) / (gamma + 1.0)
    return rho1, u1, p1


def shock_bubble_2d(nx: int, ny: int, *, length=(0.325, 0.089), center=(0.175, 0.0445), radius: float = 0.025,
                    mach: float = 1.22, shock_x: float = 0.225, air=(1.0, 1e5), helium=(0.138, 1e5)):
    """Helium bubble in air hit by a planar shock (the Haas-Sturtevant /
    Quirk-Karni geometry), cell-centred on an nx x ny grid over
    [0, Lx] x [0, Ly].

    Material 1 is helium (real inside the bubble, phi < 0), material 2 air (real
    outside; the rows start in air, which is how the 2D ghost fill's toggle
    starts, rgfmsolver.cpp:374-379).  Both materials are defined everywhere:
    helium at rest, air at rest for x < shock_x and in the Rankine-Hugoniot
    post-shock state (moving -x) beyond.  phi is the signed distance to the bubble.
    Returns (m1, m2, phi, dx, dy)."""
    lx, ly = length
    dx, dy = lx / nx, ly / ny
    x = (np.arange(nx) + 0.5) * dx
    y = (np.arange(ny) + 0.5) * dy
    X, Y = np.meshgrid(x, y)
    m1 = np.zeros((ny, nx, 6))
    m1[..., 0] = helium[0]
    m1[..., 4] = helium[1]
    m1[..., 5] = 1.0
    m2 = np.zeros((ny, nx, 6))
    m2[..., 0] = air[0]
    m2[..., 4] = air[1]
    m2[..., 5] = 1.0
    rho1, u1, p1 = post_shock_state(air[0], air[1], AIR[0], mach)
    behind = X > shock_x
    m2[behind, 0] = rho1
    m2[behind, 1] = -u1
    m2[behind, 4] = p1
    phi = np.sqrt((X - center[0]) ** 2 + (Y - center[1]) ** 2) - radius
    return m1, m2, phi, dx, dy


def shock_bubble_batch(n_inst: int = 16, seed: int = 6):
    """Per-instance (radius, mach): radius uniform in [0.015, 0.03], Mach uniform
    in [1.1, 1.5], two splitmix64 draws per instance (radius first)."""
    u = uniform01(seed, 2 * n_inst).reshape(n_inst, 2)
    return 0.015 + 0.015 * u[:, 0], 1.1 + 0.4 * u[:, 1]


# ---- two-material diffuse-interface Euler ICs ---------------------------------
# kind 0: EulerMultiphysicsStateVector primitives (alpha1, u, v, w, rho1, p1, rho2,
# p2); kind 1: EulerReducedStateVector primitives (alpha1, u, v, w, rho1, rho2, p).

def _two_material(kind: int, alpha: float, u: float, rho1: float, rho2: float, p1: float, p2: float,
                  p_reduced: float) -> list:
    if kind == 0:
        return [alpha, u, 0.0, 0.0, rho1, p1, rho2, p2]
    return [alpha, u, 0.0, 0.0, rho1, rho2, p_reduced]


def two_material_toro_1d(n: int, kind: int = 0) -> np.ndarray:
    """Toro's test 1 as a two-material problem, EulerMultiphysicsTests::solveToroTest1
    (Euler/Multiphysics/eulermultiphysicstests.cpp:7-27) and
    EulerReducedTests::solveToroTest1 (eulerreducedtests.cpp:7-27): material 1
    (alpha1 = 0.999) where i <= n/2, material 2 (0.001) elsewhere; gamma 1.4 both,
    t = 0.25, dx = 1/n."""
    rows = []
    for i in range(n):
        left = i <= 0.5 * n
        a = 0.999 if left else 0.001
        if kind == 0:
            rows.append(_two_material(0, a, 0.0, 1.0, 0.125, 1.0, 0.1, 0.0))
        else:
            rows.append(_two_material(1, a, 0.0, 1.0, 0.125, 0.0, 0.0, 1.0 if left else 0.1))
    return np.array(rows)


def two_material_fedkiw_1d(n: int, kind: int = 0) -> np.ndarray:
    """Fedkiw's shocked air/helium tube, EulerMultiphysicsTests::solveFedkiwTest
    (eulermultiphysicstests.cpp:29-59) and the reduced twin
    (eulerreducedtests.cpp:29-57): gamma 1.4 / 1.67, t = 0.0012, dx = 1/n."""
    us = 0.3535 * np.sqrt(10.0 ** 5)
    rows = []
    for i in range(n):
        if i <= 0.5 * n:
            if i <= 0.05 * n:
                rows.append(_two_material(kind, 0.999, us, 1.3333, 0.1379, 1.5e5, 1.0e5, 1.5e5))
            else:
                rows.append(_two_material(kind, 0.999, 0.0, 1.0, 0.1379, 1.0e5, 1.0e5, 1.0e5))
        else:
            rows.append(_two_material(kind, 0.001, 0.0, 1.0, 0.1379, 1.0e5, 1.0e5, 1.0e5))
    return np.array(rows)


def _radius(n: int) -> np.ndarray:
    i = np.arange(n, dtype=np.float64)[:, None] - (0.5 * n)
    j = np.arange(n, dtype=np.float64)[None, :] - (0.5 * n)
    return np.sqrt((i * i) + (j * j))


def two_material_toro_2d(n: int, kind: int = 0) -> np.ndarray:
    """The cylindrical Toro test, solve2DToroTest1 (eulermultiphysicstests.cpp:61-81,
    eulerreducedtests.cpp:59-82): material 1 inside radius 0.2 n (in cells) of the
    centre, t = 0.125."""
    inside = _radius(n) <= (0.2 * n)
    a = np.where(inside, 0.999, 0.001)
    w = np.zeros((n, n, 8 if kind == 0 else 7))
    w[..., 0] = a
    if kind == 0:
        w[..., 4], w[..., 5], w[..., 6], w[..., 7] = 1.0, 1.0, 0.125, 0.1
    else:
        w[..., 4], w[..., 5], w[..., 6] = 1.0, 0.125, np.where(inside, 1.0, 0.1)
    return w


def two_material_fedkiw_2d(n: int, kind: int = 0) -> np.ndarray:
    """The cylindrical Fedkiw test, solve2DFedkiwTest (eulermultiphysicstests.cpp:83-113,
    eulerreducedtests.cpp:84-115): a shocked core within 0.02 n of the centre, air
    to 0.2 n, helium outside; gamma 1.4 / 1.67, t = 0.0006."""
    r = _radius(n)
    w = np.zeros((n, n, 8 if kind == 0 else 7))
    for i in range(n):
        for j in range(n):
            if r[i, j] <= (0.2 * n):
                if r[i, j] <= (0.02 * n):
                    w[i, j] = _two_material(kind, 0.999, 0.3535 * np.sqrt(10.0 ** 5), 1.3333, 0.1379, 1.5e5, 1.0e5,
                                            1.5e5)
                else:
                    w[i, j] = _two_material(kind, 0.999, 0.0, 1.0, 0.1379, 1.0e5, 1.0e5, 1.0e5)
            else:
                w[i, j] = _two_material(kind, 0.001, 0.0, 1.0, 0.1379, 1.0e5, 1.0e5, 1.0e5)
    return w


# ---- two-material MHD (RGFM) ICs -------------------------------------------------
# States (rho, u, v, w, p, Bx, By, Bz, psi).

def mhd_rgfm_dumbser_1d(n: int):
    """The Brio-Wu-like two-material tube of MHDRGFMTests::solveDumbserTest1 and
    solveDumbserMultimaterialTest1 (RGFM/mhdrgfmtests.cpp:7-43): material 1
    (rho 1, p 1, B = (0.75, 1, 0)) and material 2 (rho 0.125, p 0.1,
    B = (0.75, -1, 0)) everywhere, phi_i = i dx - 0.5, dx = 1/n."""
    m1 = np.tile([1.0, 0.0, 0.0, 0.0, 1.0, 0.75, 1.0, 0.0, 0.0], (n, 1))
    m2 = np.tile([0.125, 0.0, 0.0, 0.0, 0.1, 0.75, -1.0, 0.0, 0.0], (n, 1))
    phi = ((1.0 / n) * np.arange(n, dtype=np.float64)) - 0.5
    return m1, m2, phi


def mhd_rgfm_dumbser_2d(n: int):
    """The cylindrical version, solve2DDumbserTest1 / solve2DDumbserMultimaterialTest1
    (mhdrgfmtests.cpp:45-97): the same materials, phi = r / n - 0.2 with r the
    distance in cells from (n/2, n/2)."""
    m1 = np.tile([1.0, 0.0, 0.0, 0.0, 1.0, 0.75, 1.0, 0.0, 0.0], (n, n, 1))
    m2 = np.tile([0.125, 0.0, 0.0, 0.0, 0.1, 0.75, -1.0, 0.0, 0.0], (n, n, 1))
    i = np.arange(n, dtype=np.float64)[:, None] - (0.5 * n)
    j = np.arange(n, dtype=np.float64)[None, :] - (0.5 * n)
    phi = (np.sqrt((i * i) + (j * j)) / n) - 0.2
    return m1, m2, phi


# ---- the reference's test drivers' initial conditions ---------------------------------

def _disc_radius(n: int) -> np.ndarray:
    """sqrt((i - n/2)^2 + (j - n/2)^2) over an n x n grid, as the drivers form it
    (i, j converted to double, squares summed left to right)."""
    i = np.arange(n, dtype=np.float64)[:, None] - 0.5 * n
    j = np.arange(n, dtype=np.float64)[None, :] - 0.5 * n
    return np.sqrt((i * i) + (j * j))


def euler_toro1_2d(n: int) -> np.ndarray:
    """EulerTests::solve2DToroTest1 (Euler/eulertests.cpp:29-52): (1, 0, 0, 0, 1)
    where the radius is <= 0.2 n, (0.125, 0, 0, 0, 0.1) elsewhere."""
    inside = _disc_radius(n) <= 0.2 * n
    w = np.empty((n, n, 6))
    w[inside] = (1.0, 0.0, 0.0, 0.0, 1.0, 1.0)
    w[~inside] = (0.125, 0.0, 0.0, 0.0, 0.1, 1.0)
    return w


def _euler_fill(shape, state) -> np.ndarray:
    w = np.empty(tuple(shape) + (6,))
    w[...] = tuple(state) + (1.0,)
    return w


# Fedkiw's shocked-air state of EulerRGFMTests (rgfm: 1.3333, 0.3535 sqrt(10^5),
# 0, 0, 1.5 10^5); the ambient air and the helium (eulerrgfmtests.cpp:70-83).
FEDKIW_SHOCKED = (1.3333, 0.3535 * math.sqrt(10.0 ** 5.0), 0.0, 0.0, 1.5 * 10.0 ** 5.0)
FEDKIW_AIR = (1.0, 0.0, 0.0, 0.0, 1.0 * 10.0 ** 5.0)
FEDKIW_HELIUM = (0.1379, 0.0, 0.0, 0.0, 1.0 * 10.0 ** 5.0)


def rgfm_toro1_1d(n: int):
    """EulerRGFMTests::solveToroTest1HLLC (RGFM/eulerrgfmtests.cpp:28-48): material
    1 (1, 0, 0, 0, 1) and material 2 (0.125, 0, 0, 0, 0.1) everywhere, level set
    (1/n) i - 0.5.  Returns (m1, m2, phi)."""
    h = 1.0 / n
    phi = (h * np.arange(n, dtype=np.float64)) - 0.5
    return _euler_fill((n,), (1.0, 0.0, 0.0, 0.0, 1.0)), _euler_fill((n,), (0.125, 0.0, 0.0, 0.0, 0.1)), phi


def rgfm_fedkiw_1d(n: int):
    """EulerRGFMTests::solveFedkiwTestHLLC (RGFM/eulerrgfmtests.cpp:84-111): material
    1 shocked for i <= 0.05 n and ambient beyond, material 2 helium, level set
    (1/n) i - 0.5."""
    h = 1.0 / n
    m1 = _euler_fill((n,), FEDKIW_AIR)
    m1[np.arange(n) <= n * 0.05, :5] = FEDKIW_SHOCKED
    phi = (h * np.arange(n, dtype=np.float64)) - 0.5
    return m1, _euler_fill((n,), FEDKIW_HELIUM), phi


def rgfm_toro1_2d(n: int):
    """EulerRGFMTests::solve2DToroTest1HLLC (RGFM/eulerrgfmtests.cpp:136-155): both
    materials uniform, level set radius / n - 0.2."""
    phi = (_disc_radius(n) / n) - 0.2
    return (_euler_fill((n, n), (1.0, 0.0, 0.0, 0.0, 1.0)), _euler_fill((n, n), (0.125, 0.0, 0.0, 0.0, 0.1)), phi)


def rgfm_fedkiw_2d(n: int):
    """EulerRGFMTests::solve2DFedkiwTestHLLC (RGFM/eulerrgfmtests.cpp:183-213):
    material 1 shocked where the radius is <= 0.02 n, ambient elsewhere; material
    2 helium; level set radius / n - 0.2."""
    r = _disc_radius(n)
    m1 = _euler_fill((n, n), FEDKIW_AIR)
    m1[r <= 0.02 * n, :5] = FEDKIW_SHOCKED
    return m1, _euler_fill((n, n), FEDKIW_HELIUM), (r / n) - 0.2


# MHDTests::solveDumbserTest1-3 (MHD/mhdtests.cpp:7-71): split fraction, left and
# right states, gamma, final time.
MHD_DUMBSER = {
    1: (0.5, (1.0, 0.0, 0.0, 0.0, 1.0, 0.75, 1.0, 0.0, 0.0), (0.125, 0.0, 0.0, 0.0, 0.1, 0.75, -1.0, 0.0, 0.0),
        2.0, 0.1),
    2: (0.4, (1.08, 1.2, 0.01, 0.5, 0.95, 0.564189, 1.015541, 0.564189, 0.0),
        (0.9891, -0.0131, 0.0269, 0.010037, 0.97159, 0.564189, 1.135262, 0.564923, 0.0), 5.0 / 3.0, 0.2),
    3: (0.5, (1.0, 0.0, 0.0, 0.0, 1.0, 1.3, 1.0, 0.0, 0.0), (0.4, 0.0, 0.0, 0.0, 0.4, 1.3, -1.0, 0.0, 0.0),
        5.0 / 3.0, 0.16),
}

# MHDTests::solve2DDumbserTest1-2 (MHD/mhdtests.cpp:73-120): disc radius fraction,
# inside and outside states, gamma, CFL, final time.
MHD_DUMBSER_2D = {
    1: (0.2, MHD_DUMBSER[1][1], MHD_DUMBSER[1][2], 2.0, 0.6, 0.05),
    2: (0.16, MHD_DUMBSER[2][1], MHD_DUMBSER[2][2], 5.0 / 3.0, 0.8, 0.1),
}


def mhd_dumbser_1d(n: int, test: int = 1) -> np.ndarray:
    """MHDTests::solveDumbserTest<test>: left state for i <= f n, right beyond."""
    f, left, right, _, _ = MHD_DUMBSER[test]
    w = np.empty((n, 9))
    mask = np.arange(n) <= f * n
    w[mask] = left
    w[~mask] = right
    return w


def mhd_dumbser_2d(n: int, test: int = 1) -> np.ndarray:
    """MHDTests::solve2DDumbserTest<test>: inside state where the radius is <= f n."""
    f, inside_state, outside_state, _, _, _ = MHD_DUMBSER_2D[test]
    w = np.empty((n, n, 9))
    inside = _disc_radius(n) <= f * n
    w[inside] = inside_state
    w[~inside] = outside_state
    return w


def hpr_heat_conduction_1d(n: int) -> np.ndarray:
    """HPRTests::solveHeatConductionProblem (HPR/hprtests.cpp:33-57): rho 2, p 1,
    A = 2^(1/3) I for i <= n/2; rho 0.5, p 1, A = 0.5^(1/3) I beyond; at rest."""
    a_hot, a_cold = 2.0 ** (1.0 / 3.0), 0.5 ** (1.0 / 3.0)
    return np.stack([hpr_state(2.0, 1.0, A=a_hot * np.eye(3)) if i <= 0.5 * n else
                     hpr_state(0.5, 1.0, A=a_cold * np.eye(3)) for i in range(n)])
