"""Pins the RGFM oracle (oracle/rgfm_oracle.c).

HLLC path: bit for bit against golden vectors of the compiled reference --
per-function HLLC records (different gammas: always PVRS; equal gammas: the
two-shock / two-rarefaction branches and the Y-velocity quirk of
hllcsolver.cpp:552; stiffened gas) and whole solveHLLC / solveHLLC2D runs (1D
Toro and Fedkiw tests checked against the reference's own solve, 2D discs, a
shock-bubble geometry).

Exact solver (the reference's does not compile): Toro's published exact star
states for his five tests (Toro, Riemann Solvers and Numerical Methods for
Fluid Dynamics, 3rd ed., Table 4.3), the Rankine-Hugoniot conditions across
computed shocks, the isentrope and Riemann invariant across rarefactions, and
the sampled state at x/t = s in every region of the wave fan (the conditions
that define each region, not the oracle's closed forms).
Anisotropic spacing (dx != dy): reductions to the 1D solve and to the square
grid for data constant along one axis, and a linear level set carried by a
uniform contact.
Level set and circulation: exactness of the upwind stencil on linear fields,
and the closed-form circulation of a solid-body rotation.
"""

import math

import numpy as np
import pytest

import oracle_lib as O

AIR = (1.4, 0.0, 25.0, 0.0, 0.25, 250.0)


@pytest.mark.parametrize("k", [0, 1, 2])
def test_hllc_units_match_reference(k):
    g = O.golden("hllc_units")
    out = O.hllc_units(g[f"states{k}"], tuple(g[f"p1_{k}"]), tuple(g[f"p2_{k}"]))
    O.assert_bitwise(out, g[f"out{k}"], f"hllc_units set {k}")


RUNS = ["rgfm_toro1d_hllc", "rgfm_fedkiw1d_hllc", "rgfm_2d_toro32", "rgfm_2d_fedkiw32", "rgfm_2d_fedkiw24_api",
        "rgfm_2d_bubble96"]


def run_golden(g):
    m = g["meta"]
    return O.rgfm_solve(g["m1_in"], g["m2_in"], g["phi_in"], dim=m["dim"], dx=m["dx"], cfl=m["cfl"],
                        tfinal=m["tfinal"], max_steps=m["max_steps"], bias=m["bias"], limiter=m["limiter"],
                        p1=tuple(m["p1"]), p2=tuple(m["p2"]))


@pytest.mark.parametrize("name", RUNS)
def test_rgfm_run_matches_reference(name):
    g = O.golden(name)
    o1, o2, oph, dts = run_golden(g)
    O.assert_bitwise(dts, g["dt"], f"{name} dt")
    O.assert_bitwise(o1, g["m1_out"], f"{name} material 1")
    O.assert_bitwise(o2, g["m2_out"], f"{name} material 2")
    O.assert_bitwise(oph, g["phi_out"], f"{name} level set")


# Toro's tests 1-5 (gamma 1.4): left, right, exact p*, u*, rho*_L, rho*_R.
TORO = [
    ((1.0, 0.0, 1.0), (0.125, 0.0, 0.1), 0.30313, 0.92745, 0.42632, 0.26557),
    ((1.0, -2.0, 0.4), (1.0, 2.0, 0.4), 0.00189, 0.0, 0.02185, 0.02185),
    ((1.0, 0.0, 1000.0), (1.0, 0.0, 0.01), 460.894, 19.5975, 0.57506, 5.99924),
    ((1.0, 0.0, 0.01), (1.0, 0.0, 100.0), 46.0950, -6.19633, 5.99242, 0.57511),
    ((5.99924, 19.5975, 460.894), (5.99242, -6.19633, 46.0950), 1691.64, 8.68975, 14.2823, 31.0426),
]


def pair(a, b):
    return np.array([a[0], a[1], 0.0, 0.0, a[2], 1.0, b[0], b[1], 0.0, 0.0, b[2], 1.0])


def test_exact_solver_toro_tables():
    st = np.stack([pair(a, b) for a, b, *_ in TORO])
    out = O.exact_units(st, AIR, AIR)
    # the table prints 5-6 significant digits: half a unit of the last one
    for k, (_, _, ps, us, rl, rr) in enumerate(TORO):
        for got, want, what in ((out[k, 0], ps, "p*"), (out[k, 1], us, "u*"), (out[k, 2], rl, "rho*L"),
                                (out[k, 3], rr, "rho*R")):
            assert got == pytest.approx(want, rel=6e-5, abs=5e-6), f"test {k + 1} {what}"


def test_exact_solver_jump_conditions():
    """Shocks satisfy Rankine-Hugoniot (mass and momentum flux continuous in the
    shock frame); rarefactions keep p / rho^gamma and u +- 2c/(gamma-1)."""
    rng = np.random.default_rng(5)
    g = 1.4
    for _ in range(40):
        a = (10 ** rng.uniform(-1, 1), rng.uniform(-1, 1), 10 ** rng.uniform(-1, 1))
        b = (10 ** rng.uniform(-1, 1), rng.uniform(-1, 1), 10 ** rng.uniform(-1, 1))
        ps, us, rl, rr = O.exact_units(pair(a, b)[None], AIR, AIR)[0, :4]
        for (rho, u, p), rs, side in ((a, rl, -1), (b, rr, 1)):
            if ps > p:
                S = (rs * us - rho * u) / (rs - rho)
                m_star, m_k = rs * (us - S), rho * (u - S)
                assert m_star == pytest.approx(m_k, rel=1e-7, abs=1e-9)
                assert rs * us * (us - S) + ps == pytest.approx(rho * u * (u - S) + p, rel=1e-7)
            else:
                assert ps / rs ** g == pytest.approx(p / rho ** g, rel=1e-9)
                c, cs = math.sqrt(g * p / rho), math.sqrt(g * ps / rs)
                assert us - side * 2 * cs / (g - 1) == pytest.approx(u - side * 2 * c / (g - 1), rel=1e-9, abs=1e-9)


def test_hllc_pvrs_closed_form():
    """Different gammas force the PVRS estimate: p* = (pL+pR)/2 + (uL-uR)/2 * rho_bar c_bar
    with rho_bar c_bar = (rhoL+rhoR)(cL+cR)/4."""
    he = (1.648, 0.0, 25.0, 0.0, 0.25, 250.0)
    a, b = (1.0, 0.3, 2.0), (0.2, -0.1, 1.5)
    out = O.hllc_units(pair(a, b)[None], AIR, he)[0]
    cl, cr = math.sqrt(1.4 * a[2] / a[0]), math.sqrt(1.648 * b[2] / b[0])
    prod = 0.25 * (a[0] + b[0]) * (cl + cr)
    assert out[9] == pytest.approx(0.5 * (a[2] + b[2]) + 0.5 * (a[1] - b[1]) * prod, rel=1e-14)
    assert out[8] == pytest.approx(0.5 * (a[1] + b[1]) + 0.5 * (a[2] - b[2]) / prod, rel=1e-14)


def test_level_set_exact_on_linear_fields():
    """The one-sided second-order stencil differentiates linear phi exactly, so
    phi_new = phi - dt u phi_x away from the mirrored ends, for either sign of u."""
    n = 40
    phi = 0.37 * np.arange(n) - 5.0
    for u in (0.8, -0.6):
        m = np.zeros((n, 6))
        m[:, 1] = u
        out = O.rgfm_level_set(1, 0, phi, m, m, 0.37, 0.01)
        np.testing.assert_allclose(out[3:-3], (phi - 0.01 * u)[3:-3], rtol=0, atol=1e-13)


def test_circulation_of_solid_body_rotation():
    """u = -W y, v = W x has vorticity 2W, so the circulation over the phi < 0
    interior cells is 2 W dx dy times their number."""
    ny, nx, dx, dy, W = 30, 40, 0.01, 0.02, 3.0
    x = (np.arange(nx) + 0.5) * dx
    y = (np.arange(ny) + 0.5) * dy
    X, Y = np.meshgrid(x, y)
    m1 = np.zeros((ny, nx, 6))
    m1[..., 1] = -W * Y
    m1[..., 2] = W * X
    phi = np.hypot(X - 0.2, Y - 0.3) - 0.12
    inside = (phi[1:-1, 1:-1] < 0).sum()
    assert O.rgfm_circulation(m1, phi, dx, dy) == pytest.approx(2 * W * dx * dy * inside, rel=1e-12)


# ---- the exact solver's sampled state (RGFM/exactsolver.cpp:262-446) ----------
#
# Every region of the sampled solution is fixed by conditions that do not use the
# oracle's closed forms: the data states outside the waves (copied whole), the
# Rankine-Hugoniot shock speed from mass conservation across the computed jump,
# the rarefaction head u -+ c and tail u* -+ c*, where c* follows from the star
# density by the equation of state, and inside a fan the three conditions that
# define it -- the characteristic through the sample point (u -+ c = x/t), the
# isentrope (p + p_inf) / rho^gamma and the Riemann invariant u +- 2c/(gamma - 1)
# of the data side.  Star and fan states carry zero tangential velocity and
# lambda = 1 (the reference's five-argument EulerStateVector constructor,
# Euler/eulerstatevector.cpp:14-23).

WATER = (4.4, 6.0e8, 25.0, 0.0, 0.25, 250.0)
STIFF = (3.0, 2.0, 25.0, 0.0, 0.25, 250.0)
HELIUM = (1.648, 0.0, 25.0, 0.0, 0.25, 250.0)

# (left rho, u, p), (right rho, u, p), left material, right material, the region
# x/t = 0 falls in (checked below from the wave speeds).
SAMPLED = [
    ((1.0, 0.0, 1.0), (0.125, 0.0, 0.1), AIR, AIR, "star_a"),              # Sod: left fan, right shock
    ((0.125, 0.0, 0.1), (1.0, 0.0, 1.0), AIR, AIR, "star_b"),              # its mirror
    ((1.0, 0.75, 1.0), (0.125, 0.0, 0.1), AIR, AIR, "fan_a"),              # Toro's sonic rarefaction (left)
    ((0.125, 0.0, 0.1), (1.0, -0.75, 1.0), AIR, AIR, "fan_b"),             # its mirror (right)
    ((1.0, 3.0, 1.0), (0.5, 3.0, 0.5), AIR, AIR, "state_a"),               # supersonic to the right
    ((0.5, -3.0, 0.5), (1.0, -3.0, 1.0), AIR, AIR, "state_b"),             # supersonic to the left
    ((1.0, 2.5, 1.0), (1.0, 0.0, 1.0), AIR, AIR, "state_a"),               # two shocks, both moving right
    ((1.0, 0.6, 1.0), (1.0, -0.4, 1.0), AIR, AIR, "star_a"),               # two shocks moving apart
    ((5.99924, 19.5975, 460.894), (5.99242, -6.19633, 46.0950), AIR, AIR, "state_a"),  # Toro test 5
    ((1.0, 3.0, 0.2), (0.5, 0.0, 1.0), AIR, HELIUM, "state_a"),            # left shock swept right past x/t = 0
    ((1.0, 4.0, 10.0), (0.05, 0.0, -1.5), STIFF, STIFF, "fan_a"),          # stiffened gas, sonic left fan
    ((0.05, 0.0, -1.5), (1.0, -4.0, 10.0), STIFF, STIFF, "fan_b"),         # stiffened gas, sonic right fan
    ((1000.0, 0.0, 1.0e9), (1.0, 0.0, 1.0e5), WATER, AIR, "star_a"),       # water / air shock tube
    ((1.0, 0.0, 1.0e5), (1000.0, 0.0, 1.0e9), AIR, WATER, "star_b"),
    ((1000.0, 450.0, 1.0e8), (1000.0, 0.0, 1.0e8), WATER, WATER, "star_a"),  # stiffened gas, two shocks
    ((1000.0, -20.0, 1.0e8), (1000.0, 30.0, 1.0e8), WATER, WATER, "star_a"),  # stiffened gas, two rarefactions
    # cavitation: the Newton iterate is floored at p* = 1e-8 (exactsolver.cpp:115-138) and u* = 0 exactly,
    # so x/t = 0 is sampled on the right of the contact
    ((1000.0, -200.0, 1.0e8), (1000.0, 200.0, 1.0e8), WATER, WATER, "star_b"),
]


def _c(g, pinf, rho, p):
    return math.sqrt(g * (p + pinf) / rho)


def _wave_structure(a, b, ma, mb, ps, us, ra, rb):
    """The wave speeds bounding each region, from the data, the star pressure and
    velocity, and the star densities (whose shock / isentrope conditions the
    jump-condition tests pin)."""
    out = {}
    for side, (rho, u, p), m, rs, sg in (("a", a, ma, ra, -1.0), ("b", b, mb, rb, 1.0)):
        g, pinf = m[0], m[1]
        if ps >= p:
            shock = (rs * us - rho * u) / (rs - rho)
            out[side] = ("shock", shock, shock)
        else:
            head = u + sg * _c(g, pinf, rho, p)
            tail = us + sg * _c(g, pinf, rs, ps)
            out[side] = ("fan", head, tail)
    return out


def _region(s, us, ws):
    kind_a, head_a, tail_a = ws["a"]
    kind_b, head_b, tail_b = ws["b"]
    if s < us:
        if s < head_a:
            return "state_a"
        return "star_a" if s > tail_a else "fan_a"
    if s > head_b:
        return "state_b"
    return "star_b" if s < tail_b else "fan_b"


def _check_sample(out, s, a_full, b_full, ma, mb, ps, us, ra, rb, region, what, direction=0):
    n = 1 + direction            # normal velocity slot
    t = 2 - direction            # tangential velocity slot
    if region == "state_a":
        O.assert_bitwise(out, a_full, what)
    elif region == "state_b":
        O.assert_bitwise(out, b_full, what)
    elif region.startswith("star"):
        rs = ra if region == "star_a" else rb
        want = np.zeros(6)
        want[0], want[n], want[4], want[5] = rs, us, ps, 1.0
        O.assert_bitwise(out, want, what)
    else:
        src, m, sg = (a_full, ma, -1.0) if region == "fan_a" else (b_full, mb, 1.0)
        g, pinf = m[0], m[1]
        rho, u, p = out[0], out[n], out[4]
        assert out[t] == 0.0 and out[3] == 0.0 and out[5] == 1.0, what
        c = _c(g, pinf, rho, p)
        c0 = _c(g, pinf, src[0], src[4])
        scale = abs(src[n]) + c0
        # the characteristic through x/t = s
        assert u + sg * c == pytest.approx(s, rel=1e-11, abs=1e-11 * scale), what
        # the isentrope of the data side
        assert (p + pinf) / rho ** g == pytest.approx((src[4] + pinf) / src[0] ** g, rel=1e-11), what
        # the Riemann invariant carried across the fan
        assert u - sg * 2.0 * c / (g - 1.0) == pytest.approx(src[n] - sg * 2.0 * c0 / (g - 1.0), rel=1e-11,
                                                             abs=1e-11 * scale), what


def _full(w, direction, tangential, lam):
    f = np.zeros(6)
    f[0], f[1 + direction], f[2 - direction], f[3], f[4], f[5] = w[0], w[1], tangential, 0.25, w[2], lam
    return f


@pytest.mark.parametrize("direction", [0, 1])
@pytest.mark.parametrize("k", range(len(SAMPLED)))
def test_exact_solver_sampled_state(k, direction):
    """The state orc_exact_solve returns at x/t = s lies in the region the wave
    speeds put s in and equals that region's state, across a sweep of s through
    every wave, and at s = 0 (the RGFM ghost fill's sample) in the region listed."""
    a, b, ma, mb, at_zero = SAMPLED[k]
    # tangential velocity and lambda differ per side, so copies are told apart
    af, bf = _full(a, direction, 0.3, 0.7), _full(b, direction, -0.2, 0.4)
    st = np.concatenate([af, bf])[None]
    ps, us, ra, rb = O.exact_units(st, ma, mb, 0.0, direction)[0, :4]
    ws = _wave_structure(a, b, ma, mb, ps, us, ra, rb)
    speeds = sorted([ws["a"][1], ws["a"][2], us, ws["b"][1], ws["b"][2]])
    span = speeds[-1] - speeds[0] + 1.0
    samples = list(np.linspace(speeds[0] - 0.1 * span, speeds[-1] + 0.1 * span, 241)) + [0.0]
    seen = set()
    for s in samples:
        if min(abs(s - v) for v in speeds) < 1e-9 * span:
            continue
        region = _region(s, us, ws)
        seen.add(region)
        out = O.exact_units(st, ma, mb, s, direction)[0, 4:]
        _check_sample(out, s, af, bf, ma, mb, ps, us, ra, rb, region, f"problem {k} dir {direction} s={s!r} ({region})",
                      direction)
    assert _region(0.0, us, ws) == at_zero
    # both data states and both sides of the contact are reached by the sweep
    assert {"state_a", "state_b"} <= seen and ({"star_a", "fan_a"} & seen) and ({"star_b", "fan_b"} & seen)


def test_exact_solver_direction_symmetry():
    """solveY is solveX with the velocity slots exchanged, bit for bit."""
    rng = np.random.default_rng(11)
    for _ in range(50):
        a = (10 ** rng.uniform(-1, 1), rng.uniform(-2, 2), 10 ** rng.uniform(-1, 1))
        b = (10 ** rng.uniform(-1, 1), rng.uniform(-2, 2), 10 ** rng.uniform(-1, 1))
        s = rng.uniform(-2, 2)
        x = O.exact_units(np.concatenate([_full(a, 0, 0.1, 1.0), _full(b, 0, 0.2, 1.0)])[None], AIR, STIFF, s, 0)[0]
        y = O.exact_units(np.concatenate([_full(a, 1, 0.1, 1.0), _full(b, 1, 0.2, 1.0)])[None], AIR, STIFF, s, 1)[0]
        O.assert_bitwise(y[:4], x[:4], "star values")
        O.assert_bitwise(y[4:][[0, 2, 1, 3, 4, 5]], x[4:], "sampled state")


def test_exact_solver_mirror_symmetry():
    """Reflecting x -> -x (swap the sides and materials, negate the normal
    velocities and s) reflects the sampled solution."""
    rng = np.random.default_rng(12)
    for _ in range(60):
        a = (10 ** rng.uniform(-1, 1), rng.uniform(-2, 2), 10 ** rng.uniform(-1, 1))
        b = (10 ** rng.uniform(-1, 1), rng.uniform(-2, 2), 10 ** rng.uniform(-1, 1))
        s = rng.uniform(-3, 3)
        fwd = O.exact_units(np.concatenate([_full(a, 0, 0.0, 1.0), _full(b, 0, 0.0, 1.0)])[None], AIR, STIFF, s)[0]
        ma, mb = (b[0], -b[1], b[2]), (a[0], -a[1], a[2])
        rev = O.exact_units(np.concatenate([_full(ma, 0, 0.0, 1.0), _full(mb, 0, 0.0, 1.0)])[None], STIFF, AIR, -s)[0]
        assert rev[0] == pytest.approx(fwd[0], rel=1e-10)
        assert rev[1] == pytest.approx(-fwd[1], rel=1e-10, abs=1e-12)
        st_f, st_r = fwd[4:], rev[4:]
        assert st_r[0] == pytest.approx(st_f[0], rel=1e-9)
        assert st_r[1] == pytest.approx(-st_f[1], rel=1e-9, abs=1e-11)
        assert st_r[4] == pytest.approx(st_f[4], rel=1e-9)


# ---- anisotropic spacing (dx != dy; no reference counterpart) ---------------
#
# The two-material solve on a dx != dy grid (the shock-bubble geometry's 2048 x
# 1024 cells on 0.325 x 0.089) takes dt from min(dx, dy), sweeps and advects the
# level set along y with dy and along x with dx.  These checks are fixed by the
# mathematics, not by the oracle's own code: data constant along one axis make
# every pass along that axis the identity up to round-off, so the solve reduces
# to a 1D solve along the other axis with that axis's spacing, or to the square
# grid the reference runs; and the one-sided level-set stencil is exact on linear
# fields.

SB_AIR = (1.4, 0.0, 25.0, 0.0, 0.25, 250.0)
SB_HELIUM = (1.648, 0.0, 25.0, 0.0, 0.25, 250.0)


def _line_problem(n):
    """A shock in air running into a helium slab: (air, helium, phi) along a line
    of n cells on [0, 1]; air holds phi < 0."""
    x = (np.arange(n) + 0.5) / n
    air = np.tile([1.0, 0.0, 0.0, 0.0, 1.0e5, 1.0], (n, 1))
    air[x < 0.42] = [1.3764, 124.82, 0.0, 0.0, 1.5698e5, 1.0]
    helium = np.tile([0.138, 0.0, 0.0, 0.0, 1.0e5, 1.0], (n, 1))
    return air, helium, x - 0.55


def _rel_l1(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.abs(a - b).sum() / max(np.abs(b).sum(), 1e-300)


def _extrude(line, other, axis):
    """A 2D field constant along `axis` (1: along x, rows copy the line) with the
    line's normal velocity moved into the other axis's slot."""
    n = line.shape[0]
    if line.ndim == 1:
        return np.repeat(line[:, None], other, axis=1) if axis == 1 else np.repeat(line[None, :], other, axis=0)
    w = line.copy()
    if axis == 1:                   # varies along y: the normal velocity is v
        w[:, 1], w[:, 2] = 0.0, line[:, 1]
        return np.repeat(w[:, None, :], other, axis=1)
    return np.repeat(w[None, :, :], other, axis=0)


STEPS = 40


def test_rgfm_x_uniform_anisotropic_is_the_1d_solve_along_y():
    """Data constant along x on a grid with dy < dx: every x pass is the identity,
    dt = CFL dy / S, and each column is the 1D solve along y with spacing dy.  The
    2D scan starts inside the second material (rgfmsolver.cpp:378, 564, 605 vs the
    1D :304, 503), so the materials trade places between the two runs."""
    ny, nx = 96, 7
    dy, dx = 1.0 / ny, 2.3 / ny
    air, he, phi = _line_problem(ny)
    a1, a2, aph, adt = O.rgfm_solve(air, he, phi, dim=1, dx=dy, cfl=0.8, tfinal=math.inf, max_steps=STEPS, bias=0.0,
                                    limiter=0, p1=SB_AIR, p2=SB_HELIUM)
    b1, b2, bph, bdt = O.rgfm_solve(_extrude(he, nx, 1), _extrude(air, nx, 1), _extrude(phi, nx, 1), dim=2, dx=dx,
                                    dy=dy, cfl=0.8, tfinal=math.inf, max_steps=STEPS, bias=0.0, limiter=0,
                                    p1=SB_HELIUM, p2=SB_AIR)
    np.testing.assert_allclose(bdt, adt, rtol=1e-12)
    for j in range(nx):
        col_air, col_he = b2[:, j].copy(), b1[:, j].copy()
        col_air[:, [1, 2]] = col_air[:, [2, 1]]
        col_he[:, [1, 2]] = col_he[:, [2, 1]]
        assert _rel_l1(col_air, a1) <= 1e-12, f"air, column {j}"
        assert _rel_l1(col_he, a2) <= 1e-12, f"helium, column {j}"
        np.testing.assert_allclose(bph[:, j], aph, rtol=0, atol=1e-13)
    # the waves have crossed the interface: a spacing error would be O(1), not round-off
    assert np.abs(a2[:, 4] - 1.0e5).max() > 5.0e3


def test_rgfm_x_uniform_anisotropic_is_the_square_grid():
    """The same x-uniform data on dy < dx and on the square dy x dy grid."""
    ny, nx = 64, 6
    dy = 1.0 / ny
    air, he, phi = _line_problem(ny)
    runs = []
    for dx in (dy, 1.7 * dy):
        runs.append(O.rgfm_solve(_extrude(he, nx, 1), _extrude(air, nx, 1), _extrude(phi, nx, 1), dim=2, dx=dx, dy=dy,
                                 cfl=0.8, tfinal=math.inf, max_steps=STEPS, bias=0.0, limiter=1, p1=SB_HELIUM,
                                 p2=SB_AIR))
    (s1, s2, sph, sdt), (r1, r2, rph, rdt) = runs
    np.testing.assert_allclose(rdt, sdt, rtol=1e-12)
    assert _rel_l1(r1, s1) <= 1e-12 and _rel_l1(r2, s2) <= 1e-12
    np.testing.assert_allclose(rph, sph, rtol=0, atol=1e-13)


def test_rgfm_y_uniform_anisotropic_is_the_square_grid():
    """The x-only twin: data constant along y on a grid with dx < dy equal the
    square dx x dx grid (every y pass the identity, dt = CFL dx / S)."""
    ny, nx = 6, 64
    dx = 1.0 / nx
    air, he, phi = _line_problem(nx)
    runs = []
    for dy in (dx, 1.9 * dx):
        runs.append(O.rgfm_solve(_extrude(he, ny, 0), _extrude(air, ny, 0), _extrude(phi, ny, 0), dim=2, dx=dx, dy=dy,
                                 cfl=0.8, tfinal=math.inf, max_steps=STEPS, bias=0.0, limiter=0, p1=SB_HELIUM,
                                 p2=SB_AIR))
    (s1, s2, sph, sdt), (r1, r2, rph, rdt) = runs
    np.testing.assert_allclose(rdt, sdt, rtol=1e-12)
    assert _rel_l1(r1, s1) <= 1e-12 and _rel_l1(r2, s2) <= 1e-12
    np.testing.assert_allclose(rph, sph, rtol=0, atol=1e-13)
    assert np.abs(r2[..., 4] - 1.0e5).max() > 5.0e3


@pytest.mark.parametrize("axis,vel", [(0, 37.0), (0, -41.0), (1, 29.0), (1, -33.0)])
def test_rgfm_linear_level_set_anisotropic(axis, vel):
    """phi linear along one axis, both materials at rest relative to a uniform
    velocity along it, on a dx != dy grid: the interface is a contact carried at
    that velocity, so phi_n = phi_0 - vel sum(dt) wherever the mirrored boundary
    has not reached (2 cells per step upwind of the flow), and dt = CFL min(dx, dy)
    / S, S = |vel| + the larger sound speed (x 0.2 for the first six steps)."""
    ny, nx = 40, 56
    dx, dy = 0.013, 0.0071
    steps = 6
    y = (np.arange(ny) + 0.5) * dy
    x = (np.arange(nx) + 0.5) * dx
    X, Y = np.meshgrid(x, y)
    phi = (Y - 0.5 * ny * dy) if axis == 1 else (X - 0.5 * nx * dx)
    m1 = np.zeros((ny, nx, 6))
    m2 = np.zeros((ny, nx, 6))
    m1[...] = [0.138, 0.0, 0.0, 0.0, 1.0e5, 1.0]
    m2[...] = [1.0, 0.0, 0.0, 0.0, 1.0e5, 1.0]
    m1[..., 1 + axis] = vel
    m2[..., 1 + axis] = vel
    o1, o2, oph, dts = O.rgfm_solve(m1, m2, phi, dim=2, dx=dx, dy=dy, cfl=0.8, tfinal=math.inf, max_steps=steps,
                                    bias=0.0, limiter=0, p1=SB_HELIUM, p2=SB_AIR)
    S = abs(vel) + math.sqrt(1.648 * 1.0e5 / 0.138)
    want_dt = [0.8 * min(dx, dy) / S * (0.2 if k <= 5 else 1.0) for k in range(steps)]
    np.testing.assert_allclose(dts, want_dt, rtol=1e-13)
    want = phi - vel * dts.sum()
    reach = 2 * steps
    if axis == 1:
        sl = (slice(reach, ny), slice(None)) if vel > 0 else (slice(0, ny - reach), slice(None))
    else:
        sl = (slice(None), slice(reach, nx)) if vel > 0 else (slice(None), slice(0, nx - reach))
    np.testing.assert_allclose(oph[sl], want[sl], rtol=0, atol=1e-14)
    # a displacement the check resolves: the contact moved by many round-offs
    assert abs(vel) * dts.sum() > 1e-7
