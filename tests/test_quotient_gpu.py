"""The EXACT mode's shortened division sequences (vanguard_amd/csrc/vg_quot.h)
against the compiler's IEEE division on the GPU, bit for bit, over the ranges
the solvers use them in: random operands across the exponent range, quotients
within a few units of the last place of a rounding midpoint, exact quotients,
divisors next to powers of two and all-ones mantissas, signed zeros, and the
slope limiter's operand ranges.  A failure here would make EXACT mode differ
from the reference's arithmetic."""

from fractions import Fraction

import numpy as np
import pytest

from vanguard_amd import _lib
from vanguard_amd import solvers as S

pytestmark = pytest.mark.gpu


def run(a, b):
    ctx = S.default_context()
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    outs = [np.empty_like(a) for _ in range(3)]
    rc = ctx.lib.vg_division_selfcheck(ctx.handle, a.size, _lib.as_dptr(a), _lib.as_dptr(b),
                                       *[_lib.as_dptr(o) for o in outs])
    assert rc == 0
    return outs


def bits(x):
    return np.ascontiguousarray(x).view(np.uint64)


def make(rng, n, e_lo, e_hi, signed=True):
    mant = rng.integers(0, 1 << 52, n, dtype=np.uint64)
    e = rng.integers(e_lo, e_hi + 1, n).astype(np.uint64)
    x = ((e + np.uint64(1023)) << np.uint64(52)) | mant
    if signed:
        x |= rng.integers(0, 2, n, dtype=np.uint64) << np.uint64(63)
    return x.view(np.float64)


def near_midpoints(rng, n, e_q=(-40, 40), e_b=(-100, 100)):
    """Numerators whose quotient by b lies within a few ulps of the midpoint
    between two neighbouring doubles."""
    qs = make(rng, n, *e_q)
    bs = make(rng, n, *e_b, signed=False)
    a = np.empty(n)
    for i in range(n):
        q = float(qs[i])
        mid = Fraction(q) + Fraction(np.spacing(abs(q))) / 2 * (1 if q > 0 else -1)
        a[i] = float(mid * Fraction(float(bs[i])))
    return a, bs


def assert_same(got, want, a, b, what):
    bad = bits(got) != bits(want)
    assert not bad.any(), (f"{what}: {int(bad.sum())} of {a.size} differ, first a={a[bad][0]!r} b={b[bad][0]!r} "
                           f"got={got[bad][0]!r} want={want[bad][0]!r}")


def test_shared_reciprocal_quotients_are_ieee_division():
    """Densities 2^-100 .. 2^100; numerators 0 or 2^-600 .. 2^600 in magnitude."""
    rng = np.random.default_rng(101)
    n = 1 << 22
    a = make(rng, n, -600, 599)
    b = make(rng, n, -100, 99, signed=False)
    a[:4096] = np.where(np.arange(4096) % 2 == 0, 0.0, -0.0)
    m, ma = near_midpoints(rng, 20000)
    a = np.concatenate([a, m, ma * 3.0, ma * 7.0])
    b = np.concatenate([b, ma, ma, ma])
    # divisors next to powers of two and all-ones mantissas
    e = rng.integers(-100, 100, 20000)
    k = rng.integers(0, 64, 20000)
    bb = np.ldexp(np.where(k % 2 == 0, 1.0 + k * 2.0 ** -52, 2.0 - k * 2.0 ** -52), e)
    a = np.concatenate([a, make(rng, 20000, -30, 30)])
    b = np.concatenate([b, bb])
    shared, in_range, ieee = run(a, b)
    assert_same(shared, ieee, a, b, "shared reciprocal")
    assert_same(in_range, ieee, a, b, "short sequence")
    # the signed zeros keep their signs
    assert (np.signbit(shared[:4096]) == np.signbit(a[:4096])).all()


def test_limiter_divisions_are_ieee_division():
    """The slope ratio's operands (1e-5 .. 2^600 in magnitude, either sign) and
    the limiter's own divisions (2 / (1 + r) style denominators up to 2^619,
    2r / (1 + r) down to 2^-617)."""
    rng = np.random.default_rng(102)
    n = 1 << 22
    num = make(rng, n, -17, 599)
    den = make(rng, n, -17, 599)
    num[num == 0] = 1e-5
    m, mb = near_midpoints(rng, 20000, e_q=(-300, 300), e_b=(-17, 280))
    r = make(rng, 1 << 20, -617, 617, signed=False)
    a = np.concatenate([num, m, np.full(r.size, 2.0), 2.0 * r])
    b = np.concatenate([den, mb, 1.0 + r, 1.0 + r])
    _, in_range, ieee = run(a, b)
    assert_same(in_range, ieee, a, b, "limiter division")


def test_out_of_range_divisors_fall_back():
    """Divisors outside 2^-100 .. 2^100, negative, zero, NaN and infinite ones
    take the IEEE division in the shared form."""
    rng = np.random.default_rng(103)
    b = np.concatenate([make(rng, 4096, 101, 1000, signed=False), make(rng, 4096, -1000, -101, signed=False),
                        -make(rng, 4096, -100, 99, signed=False), [0.0, -0.0, np.inf, -np.inf, np.nan]])
    a = make(rng, b.size, -600, 599)
    shared, _, ieee = run(a, b)
    assert_same(shared, ieee, a, b, "fall-back")


def test_tiny_numerators():
    """Numerators below the range the header comment of vg_quot.h states
    (2^-1000 .. 2^-600 in magnitude) by divisors within 2^-100 .. 2^100: the
    shared form is the compiler's division bit for bit.  Deeper numerators
    (2^-1074 .. 2^-1000, subnormals included) are the documented exception:
    there it may differ, by at most one unit in the last place."""
    rng = np.random.default_rng(104)
    n = 1 << 20
    a = make(rng, n, -1000, -601)
    b = make(rng, a.size, -100, 99, signed=False)
    shared, _, ieee = run(a, b)
    assert_same(shared, ieee, a, b, "numerators 2^-1000 .. 2^-600")
    deep = np.concatenate([make(rng, 1 << 18, -1022, -1001),
                           rng.integers(1, 1 << 52, 4096, dtype=np.uint64).view(np.float64)])
    deep = np.concatenate([deep, -deep[-4096:]])
    bd = make(rng, deep.size, -100, 99, signed=False)
    shared, _, ieee = run(deep, bd)
    ulps = np.abs(bits(np.abs(shared)).astype(np.int64) - bits(np.abs(ieee)).astype(np.int64))
    same_sign = np.signbit(shared) == np.signbit(ieee)
    print(f"numerators below 2^-1000: {int((ulps > 0).sum())} of {deep.size} quotients differ, "
          f"at most {int(ulps.max())} ulp", flush=True)
    assert same_sign.all() and ulps.max() <= 1
