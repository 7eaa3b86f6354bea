// Exhaustive-style check of vg::quot (vanguard_amd/csrc/vg_quot.h) against IEEE
// division on the host: random operands over the whole fast range, quotients
// next to rounding midpoints, exact quotients, signed zeros, the divisor
// mantissas next to all-ones and powers of two, and the fall-back range.
// Prints the number of cases and mismatches; exits 1 on any mismatch.
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "vg_quot.h"

static uint64_t s_state = 0x9E3779B97F4A7C15ull;
static uint64_t next_u64()
{
    uint64_t z = (s_state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static uint64_t bits(double x) { uint64_t u; memcpy(&u, &x, 8); return u; }
static double from_bits(uint64_t u) { double x; memcpy(&x, &u, 8); return x; }
// A double with a random mantissa, sign s (0/1) and unbiased exponent e.
static double make(int s, int e, uint64_t mant)
{
    return from_bits(((uint64_t)s << 63) | ((uint64_t)(e + 1023) << 52) | (mant & ((1ull << 52) - 1)));
}

static long long n_cases = 0, n_bad = 0;
static void check(double a, double b)
{
    vg::SharedDivisor d = vg::shared_divisor(b);
    double got = vg::quot(a, d), want = a / b;
    n_cases++;
    if (bits(got) != bits(want) && !(std::isnan(got) && std::isnan(want))) {
        if (n_bad < 10) printf("MISMATCH a=%a b=%a got=%a want=%a\n", a, b, got, want);
        n_bad++;
    }
}

int main(int argc, char **argv)
{
    long long n = argc > 1 ? atoll(argv[1]) : 20000000;
    // 1. random operands: divisor exponent in [-100, 99], numerator in [-800, 799]
    for (long long k = 0; k < n; k++) {
        uint64_t r = next_u64();
        double b = make(0, (int)(r % 200) - 100, next_u64());
        double a = make((int)((r >> 20) & 1), (int)((r >> 21) % 1600) - 800, next_u64());
        check(a, b);
    }
    // 2. physical magnitudes (densities 1e-3..1e3, momenta and energies 1e-12..1e9)
    for (long long k = 0; k < n; k++) {
        uint64_t r = next_u64();
        double b = make(0, (int)(r % 21) - 10, next_u64());
        double a = make((int)((r >> 8) & 1), (int)((r >> 9) % 70) - 40, next_u64());
        check(a, b);
    }
    // 3. quotients next to a rounding midpoint: a = RN(b m) with m = q + ulp(q)/2,
    //    and its neighbours, so a / b lies within a few ulp(a)/b of the midpoint
    for (long long k = 0; k < n; k++) {
        uint64_t r = next_u64();
        double b = make(0, (int)(r % 64) - 32, next_u64());
        double q = make(0, (int)((r >> 8) % 64) - 32, next_u64());
        double m = q + 0.5 * (nextafter(q, INFINITY) - q);   // exact: the midpoint needs 54 bits, rounds to even
        long double ml = (long double)q + 0.5L * ((long double)nextafter(q, INFINITY) - (long double)q);
        double a = (double)((long double)b * ml);
        (void)m;
        for (int t = -2; t <= 2; t++) {
            double aa = a;
            for (int s = 0; s < (t < 0 ? -t : t); s++) aa = nextafter(aa, t < 0 ? -INFINITY : INFINITY);
            check(aa, b);
            check(-aa, b);
        }
    }
    // 4. exact quotients (a = b k for small integers k) and powers of two
    for (long long k = 0; k < n / 10; k++) {
        uint64_t r = next_u64();
        double b = make(0, (int)(r % 200) - 100, next_u64());
        double kk = (double)((int)((r >> 10) % 2001) - 1000);
        check(b * kk, b);
        check(make(0, (int)((r >> 30) % 200) - 100, 0), b);
    }
    // 5. divisors with all-ones and near-all-ones mantissas, and powers of two
    for (int e = -60; e <= 60; e++)
        for (int t = 0; t < 64; t++) {
            double b1 = make(0, e, (1ull << 52) - 1 - (uint64_t)t);
            double b2 = make(0, e, (uint64_t)t);
            for (int k = 0; k < 2000; k++) {
                double a = make((int)(next_u64() & 1), (int)(next_u64() % 120) - 60, next_u64());
                check(a, b1);
                check(a, b2);
            }
        }
    // 6. signed zeros, infinities, NaN, and divisors outside the fast range
    // (divisors: any; numerators in the documented range or zero)
    double divs[] = {0.0, -0.0, 1.0, -1.0, INFINITY, -INFINITY, NAN, DBL_MIN, -DBL_MIN, DBL_TRUE_MIN, DBL_MAX,
                     0x1p-100, 0x1p100, 0x1.fffffffffffffp-101, 0x1.0000000000001p100, 3.0, 0.1};
    double nums[] = {0.0, -0.0, 1.0, -1.0, 0x1p-800, -0x1p-800, 0x1p800, 0x1.fffffffffffffp799, 3.0, 0.1, 1e240,
                     -1e-200};
    for (double a : nums)
        for (double b : divs) check(a, b);
    for (int k = 0; k < 100000; k++) {
        double a = make((int)(next_u64() & 1), (int)(next_u64() % 1600) - 800, next_u64());
        check(a, make(0, (int)(next_u64() % 900) + 101 - 1000 * (int)(next_u64() & 1), next_u64()));
        check(a, -make(0, (int)(next_u64() % 100) - 50, next_u64()));
        check(0.0, make(0, (int)(next_u64() % 100) - 50, next_u64()));
        check(-0.0, make(0, (int)(next_u64() % 100) - 50, next_u64()));
    }
    printf("cases %lld mismatches %lld\n", n_cases, n_bad);
    return n_bad ? 1 : 0;
}
