// Checks the refinement steps of grhd::sqrt64 and grhd::div64
// (grhd_amd/csrc/grhd_physics.hpp), which the device runs from the hardware
// estimates v_rsq_f64 and v_rcp_f64: from any estimate within a few ulps of
// 1 / sqrt(x) or 1 / b, grhd::sqrt_refine and grhd::div_refine must return the
// IEEE square root and quotient bit for bit, over the arguments the device
// path produces (square roots at or above 2^-767; quotients of operands and
// results in the normal range).  The estimates here are the exact values
// scaled by 1 + e, with e = 0, +-2^-40 and nine random values in
// (-2^-14, 2^-14): the steps return the IEEE result from estimates that far
// off (they fail from about 2^-12), while one step fewer fails at 2^-14 --
// the check pins every step.  Prints one line per part and exits non-zero on
// any mismatch.
#include "grhd_physics.hpp"

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>

static uint64_t state = 0x2545f4914f6cdd1dull;
static uint64_t next()
{
    state ^= state << 13;
    state ^= state >> 7;
    state ^= state << 17;
    return state;
}
static double d_of(uint64_t mant, int exp, int sign)
{
    const uint64_t u = (mant & 0x000fffffffffffffull) | ((uint64_t)exp << 52) | ((uint64_t)sign << 63);
    double x;
    std::memcpy(&x, &u, 8);
    return x;
}
static bool same(double x, double y) { return std::memcmp(&x, &y, 8) == 0; }
static double ulps(double x, int k)
{
    uint64_t u;
    std::memcpy(&u, &x, 8);
    u += (int64_t)k;
    std::memcpy(&x, &u, 8);
    return x;
}
// relative errors of the estimates
static const int kEst = 12;
static long double rel_err(int k)
{
    if (k == 0) return 0.0L;
    if (k == 1) return std::ldexp(1.0L, -40);
    if (k == 2) return -std::ldexp(1.0L, -40);
    const long double u = (long double)(next() >> 11) / 9007199254740992.0L;  // [0, 1)
    return (2.0L * u - 1.0L) * std::ldexp(1.0L, -14);
}

int main(int argc, char **argv)
{
    const long n = argc > 1 ? std::atol(argv[1]) : 1000000;
    long bad = 0, tried = 0;
    // square roots: random mantissas, exponents from 2^-767 to 2^1023, and the
    // values next to powers of two and to squares of short integers
    for (long t = 0; t < n; t++) {
        const uint64_t r = next();
        double x;
        if (t % 4 == 3) {
            const double m = (double)(1 + (r % 4096));
            x = ulps(m * m, (int)((r >> 20) % 5) - 2);
        } else {
            x = d_of(r, 256 + (int)((r >> 52) % 1791), 0);  // biased exponents 256 .. 2046
        }
        const long double e = 1.0L / std::sqrt((long double)x);
        for (int k = 0; k < kEst; k++) {
            const double g = grhd::sqrt_refine(x, (double)(e * (1.0L + rel_err(k))));
            tried++;
            if (!same(g, std::sqrt(x))) {
                if (bad < 5) std::printf("sqrt mismatch x=%a estimate %d got %a want %a\n", x, k, g, std::sqrt(x));
                bad++;
            }
        }
    }
    std::printf("sqrt_refine: %ld arguments x %d estimates, %ld mismatches\n", n, kEst, bad);
    long badq = 0;
    // quotients: operands of exponents within +-300 of 1 (the path's lengths,
    // densities, Lorentz factors and metric components), random and with
    // extreme mantissas
    for (long t = 0; t < n; t++) {
        const uint64_t r = next(), r2 = next();
        const uint64_t ma = (t % 8 == 7) ? 0x000fffffffffffffull : r, mb = (t % 8 == 6) ? 0 : r2;
        const double a = d_of(ma, 1023 - 300 + (int)((r >> 52) % 601), (int)(r2 >> 63));
        const double b = d_of(mb, 1023 - 300 + (int)((r2 >> 52) % 601), (int)(r >> 63));
        const long double e = 1.0L / (long double)b;
        for (int k = 0; k < kEst; k++) {
            const double q = grhd::div_refine(a, b, (double)(e * (1.0L + rel_err(k))));
            tried++;
            if (!same(q, a / b)) {
                if (badq < 5) std::printf("div mismatch a=%a b=%a estimate %d got %a want %a\n", a, b, k, q, a / b);
                badq++;
            }
        }
    }
    std::printf("div_refine: %ld operand pairs x %d estimates, %ld mismatches\n", n, kEst, badq);
    std::printf("total %ld checks, %ld mismatches\n", tried, bad + badq);
    return (bad + badq) ? 1 : 0;
}
