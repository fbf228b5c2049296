// Checks grhd::quot (grhd_amd/csrc/grhd_physics.hpp): a / b from rb = 1 / b in
// one multiply and two FMAs must equal the IEEE quotient bit for bit.
//
// 1. quot<double> itself, from the product header, on random dividends and
//    divisors (the constants 60 and 3 of MP5, and random ones), on dividends
//    with extreme mantissas, and on quotients that land next to powers of two.
// 2. The same three operations at binary32, where every dividend mantissa of a
//    binade can be tried: the theorem behind them holds at any precision, so
//    this checks the argument itself, exhaustively, for fixed divisors.
// Prints one line per part and exits non-zero on any mismatch.
#include "grhd_physics.hpp"

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>

static uint64_t state = 0x9e3779b97f4a7c15ull;
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

static float quot32(float a, float b, float rb)
{
    const float q = a * rb;
    return std::fma(std::fma(-q, b, a), rb, q);
}

int main(int argc, char **argv)
{
    const long n = argc > 1 ? std::atol(argv[1]) : 20000000L;
    long bad = 0, tot = 0;
    const double fixed[] = {60.0, 3.0, 7.0, 0.1, 1.0 / 3.0};
    for (long k = 0; k < n; k++) {
        const double a = d_of(next(), 700 + (int)(next() % 640), (int)(next() & 1));
        double b = d_of(next(), 700 + (int)(next() % 640), (int)(next() & 1));
        if (k % 3 == 0) b = fixed[k % 5];
        const double rb = 1.0 / b;
        tot++;
        if (!same(grhd::quot(a, b, rb), a / b)) {
            if (bad++ < 5) std::printf("mismatch %a / %a\n", a, b);
        }
    }
    // dividends with all-zero and all-one mantissas; quotients around 2^k
    for (int e = 1000; e < 1050; e++)
        for (uint64_t m : {0ull, 1ull, 0x000fffffffffffffull, 0x0008000000000000ull, 0x000ffffffffffffeull}) {
            const double a = d_of(m, e, 0);
            for (double b : fixed) {
                tot++;
                if (!same(grhd::quot(a, b, 1.0 / b), a / b)) bad++;
                const double c = std::ldexp(b, e - 1023);   // a / c within an ulp of 2^k
                tot++;
                if (!same(grhd::quot(a, c, 1.0 / c), a / c)) bad++;
            }
        }
    std::printf("quot<double>: %ld mismatches of %ld\n", bad, tot);
    long bad32 = 0, tot32 = 0;
    const float fixed32[] = {60.0f, 3.0f, 7.0f, 0.1f};
    for (float b : fixed32) {
        const float rb = 1.0f / b;
        for (uint32_t m = 0; m < (1u << 23); m++) {
            const uint32_t u = 0x3f800000u | m;
            float a;
            std::memcpy(&a, &u, 4);
            const float x = a / b, y = quot32(a, b, rb);
            tot32++;
            if (std::memcmp(&x, &y, 4) != 0) bad32++;
        }
    }
    std::printf("binary32, every dividend mantissa in [1, 2): %ld mismatches of %ld\n", bad32, tot32);
    return (bad || bad32) ? 1 : 0;
}
