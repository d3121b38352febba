// blink::div_rn (the gridding plans' FP64 division from a reciprocal) against
// IEEE division, on random operands: wavelengths of 50 MHz .. 2 GHz channels
// with FP32 metre coordinates, cell sizes, random exponents, and divisors
// whose significand is all ones.  Every quotient must be identical.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>

#include "../../imager_amd/csrc/blink_divrn.hpp"

static uint64_t g_state = 88172645463325252ull;
static uint64_t next_u64() {
    g_state ^= g_state << 13;
    g_state ^= g_state >> 7;
    g_state ^= g_state << 17;
    return g_state;
}
static double uniform() { return (double)(next_u64() >> 11) * (1.0 / 9007199254740992.0); }
// random sign, random 52-bit significand, exponent uniform in [lo, hi)
static double random_double(int lo, int hi) {
    const uint64_t m = next_u64() & ((1ull << 52) - 1);
    const int e = lo + (int)((hi - lo) * uniform());
    const uint64_t bits = ((uint64_t)(e + 1023) << 52) | m;
    double d;
    std::memcpy(&d, &bits, 8);
    return (next_u64() & 1) ? -d : d;
}

int main() {
    long tested = 0, bad = 0;
    for (long i = 0; i < 20000000L; i++) {
        double a, b;
        switch (i % 4) {
            case 0: {  // FP32 metres over a channel's wavelength
                a = (double)(float)random_double(-20, 12);
                b = 299792458.0 / (1e6 * (50.0 + 1950.0 * uniform()));
                break;
            }
            case 1:
                a = random_double(-99, 99);
                b = random_double(-99, 99);
                break;
            case 2:  // wavelengths over cell sizes
                a = random_double(-60, 60);
                b = std::fabs(random_double(-10, 10));
                break;
            default: {  // significand of b all ones in its top bits
                a = random_double(-5, 5);
                const uint64_t bits = ((uint64_t)1023 << 52) | (next_u64() & ((1ull << 52) - 1)) |
                                      0xFFFFFFFFFF000ull;
                std::memcpy(&b, &bits, 8);
            }
        }
        if (!blink::div_rn_range(a) || !blink::div_rn_range(b)) continue;
        tested++;
        const double q = blink::div_rn(a, b, 1.0 / b);
        if (q != a / b) {
            if (bad < 10) std::printf("mismatch: %.17g / %.17g = %.17g, div_rn %.17g\n", a, b, a / b, q);
            bad++;
        }
    }
    std::printf("%ld quotients, %ld mismatches\n", tested, bad);
    if (bad) return 1;
    std::printf("All tests passed.\n");
    return 0;
}
