// blink::div_rn (the gridding plans' FP64 division from a reciprocal) against
// IEEE division, on random operands: wavelengths of 50 MHz .. 2 GHz channels
// with FP32 metre coordinates, cell sizes, random exponents, and divisors
// whose significand is all ones.  Every quotient must be identical.
// Then blink::round_quotient_fast (a cell index from one product) against
// round((a / w) / d) with IEEE divisions: FP32 metre coordinates, channel
// wavelengths and cell sizes, and coordinates placed on and next to the
// half-integer quotients, where it must decline.
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

    // one-product cell rounding: decided cells equal the two-division cells
    // kind 0: grid coordinates, 1: random exponents, 2: exact halves, 3: next to them
    long cells[4] = {0, 0, 0, 0}, decided[4] = {0, 0, 0, 0}, wrong = 0;
    auto check = [&](double a, double w, double d, int kind) {
        if (!blink::div_rn_range(w) || !blink::div_rn_range(d)) return;
        const double k = 1.0 / (w * d);
        const double r = std::round((a / w) / d);
        int out = 0;
        cells[kind]++;
        if (!blink::round_quotient_fast(a, k, out)) return;
        decided[kind]++;
        if (!(std::fabs(r) < 536870912.0) || (double)out != r) {
            if (wrong < 10)
                std::printf("cell mismatch: a %.9g w %.17g d %.17g: %.17g, fast %d\n", a, w, d,
                            (a / w) / d, out);
            wrong++;
        }
    };
    for (long i = 0; i < 20000000L; i++) {
        const double w = 299792458.0 / (1e6 * (50.0 + 1950.0 * uniform()));
        switch (i % 4) {
            case 0: {  // metres on a grid of cell size d (wavelengths)
                const double d = std::ldexp(1.0 + uniform(), -4 + (int)(8 * uniform()));
                const float a = (float)((uniform() * 2.0 - 1.0) * 3000.0);
                check((double)a, w, d, 0);
                break;
            }
            case 1: {  // random exponents
                const float a = (float)random_double(-30, 30);
                check((double)a, std::fabs(random_double(-20, 20)),
                      std::fabs(random_double(-20, 20)), 1);
                break;
            }
            case 2: {  // on a half-integer quotient: w, d powers of two, a = (m + 1/2) w d
                const double wp = std::ldexp(1.0, -3 + (int)(6 * uniform()));
                const double dp = std::ldexp(1.0, -3 + (int)(6 * uniform()));
                const double m = std::floor((uniform() * 2.0 - 1.0) * 4096.0);
                const double a = (m + 0.5) * wp * dp;
                if ((double)(float)a == a) check(a, wp, dp, 2);
                break;
            }
            default: {  // one FP32 ulp either side of a half-integer quotient
                const double d = std::ldexp(1.0 + uniform(), -4 + (int)(8 * uniform()));
                const double m = std::floor((uniform() * 2.0 - 1.0) * 2048.0);
                const float a0 = (float)((m + 0.5) * w * d);
                check((double)std::nextafter(a0, 1e30f), w, d, 3);
                check((double)std::nextafter(a0, -1e30f), w, d, 3);
                check((double)a0, w, d, 3);
            }
        }
    }
    const char* kinds[4] = {"grid coordinates", "random exponents", "exact halves",
                            "next to halves"};
    for (int k = 0; k < 4; k++)
        std::printf("%s: %ld cells, %ld decided by one product\n", kinds[k], cells[k], decided[k]);
    std::printf("%ld mismatches\n", wrong);
    // no mismatch; exact halves always declined; ordinary coordinates almost never
    if (wrong || decided[2] || cells[2] == 0 || cells[3] == 0 ||
        decided[0] < cells[0] - cells[0] / 10000)
        return 1;
    std::printf("All tests passed.\n");
    return 0;
}
