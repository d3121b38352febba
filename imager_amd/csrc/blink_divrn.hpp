// Correctly rounded FP64 division from a precomputed reciprocal, for the
// gridding plans' cell indices (host and device, so the host tests run the
// same code).
#pragma once

#include <cmath>

#if defined(__HIPCC__)
#define BLINK_HOST_DEVICE __host__ __device__
#else
#define BLINK_HOST_DEVICE
#endif

namespace blink {

// a / b rounded to nearest from y = RN(1 / b), for operands in div_rn_range:
// q0 = RN(a y) is within 2 ulp of a / b; one correction with the exact FMA
// remainder r0 = a - b q0 brings q1 within 1 ulp, and the second, RN(q1 + r1 y),
// is the correctly rounded quotient (Markstein's theorem on the final
// correction step of an FMA division).  Five FP64 operations instead of the
// full IEEE division sequence per quotient.
BLINK_HOST_DEVICE inline double div_rn(double a, double b, double y) {
    const double q0 = a * y;
    const double q1 = std::fma(std::fma(-b, q0, a), y, q0);
    return std::fma(std::fma(-b, q1, a), y, q1);
}

// Nearest integer to q = RN(RN(a / w) / d) (halves away from zero) from the
// one product p = RN(a k), k = RN(1 / RN(w d)), when p decides it.  For w, d
// in div_rn_range and a an FP32 value no step is subnormal, so p and q are
// each within a few rounding errors of a / (w d): |p - q| <= 5.01 * 2^-53
// |a / (w d)|, under 3e-7 for |p| < 2^29.  When p lies more than 1e-6 from
// every half-integer, p and q are therefore on the same side of each
// half-integer and round to the same integer.  Returns false, leaving `out`,
// when p is within 1e-6 of a half-integer, not finite, or |p| >= 2^29 (about
// one cell in a million): the caller then takes the two divisions.
BLINK_HOST_DEVICE inline bool round_quotient_fast(double a, double k, int& out) {
    const double p = a * k;
    if (!(std::fabs(p) < 536870912.0)) return false;
    const double f = std::floor(p);
    const double t = p - f;  // in [0, 1]; exact but for -0.5 < p < 0 (error < 2^-54)
    if (std::fabs(t - 0.5) <= 1e-6) return false;
    out = (int)f + (t > 0.5 ? 1 : 0);
    return true;
}

// |x| in (1e-30, 1e30): no quotient, product or remainder of two such values
// (or of a quotient of two and a third) overflows or underflows.
BLINK_HOST_DEVICE inline bool div_rn_range(double x) {
    return std::fabs(x) > 1e-30 && std::fabs(x) < 1e30;
}

}  // namespace blink
