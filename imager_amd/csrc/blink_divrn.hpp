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

// |x| in (1e-30, 1e30): no quotient, product or remainder of two such values
// (or of a quotient of two and a third) overflows or underflows.
BLINK_HOST_DEVICE inline bool div_rn_range(double x) {
    return std::fabs(x) > 1e-30 && std::fabs(x) < 1e30;
}

}  // namespace blink
