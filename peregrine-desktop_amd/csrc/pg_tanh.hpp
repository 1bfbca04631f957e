// tanh and the traction gain, one definition for host and device.
//
// traction_gain (PeregrineSolver.jl:126-136) needs tanh of a float. Host
// libraries and device libraries round tanhf differently, so the gain of a
// command evaluated on the device (pattern-search and warm-start commands)
// would not equal the host's gain of the same command. This is one
// deterministic tanh built from IEEE double +, -, *, / only: exp by argument
// reduction (t = k ln2 + r, |r| <= ln2/2, ln2 in two parts) and its Taylor
// series, tanh(x) = expm1(2x) / (expm1(2x) + 2), expm1 of small arguments
// from its own series. Relative error below 1e-15 on [0, 30]; rounded to
// float it equals the correctly rounded tanhf for every float in
// [1e-6, 30]. Compiled without contraction, host and device agree bitwise.
#pragma once

#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdint>

namespace pgt {

__host__ __device__ inline double tanh_pos(double x) {
    const double t = 2.0 * (x < 20.0 ? x : 20.0);
    double em;  // expm1(t)
    if (t < 0.5) {
        double p = 1.0 / 355687428096000.0;  // 1/17!
        p = 1.0 / 20922789888000.0 + t * p;
        p = 1.0 / 1307674368000.0 + t * p;
        p = 1.0 / 87178291200.0 + t * p;
        p = 1.0 / 6227020800.0 + t * p;
        p = 1.0 / 479001600.0 + t * p;
        p = 1.0 / 39916800.0 + t * p;
        p = 1.0 / 3628800.0 + t * p;
        p = 1.0 / 362880.0 + t * p;
        p = 1.0 / 40320.0 + t * p;
        p = 1.0 / 5040.0 + t * p;
        p = 1.0 / 720.0 + t * p;
        p = 1.0 / 120.0 + t * p;
        p = 1.0 / 24.0 + t * p;
        p = 1.0 / 6.0 + t * p;
        p = 0.5 + t * p;
        p = 1.0 + t * p;
        em = t * p;
    } else {
        const double k = rint(t * 1.4426950408889634);
        const double r = (t - k * 6.93147180369123816490e-01) - k * 1.90821492927058770002e-10;
        double p = 1.0 / 87178291200.0;  // 1/14!
        p = 1.0 / 6227020800.0 + r * p;
        p = 1.0 / 479001600.0 + r * p;
        p = 1.0 / 39916800.0 + r * p;
        p = 1.0 / 3628800.0 + r * p;
        p = 1.0 / 362880.0 + r * p;
        p = 1.0 / 40320.0 + r * p;
        p = 1.0 / 5040.0 + r * p;
        p = 1.0 / 720.0 + r * p;
        p = 1.0 / 120.0 + r * p;
        p = 1.0 / 24.0 + r * p;
        p = 1.0 / 6.0 + r * p;
        p = 0.5 + r * p;
        p = 1.0 + r * p;
        p = 1.0 + r * p;
        em = ldexp(p, (int)k) - 1.0;
    }
    return em / (em + 2.0);
}

// traction_gain (PeregrineSolver.jl:126-136) in float: tanh(r)/r at
// r = |u| / knee, 1 when the knee is off or r < 1e-6.
__host__ __device__ inline float traction_gain_f(float u1, float u2, float u3, float knee) {
    if (knee != knee || knee <= 0.0f) return 1.0f;
    const float m = sqrtf(u1 * u1 + u2 * u2 + u3 * u3);
    const float r = m / knee;
    if (r < 1.0e-6f) return 1.0f;
    return (float)tanh_pos((double)r) / r;
}

}  // namespace pgt
