/*
 * The library's own e^x for device code, shared by the event loop (damped
 * flights, logTime) and the derivative kernels (damped tangent flights),
 * so that both round alike and the oracle's restatement of the same
 * steps gives the same bits.
 */
#ifndef EDMD_MATH_H
#define EDMD_MATH_H

#include <hip/hip_runtime.h>

/* e^x from its Taylor series: x = k ln2 + r, |r| <= ln2/2, ln2 in two
 * parts so that k ln2 is exact (Cody and Waite's reduction), e^r by the
 * nested form 1 + r(1 + r/2(1 + r/3(... (1 + r/17)))) (truncation below
 * 2e-24 relative), times 2^k; within 4 ulp.  Only +, *, rint and ldexp,
 * so the oracle's restatement of the same steps gives the same bits. */
__device__ __forceinline__ double edmd_exp(double x) {
    if (x != x) return x;
    if (x > 709.5) return __builtin_huge_val();
    if (x < -745.5) return 0.0;
    const double k = rint(x * 1.4426950408889634); /* x / ln2, to nearest */
    const double r = (x - k * 0x1.62e42fefa3800p-1) - k * 0x1.ef35793c76730p-45;
    double p = 1.0;
#pragma unroll
    for (int n = 17; n >= 1; --n) p = 1.0 + (r * p) * (1.0 / n);
    return ldexp(p, (int)k);
}

#endif
