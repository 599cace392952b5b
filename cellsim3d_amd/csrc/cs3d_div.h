// cs3d_div.h - float quotients equal to IEEE division, from a reciprocal.
//
// The force kernels divide by a bond length R three times per bond (the spring
// term stiffness*(R-R0)/R0*dX/R, propagate.cu:311-313), by the rest length R0 once,
// by squared bond lengths in the angle force and by square roots elsewhere.  A
// correctly rounded float division costs ten vector instructions.  With
// y = RN(1/b) known, the quotient RN(a/b) takes three:
//
//   q  = RN(a * y)
//   r  = RN(b * q - a)          (fma; the value b*q - a is a float, so r is exact)
//   q' = RN(q - r * y)          (fma)
//
// This is the correction step of Markstein's division (P. Markstein, IBM J. Res.
// Dev. 34(1), 1990: with y = RN(1/b) and q within one ulp of a/b, RN(q + (a-bq) y)
// is RN(a/b)).  q = RN(a * RN(1/b)) need not be within one ulp in general, so the
// sequence is not taken on faith: cs3d_check_div compares q' with IEEE division
// for every pair of significands a, b in [1, 2) (2^46 pairs) on the device, and
// the CPU tests do the same for a sample.  Both sides of every step scale exactly
// by powers of two as long as no intermediate value leaves the normal range, so
// the check covers every a, b with |a|, b in [2^-40, 2^40]: then a/b lies within
// [2^-80, 2^80] and the residual b*q - a, a multiple of 2^(e(a) - 47), stays
// normal.  The sign of a is carried through every step (RN is symmetric), so
// negative a follow from positive a; the divisors here are all positive.  For
// a = +-0 the steps give q = +-0, r = +0 and q' = -0*y + q = q, the zero of
// a / b (evaluated as written: the device check covers it; a host compiler that
// rewrites -fma(x, y, z) as -(x y) - z loses that sign).  Outside that range (and
// for infinities and NaN) the callers divide as IEEE does (quot_ok and
// quot_num_ok are false there).
#pragma once

#include <math.h>

#if defined(__HIPCC__)
#define CS3D_DIV_FN __host__ __device__ __forceinline__
#else
#define CS3D_DIV_FN static inline
#endif

// |v| within [2^-40, 2^40]: false for 0, subnormals, infinities and NaN
CS3D_DIV_FN bool quot_ok(float v) {
    const float a = fabsf(v);
    return (a >= 0x1p-40f) & (a <= 0x1p40f);  // both compares, no short circuit
}

// a numerator of quot_rn: quot_ok(v) or v = +-0
CS3D_DIV_FN bool quot_num_ok(float v) {
    const float a = fabsf(v);
    return ((a >= 0x1p-40f) | (a == 0.0f)) & (a <= 0x1p40f);
}

// RN(a / b) for b > 0 and y = RN(1 / b), when quot_num_ok(a) and quot_ok(b)
CS3D_DIV_FN float quot_rn(float a, float b, float y) {
    const float q = a * y;
    const float r = fmaf(b, q, -a);
    return fmaf(-r, y, q);
}
