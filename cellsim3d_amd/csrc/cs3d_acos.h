// cs3d_acos.h - the double acos of the angle force, shared by the device kernels
// and the CPU oracle.
//
// GetAngleForce (propagate.cu:36) takes acos of a double argument and stores the
// result into a float.  Two accurate acos implementations (a device math
// library's, a host C library's) give different doubles for some arguments, each
// within about an ulp of the true value; where acos(x) lies that close to the
// midpoint of two floats, the float angles then differ by one float ulp.  The
// kernels and the oracle both call this function instead: the same IEEE-rounded
// operations (+, -, *, fma, sqrt) in the same order on both sides, so the float
// angle is the same on both.
//
// Method, written for this project from the series of asin:
//   asin(s) = s + s z f(z),  z = s^2,  f(z) = sum_{k>=1} (2k)!/(4^k (k!)^2 (2k+1)) z^(k-1)
// with f replaced by its degree-12 Chebyshev interpolant on [0, 1/4], computed in
// 60-digit arithmetic by gen_acos_coeffs.py (relative error below 2^-52.9 with the
// coefficients rounded to double).  Then
//   |x| <= 1/2:  acos(x) = pi/2 - asin(x)              (z = x^2)
//   x > 1/2:     acos(x) = 2 asin(sqrt((1 - x)/2))      (z = (1 - |x|)/2, exact)
//   x < -1/2:    acos(x) = pi - 2 asin(sqrt((1 + x)/2))
// with pi/2 as a double-double constant.  Largest error found against long
// double acosl over 2e7 arguments (uniform on [-1, 1] and dense near -1/2, -0.3
// and +-1): 1.11 ulp.  acos(+-1) = 0 and pi (rounded), acos(+-0) = pi/2 (rounded),
// NaN for |x| > 1 and for NaN.
#pragma once

#include <math.h>

#if defined(__HIPCC__)
#define CS3D_ACOS_FN __host__ __device__ __forceinline__
#else
#define CS3D_ACOS_FN static inline
#endif

CS3D_ACOS_FN double cs3d_acos(double x) {
    const double ax = fabs(x);
    const bool big = ax > 0.5;
    const double z = big ? fma(-0.5, ax, 0.5) : x * x;  // (1 - |x|)/2 is exact
    // f(z), coefficients as printed by gen_acos_coeffs.py (constant term last here)
    double f = 0x1.d72b2bc8155f8p-6;
    f = fma(f, z, -0x1.e6aaa8a0a04ccp-7);
    f = fma(f, z, 0x1.1d189408314eep-6);
    f = fma(f, z, 0x1.65a9c4dfcf8b2p-8);
    f = fma(f, z, 0x1.52420b04b37bep-7);
    f = fma(f, z, 0x1.782651caa6547p-7);
    f = fma(f, z, 0x1.c9cf07674736ap-7);
    f = fma(f, z, 0x1.1c4d35cf95421p-6);
    f = fma(f, z, 0x1.6e8bb1c8209a2p-6);
    f = fma(f, z, 0x1.f1c71c1db0623p-6);
    f = fma(f, z, 0x1.6db6db6e31f13p-5);
    f = fma(f, z, 0x1.3333333332ecap-4);
    f = fma(f, z, 0x1.5555555555556p-3);
    const double s = big ? sqrt(z) : x;
    const double r = fma(s * z, f, s);  // asin(s)
    // acos = m pi/2 - (k r - m lo): m = 1, k = 1 for |x| <= 1/2; m = 2, k = 2 for
    // x < -1/2; m = 0, k = -2 for x > 1/2 (pi/2 = hi + lo; every product by m or k
    // is exact, so x > 1/2 gives 2 r exactly)
    const int mi = big ? (x > 0.0 ? 0 : 2) : 1;
    const int ki = big ? (x > 0.0 ? -2 : 2) : 1;
    const double m = (double)mi, k = (double)ki;
    return fma(m, 0x1.921fb54442d18p+0, -fma(k, r, -m * 0x1.1a62633145c07p-54));
}
