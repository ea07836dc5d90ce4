// Reproducible float exp / sin / cos.
//
// The reference calls the OpenCL vendor's exp (bilateral weights,
// K/bilateralFilter.cl:105) and float sin/cos (Eigen::AngleAxisf in the ICP
// update, OCL/NuiKinfuOpenCLDepthTracker.cpp:486), whose last-bit results
// differ between libraries.  These versions evaluate in double with only
// correctly rounded IEEE operations (add, mul, rint, ldexp) and round once to
// float, so host and device produce identical bits and the float result is
// the correctly rounded value in all but astronomically rare cases.
#pragma once

#include <cmath>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define NUIF_HD __host__ __device__ __forceinline__
#else
#define NUIF_HD inline
#endif

namespace nuif_det {

// e^x for float x (returns +inf / 0 outside the float range)
NUIF_HD float expf_det(float xf) {
    const double x = (double)xf;
    if (!(x < 88.8)) return x != x ? xf : INFINITY;
    if (x < -104.0) return 0.0f;
    const double k = rint(x * 1.4426950408889634074);
    // ln2 split (hi: 40 significant bits) so that k * ln2_hi is exact for |k| < 2^13
    const double r = (x - k * 0.6931471805592082) - k * 7.371002565167799e-13;
    // Taylor series to r^12 on |r| <= 0.347: truncation below 1e-17
    double p = 1.0 / 479001600.0;
    p = p * r + 1.0 / 39916800.0;
    p = p * r + 1.0 / 3628800.0;
    p = p * r + 1.0 / 362880.0;
    p = p * r + 1.0 / 40320.0;
    p = p * r + 1.0 / 5040.0;
    p = p * r + 1.0 / 720.0;
    p = p * r + 1.0 / 120.0;
    p = p * r + 1.0 / 24.0;
    p = p * r + 1.0 / 6.0;
    p = p * r + 0.5;
    p = p * r + 1.0;
    p = p * r + 1.0;
    return (float)ldexp(p, (int)k);
}

// sin and cos of a float angle: quadrant reduction by pi/2 in double, then
// Taylor polynomials on |r| <= pi/4 (truncation below 1e-17)
NUIF_HD void sincosf_det(float af, float* s_out, float* c_out) {
    const double a = (double)af;
    const double k = rint(a * 0.63661977236758134308);
    // pi/2 split (hi: 33 significant bits) so that k * hi is exact for |k| < 2^20
    const double r = (a - k * 1.5707963267341256) - k * 6.077100506506192e-11;
    const double r2 = r * r;
    double s = -1.0 / 1307674368000.0;
    s = s * r2 + 1.0 / 6227020800.0;
    s = s * r2 - 1.0 / 39916800.0;
    s = s * r2 + 1.0 / 362880.0;
    s = s * r2 - 1.0 / 5040.0;
    s = s * r2 + 1.0 / 120.0;
    s = s * r2 - 1.0 / 6.0;
    s = s * r2 * r + r;
    double c = 1.0 / 20922789888000.0;
    c = c * r2 - 1.0 / 87178291200.0;
    c = c * r2 + 1.0 / 479001600.0;
    c = c * r2 - 1.0 / 3628800.0;
    c = c * r2 + 1.0 / 40320.0;
    c = c * r2 - 1.0 / 720.0;
    c = c * r2 + 1.0 / 24.0;
    c = c * r2 - 0.5;
    c = c * r2 + 1.0;
    const long long q = ((long long)k) & 3;
    double sv, cv;
    if (q == 0) { sv = s; cv = c; }
    else if (q == 1) { sv = c; cv = -s; }
    else if (q == 2) { sv = -s; cv = -c; }
    else { sv = -c; cv = s; }
    *s_out = (float)sv;
    *c_out = (float)cv;
}

}  // namespace nuif_det
