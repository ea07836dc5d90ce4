// Reproducible float exp / sin / cos.
//
// The reference calls the OpenCL vendor's exp (bilateral weights,
// K/bilateralFilter.cl:105) and float sin/cos (Eigen::AngleAxisf in the ICP
// update, OCL/NuiKinfuOpenCLDepthTracker.cpp:486), whose last-bit results
// differ between libraries.  These versions evaluate in double with only
// correctly rounded IEEE operations (add, mul, rint, ldexp, a table of doubles) and
// round once to float, so host and device produce identical bits and the float
// result is the correctly rounded value in all but vanishingly rare cases.
#pragma once

#include <cmath>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define NUIF_HD __host__ __device__ __forceinline__
#else
#define NUIF_HD inline
#endif

namespace nuif_det {

// 2^(j/64), j = 0..63, each the double nearest the exact value (computed to 60 digits)
#define NUIF_EXP2_J64 \
    0x1.0000000000000p+0, 0x1.02c9a3e778061p+0, 0x1.059b0d3158574p+0, 0x1.0874518759bc8p+0, \
    0x1.0b5586cf9890fp+0, 0x1.0e3ec32d3d1a2p+0, 0x1.11301d0125b51p+0, 0x1.1429aaea92de0p+0, \
    0x1.172b83c7d517bp+0, 0x1.1a35beb6fcb75p+0, 0x1.1d4873168b9aap+0, 0x1.2063b88628cd6p+0, \
    0x1.2387a6e756238p+0, 0x1.26b4565e27cddp+0, 0x1.29e9df51fdee1p+0, 0x1.2d285a6e4030bp+0, \
    0x1.306fe0a31b715p+0, 0x1.33c08b26416ffp+0, 0x1.371a7373aa9cbp+0, 0x1.3a7db34e59ff7p+0, \
    0x1.3dea64c123422p+0, 0x1.4160a21f72e2ap+0, 0x1.44e086061892dp+0, 0x1.486a2b5c13cd0p+0, \
    0x1.4bfdad5362a27p+0, 0x1.4f9b2769d2ca7p+0, 0x1.5342b569d4f82p+0, 0x1.56f4736b527dap+0, \
    0x1.5ab07dd485429p+0, 0x1.5e76f15ad2148p+0, 0x1.6247eb03a5585p+0, 0x1.6623882552225p+0, \
    0x1.6a09e667f3bcdp+0, 0x1.6dfb23c651a2fp+0, 0x1.71f75e8ec5f74p+0, 0x1.75feb564267c9p+0, \
    0x1.7a11473eb0187p+0, 0x1.7e2f336cf4e62p+0, 0x1.82589994cce13p+0, 0x1.868d99b4492edp+0, \
    0x1.8ace5422aa0dbp+0, 0x1.8f1ae99157736p+0, 0x1.93737b0cdc5e5p+0, 0x1.97d829fde4e50p+0, \
    0x1.9c49182a3f090p+0, 0x1.a0c667b5de565p+0, 0x1.a5503b23e255dp+0, 0x1.a9e6b5579fdbfp+0, \
    0x1.ae89f995ad3adp+0, 0x1.b33a2b84f15fbp+0, 0x1.b7f76f2fb5e47p+0, 0x1.bcc1e904bc1d2p+0, \
    0x1.c199bdd85529cp+0, 0x1.c67f12e57d14bp+0, 0x1.cb720dcef9069p+0, 0x1.d072d4a07897cp+0, \
    0x1.d5818dcfba487p+0, 0x1.da9e603db3285p+0, 0x1.dfc97337b9b5fp+0, 0x1.e502ee78b3ff6p+0, \
    0x1.ea4afa2a490dap+0, 0x1.efa1bee615a27p+0, 0x1.f50765b6e4540p+0, 0x1.fa7c1819e90d8p+0

#if defined(__HIP_DEVICE_COMPILE__)
static __constant__ const double kExp2J64[64] = {NUIF_EXP2_J64};
#else
static const double kExp2J64[64] = {NUIF_EXP2_J64};
#endif

// e^x for float x (returns +inf / 0 outside the float range): x = (64 m + j) ln2/64 + r,
// |r| <= ln2/128, e^x = 2^m 2^(j/64) e^r with e^r to r^5 (truncation below 4e-17); the
// double result is within 3e-16 of e^x, so its rounding to float is the correctly
// rounded value except where e^x lies that close to a rounding boundary.
// T: the table kExp2J64 (or a copy of it, e.g. in LDS)
NUIF_HD float expf_det_t(float xf, const double* T) {
    const double x = (double)xf;
    if (!(x < 88.8)) return x != x ? xf : INFINITY;
    if (x < -104.0) return 0.0f;
    const double kd = rint(x * 92.332482616893658);  // 64 / ln 2
    const int k = (int)kd;
    // ln2/64 split (hi: 32 significant bits) so that kd * hi is exact for |k| < 2^21
    const double r = (x - kd * 0x1.62e42feep-7) - kd * 0x1.a39ef35793c76p-39;
    double p = 1.0 / 120.0;
    p = p * r + 1.0 / 24.0;
    p = p * r + 1.0 / 6.0;
    p = p * r + 0.5;
    p = p * r + 1.0;
    p = p * r + 1.0;
    return (float)ldexp(T[k & 63] * p, k >> 6);
}
NUIF_HD float expf_det(float xf) { return expf_det_t(xf, kExp2J64); }

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
