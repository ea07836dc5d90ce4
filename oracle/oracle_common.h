// TEST INFRASTRUCTURE ONLY.  CPU restatement ("the oracle") of the reference
// fusion path, used by tests/, __graft_entry__.smoke() and bench.py's CPU
// baseline as the checker.  Nothing in the product library links, loads or
// calls this code.
//
// Every function is templated on the arithmetic type T:
//   T = float   "compat": the reference's float arithmetic and its quirks,
//               operation by operation, compiled without FP contraction;
//   T = double  the FP64 oracle: the same algorithm evaluated in double.
#pragma once

#include <cmath>
#include <cstdlib>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>
#include <algorithm>

namespace oracle {

template <class T> inline T qnan() { return std::numeric_limits<T>::quiet_NaN(); }

template <class T> struct V3 {
    T x, y, z;
};
template <class T> inline V3<T> mk(T x, T y, T z) { return V3<T>{x, y, z}; }
template <class T> inline V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> inline V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> inline V3<T> operator*(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
// OpenCL dot(): left-to-right sum of products
template <class T> inline T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> inline V3<T> cross(V3<T> a, V3<T> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class T> inline bool isnan3(V3<T> v) { return std::isnan(v.x) || std::isnan(v.y) || std::isnan(v.z); }
template <class T> inline V3<T> load3(const T* p, size_t i) { return {p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }
template <class T> inline void store3(T* p, size_t i, V3<T> v) {
    p[3 * i] = v.x;
    p[3 * i + 1] = v.y;
    p[3 * i + 2] = v.z;
}
// fast_normalize()/normalize(): zero vector is returned unchanged (OpenCL 1.2 spec 6.12.5)
template <class T> inline V3<T> normalize3(V3<T> v) {
    T d = dot(v, v);
    if (v.x == 0 && v.y == 0 && v.z == 0) return v;
    T inv = T(1) / std::sqrt(d);
    return v * inv;
}
// OpenCL round(): half away from zero
template <class T> inline T cl_round(T v) { return std::round(v); }
// float -> int conversion of the reference's `(int)` / convert_int: round toward zero;
// out-of-range and NaN map to INT_MIN here (never inside the image).
template <class T> inline int to_int(T v) {
    if (!(v > T(-2147483648.0) && v < T(2147483647.0))) return INT32_MIN;
    return (int)v;
}

// IEEE binary16 <-> binary32.  f16_to_float is exact; f16_from_float rounds to
// nearest, ties to even (as the GPU's __float2half_rn), overflow to infinity.
inline float f16_to_float(uint16_t h) {
    const uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
    const uint32_t e = (h >> 10) & 0x1fu, m = h & 0x3ffu;
    uint32_t bits;
    if (e == 0) {
        if (m == 0) {
            bits = sign;
        } else {
            const float v = (float)m * 0x1p-24f;
            std::memcpy(&bits, &v, 4);
            bits |= sign;
        }
    } else if (e == 31) {
        bits = sign | 0x7f800000u | (m << 13);
    } else {
        bits = sign | ((e + 112u) << 23) | (m << 13);
    }
    float f;
    std::memcpy(&f, &bits, 4);
    return f;
}
inline uint16_t f16_from_float(float f) {
    uint32_t x;
    std::memcpy(&x, &f, 4);
    const uint16_t sign = (uint16_t)((x >> 16) & 0x8000u);
    const uint32_t ax = x & 0x7fffffffu;
    if (ax > 0x7f800000u) return (uint16_t)(sign | 0x7e00u | ((ax >> 13) & 0x1ffu));  // NaN (quiet)
    if (ax >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u);  // rounds to >= 65520: infinity
    if (ax < 0x38800000u) {  // below the smallest normal half (2^-14): subnormal or zero
        // value = ax as float; half subnormal unit 2^-24: q = round_even(value * 2^24)
        float v;
        std::memcpy(&v, &ax, 4);
        const float scaled = v * 0x1p24f;  // exact (power of two)
        float q = std::nearbyint(scaled);  // default rounding mode: to nearest even
        return (uint16_t)(sign | (uint16_t)q);
    }
    // normal: keep 10 mantissa bits, round to nearest even on the 13 dropped bits
    const uint32_t e = (ax >> 23) - 112u;  // half exponent
    uint32_t m = ax & 0x7fffffu;
    uint32_t h = (e << 10) | (m >> 13);
    const uint32_t rem = m & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h += 1u;  // carry may move into the exponent
    return (uint16_t)(sign | h);
}

// K/gpu_def.h:13-27 (NuiCLCameraParams), 29-34 (NuiCLRigidTransform), 5-11 (TsdfParams)
template <class T> struct Cam {
    T fx, fy, fx_inv, fy_inv, cx, cy;
    int W, H;
    T dmin, dmax;
};
template <class T> struct Rigid {
    T R[9], R_inv[9], t[3];
};
template <class T> struct Tsdf {
    int res[3];
    T dim[3];
    T cell[3];
    T trunc;
};

// Eigen-style 3x3 cofactor inverse (row-major), used for R_inv
// (OCL/NuiKinfuOpenCLCameraState.cpp:67-69 `rot.inverse()`).
template <class T> inline void inverse3(const T m[9], T out[9]) {
    T c00 = m[4] * m[8] - m[5] * m[7];
    T c01 = m[5] * m[6] - m[3] * m[8];
    T c02 = m[3] * m[7] - m[4] * m[6];
    T det = m[0] * c00 + m[1] * c01 + m[2] * c02;
    T inv = T(1) / det;
    out[0] = c00 * inv;
    out[3] = c01 * inv;
    out[6] = c02 * inv;
    out[1] = (m[2] * m[7] - m[1] * m[8]) * inv;
    out[4] = (m[0] * m[8] - m[2] * m[6]) * inv;
    out[7] = (m[1] * m[6] - m[0] * m[7]) * inv;
    out[2] = (m[1] * m[5] - m[2] * m[4]) * inv;
    out[5] = (m[2] * m[3] - m[0] * m[5]) * inv;
    out[8] = (m[0] * m[4] - m[1] * m[3]) * inv;
}
template <class T> inline void matmul3(const T a[9], const T b[9], T c[9]) {
    T r[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r[3 * i + j] = a[3 * i] * b[j] + a[3 * i + 1] * b[3 + j] + a[3 * i + 2] * b[6 + j];
    std::memcpy(c, r, sizeof(r));
}
template <class T> inline void matvec3(const T a[9], const T v[3], T out[3]) {
    T r[3];
    for (int i = 0; i < 3; ++i) r[i] = a[3 * i] * v[0] + a[3 * i + 1] * v[1] + a[3 * i + 2] * v[2];
    std::memcpy(out, r, sizeof(r));
}
template <class T> inline Rigid<T> make_rigid(const T R[9], const T t[3]) {
    Rigid<T> g;
    std::memcpy(g.R, R, sizeof(g.R));
    inverse3(R, g.R_inv);
    std::memcpy(g.t, t, sizeof(g.t));
    return g;
}
// K/utils.cl:48-60 transform() / rotation()
template <class T> inline V3<T> rotate(const T R[9], V3<T> v) {
    return {dot(mk(R[0], R[1], R[2]), v), dot(mk(R[3], R[4], R[5]), v), dot(mk(R[6], R[7], R[8]), v)};
}
template <class T> inline V3<T> transform(const Rigid<T>& g, V3<T> v) { return rotate(g.R, v) + mk(g.t[0], g.t[1], g.t[2]); }
// K/utils.cl:62-67 transformInverse()
template <class T> inline V3<T> transform_inverse(const Rigid<T>& g, V3<T> v) {
    v = v - mk(g.t[0], g.t[1], g.t[2]);
    return rotate(g.R_inv, v);
}

}  // namespace oracle
