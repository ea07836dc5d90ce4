// Device helpers shared by the depth and colour trackers: the reference's
// 128-wide tree inside one wave, the FP64 6x6 Cholesky solve standing in for
// Eigen's LLT, and the float 3x3 algebra of the Euler-angle increment.
#pragma once

#include "det_math.h"
#include "nuif_internal.h"

namespace nuif {

// Lane l (l < d) receives the 32-bit value of lane l + d, without the LDS crossbar:
// d = 32 and 16 by the lane swaps between the halves of the wave and between pairs of
// 16-lane rows (the second result of v_permlane32/16_swap with both operands x holds, in
// its low half / even rows, the high half / odd rows of x), d < 16 by a DPP row shift
// to the left (lanes past the row's end keep their own value).  Every lane of the wave
// must be active.
template <int D>
__device__ __forceinline__ uint32_t lane_from_above(uint32_t x) {
    if constexpr (D == 32)
        return __builtin_amdgcn_permlane32_swap(x, x, false, false)[1];
    else if constexpr (D == 16)
        return __builtin_amdgcn_permlane16_swap(x, x, false, false)[1];
    else
        return (uint32_t)__builtin_amdgcn_update_dpp((int)x, (int)x, 0x100 + D, 0xf, 0xf, false);
}
template <int D>
__device__ __forceinline__ float lane_from_above(float v) {
    return __uint_as_float(lane_from_above<D>(__float_as_uint(v)));
}
template <int D>
__device__ __forceinline__ double lane_from_above(double v) {
    const uint64_t u = (uint64_t)__double_as_longlong(v);
    const uint32_t lo = lane_from_above<D>((uint32_t)u), hi = lane_from_above<D>((uint32_t)(u >> 32));
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// The reference tree's steps d = 32..1 (item i takes item i + d) inside one
// wave; lane 0 ends with the sum (the other lanes' results are not defined).
// Every lane of the wave must be active.
template <class T>
__device__ __forceinline__ T wave_tree64_t(T v) {
    v += lane_from_above<32>(v);
    v += lane_from_above<16>(v);
    v += lane_from_above<8>(v);
    v += lane_from_above<4>(v);
    v += lane_from_above<2>(v);
    v += lane_from_above<1>(v);
    return v;
}
__device__ __forceinline__ float wave_tree64(float v) { return wave_tree64_t(v); }

// Cholesky A = L L^T in place (lower), as Eigen's LLT; returns false if not SPD.
__device__ inline bool llt6(double A[36], double* det) {
    double d = 1.0;
    for (int j = 0; j < 6; ++j) {
        double s = A[j * 6 + j];
        for (int k = 0; k < j; ++k) s -= A[j * 6 + k] * A[j * 6 + k];
        if (!(s > 0)) {
            *det = 0.0;
            return false;
        }
        const double l = sqrt(s);
        A[j * 6 + j] = l;
        d *= l;
        for (int i = j + 1; i < 6; ++i) {
            double u = A[i * 6 + j];
            for (int k = 0; k < j; ++k) u -= A[i * 6 + k] * A[j * 6 + k];
            A[i * 6 + j] = u / l;
        }
    }
    *det = d * d;
    return true;
}

__device__ inline void llt6_solve(const double L[36], const double b[6], double x[6]) {
    double y[6];
    for (int i = 0; i < 6; ++i) {
        double s = b[i];
        for (int k = 0; k < i; ++k) s -= L[i * 6 + k] * y[k];
        y[i] = s / L[i * 6 + i];
    }
    for (int i = 5; i >= 0; --i) {
        double s = y[i];
        for (int k = i + 1; k < 6; ++k) s -= L[k * 6 + i] * x[k];
        x[i] = s / L[i * 6 + i];
    }
}

// Float 3x3 helpers mirroring Eigen's float evaluation of the increment
__device__ inline void axis_rotf(float angle, int k, float Rm[9]) {
    float s, c;
    nuif_det::sincosf_det(angle, &s, &c);
    float ax[3] = {0.f, 0.f, 0.f};
    ax[k] = 1.f;
    const float sa[3] = {s * ax[0], s * ax[1], s * ax[2]};
    const float ca[3] = {(1.f - c) * ax[0], (1.f - c) * ax[1], (1.f - c) * ax[2]};
    Rm[1] = ca[0] * ax[1] - sa[2];
    Rm[3] = ca[0] * ax[1] + sa[2];
    Rm[2] = ca[0] * ax[2] + sa[1];
    Rm[6] = ca[0] * ax[2] - sa[1];
    Rm[5] = ca[1] * ax[2] - sa[0];
    Rm[7] = ca[1] * ax[2] + sa[0];
    Rm[0] = ca[0] * ax[0] + c;
    Rm[4] = ca[1] * ax[1] + c;
    Rm[8] = ca[2] * ax[2] + c;
}

__device__ inline void matmul3f(const float a[9], const float b[9], float c[9]) {
    float r[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r[3 * i + j] = a[3 * i] * b[j] + a[3 * i + 1] * b[3 + j] + a[3 * i + 2] * b[6 + j];
    for (int k = 0; k < 9; ++k) c[k] = r[k];
}

// Eigen-style float cofactor inverse (OCL/NuiKinfuOpenCLCameraState.cpp:67-69)
__device__ inline void inverse3f(const float m[9], float out[9]) {
    const float c00 = m[4] * m[8] - m[5] * m[7];
    const float c01 = m[5] * m[6] - m[3] * m[8];
    const float c02 = m[3] * m[7] - m[4] * m[6];
    const float det = m[0] * c00 + m[1] * c01 + m[2] * c02;
    const float inv = 1.0f / det;
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

}  // namespace nuif
