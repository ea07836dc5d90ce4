// Photometric (colour) tracker: OCL/NuiKinfuOpenCLColorTracker.cpp with the
// kernels of K/colorIcp.cl and K/transformMaps.cl:405-446,513-563.
//
//   color_half_sample_kernel   BGRA pyramid: integer mean of the window, alpha 255
//   color_gradient_kernel      Sobel / 8 per channel into uint8 (or int8) maps
//   color_difference_kernel    exact integer sums of |obs - known|^2 and the count
//                              (wave sums + one 64-bit atomic per wave; the
//                              reference adds floats atomically in any order)
//   color_icp_kernel           one launch per iteration, as the depth tracker's
//                              reference-order kernel: 128-item blocks with the
//                              reference's tree, the last block of every 128
//                              runs compute_color_sums' tree, the last group sums
//                              the groups in order as the host does and runs the
//                              host's update on one thread.
//
// The host update has two decodes of the 29 sums (nuif_color_icp_begin):
//   reference: OCL/NuiKinfuOpenCLColorTracker.cpp:469-526 as written -- the sums
//     read in the depth tracker's upper-row order while the kernel writes them
//     in lower-row order with b interleaved, result(0..2) taken as Euler angles
//     while the kernel's parameters 0..2 are translations, the camera pose
//     applied to world points, and (rotation only) the count read from a slot
//     the kernel leaves at zero;
//   corrected: the sums decoded in the kernel's order, the world -> camera
//     transform applied to the model points, signed gradients, and the
//     Gauss-Newton step -A^-1 b composed on that transform.
// A normal matrix that is not positive definite fails the track in both (the
// reference's Eigen LLT would return an unspecified solution).

#include "det_math.h"
#include "icp_math.h"
#include "nuif_internal.h"

namespace nuif {

constexpr int kCTerms = 29;        // COLOR_CORESPS_NUM (K/colorIcp.cl:3)
constexpr int kCMaxGroups2 = 256;  // 128-block groups per launch (W*H <= 2^22)

struct ColorIcpState {
    float R[9], t[3];  // the transform applied to the model points
    float error, count;
    int32_t brk[NUIF_ICP_MAX_LEVELS];
    int32_t failed, iterations, corrected, pad;
    float sums[kCTerms];  // sums of the last executed iteration (kernel slot order)
    float pad2[3];
};
static_assert(sizeof(ColorIcpState) % 4 == 0, "state is copied as 32-bit words");

struct ColorIcpSync {
    uint32_t arrive[kCMaxGroups2];
    uint32_t arrive_all;
    uint32_t pad[63];
};
constexpr size_t kCStateBytes = (sizeof(ColorIcpState) + 255) / 256 * 256;
constexpr size_t kCSyncBytes = (sizeof(ColorIcpSync) + 255) / 256 * 256;

__device__ __forceinline__ ColorIcpSync* csync_of(ColorIcpState* st) {
    return reinterpret_cast<ColorIcpSync*>(reinterpret_cast<char*>(st) + kCStateBytes);
}

// float -> int as the GPU converts: round toward zero, saturating, NaN -> 0
__device__ __forceinline__ int cvt_i32_sat(float v) {
    if (v != v) return 0;
    if (v >= 2147483648.0f) return INT32_MAX;
    if (v <= -2147483648.0f) return INT32_MIN;
    return (int)v;
}

// ---- pyramid and gradients ------------------------------------------------------

// K/transformMaps.cl:405-446 half_sample_bgra_kernel
__global__ void color_half_sample_kernel(const uchar4* __restrict__ src, uchar4* __restrict__ dst, int Wd, int Hd,
                                         int radius) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= Wd || y >= Hd) return;
    const int Ws = Wd * 2, Hs = Hd * 2, sx = 2 * x, sy = 2 * y;
    const int x0 = max(sx - radius, 0), x1 = min(sx + radius + 1, Ws);
    const int y0 = max(sy - radius, 0), y1 = min(sy + radius + 1, Hs);
    int b = 0, g = 0, r = 0, n = 0;
    for (int cy = y0; cy < y1; ++cy)
        for (int cx = x0; cx < x1; ++cx) {
            const uchar4 p = src[(size_t)cy * Ws + cx];
            b += p.x;
            g += p.y;
            r += p.z;
            ++n;
        }
    dst[(size_t)y * Wd + x] = n > 0 ? make_uchar4(b / n, g / n, r / n, 255) : make_uchar4(0, 0, 0, 255);
}

__device__ __forceinline__ unsigned char grad_u8(float v, int sign) {
    if (sign) {
        const int i = v >= 127.0f ? 127 : v <= -128.0f ? -128 : (int)v;
        return (unsigned char)(signed char)i;
    }
    return !(v > 0.0f) ? 0 : v >= 255.0f ? 255 : (unsigned char)v;
}

// K/transformMaps.cl:513-563 bgra_gradient_kernel: neighbours one row (the width)
// apart, result at the centre pixel, border pixels zero
__global__ void color_gradient_kernel(const uchar4* __restrict__ c, uchar4* __restrict__ gx, uchar4* __restrict__ gy,
                                      int W, int H, int sign) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= W || y >= H) return;
    const size_t i = (size_t)y * W + x;
    if (!(x > 0 && x < W - 1 && y > 0 && y < H - 1)) {
        gx[i] = make_uchar4(0, 0, 0, 0);
        gy[i] = make_uchar4(0, 0, 0, 0);
        return;
    }
    const uchar4 s11 = c[i - W - 1], s12 = c[i - W], s13 = c[i - W + 1];
    const uchar4 s21 = c[i - 1], s23 = c[i + 1];
    const uchar4 s31 = c[i + W - 1], s32 = c[i + W], s33 = c[i + W + 1];
    auto sob = [](float a13, float a11, float a23, float a21, float a33, float a31) {
        return ((a13 - a11) + 2 * (a23 - a21) + (a33 - a31)) / 8.0f;
    };
    const float xb = sob(s13.x, s11.x, s23.x, s21.x, s33.x, s31.x);
    const float xg = sob(s13.y, s11.y, s23.y, s21.y, s33.y, s31.y);
    const float xr = sob(s13.z, s11.z, s23.z, s21.z, s33.z, s31.z);
    const float yb = sob(s31.x, s11.x, s32.x, s12.x, s33.x, s13.x);
    const float yg = sob(s31.y, s11.y, s32.y, s12.y, s33.y, s13.y);
    const float yr = sob(s31.z, s11.z, s32.z, s12.z, s33.z, s13.z);
    gx[i] = make_uchar4(grad_u8(xb, sign), grad_u8(xg, sign), grad_u8(xr, sign), 255);
    gy[i] = make_uchar4(grad_u8(yb, sign), grad_u8(yg, sign), grad_u8(yr, sign), 255);
}

// ---- projection shared by the difference and the ICP kernels -----------------

struct LevelCam {
    float fx, fy, cx, cy;
    int W, H;
};
__device__ __forceinline__ LevelCam level_cam(const NuiCLCameraParams& cam, int div) {
    return LevelCam{cam.fx / (float)div, cam.fy / (float)div, cam.cx / (float)div, cam.cy / (float)div,
                    (int)(cam.depthImageWidth / (unsigned)div), (int)(cam.depthImageHeight / (unsigned)div)};
}
__device__ __forceinline__ bool project_round(const LevelCam& L, f3 q, int* idx) {
    const int px = cvt_i32_sat(roundf(q.x * L.fx / q.z + L.cx));
    const int py = cvt_i32_sat(roundf(q.y * L.fy / q.z + L.cy));
    if (px < 0 || px >= L.W || py < 0 || py >= L.H) return false;
    *idx = py * L.W + px;
    return true;
}

// K/colorIcp.cl:7-57 color_difference_kernel
__global__ void __launch_bounds__(256)
    color_difference_kernel(const float* __restrict__ vmap, const uchar4* __restrict__ colors,
                            const uchar4* __restrict__ lc, NuiCLCameraParams cam, int div,
                            const NuiCLRigidTransform* __restrict__ pose, unsigned long long* out2) {
    const int n = (int)(cam.depthImageWidth * cam.depthImageHeight);
    const int i = blockIdx.x * 256 + threadIdx.x;
    unsigned long long d = 0, k = 0;
    if (i < n) {
        const f3 p = ld3(vmap, i);
        const uchar4 known = colors[i];
        if (!nan3(p) && known.w >= 254) {
            const f3 q = add3(rot3(pose->R, p), mk3(pose->t[0], pose->t[1], pose->t[2]));
            int li;
            if (project_round(level_cam(cam, div), q, &li)) {
                const uchar4 obs = lc[li];
                if (obs.w >= 254) {
                    const int e0 = (int)obs.x - known.x, e1 = (int)obs.y - known.y, e2 = (int)obs.z - known.z;
                    d = (unsigned long long)(e0 * e0 + e1 * e1 + e2 * e2);
                    k = 1;
                }
            }
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        d += __shfl_down(d, o, 64);
        k += __shfl_down(k, o, 64);
    }
    if ((threadIdx.x & 63) == 0 && k) {
        atomicAdd(&out2[0], d);
        atomicAdd(&out2[1], k);
    }
}

__global__ void color_zero2_kernel(unsigned long long* out2) {
    if (threadIdx.x < 2) out2[threadIdx.x] = 0ull;
}

// ---- ICP ------------------------------------------------------------------------

// K/colorIcp.cl:92-188 color_icp_block_kernel, one model point (zeros when rejected)
__device__ __forceinline__ void color_terms(const LevelCam& L, const f3 p, const uchar4 known,
                                           const uchar4* __restrict__ lc, const uchar4* __restrict__ lgx,
                                           const uchar4* __restrict__ lgy, const float* R, const float* t,
                                           int numPara, int startPara, int sign, float out[kCTerms]) {
#pragma unroll
    for (int k = 0; k < kCTerms; ++k) out[k] = 0.f;
    if (nan3(p) || known.w != 255) return;
    const f3 q = add3(rot3(R, p), mk3(t[0], t[1], t[2]));
    int li;
    if (!project_round(L, q, &li)) return;
    const uchar4 gx = lgx[li], gy = lgy[li], obs = lc[li];
    if (!(obs.w == 255 && gx.w == 255 && gy.w == 255)) return;
    auto g = [sign](unsigned char v) { return sign ? (float)(signed char)v : (float)v; };
    const float gx0 = g(gx.x), gx1 = g(gx.y), gx2 = g(gx.z), gy0 = g(gy.x), gy1 = g(gy.y), gy2 = g(gy.z);
    const f3 diff = mk3((float)obs.x - (float)known.x, (float)obs.y - (float)known.y, (float)obs.z - (float)known.z);
    f3 d[6];
    int k = 0;
    const float zz = q.z * q.z;
#pragma unroll
    for (int para = 0; para < 6; ++para) {
        if (para >= numPara) break;
        f3 dp;
        switch (para + startPara) {
            case 0: dp = mk3(1.0f, 0.0f, 0.0f); break;
            case 1: dp = mk3(0.0f, 1.0f, 0.0f); break;
            case 2: dp = mk3(0.0f, 0.0f, 1.0f); break;
            case 3: dp = mk3(0.0f, -q.z, q.y); break;
            case 4: dp = mk3(q.z, 0.0f, -q.x); break;
            default: dp = mk3(-q.y, q.x, 0.0f); break;
        }
        const float jx = L.fx * ((q.z * dp.x - dp.z * q.x) / zz);
        const float jy = L.fy * ((q.z * dp.y - dp.z * q.y) / zz);
        d[para] = mk3(jx * gx0 + jy * gy0, jx * gx1 + jy * gy1, jx * gx2 + jy * gy2);
#pragma unroll
        for (int col = 0; col < 6; ++col)
            if (col <= para) out[k++] = dot3(d[para], d[col]);
        out[k++] = dot3(d[para], diff);
    }
    out[k++] = sqrtf(diff.x * diff.x + diff.y * diff.y + diff.z * diff.z);
    out[k++] = 1.0f;
}

__device__ __forceinline__ int kernel_slot(int p, int c) { return p * (p + 3) / 2 + c; }

// OCL/NuiKinfuOpenCLColorTracker.cpp:469-526 for one iteration, given the sums
__device__ void color_update(ColorIcpState& S, const float* sumsf, int level, int rot_only, float dist_thr,
                             float num_points) {
    for (int k = 0; k < kCTerms; ++k) S.sums[k] = sumsf[k];
    const int numPara = rot_only ? 3 : 6;
    // the reference reads |diff| and the count from slots 27, 28; the kernel writes
    // them after its terms (slots 9, 10 with three parameters)
    const int es = S.corrected ? numPara * (numPara + 3) / 2 : kCTerms - 2;
    const float err = sumsf[es], cnt = sumsf[es + 1];
    S.error = cnt > 0.f ? sqrtf(err) / cnt : 3.402823466e38f;
    S.count = cnt;
    S.iterations += 1;
    if (cnt < 1.f) {
        S.brk[level] = 1;
        return;
    }
    const float scale = num_points / cnt;
    double A[36], b[6];
    if (!S.corrected) {
        for (int i = 0, shift = 0; i < 6; ++i)
            for (int j = i; j < 7; ++j, ++shift) {
                const float value = sumsf[shift] * scale;
                if (j == 6)
                    b[i] = value;
                else
                    A[j * 6 + i] = A[i * 6 + j] = value;
            }
    } else {
        for (int i = 0; i < 36; ++i) A[i] = 0.0;
        for (int i = 0; i < 6; ++i) b[i] = 0.0;
        for (int p = 0; p < numPara; ++p) {
            for (int c = 0; c <= p; ++c) {
                const float value = sumsf[kernel_slot(p, c)] * scale;
                A[c * 6 + p] = A[p * 6 + c] = value;
            }
            b[p] = sumsf[kernel_slot(p, p + 1)] * scale;
        }
        for (int p = numPara; p < 6; ++p) A[p * 6 + p] = 1.0;
    }
    double det, x[6];
    if (!llt6(A, &det) || fabs(det) < 1e-15 || isnan(det)) {
        S.failed = 1;
        return;
    }
    llt6_solve(A, b, x);
    float rot[3], tinc[3];
    if (!S.corrected) {
        rot[0] = (float)x[0], rot[1] = (float)x[1], rot[2] = (float)x[2];
        tinc[0] = (float)x[3], tinc[1] = (float)x[4], tinc[2] = (float)x[5];
        if (fabsf(rot[0]) < dist_thr && fabsf(rot[1]) < dist_thr && fabsf(rot[2]) < dist_thr) {
            S.brk[level] = 1;
            return;
        }
    } else {
        const int r0 = rot_only ? 0 : 3;
        rot[0] = (float)x[r0], rot[1] = (float)x[r0 + 1], rot[2] = (float)x[r0 + 2];
        tinc[0] = rot_only ? 0.f : -(float)x[0];
        tinc[1] = rot_only ? 0.f : -(float)x[1];
        tinc[2] = rot_only ? 0.f : -(float)x[2];
        float m = 0.f;
        for (int k = 0; k < numPara; ++k) m = fmaxf(m, fabsf((float)x[k]));
        if (m < 1e-6f) {
            S.brk[level] = 1;
            return;
        }
    }
    float Rz[9], Ry[9], Rx[9], Rzy[9], Rinc[9], Ri[9];
    axis_rotf(rot[2], 2, Rz);
    axis_rotf(rot[1], 1, Ry);
    axis_rotf(rot[0], 0, Rx);
    matmul3f(Rz, Ry, Rzy);
    matmul3f(Rzy, Rx, Rinc);
    inverse3f(Rinc, Ri);
    float tn[3];
    for (int i = 0; i < 3; ++i) tn[i] = Ri[3 * i] * S.t[0] + Ri[3 * i + 1] * S.t[1] + Ri[3 * i + 2] * S.t[2];
    for (int i = 0; i < 3; ++i) S.t[i] = tn[i] + tinc[i];
    float Rn[9];
    matmul3f(Ri, S.R, Rn);
    for (int k = 0; k < 9; ++k) S.R[k] = Rn[k];
}

// One iteration.  Block g (128 items) takes model points 128g..128g+127; the
// partials buffer holds ngroups x 29 block sums then ceil(ngroups/128) x 29
// compute_color_sums results.  The state and the partials are written and read
// by different blocks of one launch: no __restrict__ on them.
__global__ void __launch_bounds__(128)
    color_icp_kernel(ColorIcpState* st, const float* __restrict__ vmap, const uchar4* __restrict__ colors,
                     NuiCLCameraParams cam, int level, const uchar4* __restrict__ lc, const uchar4* __restrict__ lgx,
                     const uchar4* __restrict__ lgy, int rot_only, float dist_thr, float* partials) {
    if (st->failed || st->brk[level]) return;
    __shared__ float xfer[64 * kCTerms];
    __shared__ float sumsf[kCTerms];
    __shared__ ColorIcpState S;
    __shared__ int last, last2;
    float R[9], t[3];
#pragma unroll
    for (int k = 0; k < 9; ++k) R[k] = st->R[k];
#pragma unroll
    for (int k = 0; k < 3; ++k) t[k] = st->t[k];
    const int sign = st->corrected;
    const int npx = (int)(cam.depthImageWidth * cam.depthImageHeight);
    const int lid = threadIdx.x, lane = lid & 63, wave = lid >> 6;
    const int g = blockIdx.x, ngroups = (int)gridDim.x;
    const int p = g * 128 + lid;
    float v[kCTerms];
    const LevelCam L = level_cam(cam, 1 << level);
    if (p < npx)
        color_terms(L, ld3(vmap, p), colors[p], lc, lgx, lgy, R, t, rot_only ? 3 : 6, rot_only ? 3 : 0, sign, v);
    else
#pragma unroll
        for (int k = 0; k < kCTerms; ++k) v[k] = 0.f;

    // K/colorIcp.cl:59-90 reduce(): step d = 64 through LDS, steps 32..1 in wave 0
    if (wave == 1) {
#pragma unroll
        for (int k = 0; k < kCTerms; ++k) xfer[lane * kCTerms + k] = v[k];
    }
    __syncthreads();
    if (wave == 0) {
#pragma unroll
        for (int k = 0; k < kCTerms; ++k) v[k] = wave_tree64(v[k] + xfer[lane * kCTerms + k]);
        if (lane == 0) {
#pragma unroll
            for (int k = 0; k < kCTerms; ++k) partials[(size_t)g * kCTerms + k] = v[k];
        }
    }

    // the last block of each 128-block group runs compute_color_sums' tree
    ColorIcpSync* sy = csync_of(st);
    const int grp = g >> 7;
    if (lid == 0) {
        fence_release_agent();  // release this block's partial (written by this thread)
        const uint32_t members = (uint32_t)min(128, ngroups - (grp << 7));
        last = atomicAdd(&sy->arrive[grp], 1u) == members - 1;
    }
    __syncthreads();
    if (!last) return;
    fence_acquire_agent();  // acquire the group's partials
    float* tree = partials + (size_t)ngroups * kCTerms;
    for (int k = wave; k < kCTerms; k += 2) {
        const int g0 = (grp << 7) + lane, g1 = g0 + 64;
        const float x = g0 < ngroups ? partials[(size_t)g0 * kCTerms + k] : 0.f;
        const float y = g1 < ngroups ? partials[(size_t)g1 * kCTerms + k] : 0.f;
        const float sum = wave_tree64(x + y);
        if (lane == 0) tree[grp * kCTerms + k] = sum;
    }
    __syncthreads();  // the group result is written
    const int ngroups2 = (ngroups + 127) >> 7;
    if (lid == 0) {
        sy->arrive[grp] = 0u;  // re-arm
        fence_release_agent();  // release the group result
        last2 = atomicAdd(&sy->arrive_all, 1u) == (uint32_t)(ngroups2 - 1);
    }
    __syncthreads();
    if (!last2) return;
    fence_acquire_agent();  // acquire every group result

    // the host's sequential float sums over the compute_color_sums blocks
    if (lid < kCTerms) {
        float sum = 0.f;
        for (int c = 0; c < ngroups2; ++c) sum += tree[c * kCTerms + lid];
        sumsf[lid] = sum;
    }
    if (lid == 0) sy->arrive_all = 0u;
    constexpr int kWords = (int)(sizeof(ColorIcpState) / 4);
    const uint32_t* src = reinterpret_cast<const uint32_t*>(st);
    uint32_t* lds = reinterpret_cast<uint32_t*>(&S);
    for (int i = lid; i < kWords; i += 128) lds[i] = src[i];
    __syncthreads();
    if (lid == 0) color_update(S, sumsf, level, rot_only, dist_thr, (float)npx);
    __syncthreads();
    uint32_t* dst = reinterpret_cast<uint32_t*>(st);
    for (int i = lid; i < kWords; i += 128) dst[i] = lds[i];
}

// The transform applied to the model points: the camera pose (reference decode)
// or its inverse, world -> camera (corrected)
__global__ void color_icp_begin_kernel(ColorIcpState* st, const NuiCLRigidTransform* pose, int corrected) {
    ColorIcpSync* sy = csync_of(st);
    for (int i = threadIdx.x; i < kCMaxGroups2; i += blockDim.x) sy->arrive[i] = 0u;
    if (threadIdx.x != 0) return;
    sy->arrive_all = 0u;
    if (corrected) {
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) st->R[3 * i + j] = pose->R[3 * j + i];
        for (int i = 0; i < 3; ++i)
            st->t[i] = -(st->R[3 * i] * pose->t[0] + st->R[3 * i + 1] * pose->t[1] + st->R[3 * i + 2] * pose->t[2]);
    } else {
        for (int k = 0; k < 9; ++k) st->R[k] = pose->R[k];
        for (int k = 0; k < 3; ++k) st->t[k] = pose->t[k];
    }
    for (int l = 0; l < NUIF_ICP_MAX_LEVELS; ++l) st->brk[l] = 0;
    st->failed = 0;
    st->iterations = 0;
    st->corrected = corrected;
    st->error = 0.f;
    st->count = 0.f;
}

// UpdateCameraTransform(Rcurr, tcurr) when the track holds
__global__ void color_icp_end_kernel(const ColorIcpState* st, NuiCLRigidTransform* pose) {
    if (threadIdx.x != 0 || st->failed) return;
    float R[9], t[3];
    if (st->corrected) {
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) R[3 * i + j] = st->R[3 * j + i];
        for (int i = 0; i < 3; ++i) t[i] = -(R[3 * i] * st->t[0] + R[3 * i + 1] * st->t[1] + R[3 * i + 2] * st->t[2]);
    } else {
        for (int k = 0; k < 9; ++k) R[k] = st->R[k];
        for (int k = 0; k < 3; ++k) t[k] = st->t[k];
    }
    for (int k = 0; k < 9; ++k) pose->R[k] = R[k];
    for (int k = 0; k < 3; ++k) pose->t[k] = t[k];
    inverse3f(R, pose->R_inv);
}

// ---- joint depth + intensity ICP (K/icp.cl:328-480; no host in the reference) ------

// K/transformMaps.cl:372-384 bgra_to_intensity_kernel, one pixel
__device__ __forceinline__ float intensity_of(uchar4 c) {
    return c.w > 0 ? 0.299f * ((float)c.z / 255.0f) + 0.587f * ((float)c.y / 255.0f) + 0.114f * ((float)c.x / 255.0f)
                   : kNaN;
}

__global__ void bgra_to_intensity_kernel(const uchar4* __restrict__ bgra, float* __restrict__ out, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = intensity_of(bgra[i]);
}

// K/transformMaps.cl:448-511 intensity_derivatives_kernel, one pixel: I(dx, dy) the
// intensity at the offset from the centre; NaN where the reference writes nothing
template <class F>
__device__ __forceinline__ float2 sobel8(F I, bool inside) {
    float2 r = make_float2(kNaN, kNaN);
    if (!isnan(I(0, 0)) && inside) {
        const float i00 = I(-1, -1), i10 = I(0, -1), i20 = I(1, -1);
        const float i01 = I(-1, 0), i21 = I(1, 0);
        const float i02 = I(-1, 1), i12 = I(0, 1), i22 = I(1, 1);
        if (!(isnan(i00) || isnan(i10) || isnan(i20) || isnan(i01) || isnan(i21) || isnan(i02) || isnan(i12) ||
              isnan(i22))) {
            const float dx = (-1.0f) * i00 + (1.0f) * i20 + (-2.0f) * i01 + (2.0f) * i21 + (-1.0f) * i02 + (1.0f) * i22;
            const float dy = (-1.0f) * i00 + (-2.0f) * i10 + (-1.0f) * i20 + (1.0f) * i02 + (2.0f) * i12 + (1.0f) * i22;
            r = make_float2(dx / 8.0f, dy / 8.0f);
        }
    }
    return r;
}

// neighbours one row (the width) apart, result at the centre
__global__ void intensity_derivatives_kernel(const float* __restrict__ I, float2* __restrict__ out, int W, int H) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= W || y >= H) return;
    const size_t c = (size_t)y * W + x;
    out[c] = sobel8([&](int dx, int dy) { return I[(long long)c + dy * W + dx]; },
                    x > 0 && x < W - 1 && y > 0 && y < H - 1);
}

// The mean of the valid ones of four intensities, summed in order; NaN when none is
__device__ __forceinline__ float mean_valid4(const float v[4]) {
    float sum = 0.f;
    int cnt = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (!isnan(v[k])) {
            sum += v[k];
            ++cnt;
        }
    return cnt ? sum / (float)cnt : kNaN;
}

// Intensities one level down: the mean of the valid ones of the 2 x 2 source pixels in
// row order (no counterpart in the reference)
__global__ void intensity_half_sample_kernel(const float* __restrict__ src, float* __restrict__ dst, int Wd, int Hd) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= Wd || y >= Hd) return;
    const int Ws = 2 * Wd;
    const size_t i = (size_t)(2 * y) * Ws + 2 * x;
    const float v[4] = {src[i], src[i + 1], src[i + Ws], src[i + Ws + 1]};
    dst[(size_t)y * Wd + x] = mean_valid4(v);
}

// One frame's joint-order inputs (nuif_icp_photometric) in 16 x 16 tiles of level-0
// pixels numbered from block0.
struct PhotoJob {
    const uchar4* bgra;
    const uchar4* ref;
    float* pyr[NUIF_ICP_MAX_LEVELS];
    float4* model;
    uchar4* keep;  // photometric_keep_kernel: receives the frame's colours (the next frame's ref)
    int W, H, levels, gx;
    unsigned block0;
};

// All sessions' inputs in one launch (jobs == NULL: the one job `single`), one 16 x 16
// tile of level-0 pixels per block.  The intensities come from per-channel tables of
// intensity_of's three products (the same roundings, added in the same order); the
// tile's intensities stay in LDS for the pyramid (a 16-pixel tile holds whole 2^l x 2^l
// blocks for l <= 3, the image size being a multiple of 2^(levels - 1)), and the
// reference intensities of the tile and its one-pixel halo for the Sobel gradient.
__global__ void __launch_bounds__(256)
    photometric_inputs_kernel(const PhotoJob* __restrict__ jobs, int n, PhotoJob single) {
    __shared__ float tab[3][256];
    __shared__ float lev[16 * 16 + 8 * 8 + 4 * 4 + 2 * 2];
    __shared__ float refi[18][19];
    const int tid = threadIdx.x;
    tab[0][tid] = 0.299f * ((float)tid / 255.0f);  // red
    tab[1][tid] = 0.587f * ((float)tid / 255.0f);  // green
    tab[2][tid] = 0.114f * ((float)tid / 255.0f);  // blue
    int j = 0;
    if (jobs) {  // the last job whose first block is at or below this block
        int lo = 0, hi = n - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (jobs[mid].block0 <= blockIdx.x) lo = mid;
            else hi = mid - 1;
        }
        j = lo;
    }
    const PhotoJob& J = jobs ? jobs[j] : single;
    const unsigned b = blockIdx.x - J.block0;
    const int x0 = (int)(b % (unsigned)J.gx) * 16, y0 = (int)(b / (unsigned)J.gx) * 16;
    const int W = J.W, H = J.H;
    __syncthreads();
    auto inten = [&](uchar4 c) { return c.w > 0 ? (tab[0][c.z] + tab[1][c.y]) + tab[2][c.x] : kNaN; };
    // the reference intensities of the tile and its halo (18 x 18, NaN outside the image)
    const uchar4* __restrict__ ref = J.ref;
    for (int k = tid; k < 18 * 18; k += 256) {
        const int hx = x0 - 1 + k % 18, hy = y0 - 1 + k / 18;
        refi[k / 18][k % 18] =
            hx >= 0 && hx < W && hy >= 0 && hy < H ? inten(ref[(size_t)hy * W + hx]) : kNaN;
    }
    const int tx = tid & 15, ty = tid >> 4;
    const int x = x0 + tx, y = y0 + ty;
    const bool in = x < W && y < H;
    const size_t i = (size_t)y * W + x;
    const float i0 = in ? inten(J.bgra[i]) : kNaN;
    lev[tid] = i0;
    __syncthreads();
    if (in) {
        J.pyr[0][i] = i0;
        const float ir = refi[ty + 1][tx + 1];
        const float2 g = sobel8([&](int dx, int dy) { return refi[ty + 1 + dy][tx + 1 + dx]; },
                                x > 0 && x < W - 1 && y > 0 && y < H - 1);
        J.model[i] = make_float4(ir, g.x, g.y, 0.f);
    }
    // levels 1..3 from the tile's own intensities, each (16 >> l)^2 after the previous
    float* src = lev;
    float* dst = lev + 256;
    for (int l = 1; l < J.levels; ++l) {
        const int w = 16 >> l;
        if (tid < w * w) {
            const int lx = tid % w, ly = tid / w;
            const int sw = 2 * w;
            const float v[4] = {src[(2 * ly) * sw + 2 * lx], src[(2 * ly) * sw + 2 * lx + 1],
                                src[(2 * ly + 1) * sw + 2 * lx], src[(2 * ly + 1) * sw + 2 * lx + 1]};
            const float m = mean_valid4(v);
            dst[tid] = m;
            const int gx = (x0 >> l) + lx, gy = (y0 >> l) + ly;
            if (gx < (W >> l) && gy < (H >> l)) J.pyr[l][(size_t)gy * (W >> l) + gx] = m;
        }
        __syncthreads();
        src = dst;
        dst += w * w;
    }
}

// The frame's colours into `keep` of every job, after photometric_inputs_kernel has read
// the reference colours there (the next frame's model).
__global__ void __launch_bounds__(256) photometric_keep_kernel(const PhotoJob* __restrict__ jobs, int n) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].block0 <= blockIdx.x) lo = mid;
        else hi = mid - 1;
    }
    const PhotoJob& J = jobs[lo];
    const unsigned b = blockIdx.x - J.block0;
    const int x = (int)(b % (unsigned)J.gx) * 16 + (int)(threadIdx.x & 15);
    const int y = (int)(b / (unsigned)J.gx) * 16 + (int)(threadIdx.x >> 4);
    if (J.keep && x < J.W && y < J.H) J.keep[(size_t)y * J.W + x] = J.bgra[(size_t)y * J.W + x];
}

int make_photo_job(const uint8_t* bgra, const uint8_t* ref_bgra, const nuif_icp_photometric& ph, int W, int H,
                   int levels, unsigned block0, PhotoJob* out) {
    NUIF_CHECK_ARG(bgra && ref_bgra && ph.ref_model && W > 0 && H > 0 && levels >= 1 && levels <= NUIF_ICP_MAX_LEVELS,
                   "joint ICP inputs: bad arguments");
    NUIF_CHECK_ARG(W % (1 << (levels - 1)) == 0 && H % (1 << (levels - 1)) == 0,
                   "joint ICP inputs: %d x %d is not divisible by 2^(levels - 1)", W, H);
    PhotoJob J{};
    J.bgra = reinterpret_cast<const uchar4*>(bgra);
    J.ref = reinterpret_cast<const uchar4*>(ref_bgra);
    for (int l = 0; l < levels; ++l) {
        NUIF_CHECK_ARG(ph.intensity_pyr[l], "joint ICP inputs: intensity level %d is NULL", l);
        J.pyr[l] = const_cast<float*>(ph.intensity_pyr[l]);
    }
    J.model = reinterpret_cast<float4*>(const_cast<float*>(ph.ref_model));
    J.W = W;
    J.H = H;
    J.levels = levels;
    J.gx = div_up(W, 16);
    J.block0 = block0;
    *out = J;
    return NUIF_OK;
}

unsigned photo_job_blocks(const PhotoJob& J) { return (unsigned)J.gx * (unsigned)div_up(J.H, 16); }

size_t photo_job_bytes() { return sizeof(PhotoJob); }

int make_photo_job_at(void* hjobs, int k, const uint8_t* bgra, uint8_t* ref_bgra, const nuif_icp_photometric& ph,
                      int W, int H, int levels, unsigned* blocks) {
    PhotoJob* J = static_cast<PhotoJob*>(hjobs) + k;
    const int rc = make_photo_job(bgra, ref_bgra, ph, W, H, levels, *blocks, J);
    if (rc) return rc;
    J->keep = reinterpret_cast<uchar4*>(ref_bgra);
    *blocks += photo_job_blocks(*J);
    return NUIF_OK;
}

int launch_photo_jobs(const void* djobs, int n, unsigned blocks, hipStream_t st) {
    if (n <= 0) return NUIF_OK;
    photometric_inputs_kernel<<<blocks, 256, 0, st>>>(static_cast<const PhotoJob*>(djobs), n, PhotoJob{});
    photometric_keep_kernel<<<blocks, 256, 0, st>>>(static_cast<const PhotoJob*>(djobs), n);
    NUIF_LAUNCHED("photometric_inputs_kernel / photometric_keep_kernel");
    return NUIF_OK;
}

// K/icp.cl:328-350 evalR_dAlpha / evalR_dBeta / evalR_dGamma as written
__device__ __forceinline__ void eval_phis(const float s[3], const float c[3], f3 p, f3* pa, f3* pb, f3* pg) {
    const float sx = s[0], cx = c[0], sy = s[1], cy = c[1], sz = s[2], cz = c[2];
    *pa = mk3(dot3(mk3(-sx * cy, -cx * cz - sx * sy * sz, cx * sz - sx * sy * cz), p),
              dot3(mk3(cx * cy, -sx * cz + cx * sy * sz, sx * sz + cx * sy * cz), p), 0.0f);
    *pb = mk3(dot3(mk3(-cx * sy, cx * cy * sz, cx * cy * cz), p), dot3(mk3(-sx * sy, sx * cy * sz, sx * cy * sz), p),
              dot3(mk3(-cy, -sy * sz, -sy * cz), p));
    *pg = mk3(dot3(mk3(0.0f, sx * sz + sx * cy * sz, sx * cz - cx * sy * sz), p),
              dot3(mk3(0.0f, -cx * sz + sx * sy * cz, -cx * cz - sx * sy * sz), p), dot3(mk3(0.0f, cy * cz, -cy * sz), p));
}

struct IntensityIcpArgs {
    const float *v, *n, *I, *vp, *np, *Ip;
    const float2* dp;
    float es[3], ec[3];  // sin, cos of the Euler angles
    float distThr, angleThr, colorThr, gradMin;
};

// K/icp.cl:352-480 intensity_icp_block_kernel, one pixel (zeros when rejected)
__device__ __forceinline__ void intensity_terms(const IntensityIcpArgs& a, const LevelCam& L, const float* R,
                                                const float* t, const NuiCLRigidTransform& prev, int gid,
                                                float out[kCTerms]) {
#pragma unroll
    for (int k = 0; k < kCTerms; ++k) out[k] = 0.f;
    const f3 nc = ld3(a.n, gid);
    if (nan3(nc)) return;
    const f3 vc = ld3(a.v, gid);
    if (nan3(vc)) return;
    const f3 pv = add3(rot3(R, vc), mk3(t[0], t[1], t[2]));
    const f3 pp = rot3(prev.R_inv, sub3(pv, mk3(prev.t[0], prev.t[1], prev.t[2])));
    int ref;
    if (!project_round(L, pp, &ref)) return;
    const f3 rn = ld3(a.np, ref), rv = ld3(a.vp, ref);
    const float ri = a.Ip[ref];
    const float2 rd = a.dp[ref];
    if (nan3(rn) || isnan(ri) || nan3(rv)) return;
    const f3 diff = mk3(fabsf(rv.x - pv.x), fabsf(rv.y - pv.y), fabsf(rv.z - pv.z));
    const float dist = sqrtf(diff.x * diff.x + diff.y * diff.y + diff.z * diff.z);
    if (!(dist < a.distThr)) return;
    f3 pa, pb, pg;
    eval_phis(a.es, a.ec, pp, &pa, &pb, &pg);
    const f3 cr = cross3(rot3(R, nc), rn);
    const float sine = sqrtf(cr.x * cr.x + cr.y * cr.y + cr.z * cr.z);
    if (!(sine < a.angleThr)) return;
    const float wd = fmaxf(0.0f, 1.0f - dist / a.distThr);
    const f3 mn = mk3(-rn.x, -rn.y, -rn.z);
    float c[7] = {wd * dot3(mn, pa), wd * dot3(mn, pb), wd * dot3(mn, pg), wd * rn.x, wd * rn.y, wd * rn.z,
                  wd * dot3(rn, diff)};
    const float ic = a.I[gid];
    if (!isnan(ic)) {
        const float di = fabsf(ic - ri);
        if (di < a.colorThr && !isnan(rd.x) && !isnan(rd.y) && sqrtf(rd.x * rd.x + rd.y * rd.y) > a.gradMin) {
            const float wc = fmaxf(0.0f, 1.0f - di / a.colorThr);
            const f3 t0 = mk3(rd.x * L.fx / pp.z, rd.y * L.fy / pp.z,
                              (rd.x * L.cx + rd.y * L.cy) / pp.z - (rd.x * pp.x + rd.y * pp.y) / (pp.z * pp.z));
            c[0] += wc * dot3(t0, pa);
            c[1] += wc * dot3(t0, pb);
            c[2] += wc * dot3(t0, pg);
            c[3] += wc * t0.x;
            c[4] += wc * t0.y;
            c[5] += wc * t0.z;
            c[6] += wc * di;
        }
    }
    int k = 0;
#pragma unroll
    for (int i = 0; i < 6; ++i)
#pragma unroll
        for (int j = i; j < 7; ++j) out[k++] = c[i] * c[j];
    if (dist > a.distThr / 100.0f && sine > a.angleThr / 100.0f) {
        out[27] = c[6] * c[6];
        out[28] = 1.0f;
    }
}

// The joint kernel's sums in one launch: kIntGroups 128-item groups per block, each
// reduced with the reference's 128-wide tree; the last block of every 128 groups runs
// compute_sums' tree, the last of those sums the results in order into out29.
constexpr int kIntGroups = 4;
__global__ void __launch_bounds__(128 * kIntGroups)
    intensity_icp_kernel(IntensityIcpArgs a, NuiCLCameraParams cam, int level, const NuiCLRigidTransform* pose,
                         const NuiCLRigidTransform* prev_pose, float* partials, ColorIcpSync* sy, float* out29) {
    __shared__ float xfer[kIntGroups][64 * kCTerms];
    __shared__ int last, last2;
    const LevelCam L = level_cam(cam, 1 << level);
    const int npx = L.W * L.H, ngroups = (npx + 127) >> 7;
    float R[9], t[3];
#pragma unroll
    for (int k = 0; k < 9; ++k) R[k] = pose->R[k];
#pragma unroll
    for (int k = 0; k < 3; ++k) t[k] = pose->t[k];
    const NuiCLRigidTransform prev = *prev_pose;
    const int tid = threadIdx.x, sub = tid >> 7, l = tid & 127, lane = l & 63, wave = l >> 6;
    const int g = blockIdx.x * kIntGroups + sub, p = g * 128 + l;
    float v[kCTerms];
    if (p < npx)
        intensity_terms(a, L, R, t, prev, p, v);
    else
#pragma unroll
        for (int k = 0; k < kCTerms; ++k) v[k] = 0.f;
    if (wave == 1) {
#pragma unroll
        for (int k = 0; k < kCTerms; ++k) xfer[sub][lane * kCTerms + k] = v[k];
    }
    __syncthreads();
    if (wave == 0) {
#pragma unroll
        for (int k = 0; k < kCTerms; ++k) v[k] = wave_tree64(v[k] + xfer[sub][lane * kCTerms + k]);
        if (lane == 0 && g < ngroups) {
#pragma unroll
            for (int k = 0; k < kCTerms; ++k) partials[(size_t)g * kCTerms + k] = v[k];
        }
    }
    __syncthreads();
    const int chunk = (blockIdx.x * kIntGroups) >> 7;
    if (tid == 0) {
        fence_release_agent();
        const int in_chunk = min(128, ngroups - (chunk << 7));
        last = atomicAdd(&sy->arrive[chunk], 1u) == (uint32_t)((in_chunk + kIntGroups - 1) / kIntGroups) - 1;
    }
    __syncthreads();
    if (!last) return;
    fence_acquire_agent();
    float* tree = partials + (size_t)ngroups * kCTerms;
    const int wid = tid >> 6, nwaves = (128 * kIntGroups) >> 6;
    for (int k = wid; k < kCTerms; k += nwaves) {
        const int g0 = (chunk << 7) + (tid & 63), g1 = g0 + 64;
        const float x = g0 < ngroups ? partials[(size_t)g0 * kCTerms + k] : 0.f;
        const float y = g1 < ngroups ? partials[(size_t)g1 * kCTerms + k] : 0.f;
        const float sum = wave_tree64(x + y);
        if ((tid & 63) == 0) tree[chunk * kCTerms + k] = sum;
    }
    __syncthreads();
    const int nchunks = (ngroups + 127) >> 7;
    if (tid == 0) {
        sy->arrive[chunk] = 0u;
        fence_release_agent();
        last2 = atomicAdd(&sy->arrive_all, 1u) == (uint32_t)(nchunks - 1);
    }
    __syncthreads();
    if (!last2) return;
    fence_acquire_agent();
    if (tid < kCTerms) {
        float sum = 0.f;
        for (int c = 0; c < nchunks; ++c) sum += tree[c * kCTerms + tid];
        out29[tid] = sum;
    }
    if (tid == 0) sy->arrive_all = 0u;
}

__global__ void intensity_sync_clear_kernel(ColorIcpSync* sy) {
    for (int i = threadIdx.x; i < kCMaxGroups2; i += blockDim.x) sy->arrive[i] = 0u;
    if (threadIdx.x == 0) sy->arrive_all = 0u;
}

}  // namespace nuif

using namespace nuif;

extern "C" {

int nuif_color_half_sample(const uint8_t* src, uint8_t* dst, int dst_width, int dst_height, int radius,
                           nuif_stream_t s) {
    NUIF_REQUIRE_INIT();
    NUIF_CHECK_ARG(src && dst && dst_width > 0 && dst_height > 0 && radius >= 0 && src != dst,
                   "nuif_color_half_sample: bad arguments");
    const dim3 block(16, 16), grid(div_up(dst_width, 16), div_up(dst_height, 16));
    color_half_sample_kernel<<<grid, block, 0, stream_of(s)>>>(reinterpret_cast<const uchar4*>(src),
                                                               reinterpret_cast<uchar4*>(dst), dst_width, dst_height,
                                                               radius);
    NUIF_LAUNCHED("color_half_sample_kernel");
    return NUIF_OK;
}

int nuif_color_gradient(const uint8_t* bgra, uint8_t* gx, uint8_t* gy, int width, int height, int sign,
                        nuif_stream_t s) {
    NUIF_REQUIRE_INIT();
    NUIF_CHECK_ARG(bgra && gx && gy && width > 0 && height > 0 && gx != gy && bgra != gx && bgra != gy,
                   "nuif_color_gradient: bad arguments");
    const dim3 block(16, 16), grid(div_up(width, 16), div_up(height, 16));
    color_gradient_kernel<<<grid, block, 0, stream_of(s)>>>(reinterpret_cast<const uchar4*>(bgra),
                                                            reinterpret_cast<uchar4*>(gx),
                                                            reinterpret_cast<uchar4*>(gy), width, height, sign);
    NUIF_LAUNCHED("color_gradient_kernel");
    return NUIF_OK;
}

int nuif_color_pyramid(const uint8_t* bgra, uint8_t* const* colors, uint8_t* const* gx, uint8_t* const* gy,
                       int width, int height, int levels, int sign, nuif_stream_t s) {
    NUIF_REQUIRE_INIT();
    NUIF_CHECK_ARG(bgra && colors && gx && gy && width > 0 && height > 0 && levels >= 1 &&
                       levels <= NUIF_ICP_MAX_LEVELS,
                   "nuif_color_pyramid: bad arguments");
    for (int l = 0; l < levels; ++l)
        NUIF_CHECK_ARG(colors[l] && gx[l] && gy[l], "nuif_color_pyramid: null level buffer");
    hipStream_t st = stream_of(s);
    if (colors[0] != bgra)
        NUIF_HIP(hipMemcpyAsync(colors[0], bgra, (size_t)width * height * 4, hipMemcpyDeviceToDevice, st));
    for (int l = 1; l < levels; ++l) {
        const int rc = nuif_color_half_sample(colors[l - 1], colors[l], width >> l, height >> l, 1, s);
        if (rc) return rc;
    }
    for (int l = 0; l < levels; ++l) {
        const int rc = nuif_color_gradient(colors[l], gx[l], gy[l], width >> l, height >> l, sign, s);
        if (rc) return rc;
    }
    return NUIF_OK;
}

int nuif_color_difference(const float* vmap, const uint8_t* colors, const uint8_t* level_colors,
                          const NuiCLCameraParams* cam, int level, const NuiCLRigidTransform* pose,
                          unsigned long long* out2, nuif_stream_t s) {
    NUIF_REQUIRE_INIT();
    NUIF_CHECK_ARG(vmap && colors && level_colors && cam && pose && out2 && level >= 0 && level < 16,
                   "nuif_color_difference: bad arguments");
    const int n = (int)(cam->depthImageWidth * cam->depthImageHeight);
    NUIF_CHECK_ARG(n > 0, "nuif_color_difference: empty image");
    hipStream_t st = stream_of(s);
    color_zero2_kernel<<<1, 64, 0, st>>>(out2);
    NUIF_LAUNCHED("color_zero2_kernel");
    color_difference_kernel<<<div_up((size_t)n, 256), 256, 0, st>>>(
        vmap, reinterpret_cast<const uchar4*>(colors), reinterpret_cast<const uchar4*>(level_colors), *cam,
        1 << level, pose, out2);
    NUIF_LAUNCHED("color_difference_kernel");
    return NUIF_OK;
}

size_t nuif_color_icp_state_bytes(void) { return kCStateBytes + kCSyncBytes; }

size_t nuif_color_icp_partials_bytes(int width, int height) {
    if (width <= 0 || height <= 0) return 0;
    const size_t groups = div_up((size_t)width * height, 128);
    return (groups + div_up(groups, 128)) * kCTerms * sizeof(float);
}

int nuif_color_icp_begin(void* state, const NuiCLRigidTransform* pose, int corrected, nuif_stream_t s) {
    NUIF_REQUIRE_INIT();
    NUIF_CHECK_ARG(state && pose, "nuif_color_icp_begin: null argument");
    color_icp_begin_kernel<<<1, 64, 0, stream_of(s)>>>(static_cast<ColorIcpState*>(state), pose, corrected ? 1 : 0);
    NUIF_LAUNCHED("color_icp_begin_kernel");
    return NUIF_OK;
}

int nuif_color_icp_iteration(void* state, float* partials, const NuiCLCameraParams* cam, int level,
                             int rotation_only, float dist_threshold, const float* vmap, const uint8_t* colors,
                             const uint8_t* level_colors, const uint8_t* level_gx, const uint8_t* level_gy,
                             nuif_stream_t s) {
    NUIF_REQUIRE_INIT();
    NUIF_CHECK_ARG(state && partials && cam && vmap && colors && level_colors && level_gx && level_gy,
                   "nuif_color_icp_iteration: null argument");
    NUIF_CHECK_ARG(level >= 0 && level < NUIF_ICP_MAX_LEVELS, "nuif_color_icp_iteration: bad level");
    const size_t n = (size_t)cam->depthImageWidth * cam->depthImageHeight;
    NUIF_CHECK_ARG(n > 0, "nuif_color_icp_iteration: empty image");
    const size_t groups = div_up(n, 128);
    NUIF_CHECK_ARG(div_up(groups, 128) <= (size_t)kCMaxGroups2, "nuif_color_icp_iteration: image too large");
    NUIF_CHECK_ARG((cam->depthImageWidth >> level) > 0 && (cam->depthImageHeight >> level) > 0,
                   "nuif_color_icp_iteration: level past the pyramid");
    color_icp_kernel<<<(unsigned)groups, 128, 0, stream_of(s)>>>(
        static_cast<ColorIcpState*>(state), vmap, reinterpret_cast<const uchar4*>(colors), *cam, level,
        reinterpret_cast<const uchar4*>(level_colors), reinterpret_cast<const uchar4*>(level_gx),
        reinterpret_cast<const uchar4*>(level_gy), rotation_only ? 1 : 0, dist_threshold, partials);
    NUIF_LAUNCHED("color_icp_kernel");
    return NUIF_OK;
}

int nuif_color_icp_end(const void* state, NuiCLRigidTransform* pose, nuif_stream_t s) {
    NUIF_REQUIRE_INIT();
    NUIF_CHECK_ARG(state && pose, "nuif_color_icp_end: null argument");
    color_icp_end_kernel<<<1, 64, 0, stream_of(s)>>>(static_cast<const ColorIcpState*>(state), pose);
    NUIF_LAUNCHED("color_icp_end_kernel");
    return NUIF_OK;
}

int nuif_color_icp_estimate(void* state, float* partials, const NuiCLCameraParams* cam, int levels,
                            const int* iterations, const int* rotation_only, float dist_threshold, const float* vmap,
                            const uint8_t* colors, const uint8_t* const* level_colors,
                            const uint8_t* const* level_gx, const uint8_t* const* level_gy,
                            NuiCLRigidTransform* pose, int corrected, nuif_stream_t s) {
    NUIF_CHECK_ARG(iterations && level_colors && level_gx && level_gy && levels >= 1 &&
                       levels <= NUIF_ICP_MAX_LEVELS,
                   "nuif_color_icp_estimate: bad arguments");
    int rc = nuif_color_icp_begin(state, pose, corrected, s);
    for (int level = levels - 1; level >= 0 && !rc; --level)
        for (int it = 0; it < iterations[level] && !rc; ++it)
            rc = nuif_color_icp_iteration(state, partials, cam, level, rotation_only ? rotation_only[level] : 0,
                                          dist_threshold, vmap, colors, level_colors[level], level_gx[level],
                                          level_gy[level], s);
    if (!rc) rc = nuif_color_icp_end(state, pose, s);
    return rc;
}

int nuif_bgra_to_intensity(const uint8_t* bgra, float* intensities, size_t n, nuif_stream_t s) {
    NUIF_REQUIRE_INIT();
    NUIF_CHECK_ARG(bgra && intensities && n > 0, "nuif_bgra_to_intensity: bad arguments");
    bgra_to_intensity_kernel<<<div_up(n, 256), 256, 0, stream_of(s)>>>(reinterpret_cast<const uchar4*>(bgra),
                                                                      intensities, n);
    NUIF_LAUNCHED("bgra_to_intensity_kernel");
    return NUIF_OK;
}

int nuif_intensity_derivatives(const float* intensities, float* derivatives, int width, int height,
                               nuif_stream_t s) {
    NUIF_REQUIRE_INIT();
    NUIF_CHECK_ARG(intensities && derivatives && width > 0 && height > 0 && (const void*)intensities != derivatives,
                   "nuif_intensity_derivatives: bad arguments");
    const dim3 block(16, 16), grid(div_up(width, 16), div_up(height, 16));
    intensity_derivatives_kernel<<<grid, block, 0, stream_of(s)>>>(intensities, reinterpret_cast<float2*>(derivatives),
                                                                   width, height);
    NUIF_LAUNCHED("intensity_derivatives_kernel");
    return NUIF_OK;
}

int nuif_intensity_half_sample(const float* src, float* dst, int dst_width, int dst_height, nuif_stream_t s) {
    NUIF_REQUIRE_INIT();
    NUIF_CHECK_ARG(src && dst && dst_width > 0 && dst_height > 0 && (const void*)src != dst,
                   "nuif_intensity_half_sample: bad arguments");
    const dim3 block(16, 16), grid(div_up(dst_width, 16), div_up(dst_height, 16));
    intensity_half_sample_kernel<<<grid, block, 0, stream_of(s)>>>(src, dst, dst_width, dst_height);
    NUIF_LAUNCHED("intensity_half_sample_kernel");
    return NUIF_OK;
}

int nuif_icp_photometric_inputs(const uint8_t* bgra, const uint8_t* ref_bgra, int width, int height, int levels,
                                const nuif_icp_photometric* photometric, nuif_stream_t s) {
    NUIF_REQUIRE_INIT();
    NUIF_CHECK_ARG(photometric, "nuif_icp_photometric_inputs: NULL inputs");
    PhotoJob J;
    const int rc = make_photo_job(bgra, ref_bgra, *photometric, width, height, levels, 0, &J);
    if (rc) return rc;
    photometric_inputs_kernel<<<photo_job_blocks(J), 256, 0, stream_of(s)>>>(nullptr, 1, J);
    NUIF_LAUNCHED("photometric_inputs_kernel");
    return NUIF_OK;
}

size_t nuif_intensity_icp_workspace_bytes(int width, int height) {
    if (width <= 0 || height <= 0) return 0;
    const size_t groups = div_up((size_t)width * height, 128);
    return kCSyncBytes + (groups + div_up(groups, 128)) * kCTerms * sizeof(float);
}

int nuif_intensity_icp_sums(const NuiCLCameraParams* cam, int level, const float* vmap, const float* nmap,
                            const float* intensities, const NuiCLRigidTransform* pose, const float* vmap_prev,
                            const float* nmap_prev, const float* intensities_prev, const float* derivatives_prev,
                            const NuiCLRigidTransform* prev_pose, const float* euler3, const float* thresholds4,
                            void* workspace, float* sums29, nuif_stream_t s) {
    NUIF_REQUIRE_INIT();
    NUIF_CHECK_ARG(cam && vmap && nmap && intensities && pose && vmap_prev && nmap_prev && intensities_prev &&
                       derivatives_prev && prev_pose && euler3 && thresholds4 && workspace && sums29,
                   "nuif_intensity_icp_sums: null argument");
    NUIF_CHECK_ARG(level >= 0 && level < 16, "nuif_intensity_icp_sums: bad level");
    const int Wl = (int)(cam->depthImageWidth >> level), Hl = (int)(cam->depthImageHeight >> level);
    NUIF_CHECK_ARG(Wl > 0 && Hl > 0, "nuif_intensity_icp_sums: empty level");
    const size_t groups = div_up((size_t)Wl * Hl, 128);
    NUIF_CHECK_ARG(div_up(groups, 128) <= (size_t)kCMaxGroups2, "nuif_intensity_icp_sums: image too large");
    IntensityIcpArgs a;
    a.v = vmap, a.n = nmap, a.I = intensities, a.vp = vmap_prev, a.np = nmap_prev, a.Ip = intensities_prev;
    a.dp = reinterpret_cast<const float2*>(derivatives_prev);
    for (int k = 0; k < 3; ++k) nuif_det::sincosf_det(euler3[k], &a.es[k], &a.ec[k]);
    a.distThr = thresholds4[0], a.angleThr = thresholds4[1], a.colorThr = thresholds4[2], a.gradMin = thresholds4[3];
    ColorIcpSync* sy = reinterpret_cast<ColorIcpSync*>(workspace);
    float* partials = reinterpret_cast<float*>(static_cast<char*>(workspace) + kCSyncBytes);
    hipStream_t st = stream_of(s);
    intensity_sync_clear_kernel<<<1, 256, 0, st>>>(sy);
    NUIF_LAUNCHED("intensity_sync_clear_kernel");
    intensity_icp_kernel<<<(unsigned)div_up(groups, kIntGroups), 128 * kIntGroups, 0, st>>>(a, *cam, level, pose,
                                                                                            prev_pose, partials, sy,
                                                                                            sums29);
    NUIF_LAUNCHED("intensity_icp_kernel");
    return NUIF_OK;
}

int nuif_color_icp_read(const void* state, nuif_color_icp_result* out, nuif_stream_t s) {
    NUIF_REQUIRE_INIT();
    NUIF_CHECK_ARG(state && out, "nuif_color_icp_read: null argument");
    ColorIcpState h;
    NUIF_HIP(hipMemcpyAsync(&h, state, sizeof(h), hipMemcpyDeviceToHost, stream_of(s)));
    NUIF_HIP(hipStreamSynchronize(stream_of(s)));
    out->ok = h.failed ? 0 : 1;
    out->iterations = h.iterations;
    out->error = h.error;
    out->count = h.count;
    for (int k = 0; k < 9; ++k) out->R[k] = h.R[k];
    for (int k = 0; k < 3; ++k) out->t[k] = h.t[k];
    for (int k = 0; k < kCTerms; ++k) out->sums[k] = h.sums[k];
    return NUIF_OK;
}

}  // extern "C"
