// TEST INFRASTRUCTURE ONLY (see oracle_common.h).  CPU restatement of the
// reference's photometric (colour) tracker, OCL/NuiKinfuOpenCLColorTracker.cpp
// with the kernels of K/colorIcp.cl and K/transformMaps.cl:405-446,513-563, in
// float arithmetic that follows the kernels operation by operation.
//
// Where the reference is undefined, the defined result used here:
//   * float -> uchar conversions of the gradients saturate (negative -> 0), as
//     the GPU's conversion does (convert_uchar of an out-of-range value is
//     undefined in OpenCL C);
//   * the gradient kernel's neighbours are taken one row (the image width)
//     apart and the result is stored at the centre pixel (the reference steps
//     by the image height and stores at its last neighbour, racing with the
//     border writes);
//   * the colour difference is summed exactly (integers) instead of with
//     float atomics in an arbitrary order;
//   * a normal matrix that is not positive definite fails the track, as a
//     zero determinant does (Eigen's LLT of such a matrix returns an
//     unspecified solution).
// Two host decodes of the 29 sums:
//   * reference (OCL/NuiKinfuOpenCLColorTracker.cpp:469-526): the sums read
//     in the upper-row order of the depth tracker although the kernel writes
//     them in lower-row order with b interleaved; result(0..2) taken as the
//     Euler angles although the kernel's parameters 0..2 are translations;
//     the pose (camera -> world) applied to world points;
//     with three parameters (rotation only) the count is read from a slot the
//     kernel leaves at zero, so that mode stops at once;
//   * corrected: the sums decoded in the kernel's order, the pose applied
//     world -> camera, signed gradients, the Gauss-Newton step -A^-1 b in the
//     kernel's parameter order, composed on the world -> camera transform.
// fast_length is evaluated correctly rounded: parity at its last bit is
// unpinned.

#include "kinfu_core.h"

#include <cfloat>

namespace coloror {
using namespace oracle;

// float -> int as the GPU converts: round toward zero, saturating, NaN -> 0
inline int cvt_sat(float v) {
    if (v != v) return 0;
    if (v >= 2147483648.0f) return INT32_MAX;
    if (v <= -2147483648.0f) return INT32_MIN;
    return (int)v;
}
inline uint8_t u8_sat(float v) {
    if (!(v > 0.0f)) return 0;
    if (v >= 255.0f) return 255;
    return (uint8_t)v;
}
inline int8_t i8_sat(float v) {
    if (v != v) return 0;
    if (v >= 127.0f) return 127;
    if (v <= -128.0f) return -128;
    return (int8_t)v;
}

// K/transformMaps.cl:405-446 half_sample_bgra_kernel: integer mean of the
// (2r+1)^2 window around (2x, 2y) clipped to the image, alpha 255
void half_sample_bgra(const uint8_t* src, uint8_t* dst, int Wd, int Hd, int radius) {
    const int Ws = Wd * 2, Hs = Hd * 2;
    for (int y = 0; y < Hd; ++y)
        for (int x = 0; x < Wd; ++x) {
            const int sx = 2 * x, sy = 2 * y;
            const int x0 = std::max(sx - radius, 0), x1 = std::min(sx + radius + 1, Ws);
            const int y0 = std::max(sy - radius, 0), y1 = std::min(sy + radius + 1, Hs);
            int sum[4] = {0, 0, 0, 0}, n = 0;
            for (int cy = y0; cy < y1; ++cy)
                for (int cx = x0; cx < x1; ++cx) {
                    const uint8_t* p = src + 4 * ((size_t)cy * Ws + cx);
                    for (int c = 0; c < 4; ++c) sum[c] += p[c];
                    ++n;
                }
            uint8_t* o = dst + 4 * ((size_t)y * Wd + x);
            for (int c = 0; c < 3; ++c) o[c] = (uint8_t)(n > 0 ? sum[c] / n : 0);
            o[3] = 255;
        }
}

// K/transformMaps.cl:513-563 bgra_gradient_kernel: Sobel / 8 per channel,
// alpha 255 inside, zero on the border.  sign: int8 instead of uint8 values.
void bgra_gradient(const uint8_t* c, uint8_t* gx, uint8_t* gy, int W, int H, bool sign) {
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            const size_t i = (size_t)y * W + x;
            uint8_t* ox = gx + 4 * i;
            uint8_t* oy = gy + 4 * i;
            if (!(x > 0 && x < W - 1 && y > 0 && y < H - 1)) {
                std::memset(ox, 0, 4);
                std::memset(oy, 0, 4);
                continue;
            }
            auto at = [&](int dx, int dy, int ch) { return (float)c[4 * ((size_t)(y + dy) * W + x + dx) + ch]; };
            for (int ch = 0; ch < 3; ++ch) {
                float tx = (at(1, -1, ch) - at(-1, -1, ch)) + 2 * (at(1, 0, ch) - at(-1, 0, ch)) +
                           (at(1, 1, ch) - at(-1, 1, ch));
                float ty = (at(-1, 1, ch) - at(-1, -1, ch)) + 2 * (at(0, 1, ch) - at(0, -1, ch)) +
                           (at(1, 1, ch) - at(1, -1, ch));
                tx /= 8.0f;
                ty /= 8.0f;
                ox[ch] = sign ? (uint8_t)i8_sat(tx) : u8_sat(tx);
                oy[ch] = sign ? (uint8_t)i8_sat(ty) : u8_sat(ty);
            }
            ox[3] = oy[3] = 255;
        }
}

struct Level {
    float fx, fy, cx, cy;
    int W, H;
};
inline Level level_of(const Cam<float>& cam, int div) {
    return Level{cam.fx / (float)div, cam.fy / (float)div, cam.cx / (float)div, cam.cy / (float)div, cam.W / div,
                 cam.H / div};
}
// round() then the int conversion of the projected pixel; false outside the image
inline bool project(const Level& L, V3<float> q, int* idx) {
    const int px = cvt_sat(std::round(q.x * L.fx / q.z + L.cx));
    const int py = cvt_sat(std::round(q.y * L.fy / q.z + L.cy));
    if (px < 0 || px >= L.W || py < 0 || py >= L.H) return false;
    *idx = py * L.W + px;
    return true;
}

// K/colorIcp.cl:7-57 color_difference_kernel over all points: exact sums of
// |obs - known|^2 and of the valid count
void color_difference(const Cam<float>& cam, int div, const float* vmap, const uint8_t* colors, const uint8_t* lc,
                      const float R[9], const float t[3], long long* dist, long long* count) {
    const Level L = level_of(cam, div);
    const size_t n = (size_t)cam.W * cam.H;
    long long d = 0, k = 0;
    for (size_t i = 0; i < n; ++i) {
        const V3<float> p = load3(vmap, i);
        const uint8_t* known = colors + 4 * i;
        if (isnan3(p) || known[3] < 254) continue;
        const V3<float> q = rotate(R, p) + mk(t[0], t[1], t[2]);
        int li;
        if (!project(L, q, &li)) continue;
        const uint8_t* obs = lc + 4 * (size_t)li;
        if (obs[3] < 254) continue;
        for (int c = 0; c < 3; ++c) {
            const long long e = (long long)obs[c] - known[c];
            d += e * e;
        }
        ++k;
    }
    *dist = d;
    *count = k;
}

// K/colorIcp.cl:92-188 color_icp_block_kernel, one point: the terms in the
// kernel's order (per parameter: its products with the parameters before it
// and itself, then its product with the colour difference), |diff|, 1
void color_terms(const Level& L, V3<float> p, const uint8_t* known, const uint8_t* lc, const uint8_t* lgx,
                 const uint8_t* lgy, const float R[9], const float t[3], int numPara, int startPara, bool sign,
                 float out[kTerms]) {
    for (int k = 0; k < kTerms; ++k) out[k] = 0.f;
    if (isnan3(p) || known[3] != 255) return;
    const V3<float> q = rotate(R, p) + mk(t[0], t[1], t[2]);
    int li;
    if (!project(L, q, &li)) return;
    const uint8_t* gx = lgx + 4 * (size_t)li;
    const uint8_t* gy = lgy + 4 * (size_t)li;
    const uint8_t* obs = lc + 4 * (size_t)li;
    if (!(obs[3] == 255 && gx[3] == 255 && gy[3] == 255)) return;
    auto g = [&](const uint8_t* v, int c) { return sign ? (float)(int8_t)v[c] : (float)v[c]; };
    const V3<float> diff = mk((float)obs[0] - (float)known[0], (float)obs[1] - (float)known[1],
                              (float)obs[2] - (float)known[2]);
    V3<float> d[6];
    int k = 0;
    for (int para = 0; para < numPara; ++para) {
        V3<float> dp;
        switch (para + startPara) {
            case 0: dp = mk(1.0f, 0.0f, 0.0f); break;
            case 1: dp = mk(0.0f, 1.0f, 0.0f); break;
            case 2: dp = mk(0.0f, 0.0f, 1.0f); break;
            case 3: dp = mk(0.0f, -q.z, q.y); break;
            case 4: dp = mk(q.z, 0.0f, -q.x); break;
            default: dp = mk(-q.y, q.x, 0.0f); break;
        }
        const float zz = q.z * q.z;
        const float jx = L.fx * ((q.z * dp.x - dp.z * q.x) / zz);
        const float jy = L.fy * ((q.z * dp.y - dp.z * q.y) / zz);
        d[para] = mk(jx * g(gx, 0) + jy * g(gy, 0), jx * g(gx, 1) + jy * g(gy, 1), jx * g(gx, 2) + jy * g(gy, 2));
        for (int col = 0; col <= para; ++col) out[k++] = dot(d[para], d[col]);
        out[k++] = dot(d[para], diff);
    }
    out[k++] = std::sqrt(diff.x * diff.x + diff.y * diff.y + diff.z * diff.z);
    out[k++] = 1.0f;
}

// color_icp_block_kernel (128-wide trees) -> compute_color_sums (128-wide trees
// over groups); returns the second-level blocks that the host sums in order
int color_reduce(const Cam<float>& cam, int div, const float* vmap, const uint8_t* colors, const uint8_t* lc,
                 const uint8_t* lgx, const uint8_t* lgy, const float R[9], const float t[3], int numPara,
                 int startPara, bool sign, std::vector<float>& blocks) {
    const Level L = level_of(cam, div);
    const int npx = cam.W * cam.H;
    const int groups = (npx + 127) / 128;
    std::vector<float> gsum((size_t)groups * kTerms);
#pragma omp parallel
    {
        std::vector<float> buf(128 * kTerms);
#pragma omp for schedule(static)
        for (int gi = 0; gi < groups; ++gi) {
            for (int l = 0; l < 128; ++l) {
                const int p = gi * 128 + l;
                float* o = &buf[l * kTerms];
                if (p < npx)
                    color_terms(L, load3(vmap, p), colors + 4 * (size_t)p, lc, lgx, lgy, R, t, numPara, startPara,
                                sign, o);
                else
                    for (int k = 0; k < kTerms; ++k) o[k] = 0.f;
            }
            tree128(buf.data());
            std::memcpy(&gsum[(size_t)gi * kTerms], buf.data(), kTerms * sizeof(float));
        }
    }
    const int nblocks = (groups + 127) / 128;
    blocks.assign((size_t)nblocks * kTerms, 0.f);
    std::vector<float> buf(128 * kTerms);
    for (int b = 0; b < nblocks; ++b) {
        for (int l = 0; l < 128; ++l) {
            const int gi = b * 128 + l;
            for (int k = 0; k < kTerms; ++k) buf[l * kTerms + k] = gi < groups ? gsum[(size_t)gi * kTerms + k] : 0.f;
        }
        tree128(buf.data());
        std::memcpy(&blocks[(size_t)b * kTerms], buf.data(), kTerms * sizeof(float));
    }
    return nblocks;
}

// Slot of the kernel's term (para p, col c <= p; c = p + 1 is the product with the
// colour difference)
inline int kernel_slot(int p, int c) { return p * (p + 3) / 2 + c; }

struct ColorIcpResult {
    int ok, iterations;
    float error, count;
};

// OCL/NuiKinfuOpenCLColorTracker.cpp:341-578 ColorIterativeClosestPoint from
// the pose (R, t), which it updates.  pyr: per level colours, gradient x, y.
ColorIcpResult color_icp(const Cam<float>& cam, int levels, const int* iters, const int* rot_only, float dist_thr,
                         const float* vmap, const uint8_t* colors, const uint8_t* const* pc,
                         const uint8_t* const* pgx, const uint8_t* const* pgy, float R[9], float t[3],
                         bool corrected) {
    ColorIcpResult res{1, 0, 0.f, 0.f};
    const float numPoints = (float)(cam.W * cam.H);
    float Rk[9], tk[3];  // the transform the kernel applies to the model points
    if (corrected) {
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) Rk[3 * i + j] = R[3 * j + i];
        float tn[3];
        matvec3(Rk, t, tn);
        for (int i = 0; i < 3; ++i) tk[i] = -tn[i];
    } else {
        std::memcpy(Rk, R, sizeof(Rk));
        std::memcpy(tk, t, sizeof(tk));
    }
    std::vector<float> blocks;
    for (int level = levels - 1; level >= 0; --level) {
        const int div = 1 << level;
        const bool ro = rot_only[level] != 0;
        const int numPara = ro ? 3 : 6, startPara = ro ? 3 : 0;
        for (int iter = 0; iter < iters[level]; ++iter) {
            const int nb = color_reduce(cam, div, vmap, colors, pc[level], pgx[level], pgy[level], Rk, tk, numPara,
                                        startPara, corrected, blocks);
            res.iterations++;
            // the reference reads |diff| and the count from slots 27, 28; the kernel
            // writes them after its terms (slots 9, 10 with three parameters)
            const int es = corrected ? numPara * (numPara + 3) / 2 : kTerms - 2;
            float err = 0.f, cnt = 0.f;
            for (int n = 0; n < nb; ++n) {
                err += blocks[(size_t)n * kTerms + es];
                cnt += blocks[(size_t)n * kTerms + es + 1];
            }
            res.error = cnt > 0.f ? std::sqrt(err) / cnt : FLT_MAX;
            res.count = cnt;
            if (cnt < 1.f) break;
            const float scale = numPoints / cnt;
            auto slot_sum = [&](int s) {
                float v = 0.f;
                for (int n = 0; n < nb; ++n) v += blocks[(size_t)n * kTerms + s];
                return v;
            };
            double A[36], b[6];
            if (!corrected) {
                for (int i = 0, shift = 0; i < 6; ++i)
                    for (int j = i; j < 7; ++j, ++shift) {
                        const float value = slot_sum(shift) * scale;
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
                        const float value = slot_sum(kernel_slot(p, c)) * scale;
                        A[c * 6 + p] = A[p * 6 + c] = value;
                    }
                    b[p] = slot_sum(kernel_slot(p, p + 1)) * scale;
                }
                for (int p = numPara; p < 6; ++p) A[p * 6 + p] = 1.0;  // unused parameters: x = 0
            }
            double x[6], det;
            const bool spd = cholesky6(A, b, x, &det);
            if (!spd) det = det6(A);
            if (!spd || std::fabs(det) < 1e-15 || std::isnan(det)) {
                res.ok = 0;
                return res;
            }
            float rot[3], tinc[3];
            if (!corrected) {
                rot[0] = (float)x[0], rot[1] = (float)x[1], rot[2] = (float)x[2];
                tinc[0] = (float)x[3], tinc[1] = (float)x[4], tinc[2] = (float)x[5];
                if (std::fabs(rot[0]) < dist_thr && std::fabs(rot[1]) < dist_thr && std::fabs(rot[2]) < dist_thr)
                    break;
            } else {
                // Gauss-Newton step -A^-1 b in the kernel's order (translation, rotation)
                const int r0 = ro ? 0 : 3;
                rot[0] = (float)x[r0], rot[1] = (float)x[r0 + 1], rot[2] = (float)x[r0 + 2];
                tinc[0] = ro ? 0.f : -(float)x[0];
                tinc[1] = ro ? 0.f : -(float)x[1];
                tinc[2] = ro ? 0.f : -(float)x[2];
                float m = 0.f;
                for (int k = 0; k < numPara; ++k) m = std::max(m, std::fabs((float)x[k]));
                if (m < 1e-6f) break;
            }
            float Rz[9], Ry[9], Rx[9], Rzy[9], Rinc[9], Ri[9], tn[3];
            axis_rot(rot[2], 2, Rz);
            axis_rot(rot[1], 1, Ry);
            axis_rot(rot[0], 0, Rx);
            matmul3(Rz, Ry, Rzy);
            matmul3(Rzy, Rx, Rinc);
            inverse3(Rinc, Ri);
            matvec3(Ri, tk, tn);
            for (int i = 0; i < 3; ++i) tk[i] = tn[i] + tinc[i];
            matmul3(Ri, Rk, Rk);
        }
    }
    if (corrected) {
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) R[3 * i + j] = Rk[3 * j + i];
        float tn[3];
        matvec3(R, tk, tn);
        for (int i = 0; i < 3; ++i) t[i] = -tn[i];
    } else {
        std::memcpy(R, Rk, sizeof(Rk));
        std::memcpy(t, tk, sizeof(tk));
    }
    return res;
}

// ---- joint depth + intensity ICP (K/icp.cl:328-480, no host in the reference) ----

// K/transformMaps.cl:372-384 bgra_to_intensity_kernel
void bgra_to_intensity(const uint8_t* bgra, float* out, size_t n) {
    for (size_t i = 0; i < n; ++i) {
        const uint8_t* c = bgra + 4 * i;
        out[i] = c[3] > 0 ? 0.299f * ((float)c[2] / 255.0f) + 0.587f * ((float)c[1] / 255.0f) +
                                0.114f * ((float)c[0] / 255.0f)
                          : qnan<float>();
    }
}

// K/transformMaps.cl:448-511 intensity_derivatives_kernel with the neighbours one row
// (the width) apart and the result at the centre; NaN where the reference writes
// nothing (border, NaN centre or neighbour).  out: float2 per pixel.
void intensity_derivatives(const float* I, float* out, int W, int H) {
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            const size_t c = (size_t)y * W + x;
            out[2 * c] = out[2 * c + 1] = qnan<float>();
            if (std::isnan(I[c]) || !(x > 0 && x < W - 1 && y > 0 && y < H - 1)) continue;
            auto at = [&](int dx, int dy) { return I[(size_t)(y + dy) * W + x + dx]; };
            const float i00 = at(-1, -1), i10 = at(0, -1), i20 = at(1, -1);
            const float i01 = at(-1, 0), i21 = at(1, 0);
            const float i02 = at(-1, 1), i12 = at(0, 1), i22 = at(1, 1);
            if (std::isnan(i00) || std::isnan(i10) || std::isnan(i20) || std::isnan(i01) || std::isnan(i21) ||
                std::isnan(i02) || std::isnan(i12) || std::isnan(i22))
                continue;
            float dx = (-1.0f) * i00 + (1.0f) * i20 + (-2.0f) * i01 + (2.0f) * i21 + (-1.0f) * i02 + (1.0f) * i22;
            float dy = (-1.0f) * i00 + (-2.0f) * i10 + (-1.0f) * i20 + (1.0f) * i02 + (2.0f) * i12 + (1.0f) * i22;
            out[2 * c] = dx / 8.0f;
            out[2 * c + 1] = dy / 8.0f;
        }
}

// K/icp.cl:328-350 evalR_dAlpha / evalR_dBeta / evalR_dGamma as written
void eval_phis(const float e[3], V3<float> p, V3<float>* pa, V3<float>* pb, V3<float>* pg) {
    float sx, cx, sy, cy, sz, cz;
    sincos_t(e[0], &sx, &cx);
    sincos_t(e[1], &sy, &cy);
    sincos_t(e[2], &sz, &cz);
    *pa = mk(dot(mk(-sx * cy, -cx * cz - sx * sy * sz, cx * sz - sx * sy * cz), p),
             dot(mk(cx * cy, -sx * cz + cx * sy * sz, sx * sz + cx * sy * cz), p), 0.0f);
    *pb = mk(dot(mk(-cx * sy, cx * cy * sz, cx * cy * cz), p), dot(mk(-sx * sy, sx * cy * sz, sx * cy * sz), p),
             dot(mk(-cy, -sy * sz, -sy * cz), p));
    *pg = mk(dot(mk(0.0f, sx * sz + sx * cy * sz, sx * cz - cx * sy * sz), p),
             dot(mk(0.0f, -cx * sz + sx * sy * cz, -cx * cz - sx * sy * sz), p), dot(mk(0.0f, cy * cz, -cy * sz), p));
}

struct IntensityArgs {
    Level L;
    const float *v, *n, *I;              // current level maps
    const float *vp, *np, *Ip, *dp;      // previous level maps (dp: float2)
    Rigid<float> cur, prev;
    float euler[3];
    float distThr, angleThr, colorThr, gradMin;
};

// K/icp.cl:352-480 intensity_icp_block_kernel, one pixel
void intensity_terms(const IntensityArgs& a, size_t gid, float out[kTerms]) {
    for (int k = 0; k < kTerms; ++k) out[k] = 0.f;
    const V3<float> nc = load3(a.n, gid);
    if (isnan3(nc)) return;
    const V3<float> vc = load3(a.v, gid);
    if (isnan3(vc)) return;
    const V3<float> pv = transform(a.cur, vc);
    const V3<float> pp = transform_inverse(a.prev, pv);
    int ref;
    if (!project(a.L, pp, &ref)) return;
    const V3<float> rn = load3(a.np, (size_t)ref);
    const float ri = a.Ip[ref];
    if (isnan3(rn) || std::isnan(ri)) return;
    const V3<float> rv = load3(a.vp, (size_t)ref);
    if (isnan3(rv)) return;
    const V3<float> diff = mk(std::fabs(rv.x - pv.x), std::fabs(rv.y - pv.y), std::fabs(rv.z - pv.z));
    const float dist = std::sqrt(diff.x * diff.x + diff.y * diff.y + diff.z * diff.z);
    if (!(dist < a.distThr)) return;
    V3<float> pa, pb, pg;
    eval_phis(a.euler, pp, &pa, &pb, &pg);
    const V3<float> pn = rotate(a.cur.R, nc);
    const V3<float> cr = cross(pn, rn);
    const float sine = std::sqrt(cr.x * cr.x + cr.y * cr.y + cr.z * cr.z);
    if (!(sine < a.angleThr)) return;
    const float wd = std::max(0.0f, 1.0f - dist / a.distThr);
    const V3<float> mn = mk(-rn.x, -rn.y, -rn.z);
    float c[7] = {wd * dot(mn, pa), wd * dot(mn, pb), wd * dot(mn, pg), wd * rn.x, wd * rn.y, wd * rn.z,
                  wd * dot(rn, diff)};
    const float ic = a.I[gid];
    if (!std::isnan(ic)) {
        const float di = std::fabs(ic - ri);
        const float gx = a.dp[2 * (size_t)ref], gy = a.dp[2 * (size_t)ref + 1];
        if (di < a.colorThr && !std::isnan(gx) && !std::isnan(gy) && std::sqrt(gx * gx + gy * gy) > a.gradMin) {
            const float wc = std::max(0.0f, 1.0f - di / a.colorThr);
            const V3<float> t0 = mk(gx * a.L.fx / pp.z, gy * a.L.fy / pp.z,
                                    (gx * a.L.cx + gy * a.L.cy) / pp.z - (gx * pp.x + gy * pp.y) / (pp.z * pp.z));
            c[0] += wc * dot(t0, pa);
            c[1] += wc * dot(t0, pb);
            c[2] += wc * dot(t0, pg);
            c[3] += wc * t0.x;
            c[4] += wc * t0.y;
            c[5] += wc * t0.z;
            c[6] += wc * di;
        }
    }
    int k = 0;
    for (int i = 0; i < 6; ++i)
        for (int j = i; j < 7; ++j) out[k++] = c[i] * c[j];
    if (dist > a.distThr / 100.0f && sine > a.angleThr / 100.0f) {
        out[k++] = c[6] * c[6];
        out[k++] = 1.0f;
    }
}

// The 128-wide block trees, compute_sums' trees over 128 groups and the host's
// sequential sum (the depth tracker's reduction, K/icp.cl:162-190, 294-326)
void intensity_sums(const IntensityArgs& a, int npx, float out[kTerms]) {
    const int groups = (npx + 127) / 128;
    std::vector<float> gsum((size_t)groups * kTerms);
#pragma omp parallel
    {
        std::vector<float> buf(128 * kTerms);
#pragma omp for schedule(static)
        for (int gi = 0; gi < groups; ++gi) {
            for (int l = 0; l < 128; ++l) {
                const int p = gi * 128 + l;
                float* o = &buf[l * kTerms];
                if (p < npx)
                    intensity_terms(a, (size_t)p, o);
                else
                    for (int k = 0; k < kTerms; ++k) o[k] = 0.f;
            }
            tree128(buf.data());
            std::memcpy(&gsum[(size_t)gi * kTerms], buf.data(), kTerms * sizeof(float));
        }
    }
    const int nblocks = (groups + 127) / 128;
    std::vector<float> buf(128 * kTerms);
    for (int k = 0; k < kTerms; ++k) out[k] = 0.f;
    for (int b = 0; b < nblocks; ++b) {
        for (int l = 0; l < 128; ++l) {
            const int gi = b * 128 + l;
            for (int k = 0; k < kTerms; ++k) buf[l * kTerms + k] = gi < groups ? gsum[(size_t)gi * kTerms + k] : 0.f;
        }
        tree128(buf.data());
        for (int k = 0; k < kTerms; ++k) out[k] += buf[k];
    }
}

}  // namespace coloror

using namespace coloror;

#define OR_EXPORT extern "C" __attribute__((visibility("default")))

namespace {
Cam<float> cam_of(const float* cam6, int W, int H) {
    Cam<float> c;
    c.fx = cam6[0];
    c.fy = cam6[1];
    c.fx_inv = 1.0f / c.fx;
    c.fy_inv = 1.0f / c.fy;
    c.cx = cam6[2];
    c.cy = cam6[3];
    c.W = W;
    c.H = H;
    c.dmin = cam6[4];
    c.dmax = cam6[5];
    return c;
}
}  // namespace

OR_EXPORT void or_color_half_sample(const uint8_t* src, uint8_t* dst, int Wd, int Hd, int radius) {
    half_sample_bgra(src, dst, Wd, Hd, radius);
}

OR_EXPORT void or_color_gradient(const uint8_t* c, uint8_t* gx, uint8_t* gy, int W, int H, int sign) {
    bgra_gradient(c, gx, gy, W, H, sign != 0);
}

// GetColorsDifference's sums: out2 = {sum |obs - known|^2, valid count}
OR_EXPORT void or_color_difference(int W, int H, const float* cam6, int level, const float* vmap,
                                   const uint8_t* colors, const uint8_t* lc, const float* R9, const float* t3,
                                   long long* out2) {
    color_difference(cam_of(cam6, W, H), 1 << level, vmap, colors, lc, R9, t3, &out2[0], &out2[1]);
}

// One pixel's 29 terms (kernel order) of color_icp_block_kernel.
OR_EXPORT void or_color_terms(int W, int H, const float* cam6, int level, const float* p3, const uint8_t* known4,
                              const uint8_t* lc, const uint8_t* lgx, const uint8_t* lgy, const float* R9,
                              const float* t3, int rot_only, int sign, float* out29) {
    const Level L = level_of(cam_of(cam6, W, H), 1 << level);
    color_terms(L, mk(p3[0], p3[1], p3[2]), known4, lc, lgx, lgy, R9, t3, rot_only ? 3 : 6, rot_only ? 3 : 0,
                sign != 0, out29);
}

// One iteration's reduction: the second-level blocks (nblocks x 29); returns nblocks.
OR_EXPORT int or_color_reduce(int W, int H, const float* cam6, int level, const float* vmap, const uint8_t* colors,
                              const uint8_t* lc, const uint8_t* lgx, const uint8_t* lgy, const float* R9,
                              const float* t3, int rot_only, int sign, float* blocks_out, int cap) {
    std::vector<float> blocks;
    const int nb = color_reduce(cam_of(cam6, W, H), 1 << level, vmap, colors, lc, lgx, lgy, R9, t3,
                                rot_only ? 3 : 6, rot_only ? 3 : 0, sign != 0, blocks);
    if (blocks_out) std::memcpy(blocks_out, blocks.data(), sizeof(float) * std::min<size_t>(blocks.size(), (size_t)cap * kTerms));
    return nb;
}

// The whole tracker from a BGRA frame: pyramid (radius 1), gradients, and the
// level loop from pose (R9, t3), updated in place.  info4: ok, iterations, error, count.
OR_EXPORT int or_color_icp(int W, int H, const float* cam6, int levels, const int* iters, const int* rot_only,
                           float dist_thr, const float* vmap, const uint8_t* colors, const uint8_t* frame,
                           float* R9, float* t3, int corrected, double* info4) {
    const Cam<float> cam = cam_of(cam6, W, H);
    std::vector<std::vector<uint8_t>> pc(levels), pgx(levels), pgy(levels);
    for (int l = 0; l < levels; ++l) {
        const size_t n = (size_t)(W >> l) * (H >> l) * 4;
        pc[l].resize(n);
        pgx[l].resize(n);
        pgy[l].resize(n);
    }
    std::memcpy(pc[0].data(), frame, pc[0].size());
    for (int l = 1; l < levels; ++l) half_sample_bgra(pc[l - 1].data(), pc[l].data(), W >> l, H >> l, 1);
    for (int l = 0; l < levels; ++l)
        bgra_gradient(pc[l].data(), pgx[l].data(), pgy[l].data(), W >> l, H >> l, corrected != 0);
    std::vector<const uint8_t*> c(levels), gx(levels), gy(levels);
    for (int l = 0; l < levels; ++l) {
        c[l] = pc[l].data();
        gx[l] = pgx[l].data();
        gy[l] = pgy[l].data();
    }
    const ColorIcpResult r = color_icp(cam, levels, iters, rot_only, dist_thr, vmap, colors, c.data(), gx.data(),
                                       gy.data(), R9, t3, corrected != 0);
    if (info4) {
        info4[0] = r.ok;
        info4[1] = r.iterations;
        info4[2] = r.error;
        info4[3] = r.count;
    }
    return r.ok;
}

OR_EXPORT void or_bgra_to_intensity(const uint8_t* bgra, float* out, long long n) {
    bgra_to_intensity(bgra, out, (size_t)n);
}

OR_EXPORT void or_intensity_derivatives(const float* I, float* out, int W, int H) {
    intensity_derivatives(I, out, W, H);
}

// evalR_dAlpha / dBeta / dGamma at (euler, p): out9 = the three vectors
OR_EXPORT void or_eval_phis(const float* euler3, const float* p3, float* out9) {
    V3<float> a, b, g;
    eval_phis(euler3, mk(p3[0], p3[1], p3[2]), &a, &b, &g);
    const float v[9] = {a.x, a.y, a.z, b.x, b.y, b.z, g.x, g.y, g.z};
    std::memcpy(out9, v, sizeof(v));
}

// The joint kernel's 29 sums at `level` (maps of that level, W >> level by H >> level);
// pose12 / prev12: R (row-major) + t of the current and previous camera poses;
// thr4: dist, angle, colour thresholds and the gradient minimum.
OR_EXPORT void or_intensity_icp_sums(int W, int H, const float* cam6, int level, const float* v, const float* n,
                                     const float* I, const float* vp, const float* np, const float* Ip,
                                     const float* dp, const float* pose12, const float* prev12, const float* euler3,
                                     const float* thr4, float* out29) {
    IntensityArgs a;
    a.L = level_of(cam_of(cam6, W, H), 1 << level);
    a.v = v, a.n = n, a.I = I, a.vp = vp, a.np = np, a.Ip = Ip, a.dp = dp;
    a.cur = make_rigid<float>(pose12, pose12 + 9);
    a.prev = make_rigid<float>(prev12, prev12 + 9);
    std::memcpy(a.euler, euler3, sizeof(a.euler));
    a.distThr = thr4[0], a.angleThr = thr4[1], a.colorThr = thr4[2], a.gradMin = thr4[3];
    intensity_sums(a, a.L.W * a.L.H, out29);
}
