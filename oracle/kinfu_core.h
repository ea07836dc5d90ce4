// TEST INFRASTRUCTURE ONLY (see oracle_common.h).  The shared part of the CPU
// restatement of the reference's fusion path: frame preprocessing, projective
// ICP with its Levenberg-Marquardt host loop, dense TSDF integration and
// raycast.  Templated on the arithmetic type: float follows the reference's
// float arithmetic operation by operation, double is the FP64 oracle.  Each
// function names the reference file:line it follows.
#pragma once

#include "oracle_common.h"

#include <cstdio>

#include "../natureuserinterfacestudio_amd/csrc/det_math.h"
#include "../natureuserinterfacestudio_amd/csrc/se3_math.h"

namespace oracle {

// exp / sin / cos: the float flavour uses the reproducible float versions (the
// vendor functions the reference calls differ in their last bit between
// libraries); the FP64 oracle uses the C library in double.
inline float exp_t(float x) { return nuif_det::expf_det(x); }
inline double exp_t(double x) { return std::exp(x); }
inline void sincos_t(float a, float* s, float* c) { nuif_det::sincosf_det(a, s, c); }
inline void sincos_t(double a, double* s, double* c) {
    *s = std::sin(a);
    *c = std::cos(a);
}

// ---------------------------------------------------------------- preprocess

// K/filters.cl:9-23 depth_passing_filter_kernel: u16 mm -> m, NaN unless min < d < max
template <class T> void passing_filter(const uint16_t* raw, T* out, size_t n, T dmin, T dmax) {
    for (size_t i = 0; i < n; ++i) {
        T d = (T)raw[i] * (T)0.001;
        if (!(d > dmin && d < dmax)) d = qnan<T>();
        out[i] = d;
    }
}

// OCL/NuiKinfuOpenCLFrame.cpp:165-189 GenerateGaussianBuffer: g[i] = exp(-(i-r)^2 * s)
template <class T> void gaussian_table(int radius, T sigma_space, T* g) {
    for (int i = 0; i < 2 * radius + 1; ++i) {
        int x = i - radius;
        g[i] = std::exp(-(T)(x * x) * sigma_space);
    }
}

// K/bilateralFilter.cl:68-119 bilateral_filter_depth_kernel
template <class T>
void bilateral(const T* in, T* out, int W, int H, const T* g, int radius, T sigma_depth, T thr) {
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            T c = in[y * W + x];
            if (std::isnan(c)) {
                out[y * W + x] = qnan<T>();
                continue;
            }
            int x0 = std::max(x - radius, 0), x1 = std::min(x + radius + 1, W);
            int y0 = std::max(y - radius, 0), y1 = std::min(y + radius + 1, H);
            T sd = 0, sw = 0;
            for (int cy = y0; cy < y1; ++cy)
                for (int cx = x0; cx < x1; ++cx) {
                    T n = in[cy * W + cx];
                    if (std::isnan(n)) continue;
                    T diff = std::fabs(c - n);
                    if (n > T(0) && diff < thr) {
                        T d2 = diff * diff;
                        T w = g[cx + radius - x] * g[cy + radius - y] * exp_t(-(d2 * sigma_depth));
                        sd += n * w;
                        sw += w;
                    }
                }
            out[y * W + x] = sd / sw;
        }
}

// K/transformMaps.cl:44-88 half_sample_kernel (source size taken as 2x the destination)
template <class T> void half_sample(const T* src, T* dst, int Wd, int Hd, int radius, T thr) {
    const int Ws = Wd << 1, Hs = Hd << 1;
    for (int y = 0; y < Hd; ++y)
        for (int x = 0; x < Wd; ++x) {
            int sx = x << 1, sy = y << 1;
            T c = src[sy * Ws + sx];
            int x0 = std::max(sx - radius, 0), x1 = std::min(sx + radius + 1, Ws);
            int y0 = std::max(sy - radius, 0), y1 = std::min(sy + radius + 1, Hs);
            T s = 0;
            int cnt = 0;
            for (int cy = y0; cy < y1; ++cy)
                for (int cx = x0; cx < x1; ++cx) {
                    T n = src[cy * Ws + cx];
                    if (!std::isnan(n) && (std::isnan(c) || std::fabs(c - n) < thr)) {
                        s += n;
                        cnt += 1;
                    }
                }
            dst[y * Wd + x] = cnt > 0 ? s / (T)cnt : qnan<T>();
        }
}

// K/transformMaps.cl:18-42 depth2vertex_kernel with div = 2^level
template <class T> void depth2vertex(const T* depth, T* vmap, int W, int H, const Cam<T>& cam, int div) {
    const T fxi = (T)div * cam.fx_inv, fyi = (T)div * cam.fy_inv;
    const T cx = cam.cx / (T)div, cy = cam.cy / (T)div;
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            size_t i = (size_t)y * W + x;
            T d = depth[i];
            V3<T> v = mk(qnan<T>(), qnan<T>(), qnan<T>());
            if (!std::isnan(d)) v = mk(((T)x - cx) * fxi, ((T)y - cy) * fyi, T(1)) * d;
            store3(vmap, i, v);
        }
}

// K/normalEstimation.cl:770-822 estimate_normals_simple_kernel
template <class T> void normals_simple(const T* vmap, T* nmap, int W, int H, T thr) {
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            size_t i = (size_t)y * W + x;
            V3<T> c = load3(vmap, i);
            if (isnan3(c)) {
                store3(nmap, i, mk(qnan<T>(), qnan<T>(), qnan<T>()));
                continue;
            }
            auto pick = [&](bool ok, size_t j) {
                if (!ok) return c;
                V3<T> n = load3(vmap, j);
                if (isnan3(n) || std::fabs(c.z - n.z) > thr) return c;
                return n;
            };
            V3<T> l = pick(x > 0, i - 1), r = pick(x < W - 1, i + 1);
            V3<T> u = pick(y > 0, i - W), d = pick(y < H - 1, i + W);
            store3(nmap, i, normalize3(cross(r - l, d - u)));
        }
}

// K/transformMaps.cl:91-127 transform_maps_kernel
template <class T> void transform_maps(const T* v, const T* n, T* vd, T* nd, size_t npx, const Rigid<T>& g) {
    const V3<T> NaN3 = mk(qnan<T>(), qnan<T>(), qnan<T>());
    for (size_t i = 0; i < npx; ++i) {
        V3<T> vs = load3(v, i);
        if (isnan3(vs)) {
            store3(vd, i, NaN3);
            store3(nd, i, NaN3);
            continue;
        }
        store3(vd, i, transform(g, vs));
        V3<T> ns = load3(n, i);
        store3(nd, i, isnan3(ns) ? NaN3 : rotate(g.R, ns));
    }
}

// K/tsdfVolume.cl:8-29 scaleDepthsKernel (in place): depth -> ray length
template <class T> void scale_depths(T* depth, int W, int H, const Cam<T>& cam) {
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            T d = depth[(size_t)y * W + x];
            if (std::isnan(d)) continue;
            T xl = ((T)x - cam.cx) * cam.fx_inv, yl = ((T)y - cam.cy) * cam.fy_inv;
            T lambda = std::sqrt(xl * xl + yl * yl + T(1));
            depth[(size_t)y * W + x] = d * lambda;
        }
}

// ---------------------------------------------------------------- ICP

constexpr int kTerms = 29;  // 27 unique normal-equation terms + r^2 + count (K/icp.cl:130-160)

// K/icp.cl:192-220 interpolateBilinear_withHoles.  The reference reads p.x+1 /
// p.y+1 without bounds; a linear index inside the buffer reads the
// same element as the reference, one past the buffer is treated as a hole.
template <class T> V3<T> bilinear_holes(const T* src, T px, T py, int W, int H) {
    const V3<T> NaN3 = mk(qnan<T>(), qnan<T>(), qnan<T>());
    int x = (int)std::floor(px), y = (int)std::floor(py);
    T dx = px - (T)x, dy = py - (T)y;
    const long long n = (long long)W * H;
    long long ia = x + (long long)y * W, ib = (x + 1) + (long long)y * W;
    long long ic = x + (long long)(y + 1) * W, id = (x + 1) + (long long)(y + 1) * W;
    if (ia >= n || ib >= n || ic >= n || id >= n) return NaN3;
    V3<T> a = load3(src, ia), b = load3(src, ib), c = load3(src, ic), d = load3(src, id);
    if (isnan3(a) || isnan3(b) || isnan3(c) || isnan3(d)) return NaN3;
    V3<T> r;
    r.x = a.x * (T(1) - dx) * (T(1) - dy) + b.x * dx * (T(1) - dy) + c.x * (T(1) - dx) * dy + d.x * dx * dy;
    r.y = a.y * (T(1) - dx) * (T(1) - dy) + b.y * dx * (T(1) - dy) + c.y * (T(1) - dx) * dy + d.y * dx * dy;
    r.z = a.z * (T(1) - dx) * (T(1) - dy) + b.z * dx * (T(1) - dy) + c.z * (T(1) - dx) * dy + d.z * dx * dy;
    return r;
}

// Where a pixel's association lands: the current vertex at the estimate (pv), in the
// previous camera (pp), and its full-resolution projection (px, py).
template <class T> struct AssocAt {
    V3<T> pv, pp;
    T px, py;
};

// K/icp.cl:222-292 icp_block_kernel, one pixel: fills the 29 terms (zeros if rejected);
// `at` (optional) receives where the association lands once it is projected.
template <class T>
void icp_pixel(const Cam<T>& cam, V3<T> vc, const T R[9], const T t[3], const T* vref, const T* nref,
               const Rigid<T>& prev, T distThr, T out[kTerms], AssocAt<T>* at = nullptr, T* coef = nullptr) {
    for (int k = 0; k < kTerms; ++k) out[k] = 0;
    if (isnan3(vc)) return;
    V3<T> pv = rotate(R, vc) + mk(t[0], t[1], t[2]);
    V3<T> pp = transform_inverse(prev, pv);
    T px = pp.x * cam.fx / pp.z + cam.cx, py = pp.y * cam.fy / pp.z + cam.cy;
    if (at) *at = AssocAt<T>{pv, pp, px, py};
    if (!(px >= 0 && px < (T)cam.W && py >= 0 && py < (T)cam.H)) return;
    V3<T> rn = bilinear_holes(nref, px, py, cam.W, cam.H);
    if (isnan3(rn)) return;
    V3<T> rv = bilinear_holes(vref, px, py, cam.W, cam.H);
    if (isnan3(rv)) return;
    V3<T> diff = rv - pv;
    T dist = std::sqrt(dot(diff, diff));
    if (!(dist < distThr)) return;
    V3<T> row0 = cross(rn, pv);
    T c[7] = {row0.x, row0.y, row0.z, rn.x, rn.y, rn.z, dot(rn, diff)};
    if (coef)
        for (int k = 0; k < 7; ++k) coef[k] = c[k];
    int s = 0;
    for (int i = 0; i < 6; ++i)
        for (int j = i; j < 7; ++j) out[s++] = c[i] * c[j];
    out[s++] = c[6] * c[6];
    out[s++] = T(1);
}

// In-place 128-wide tree of K/icp.cl:162-190 reduce(): buf holds 128 x 29 values
template <class T> void tree128(T* buf) {
    for (int d = 64; d > 0; d >>= 1)
        for (int l = 0; l < d; ++l)
            for (int i = 0; i < kTerms; ++i) buf[l * kTerms + i] += buf[(l + d) * kTerms + i];
}

// One ICP iteration's reduction exactly as the reference orders it:
// icp_block_kernel (128-wide trees per group) -> compute_sums (128-wide trees
// over groups) -> host sequential sum over the second-level blocks
// (OCL/NuiKinfuOpenCLDepthTracker.cpp:318-404).  Returns the per-block
// results in `blocks` (nblocks x 29) for the host loop.
template <class T>
int icp_reduce(const Cam<T>& cam, const T* vcur, int Wl, int Hl, const T R[9], const T t[3], const T* vref,
               const T* nref, const Rigid<T>& prev, T distThr, std::vector<T>& blocks) {
    const int npx = Wl * Hl;
    const int groups = (npx + 127) / 128;
    std::vector<T> gsum((size_t)groups * kTerms);
#pragma omp parallel
    {
        std::vector<T> buf(128 * kTerms);
#pragma omp for schedule(static)
        for (int gi = 0; gi < groups; ++gi) {
            for (int l = 0; l < 128; ++l) {
                int p = gi * 128 + l;
                T* o = &buf[l * kTerms];
                if (p < npx)
                    icp_pixel(cam, load3(vcur, p), R, t, vref, nref, prev, distThr, o);
                else
                    for (int k = 0; k < kTerms; ++k) o[k] = 0;  // a partial group counts its missing items as zeros
            }
            tree128(buf.data());
            std::memcpy(&gsum[(size_t)gi * kTerms], buf.data(), kTerms * sizeof(T));
        }
    }
    const int nblocks = (groups + 127) / 128;
    blocks.assign((size_t)nblocks * kTerms, T(0));
    std::vector<T> buf(128 * kTerms);
    for (int b = 0; b < nblocks; ++b) {
        for (int l = 0; l < 128; ++l) {
            int gi = b * 128 + l;
            for (int k = 0; k < kTerms; ++k) buf[l * kTerms + k] = gi < groups ? gsum[(size_t)gi * kTerms + k] : T(0);
        }
        tree128(buf.data());
        std::memcpy(&blocks[(size_t)b * kTerms], buf.data(), kTerms * sizeof(T));
    }
    return nblocks;
}

// Eigen::AngleAxis<T>(angle, unit axis e_k).toRotationMatrix()
template <class T> void axis_rot(T angle, int k, T Rm[9]) {
    T s, c;
    sincos_t(angle, &s, &c);
    T ax[3] = {0, 0, 0};
    ax[k] = 1;
    T sa[3] = {s * ax[0], s * ax[1], s * ax[2]};
    T ca[3] = {(T(1) - c) * ax[0], (T(1) - c) * ax[1], (T(1) - c) * ax[2]};
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

// 6x6 SPD solve by Cholesky (A = L L^T) in double; det = prod(L_ii)^2.
// Stands in for Eigen's A.determinant() / A.llt().solve(b)
// (OCL/NuiKinfuOpenCLDepthTracker.cpp:465-475).  Returns false if not SPD.
inline bool cholesky6(const double A[36], const double b[6], double x[6], double* det) {
    double L[36] = {0};
    for (int j = 0; j < 6; ++j) {
        double s = A[j * 6 + j];
        for (int k = 0; k < j; ++k) s -= L[j * 6 + k] * L[j * 6 + k];
        if (!(s > 0)) {
            *det = 0;
            return false;
        }
        L[j * 6 + j] = std::sqrt(s);
        for (int i = j + 1; i < 6; ++i) {
            double t = A[i * 6 + j];
            for (int k = 0; k < j; ++k) t -= L[i * 6 + k] * L[j * 6 + k];
            L[i * 6 + j] = t / L[j * 6 + j];
        }
    }
    double d = 1;
    for (int j = 0; j < 6; ++j) d *= L[j * 6 + j];
    *det = d * d;
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
    return true;
}

// Determinant by partially pivoted LU (for the nullspace check when the
// damped matrix is not positive definite).
inline double det6(const double Ain[36]) {
    double A[36];
    std::memcpy(A, Ain, sizeof(A));
    double det = 1;
    for (int c = 0; c < 6; ++c) {
        int p = c;
        for (int r = c + 1; r < 6; ++r)
            if (std::fabs(A[r * 6 + c]) > std::fabs(A[p * 6 + c])) p = r;
        if (A[p * 6 + c] == 0) return 0;
        if (p != c) {
            for (int k = 0; k < 6; ++k) std::swap(A[c * 6 + k], A[p * 6 + k]);
            det = -det;
        }
        det *= A[c * 6 + c];
        for (int r = c + 1; r < 6; ++r) {
            double f = A[r * 6 + c] / A[c * 6 + c];
            for (int k = c; k < 6; ++k) A[r * 6 + k] -= f * A[c * 6 + k];
        }
    }
    return det;
}

struct IcpConfig {
    int levels;
    int iters[8];
    float dist_threshold;
};

struct IcpResult {
    int ok;          // 0: tracking failure (singular system)
    float error;     // m_error of the last iteration
    float count;     // m_count of the last iteration
    int total_iters; // iterations executed (kernel launches)
};

// OCL/NuiKinfuOpenCLDepthTracker.cpp:240-504 IterativeClosestPoint.
// vpyr[l] / (Wl,Hl) are the current frame's vertex pyramid; vref/nref the
// world-space reference maps at full resolution; prev the camera transform
// of the previous frame; R,t in/out the pose estimate.
template <class T>
IcpResult icp(const Cam<T>& cam, const IcpConfig& cfg, const T* const* vpyr, const T* vref, const T* nref,
              const Rigid<T>& prev, T R[9], T t[3]) {
    IcpResult res{1, 0, 0, 0};
    const int L = cfg.levels;
    const T step = (T)cfg.dist_threshold / (T)L;
    const T distThr = (T)cfg.dist_threshold + step;
    size_t kgs = (size_t)cam.W * cam.H;  // kernelGlobalSize[0]; stale across levels
    T goodCores[kTerms] = {0};
    T err = 0, cnt = 0;
    std::vector<T> blocks;
    for (int level = L - 1; level >= 0; --level) {
        const unsigned thr = (unsigned)((double)(kgs >> (2 * level)) * 0.1);
        T lambda = 1;
        T Rg[9], tg[3];
        std::memcpy(Rg, R, sizeof(Rg));
        std::memcpy(tg, t, sizeof(tg));
        T errOld = (T)1e20f;
        const int Wl = cam.W >> level, Hl = cam.H >> level;
        for (int it = 0; it < cfg.iters[level]; ++it) {
            int nb = icp_reduce(cam, vpyr[level], Wl, Hl, R, t, vref, nref, prev, distThr, blocks);
            res.total_iters++;
            const int groups = (Wl * Hl + 127) / 128;
            kgs = (size_t)((groups + 127) / 128) * 128;
            err = 0;
            cnt = 0;
            for (int n = 0; n < nb; ++n) {
                err += blocks[(size_t)n * kTerms + kTerms - 2];
                cnt += blocks[(size_t)n * kTerms + kTerms - 1];
            }
            err = cnt > 0 ? std::sqrt(err) / cnt : std::numeric_limits<T>::max();
            if (cnt < (T)thr || err > errOld) {
                lambda *= T(10);
                std::memcpy(R, Rg, sizeof(Rg));
                std::memcpy(t, tg, sizeof(tg));
                if (it == 0) break;
            } else {
                lambda /= T(10);
                std::memcpy(Rg, R, sizeof(Rg));
                std::memcpy(tg, t, sizeof(tg));
                errOld = err;
                for (int s = 0; s < 27; ++s) {
                    T acc = 0;
                    for (int n = 0; n < nb; ++n) acc += blocks[(size_t)n * kTerms + s];
                    goodCores[s] = acc;
                }
            }
            double A[36], b[6];
            for (int i = 0, s = 0; i < 6; ++i) {
                for (int j = i; j < 7; ++j) {
                    double v = (double)(goodCores[s] / cnt);
                    if (j == 6)
                        b[i] = v;
                    else
                        A[j * 6 + i] = A[i * 6 + j] = v;
                    s++;
                }
                A[i * 6 + i] *= (double)(T(1) + lambda);
            }
            double x[6], det;
            bool spd = cholesky6(A, b, x, &det);
            if (!spd) det = det6(A);
            if (std::fabs(det) < 1e-15 || std::isnan(det) || !spd) {
                res.ok = 0;
                res.error = (float)err;
                res.count = (float)cnt;
                return res;
            }
            T alpha = (T)x[0], beta = (T)x[1], gamma = (T)x[2];
            if (alpha * alpha + beta * beta + gamma * gamma < (T)1e-6f) break;
            T Rz[9], Ry[9], Rx[9], Rzy[9], Rinc[9], RincInv[9];
            axis_rot(gamma, 2, Rz);
            axis_rot(beta, 1, Ry);
            axis_rot(alpha, 0, Rx);
            matmul3(Rz, Ry, Rzy);
            matmul3(Rzy, Rx, Rinc);
            inverse3(Rinc, RincInv);
            T tn[3];
            matvec3(RincInv, t, tn);
            t[0] = tn[0] + (T)x[3];
            t[1] = tn[1] + (T)x[4];
            t[2] = tn[2] + (T)x[5];
            matmul3(RincInv, R, R);
        }
    }
    res.error = (float)err;
    res.count = (float)cnt;
    return res;
}

// ---------------------------------------------------------------- ICP, compensated mode
//
// The GPU's compensated mode (parity unpinned against the reference, which has no
// such mode; its parts are pinned by closed forms in the tests) replaces the
// reference's Levenberg-Marquardt host loop (OCL/NuiKinfuOpenCLDepthTracker.cpp:
// 240-504) by Gauss-Newton with an FP64 pose:
//   * the same association and 29 terms per pixel (icp_pixel), evaluated in T at
//     the pose rounded to T, against the same maps and distance threshold;
//   * the sums accumulated with compensation (float instance: exactly the GPU's
//     order, icp_sums_compensated; FP64 instance: in double, pixel order);
//   * per level, from the coarsest: an iteration whose count is below the
//     reference's (stale) count threshold ends the level without a step; else the
//     undamped normal equations are solved in FP64 (LDL^T) and the increment
//     applied as an SE(3) exponential in FP64 (nuif_se3::gn_step); a step below
//     1e-6 ends the level; a singular system is a tracking failure.

// ---------------------------------------------------------------- ICP, joint order
//
// The joint order (no counterpart in the reference: its photometric terms,
// K/icp.cl:352-480, have no host loop) is the compensated order with a photometric row
// per associated pixel added to the geometric row's products.  Parity unpinned against
// the reference; the rows are pinned by closed forms in the tests (a constant image
// gives no rows, a linear image's row against finite differences of the projection).

// Intensities one level down: the mean of the valid ones of the 2 x 2 source pixels,
// summed in row order, NaN when none is valid (no counterpart in the reference).
template <class T> void intensity_half_sample(const T* src, T* dst, int Wd, int Hd) {
    const int Ws = 2 * Wd;
    for (int y = 0; y < Hd; ++y)
        for (int x = 0; x < Wd; ++x) {
            const size_t i = (size_t)(2 * y) * Ws + 2 * x;
            const T v[4] = {src[i], src[i + 1], src[i + Ws], src[i + Ws + 1]};
            T sum = 0;
            int cnt = 0;
            for (int k = 0; k < 4; ++k)
                if (!std::isnan(v[k])) {
                    sum += v[k];
                    ++cnt;
                }
            dst[(size_t)y * Wd + x] = cnt ? sum / (T)cnt : qnan<T>();
        }
}

// Bilinear samples of a scalar / float2 image with the holes rule of bilinear_holes.
template <class T> T bilinear1(const T* src, T px, T py, int W, int H) {
    const int x = (int)std::floor(px), y = (int)std::floor(py);
    const T dx = px - (T)x, dy = py - (T)y;
    const long long n = (long long)W * H;
    const long long ia = x + (long long)y * W, ib = ia + 1, ic = ia + W, id = ic + 1;
    if (ia < 0 || id >= n) return qnan<T>();
    const T a = src[ia], b = src[ib], c = src[ic], d = src[id];
    if (std::isnan(a) || std::isnan(b) || std::isnan(c) || std::isnan(d)) return qnan<T>();
    return a * (T(1) - dx) * (T(1) - dy) + b * dx * (T(1) - dy) + c * (T(1) - dx) * dy + d * dx * dy;
}
template <class T> bool bilinear2(const T* src, T px, T py, int W, int H, T g[2]) {
    const int x = (int)std::floor(px), y = (int)std::floor(py);
    const T dx = px - (T)x, dy = py - (T)y;
    const long long n = (long long)W * H;
    const long long ia = x + (long long)y * W, ib = ia + 1, ic = ia + W, id = ic + 1;
    if (ia < 0 || id >= n) return false;
    for (int k = 0; k < 2; ++k) {
        const T a = src[2 * ia + k], b = src[2 * ib + k], c = src[2 * ic + k], d = src[2 * id + k];
        if (std::isnan(a) || std::isnan(b) || std::isnan(c) || std::isnan(d)) return false;
    }
    for (int k = 0; k < 2; ++k) {
        const T a = src[2 * ia + k], b = src[2 * ib + k], c = src[2 * ic + k], d = src[2 * id + k];
        g[k] = a * (T(1) - dx) * (T(1) - dy) + b * dx * (T(1) - dy) + c * (T(1) - dx) * dy + d * dx * dy;
    }
    return true;
}

// The photometric inputs of one level: the current intensities at the level, the
// model's intensities and gradients (full resolution), sqrt(weight), the largest
// intensity difference and the squared minimum gradient.
template <class T> struct PhotoIn {
    const T* icur;
    const T* iref;
    const T* gref;
    T psw, pdiff, pgrad2;
};

// The photometric row of an associated pixel: false when it gives none.  ic: the
// current intensity at the pixel.  I(u, v) with u = fx x / z + cx, v = fy y / z + cy in
// the previous camera (p' = Ri (pv - tp)): dI/dp' = (gx fx / z, gy fy / z,
// -(gx fx x + gy fy y) / z^2), carried to world coordinates by Ri^T; the row is
// sqrt(w) [J x pv, J, I - I_ref].
template <class T>
bool photometric_row(T ic, const AssocAt<T>& at, const Rigid<T>& prev, const Cam<T>& cam, const PhotoIn<T>& ph,
                     T c[7]) {
    if (std::isnan(ic)) return false;
    const T ir = bilinear1(ph.iref, at.px, at.py, cam.W, cam.H);
    T g[2];
    const bool gok = bilinear2(ph.gref, at.px, at.py, cam.W, cam.H, g);
    if (std::isnan(ir) || !gok) return false;
    const T di = ic - ir;
    if (!(std::fabs(di) < ph.pdiff) || !(g[0] * g[0] + g[1] * g[1] > ph.pgrad2)) return false;
    const T iz = T(1) / at.pp.z;
    const T ax = g[0] * cam.fx, ay = g[1] * cam.fy;
    const T jc[3] = {ax * iz, ay * iz, -((ax * at.pp.x + ay * at.pp.y) * iz) * iz};
    const T* Ri = prev.R_inv;
    const V3<T> jw = mk((jc[0] * Ri[0] + jc[1] * Ri[3]) + jc[2] * Ri[6], (jc[0] * Ri[1] + jc[1] * Ri[4]) + jc[2] * Ri[7],
                        (jc[0] * Ri[2] + jc[1] * Ri[5]) + jc[2] * Ri[8]);
    const V3<T> a = cross(jw, at.pv);
    const T r[7] = {a.x, a.y, a.z, jw.x, jw.y, jw.z, di};
    for (int k = 0; k < 7; ++k) c[k] = ph.psw * r[k];
    return true;
}

// A pixel's 29 values in the joint order: the geometric products plus, where the
// photometric row exists, its products (the count only for the geometric row).
template <class T>
void joint_pixel(const Cam<T>& cam, V3<T> vc, const T R[9], const T t[3], const T* vref, const T* nref,
                 const Rigid<T>& prev, T distThr, const PhotoIn<T>& ph, int p, T out[kTerms]) {
    AssocAt<T> at;
    T cf[7];
    icp_pixel(cam, vc, R, t, vref, nref, prev, distThr, out, &at, cf);
    if (out[kTerms - 1] == T(0)) return;
    T pc[7];
    if (!photometric_row(ph.icur[p], at, prev, cam, ph, pc)) return;
    int s = 0;
    for (int i = 0; i < 6; ++i)
        for (int j = i; j < 7; ++j, ++s) out[s] = out[s] + pc[i] * pc[j];
    out[s] = out[s] + pc[6] * pc[6];
}

// The GPU's compensated reduction of one iteration (nuif ICP kernel): tiles of T =
// kCompTilePixels pixels; in tile k thread j (256 threads) takes pixels Tk + j + 256i,
// i = 0..T/256-1, with Kahan-compensated FP32 sums (the count plainly); each thread's value s - c
// in double; per term, the values of threads 16g .. 16g + 15 added in thread order
// (g = 0..15), then a tree over the 16 run sums (item l takes item l + d for d = 8, 4,
// 2, 1); then term by term the tile sums k = j, j + 32, j + 64, .. in order for
// j = 0..31, and those 32 in order.
inline void icp_sums_compensated(const Cam<float>& cam, const float* vcur, int npx, const float R[9], const float t[3],
                                 const float* vref, const float* nref, const Rigid<float>& prev, float distThr,
                                 double out[kTerms], const PhotoIn<float>* ph = nullptr) {
    constexpr int T = nuif_se3::kCompTilePixels, PER = T / 256;
    const int ntiles = (npx + T - 1) / T;
    std::vector<double> tiles((size_t)ntiles * kTerms);
#pragma omp parallel for schedule(static)
    for (int k = 0; k < ntiles; ++k) {
        std::vector<double> v(256 * kTerms);
        float o[kTerms];
        for (int j = 0; j < 256; ++j) {
            float sm[kTerms] = {0}, cp[kTerms] = {0};
            for (int i = 0; i < PER; ++i) {
                const int p = T * k + j + 256 * i;
                if (p >= npx) continue;
                if (ph)
                    joint_pixel(cam, load3(vcur, p), R, t, vref, nref, prev, distThr, *ph, p, o);
                else
                    icp_pixel(cam, load3(vcur, p), R, t, vref, nref, prev, distThr, o);
                if (o[kTerms - 1] == 0.f) continue;  // rejected: every term is zero
                for (int q = 0; q < kTerms - 1; ++q) {
                    const float y = o[q] - cp[q];
                    const float tt = sm[q] + y;
                    cp[q] = (tt - sm[q]) - y;
                    sm[q] = tt;
                }
                sm[kTerms - 1] += 1.0f;
            }
            for (int q = 0; q < kTerms; ++q) v[(size_t)j * kTerms + q] = (double)sm[q] - (double)cp[q];
        }
        for (int q = 0; q < kTerms; ++q) {
            double run[16];
            for (int g = 0; g < 16; ++g) {
                double sacc = v[(size_t)(16 * g) * kTerms + q];
                for (int j = 1; j < 16; ++j) sacc += v[(size_t)(16 * g + j) * kTerms + q];
                run[g] = sacc;
            }
            for (int d = 8; d > 0; d >>= 1)
                for (int l = 0; l < d; ++l) run[l] = run[l] + run[l + d];
            tiles[(size_t)k * kTerms + q] = run[0];
        }
    }
    for (int q = 0; q < kTerms; ++q) {
        double seg[32];
        for (int j = 0; j < 32; ++j) {
            double sacc = 0.0;
            for (int k = j; k < ntiles; k += 32) sacc += tiles[(size_t)k * kTerms + q];
            seg[j] = sacc;
        }
        double tot = 0.0;
        for (int j = 0; j < 32; ++j) tot += seg[j];
        out[q] = tot;
    }
}

inline void icp_sums_compensated(const Cam<double>& cam, const double* vcur, int npx, const double R[9],
                                 const double t[3], const double* vref, const double* nref, const Rigid<double>& prev,
                                 double distThr, double out[kTerms], const PhotoIn<double>* ph = nullptr) {
    for (int q = 0; q < kTerms; ++q) out[q] = 0.0;
    double o[kTerms];
    for (int p = 0; p < npx; ++p) {
        if (ph)
            joint_pixel(cam, load3(vcur, p), R, t, vref, nref, prev, distThr, *ph, p, o);
        else
            icp_pixel(cam, load3(vcur, p), R, t, vref, nref, prev, distThr, o);
        for (int q = 0; q < kTerms; ++q) out[q] += o[q];
    }
}

// The compensated mode's iteration loop; (Rd, td) in/out, the pose in FP64.
template <class T>
IcpResult icp_compensated(const Cam<T>& cam, const IcpConfig& cfg, const T* const* vpyr, const T* vref,
                          const T* nref, const Rigid<T>& prev, double Rd[9], double td[3],
                          const PhotoIn<T>* photo = nullptr, const T* const* ipyr = nullptr) {
    IcpResult res{1, 0, 0, 0};
    const int L = cfg.levels;
    const T step = (T)cfg.dist_threshold / (T)L;
    const T distThr = (T)cfg.dist_threshold + step;
    size_t kgs = (size_t)cam.W * cam.H;
    for (int level = L - 1; level >= 0; --level) {
        const unsigned thr = (unsigned)((double)(kgs >> (2 * level)) * 0.1);
        const int Wl = cam.W >> level, Hl = cam.H >> level, npx = Wl * Hl;
        for (int it = 0; it < cfg.iters[level]; ++it) {
            T R[9], t[3];
            for (int k = 0; k < 9; ++k) R[k] = (T)Rd[k];
            for (int k = 0; k < 3; ++k) t[k] = (T)td[k];
            double sums[kTerms];
            PhotoIn<T> ph{};
            if (photo) {  // the joint order: the level's current intensities
                ph = *photo;
                ph.icur = ipyr[level];
            }
            icp_sums_compensated(cam, vpyr[level], npx, R, t, vref, nref, prev, distThr, sums, photo ? &ph : nullptr);
            res.total_iters++;
            const int groups = (npx + 127) / 128;
            kgs = (size_t)((groups + 127) / 128) * 128;
            const double cnt = sums[kTerms - 1];
            res.count = (float)cnt;
            res.error = cnt > 0 ? (float)(std::sqrt(sums[kTerms - 2]) / cnt) : 3.402823466e38f;
            if (std::getenv("OR_ICP_TRACE"))
                std::fprintf(stderr, "  L%d it%d err %.9g cnt %.0f\n", level, it, (double)res.error, cnt);
            if (cnt < (double)thr) break;
            const int r = nuif_se3::gn_step(sums, nuif_se3::kGnStep2, Rd, td);
            if (r == 2) {
                res.ok = 0;
                return res;
            }
            if (r == 1) break;
        }
    }
    return res;
}

// src/Foundation/NuiMatrixUtilities.cpp:7-52 rodrigues2(...).norm(), for an
// (already orthonormal) rotation: 0 for tiny rotations, else the angle.
template <class T> T rodrigues_norm(const T Rm[9]) {
    double rx = Rm[7] - Rm[5], ry = Rm[2] - Rm[6], rz = Rm[3] - Rm[1];
    double s = std::sqrt((rx * rx + ry * ry + rz * rz) * 0.25);
    double c = ((double)Rm[0] + Rm[4] + Rm[8] - 1) * 0.5;
    c = c > 1. ? 1. : c < -1. ? -1. : c;
    double theta = std::acos(c);
    if (s < 1e-5 && c > 0) return 0;
    return (T)theta;
}

// ---------------------------------------------------------------- dense TSDF

// Voxel storage policies.  S16: the reference's short2 (int16 tsdf x 32766 with
// round-toward-zero re-quantisation, int16 weight; K/utils.cl:69-84).  F64:
// the FP64 oracle's exact storage.
struct VolS16 {
    std::vector<int16_t> v;  // 2 per voxel
    void init(size_t n) { v.assign(2 * n, 0); for (size_t i = 0; i < n; ++i) v[2 * i] = 32767; }
    template <class T> T tsdf(size_t i) const { return (T)v[2 * i] / (T)32766; }
    int weight(size_t i) const { return v[2 * i + 1]; }
    template <class T> void set(size_t i, T f, int w) {
        v[2 * i] = (int16_t)(int)(f * (T)32766);  // convert_short: RTZ
        v[2 * i + 1] = (int16_t)w;
    }
    void reset_voxel(size_t i) { v[2 * i] = 32767; v[2 * i + 1] = 0; }  // (SHORT_NAN, 0)
    const void* words() const { return v.data(); }
};
// FP16 tsdf (round to nearest even) + uint16 weight, in the GPU's 32-bit word
// (half bits | weight << 16); FP32 arithmetic.
struct VolF16 {
    std::vector<uint32_t> v;
    void init(size_t n) { v.assign(n, (uint32_t)f16_from_float(1.0f)); }
    template <class T> T tsdf(size_t i) const { return (T)f16_to_float((uint16_t)(v[i] & 0xffffu)); }
    int weight(size_t i) const { return (int)(v[i] >> 16); }
    template <class T> void set(size_t i, T f, int w) {
        v[i] = (uint32_t)f16_from_float((float)f) | ((uint32_t)(uint16_t)w << 16);
    }
    void reset_voxel(size_t i) { v[i] = (uint32_t)f16_from_float(1.0f); }
    const void* words() const { return v.data(); }
};
// FP32 tsdf + FP32 weight (the GPU's float pair).
struct VolF32 {
    std::vector<float> v;  // 2 per voxel
    void init(size_t n) {
        v.assign(2 * n, 0.f);
        for (size_t i = 0; i < n; ++i) v[2 * i] = 1.0f;
    }
    template <class T> T tsdf(size_t i) const { return (T)v[2 * i]; }
    int weight(size_t i) const { return (int)v[2 * i + 1]; }
    template <class T> void set(size_t i, T f, int w) {
        v[2 * i] = (float)f;
        v[2 * i + 1] = (float)w;
    }
    void reset_voxel(size_t i) {
        v[2 * i] = 1.0f;
        v[2 * i + 1] = 0.f;
    }
    const void* words() const { return v.data(); }
};
struct VolF64 {
    std::vector<double> f;
    std::vector<int> w;
    void init(size_t n) { f.assign(n, 32767.0 / 32766.0); w.assign(n, 0); }
    template <class T> T tsdf(size_t i) const { return (T)f[i]; }
    int weight(size_t i) const { return w[i]; }
    template <class T> void set(size_t i, T v, int wt) { f[i] = (double)v; w[i] = wt; }
    void reset_voxel(size_t i) { f[i] = 32767.0 / 32766.0; w[i] = 0; }
    const void* words() const { return nullptr; }
};

// Voxel wrap of a shifting volume (OCL/NuiKinfuOpenCLShiftScene.cpp:44-56 getVoxelWrap): the
// voxel at logical (x, y, z) is stored at ((x + w0) % Nx, (y + w1) % Ny, (z + w2) % Nz)
// (K/tsdfVolume.cl:71,104, K/raycast.cl:62,184).  w >= 0.
struct Wrap3 {
    int w[3] = {0, 0, 0};
};
inline size_t stored_index(int x, int y, int z, int Nx, int Ny, int Nz, const Wrap3& wr) {
    return ((size_t)((y + wr.w[1]) % Ny) * Nx + (size_t)((x + wr.w[0]) % Nx)) * Nz + (size_t)((z + wr.w[2]) % Nz);
}

// K/tsdfVolume.cl:31-190 integrateTsdfVolumeKernel, one column per (x,y) walking z
// with the reference's incremental updates.  depth: ray-length depth (after
// scale_depths); bgra may be NULL.
template <class T, class Vol>
long long integrate(const T* depth, const uint8_t* bgra, Vol& vol, uint8_t* color_vol, const Tsdf<T>& p,
                    const Cam<T>& cam, const Rigid<T>& g, int max_color_weight, const Wrap3& wr = Wrap3()) {
    const int Nx = p.res[0], Ny = p.res[1], Nz = p.res[2];
    const T tr_inv = T(1) / p.trunc;
    long long nupd = 0;
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : nupd)
    for (int y = 0; y < Ny; ++y)
        for (int x = 0; x < Nx; ++x) {
            const T vgx = ((T)(x - Nx / 2) + T(0.5)) * p.cell[0] - g.t[0];
            const T vgy = ((T)(y - Ny / 2) + T(0.5)) * p.cell[1] - g.t[1];
            T vgz = T(0.5) * p.cell[2] - g.t[2];
            const T part = vgx * vgx + vgy * vgy;
            T vx = (g.R_inv[0] * vgx + g.R_inv[1] * vgy + g.R_inv[2] * vgz) * cam.fx;
            T vy = (g.R_inv[3] * vgx + g.R_inv[4] * vgy + g.R_inv[5] * vgz) * cam.fy;
            const T vz = g.R_inv[6] * vgx + g.R_inv[7] * vgy + g.R_inv[8] * vgz;
            T zs = 0;
            const T incx = g.R_inv[2] * p.cell[2] * cam.fx;
            const T incy = g.R_inv[5] * p.cell[2] * cam.fy;
            const size_t base = stored_index(x, y, 0, Nx, Ny, Nz, Wrap3{{wr.w[0], wr.w[1], 0}});
            for (int z = 0; z < Nz; ++z, vgz += p.cell[2], zs += p.cell[2], vx += incx, vy += incy) {
                const T inv_z = T(1) / (vz + g.R_inv[8] * zs);
                if (inv_z < 0) continue;
                const int u = to_int(cl_round(vx * inv_z + cam.cx));
                const int v = to_int(cl_round(vy * inv_z + cam.cy));
                if (!(u >= 0 && v >= 0 && u < cam.W && v < cam.H)) continue;
                const size_t pix = (size_t)v * cam.W + u;
                const T Dp = depth[pix];
                const T sdf = Dp - std::sqrt(vgz * vgz + part);
                if (std::isnan(Dp) || !(sdf >= -p.trunc)) continue;
                const T tsdf = std::min(T(1), sdf * tr_inv);
                const size_t idx = base + (size_t)((z + wr.w[2]) % Nz);
                const T prev = vol.template tsdf<T>(idx);
                const int w = vol.weight(idx);
                const T fnew = (prev * (T)w + T(1) * tsdf) / (T)(w + 1);
                vol.set(idx, fnew, std::min(w + 1, 128));
                ++nupd;
                if (color_vol && bgra && std::fabs(sdf) < p.trunc) {
                    const uint8_t* nc = bgra + 4 * pix;
                    if (nc[3] != 0) {
                        uint8_t* vc = color_vol + 4 * idx;
                        const float Wrk = 1.f;
                        float c[3];
                        for (int k = 0; k < 3; ++k)
                            c[k] = ((float)(vc[k] * vc[3]) + Wrk * (float)nc[k]) / ((float)vc[3] + Wrk);
                        uint8_t wn = (uint8_t)(vc[3] + 1);
                        for (int k = 0; k < 3; ++k) vc[k] = (uint8_t)(int)std::min(std::max(c[k], 0.f), 255.f);
                        vc[3] = (uint8_t)std::min(max_color_weight, (int)wn);
                    }
                }
            }
        }
    return nupd;
}

// K/raycast.cl:22-29 getVoxel: round(p / cell)
template <class T> inline void get_voxel(V3<T> p, const Tsdf<T>& prm, int g[3]) {
    g[0] = to_int(cl_round(p.x / prm.cell[0]));
    g[1] = to_int(cl_round(p.y / prm.cell[1]));
    g[2] = to_int(cl_round(p.z / prm.cell[2]));
}

template <class T, class Vol> inline T sample(const Vol& vol, size_t idx) {
    return vol.weight(idx) > 0 ? vol.template tsdf<T>(idx) : qnan<T>();
}

// K/raycast.cl:36-120 interpolateTrilineary.  The reference names its eight
// corners for an x-major layout while the volume is y-major
// ((y*Nx + x)*Nz + z), so the x and y fractions weight the swapped corners;
// this is reproduced as written.  In a wrapped volume every corner is wrapped
// on its own (intended; the reference wraps the first corner only and steps to
// the others by unwrapped strides, which leaves the cell where it straddles
// the wrap and can read past the end of the volume).
template <class T, class Vol> T trilinear(V3<T> pt, const Vol& vol, const Tsdf<T>& prm, const Wrap3& wr = Wrap3()) {
    int g[3];
    get_voxel(pt, prm, g);
    const int Nx = prm.res[0], Ny = prm.res[1], Nz = prm.res[2];
    if (g[0] <= 0 || g[0] >= Nx - 1) return qnan<T>();
    if (g[1] <= 0 || g[1] >= Ny - 1) return qnan<T>();
    if (g[2] <= 0 || g[2] >= Nz - 1) return qnan<T>();
    const T vx = ((T)g[0] + T(0.5)) * prm.cell[0];
    const T vy = ((T)g[1] + T(0.5)) * prm.cell[1];
    const T vz = ((T)g[2] + T(0.5)) * prm.cell[2];
    g[0] = pt.x < vx ? g[0] - 1 : g[0];
    g[1] = pt.y < vy ? g[1] - 1 : g[1];
    g[2] = pt.z < vz ? g[2] - 1 : g[2];
    const T a = (pt.x - ((T)g[0] + T(0.5)) * prm.cell[0]) / prm.cell[0];
    const T b = (pt.y - ((T)g[1] + T(0.5)) * prm.cell[1]) / prm.cell[1];
    const T c = (pt.z - ((T)g[2] + T(0.5)) * prm.cell[2]) / prm.cell[2];
    // the corner (g0 + dx, g1 + dy, g2 + dz); the reference reaches it from idx by
    // +1 (z), +Nz (x) and +Nx*Nz (y)
    auto at = [&](int dx, int dy, int dz) {
        return sample<T>(vol, stored_index(g[0] + dx, g[1] + dy, g[2] + dz, Nx, Ny, Nz, wr));
    };
    T f000 = at(0, 0, 0); if (std::isnan(f000)) return qnan<T>();
    T f001 = at(0, 0, 1); if (std::isnan(f001)) return qnan<T>();
    T f011 = at(1, 0, 1); if (std::isnan(f011)) return qnan<T>();
    T f010 = at(1, 0, 0); if (std::isnan(f010)) return qnan<T>();
    T f110 = at(1, 1, 0); if (std::isnan(f110)) return qnan<T>();
    T f111 = at(1, 1, 1); if (std::isnan(f111)) return qnan<T>();
    T f101 = at(0, 1, 1); if (std::isnan(f101)) return qnan<T>();
    T f100 = at(0, 1, 0); if (std::isnan(f100)) return qnan<T>();
    return f000 * (1 - a) * (1 - b) * (1 - c) + f001 * (1 - a) * (1 - b) * c + f010 * (1 - a) * b * (1 - c) +
           f011 * (1 - a) * b * c + f100 * a * (1 - b) * (1 - c) + f101 * a * (1 - b) * c +
           f110 * a * b * (1 - c) + f111 * a * b * c;
}

template <class T> inline T min_time(const T* vmax, V3<T> o, V3<T> d) {
    T tx = std::fabs(d.x) < 1e-14 ? T(0) : ((d.x > 0 ? T(0) : vmax[0]) - o.x) / d.x;
    T ty = std::fabs(d.y) < 1e-14 ? T(0) : ((d.y > 0 ? T(0) : vmax[1]) - o.y) / d.y;
    T tz = std::fabs(d.z) < 1e-14 ? T(0) : ((d.z > 0 ? T(0) : vmax[2]) - o.z) / d.z;
    return std::max(std::max(tx, ty), tz);
}
template <class T> inline T max_time(const T* vmax, V3<T> o, V3<T> d) {
    T tx = std::fabs(d.x) < 1e-14 ? T(1e14) : ((d.x > 0 ? vmax[0] : T(0)) - o.x) / d.x;
    T ty = std::fabs(d.y) < 1e-14 ? T(1e14) : ((d.y > 0 ? vmax[1] : T(0)) - o.y) / d.y;
    T tz = std::fabs(d.z) < 1e-14 ? T(1e14) : ((d.z > 0 ? vmax[2] : T(0)) - o.z) / d.z;
    return std::min(std::min(tx, ty), tz);
}

// K/raycast.cl:128-277 raycastKernel.  Outputs world-space vmap/nmap (NaN where
// no surface) and, if both colour buffers are given, cmap.
template <class T, class Vol>
void raycast(const Vol& vol, const uint8_t* color_vol, const Tsdf<T>& prm, const Cam<T>& cam, const Rigid<T>& g,
             T* vmap, T* nmap, uint8_t* cmap, const Wrap3& wr = Wrap3()) {
    const int W = cam.W, H = cam.H;
    const int Nx = prm.res[0], Ny = prm.res[1], Nz = prm.res[2];
#pragma omp parallel for schedule(dynamic, 2)
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            const size_t id = (size_t)y * W + x;
            const V3<T> NaN3 = mk(qnan<T>(), qnan<T>(), qnan<T>());
            store3(vmap, id, NaN3);
            store3(nmap, id, NaN3);
            if (cmap) std::memset(cmap + 4 * id, 0, 4);
            const V3<T> off = mk((T)(prm.dim[0] / 2.0), (T)(prm.dim[1] / 2.0), T(0));
            V3<T> rs = mk(g.t[0], g.t[1], g.t[2]);
            rs = rs + off;
            V3<T> rn = mk(((T)x - cam.cx) * cam.fx_inv, ((T)y - cam.cy) * cam.fy_inv, T(1));
            rn = transform(g, rn);
            rn = rn + off;
            V3<T> dir = normalize3(rn - rs);
            T tcur = min_time(prm.dim, rs, dir);
            const T texit = max_time(prm.dim, rs, dir);
            tcur = std::max(tcur, T(0));
            if (tcur >= texit) continue;
            int gv[3];
            get_voxel(rs + dir * tcur, prm, gv);
            gv[0] = std::max(0, std::min(gv[0], Nx - 1));
            gv[1] = std::max(0, std::min(gv[1], Ny - 1));
            gv[2] = std::max(0, std::min(gv[2], Nz - 1));
            size_t idx = stored_index(gv[0], gv[1], gv[2], Nx, Ny, Nz, wr);
            T tsdf = sample<T>(vol, idx);
            const T max_t = T(3) * (prm.dim[0] + prm.dim[1] + prm.dim[2]);
            const T step = prm.trunc * T(0.8);
            for (; tcur < max_t; tcur += step) {
                const T tprev = tsdf;
                get_voxel(rs + dir * (tcur + step), prm, gv);
                if (!(gv[0] >= 0 && gv[1] >= 0 && gv[2] >= 0 && gv[0] < Nx && gv[1] < Ny && gv[2] < Nz)) break;
                idx = stored_index(gv[0], gv[1], gv[2], Nx, Ny, Nz, wr);
                tsdf = sample<T>(vol, idx);
                if (std::isnan(tsdf) || std::isnan(tprev)) continue;
                if (tprev < 0 && tsdf > 0) break;
                if (tprev >= 0 && tsdf <= 0) {
                    const T Ftdt = trilinear(rs + dir * (tcur + step), vol, prm, wr);
                    if (std::isnan(Ftdt)) break;
                    const T Ft = trilinear(rs + dir * tcur, vol, prm, wr);
                    if (std::isnan(Ft)) break;
                    const T Ts = tcur - step * Ft / (Ftdt - Ft);
                    const V3<T> vf = rs + dir * Ts;
                    store3(vmap, id, vf - off);
                    if (color_vol && cmap) std::memcpy(cmap + 4 * id, color_vol + 4 * idx, 4);
                    int gg[3];
                    get_voxel(rs + dir * tcur, prm, gg);
                    if (gg[0] > 1 && gg[1] > 1 && gg[2] > 1 && gg[0] < Nx - 2 && gg[1] < Ny - 2 && gg[2] < Nz - 2) {
                        V3<T> n, q;
                        q = vf; q.x += prm.cell[0]; T fx1 = trilinear(q, vol, prm, wr);
                        q = vf; q.x -= prm.cell[0]; T fx2 = trilinear(q, vol, prm, wr);
                        n.x = fx1 - fx2;
                        q = vf; q.y += prm.cell[1]; T fy1 = trilinear(q, vol, prm, wr);
                        q = vf; q.y -= prm.cell[1]; T fy2 = trilinear(q, vol, prm, wr);
                        n.y = fy1 - fy2;
                        q = vf; q.z += prm.cell[2]; T fz1 = trilinear(q, vol, prm, wr);
                        q = vf; q.z -= prm.cell[2]; T fz2 = trilinear(q, vol, prm, wr);
                        n.z = fz1 - fz2;
                        store3(nmap, id, normalize3(n));
                    }
                    break;
                }
            }
        }
}

}  // namespace oracle
