// Projective point-to-plane ICP: every iteration of every session of a frame,
// without host synchronisation.
//
// The reference runs per iteration: icp_block_kernel (128-wide LDS trees of 29
// floats), compute_sums, a blocking read-back, and the Levenberg-Marquardt
// bookkeeping + Eigen FP64 6x6 solve on the host
// (OCL/NuiKinfuOpenCLDepthTracker.cpp:240-504), for one session on one queue.
// Here an iteration ("epoch": level * 64 + iteration) is two launches for any
// number of sessions:
//   tiles   every session's tiles of 512 pixels at its current epoch, spread over
//           the grid: the 29 terms of each pixel and the tile's partial sums
//           (plain stores: no atomics, no fences; the launch boundary orders them);
//   finish  one block per session: the partials reduced in a fixed order, the
//           update (the reference's LM step, or the compensated Gauss-Newton step)
//           and the next epoch (the next iteration, the next level after a break or
//           the last iteration, or "finished").
// A frame is begin, the pairs of launches, end.  One session tracked alone runs its
// planned epochs in order, each pair launched with exactly its level's tiles, and a
// pair whose epoch is not the session's current one (after a break or a failure)
// returns at once.  A batch runs in lockstep steps: step k carries every session's
// k-th epoch, whatever its level, so the steps are as many as the most iterations
// any session is configured for, and the tile launch is sized for the most tiles a
// session can have at that step.
//
// Reduction orders (nuif_icp_config::reduction):
//   reference    a tile is four 128-pixel groups, each reduced with the
//                reference's tree (item i takes item i + d for d = 64..1); the
//                finish block runs compute_sums' tree over each chunk of 128
//                group sums and sums the chunk results in order, as the host does,
//                then the LM step: float sums, Eigen-like LLT in FP64, float Euler
//                increment.  Bit for bit the reference's arithmetic.
//   compensated  thread j of tile k takes pixels 512k + j + 256i (i < 2) with
//                Kahan-compensated FP32 sums, FP64 across threads (runs of 16
//                threads in order, then a 16-way tree), tiles and segments in a
//                fixed order; Gauss-Newton on the undamped normal
//                equations (FP64 LDL^T), increment applied as an SE(3)
//                exponential to an FP64 pose (se3_math.h).
// Every sum is taken in a fixed order: results do not depend on which block
// processed which tile, nor on how many sessions share the launches.

#include "batch_jobs.h"
#include "det_math.h"
#include "icp_math.h"
#include "nuif_internal.h"
#include "schedule.h"
#include "se3_math.h"

#include <algorithm>
#include <type_traits>
#include <mutex>

namespace nuif {

constexpr int kTerms = NUIF_ICP_TERMS;
constexpr int kBlock = 256;
constexpr int kFinBlock = 1024;              // threads of a finish block
constexpr int kTilePx = nuif_se3::kCompTilePixels;  // pixels per tile (512)
constexpr int kPxPerThread = kTilePx / kBlock;      // compensated: pixels per thread of a tile
constexpr int kTileGroups = kTilePx / 128;   // reference order: 128-item groups per tile
constexpr int kMaxChunks = 256;              // reference order: 128-group chunks per level (W*H <= 2^22)
constexpr int kEpochIters = 64;              // iterations per level are below this
constexpr int kMaxBatch = 4096;              // sessions per batch
constexpr uint32_t kFinished = 0xffffffffu;  // epoch of a session whose frame is over
constexpr uint32_t kAnyEpoch = 0xfffffffeu;  // batch launches: each session's current epoch
static_assert(128 % kTileGroups == 0, "a tile lies in one compute_sums chunk");

// Reduction orders (nuif_icp_config::reduction)
constexpr int kRefOrder = 0;      // the reference's float trees and LM loop, bit for bit
constexpr int kCompensated = 1;   // Kahan FP32 + FP64, Gauss-Newton with an FP64 pose
constexpr int kJoint = 2;         // the compensated order with photometric rows

struct IcpState {
    float R[9], t[3];          // current estimate (compensated: the FP64 pose rounded)
    float Rg[9], tg[3];        // reference order: last accepted (Rgood, tgood)
    float Rlast[9], tlast[3];  // pose of the last integration
    float lambda, err_old;
    float error, count;
    int32_t brk[NUIF_ICP_MAX_LEVELS];
    int32_t failed, iterations, gate, frame, ok, timeout;
    double good[27];       // unused by every order today; kept for the layout
    double sums[kTerms];   // sums of the last executed iteration (every order)
    float goodf[27];       // reference order: goodCores as the reference keeps them (float)
    float pad2;
    float off[3];          // shifting volume: offset of the local frame (metres), 0 otherwise
    float pad3;
    double Rd[9], td[3];   // compensated order: the pose in FP64
    int32_t level_iters[NUIF_ICP_MAX_LEVELS];  // iterations that ran per level this frame
};
static_assert(sizeof(IcpState) % 8 == 0, "state is copied as 32-bit words");

struct IcpSync {
    uint32_t epoch;  // the current epoch (kFinished: the frame's iterations are over)
};
constexpr size_t kStateBytes = (sizeof(IcpState) + 255) / 256 * 256;
constexpr size_t kSyncBytes = (sizeof(IcpSync) + 255) / 256 * 256;

__device__ __forceinline__ IcpSync* sync_of(IcpState* st) {
    return reinterpret_cast<IcpSync*>(reinterpret_cast<char*>(st) + kStateBytes);
}

// One session of a launch (device copy of nuif_icp_session).
struct IcpJob {
    IcpState* st;
    void* partials;
    const float* vmap[NUIF_ICP_MAX_LEVELS];
    const float* vref;
    const float* nref;
    NuiCLRigidTransform* pose;  // the camera pose: the previous frame's during the iterations
    float* history;
    int capacity;
    NuiCLCameraParams cam;
    float dist;                 // constant distance threshold (:300-304)
    float gate_threshold;
    uint32_t thr[NUIF_ICP_MAX_LEVELS];
    int32_t iters[NUIF_ICP_MAX_LEVELS];
    int32_t levels;
    int32_t mode;
    int32_t single;             // one epoch only (nuif_icp_iteration): finished after it
    int32_t first_epoch;        // single: the epoch to run
    int32_t* outcome;           // NULL, or receives 1 (tracked) / 0 (failed) at the frame's end
    // joint order: the photometric inputs (nuif_icp_photometric)
    const float* icur[NUIF_ICP_MAX_LEVELS];
    const float4* model;        // the model's intensity and gradient per pixel
    float psw, pdiff, pgrad2;   // sqrt(weight), max |I - I_ref|, min |g|^2
};

__host__ __device__ __forceinline__ int level_pixels(const NuiCLCameraParams& c, int level) {
    return ((int)c.depthImageWidth >> level) * ((int)c.depthImageHeight >> level);
}
__host__ __device__ __forceinline__ int level_tiles(const NuiCLCameraParams& c, int level) {
    return (level_pixels(c, level) + kTilePx - 1) / kTilePx;
}

__device__ __forceinline__ void kahan_add(float& s, float& c, float v) {
    const float y = v - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
}


// K/icp.cl:192-220 interpolateBilinear_withHoles.  Reads past the row end as
// the reference does; a read past the buffer end is a hole.
__device__ __forceinline__ f3 bilinear_holes(const float* __restrict__ src, float px, float py, int W, int H) {
    const int x = (int)floorf(px), y = (int)floorf(py);
    const float dx = px - (float)x, dy = py - (float)y;
    const long long n = (long long)W * H;
    const long long ia = x + (long long)y * W, ib = ia + 1, ic = ia + W, id = ic + 1;
    const f3 NaN3 = mk3(kNaN, kNaN, kNaN);
    if (ia < 0 || id >= n) return NaN3;
    const f3 a = ld3(src, ia), b = ld3(src, ib), c = ld3(src, ic), d = ld3(src, id);
    if (nan3(a) || nan3(b) || nan3(c) || nan3(d)) return NaN3;
    f3 r;
    r.x = a.x * (1.0f - dx) * (1.0f - dy) + b.x * dx * (1.0f - dy) + c.x * (1.0f - dx) * dy + d.x * dx * dy;
    r.y = a.y * (1.0f - dx) * (1.0f - dy) + b.y * dx * (1.0f - dy) + c.y * (1.0f - dx) * dy + d.y * dx * dy;
    r.z = a.z * (1.0f - dx) * (1.0f - dy) + b.z * dx * (1.0f - dy) + c.z * (1.0f - dx) * dy + d.z * dx * dy;
    return r;
}

// The pose terms of a tile: current estimate (R, t) and the previous camera
// (R_inv, t) of transformInverse (K/utils.cl:62-67).
struct TilePose {
    float R[9], t[3], Ri[9], tp[3];
};

// Where a pixel's association lands: the current vertex at the estimate (pv), in the
// previous camera (pp), and its full-resolution projection (px, py).
struct AssocAt {
    f3 pv, pp;
    float px, py;
};

// One pixel of K/icp.cl:243-289: its association and its 7 coefficients
// c = (n x p, n, n . (q - p)); false when rejected.
__device__ __forceinline__ bool icp_assoc(const f3 vc, const TilePose& P, const NuiCLCameraParams& cam,
                                          const float* __restrict__ vref, const float* __restrict__ nref,
                                          float distThr, float c[7]) {
    if (nan3(vc)) return false;
    const int W = (int)cam.depthImageWidth, H = (int)cam.depthImageHeight;
    const f3 pv = add3(rot3(P.R, vc), mk3(P.t[0], P.t[1], P.t[2]));
    const f3 pp = rot3(P.Ri, sub3(pv, mk3(P.tp[0], P.tp[1], P.tp[2])));
    const float px = pp.x * cam.fx / pp.z + cam.cx, py = pp.y * cam.fy / pp.z + cam.cy;
    if (!(px >= 0.0f && px < (float)W && py >= 0.0f && py < (float)H)) return false;
    // both lookups are issued before either is tested (pure reads)
    const f3 rn = bilinear_holes(nref, px, py, W, H);
    const f3 rv = bilinear_holes(vref, px, py, W, H);
    if (nan3(rn) || nan3(rv)) return false;
    const f3 diff = sub3(rv, pv);
    const float dist = sqrtf(dot3(diff, diff));
    if (!(dist < distThr)) return false;
    const f3 row0 = cross3(rn, pv);
    c[0] = row0.x;
    c[1] = row0.y;
    c[2] = row0.z;
    c[3] = rn.x;
    c[4] = rn.y;
    c[5] = rn.z;
    c[6] = dot3(rn, diff);
    return true;
}

// The joint order's photometric row of an associated pixel (nuif_icp_photometric) from
// the current intensity ic and the model's intensity ir and gradient g at the
// projection: false when the pixel gives none.
__device__ __forceinline__ bool photometric_row(float ic, float ir, float2 g, const AssocAt& at, const TilePose& P,
                                                const NuiCLCameraParams& cam, float psw, float pdiff, float pgrad2,
                                                float c[7]) {
    const float di = ic - ir;
    if (!(fabsf(di) < pdiff) || !(g.x * g.x + g.y * g.y > pgrad2)) return false;
    // d(intensity)/d(point in the previous camera), then to world coordinates: pp = Ri (pv - tp)
    const float iz = 1.0f / at.pp.z;
    const float ax = g.x * cam.fx, ay = g.y * cam.fy;
    const f3 jc = mk3(ax * iz, ay * iz, -((ax * at.pp.x + ay * at.pp.y) * iz) * iz);
    const f3 jw = mk3((jc.x * P.Ri[0] + jc.y * P.Ri[3]) + jc.z * P.Ri[6], (jc.x * P.Ri[1] + jc.y * P.Ri[4]) + jc.z * P.Ri[7],
                      (jc.x * P.Ri[2] + jc.y * P.Ri[5]) + jc.z * P.Ri[8]);
    const f3 a = cross3(jw, at.pv);
    c[0] = psw * a.x;
    c[1] = psw * a.y;
    c[2] = psw * a.z;
    c[3] = psw * jw.x;
    c[4] = psw * jw.y;
    c[5] = psw * jw.z;
    c[6] = psw * di;
    return true;
}

// One pixel of the joint order: icp_assoc's geometric row (ok) and, for an associated
// pixel, its photometric row (ph), with every gather -- the reference normals and
// vertices, the model's texels (intensity, dI/du, dI/dv; bilinear with the holes rule
// of bilinear_holes) and the current intensity -- issued before any test.
__device__ __forceinline__ void joint_assoc(const f3 vc, const TilePose& P, const NuiCLCameraParams& cam,
                                            const float* __restrict__ vref, const float* __restrict__ nref,
                                            const float4* __restrict__ model, const float* __restrict__ icur, int p,
                                            float distThr, float psw, float pdiff, float pgrad2, float c[7],
                                            float pc[7], bool& ok, bool& ph) {
    ok = ph = false;
    if (nan3(vc)) return;
    const int W = (int)cam.depthImageWidth, H = (int)cam.depthImageHeight;
    const f3 pv = add3(rot3(P.R, vc), mk3(P.t[0], P.t[1], P.t[2]));
    const f3 pp = rot3(P.Ri, sub3(pv, mk3(P.tp[0], P.tp[1], P.tp[2])));
    const float px = pp.x * cam.fx / pp.z + cam.cx, py = pp.y * cam.fy / pp.z + cam.cy;
    if (!(px >= 0.0f && px < (float)W && py >= 0.0f && py < (float)H)) return;
    const int x = (int)floorf(px), y = (int)floorf(py);
    const float dx = px - (float)x, dy = py - (float)y;
    const long long n = (long long)W * H;
    const long long ia = x + (long long)y * W, ib = ia + 1, ic = ia + W, id = ic + 1;
    if (ia < 0 || id >= n) return;
    const f3 na = ld3(nref, ia), nb = ld3(nref, ib), nc = ld3(nref, ic), nd = ld3(nref, id);
    const f3 va = ld3(vref, ia), vb = ld3(vref, ib), vcn = ld3(vref, ic), vd = ld3(vref, id);
    const float4 ma = model[ia], mb = model[ib], mc = model[ic], md = model[id];
    const float ival = icur[p];
    auto mix = [&](float a, float b, float c_, float d) {
        return a * (1.0f - dx) * (1.0f - dy) + b * dx * (1.0f - dy) + c_ * (1.0f - dx) * dy + d * dx * dy;
    };
    if (nan3(na) || nan3(nb) || nan3(nc) || nan3(nd) || nan3(va) || nan3(vb) || nan3(vcn) || nan3(vd)) return;
    const f3 rn = mk3(mix(na.x, nb.x, nc.x, nd.x), mix(na.y, nb.y, nc.y, nd.y), mix(na.z, nb.z, nc.z, nd.z));
    const f3 rv = mk3(mix(va.x, vb.x, vcn.x, vd.x), mix(va.y, vb.y, vcn.y, vd.y), mix(va.z, vb.z, vcn.z, vd.z));
    const f3 diff = sub3(rv, pv);
    const float dist = sqrtf(dot3(diff, diff));
    if (!(dist < distThr)) return;
    const f3 row0 = cross3(rn, pv);
    c[0] = row0.x;
    c[1] = row0.y;
    c[2] = row0.z;
    c[3] = rn.x;
    c[4] = rn.y;
    c[5] = rn.z;
    c[6] = dot3(rn, diff);
    ok = true;
    if (isnan(ival) || isnan(ma.x) || isnan(mb.x) || isnan(mc.x) || isnan(md.x) || isnan(ma.y) || isnan(mb.y) ||
        isnan(mc.y) || isnan(md.y) || isnan(ma.z) || isnan(mb.z) || isnan(mc.z) || isnan(md.z))
        return;
    const float ir = mix(ma.x, mb.x, mc.x, md.x);
    const float2 g = make_float2(mix(ma.y, mb.y, mc.y, md.y), mix(ma.z, mb.z, mc.z, md.z));
    ph = photometric_row(ival, ir, g, AssocAt{pv, pp, px, py}, P, cam, psw, pdiff, pgrad2, pc);
}

// The 29 terms of a pixel (zeros when rejected)
__device__ __forceinline__ void icp_terms(const f3 vc, const TilePose& P, const NuiCLCameraParams& cam,
                                          const float* __restrict__ vref, const float* __restrict__ nref,
                                          float distThr, float out[kTerms]) {
    float c[7];
#pragma unroll
    for (int k = 0; k < kTerms; ++k) out[k] = 0.f;
    if (!icp_assoc(vc, P, cam, vref, nref, distThr, c)) return;
    int s = 0;
#pragma unroll
    for (int i = 0; i < 6; ++i)
#pragma unroll
        for (int j = i; j < 7; ++j) out[s++] = c[i] * c[j];
    out[27] = c[6] * c[6];
    out[28] = 1.0f;
}


// ---- the updates ------------------------------------------------------------------

__device__ void matmul3d(const double a[9], const double b[9], double c[9]) {
    double r[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r[3 * i + j] = a[3 * i] * b[j] + a[3 * i + 1] * b[3 + j] + a[3 * i + 2] * b[6 + j];
    for (int k = 0; k < 9; ++k) c[k] = r[k];
}

// OCL/NuiKinfuOpenCLDepthTracker.cpp:402-458 for one iteration: error / count,
// accept or reject.  Returns whether a system is solved in this iteration.
__device__ bool ref_accept(IcpState& S, const float* sumsf, int level, int iter, uint32_t count_threshold) {
    for (int k = 0; k < kTerms; ++k) S.sums[k] = (double)sumsf[k];
    const float cnt = sumsf[28];
    const float err = cnt > 0.f ? sqrtf(sumsf[27]) / cnt : 3.402823466e38f;
    S.error = err;
    S.count = cnt;
    S.iterations += 1;
    S.level_iters[level] += 1;
    if (iter == 0) {  // per-level initialisation (:299-306)
        S.lambda = 1.0f;
        S.err_old = 1e20f;
        for (int k = 0; k < 9; ++k) S.Rg[k] = S.R[k];
        for (int k = 0; k < 3; ++k) S.tg[k] = S.t[k];
    }
    float lambda = S.lambda;
    if (cnt < (float)count_threshold || err > S.err_old) {
        lambda *= 10.0f;
        for (int k = 0; k < 9; ++k) S.R[k] = S.Rg[k];
        for (int k = 0; k < 3; ++k) S.t[k] = S.tg[k];
        S.lambda = lambda;
        if (iter == 0) {
            S.brk[level] = 1;
            return false;
        }
    } else {
        lambda /= 10.0f;
        S.lambda = lambda;
        for (int k = 0; k < 9; ++k) S.Rg[k] = S.R[k];
        for (int k = 0; k < 3; ++k) S.tg[k] = S.t[k];
        S.err_old = err;
        for (int k = 0; k < 27; ++k) S.goodf[k] = sumsf[k];
    }
    return true;
}

// :459-484 the damped system normalised by the current count (Q12), LLT in
// double, the rotation-only convergence test.  Returns whether the increment
// (angles, translation) is applied.
__device__ bool ref_solve(IcpState& S, const double* Av, int level, float ang[3], float tinc[3]) {
    const float lambda = S.lambda;
    double A[36], b[6];
    for (int i = 0, s = 0; i < 6; ++i) {
        for (int j = i; j < 7; ++j, ++s) {
            const double v = Av[s];
            if (j == 6)
                b[i] = v;
            else
                A[j * 6 + i] = A[i * 6 + j] = v;
        }
        A[i * 6 + i] *= (double)(1.0f + lambda);
    }
    double det, x[6];
    if (!llt6(A, &det) || fabs(det) < 1e-15 || isnan(det)) {
        S.failed = 1;
        return false;
    }
    llt6_solve(A, b, x);
    const float al = (float)x[0], be = (float)x[1], ga = (float)x[2];
    if (al * al + be * be + ga * ga < 1e-6f) {  // converged: break before applying (:482-484)
        S.brk[level] = 1;
        return false;
    }
    ang[0] = al;
    ang[1] = be;
    ang[2] = ga;
    for (int k = 0; k < 3; ++k) tinc[k] = (float)x[3 + k];
    return true;
}

// Rinc = Rz(gamma) Ry(beta) Rx(alpha); t <- Rinc^-1 t + t_inc; R <- Rinc^-1 R   (:486-497)
__device__ void ref_compose(IcpState& S, const float ang[3], const float tinc[3]) {
    float Rz[9], Ry[9], Rx[9], Rzy[9], Rinc[9], Ri[9];
    axis_rotf(ang[2], 2, Rz);
    axis_rotf(ang[1], 1, Ry);
    axis_rotf(ang[0], 0, Rx);
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

// The compensated order's iteration from its FP64 sums (thread 0): the count
// threshold, the Gauss-Newton step on the FP64 pose (nuif_se3::gn_step).
__device__ void comp_update(IcpState& S, const double* sums, int level, uint32_t count_threshold, double* ws) {
    for (int k = 0; k < kTerms; ++k) S.sums[k] = sums[k];
    const double cnt = sums[28];
    S.count = (float)cnt;
    S.error = cnt > 0.0 ? (float)(sqrt(sums[27]) / cnt) : 3.402823466e38f;
    S.iterations += 1;
    S.level_iters[level] += 1;
    if (cnt < (double)count_threshold) {
        S.brk[level] = 1;
        return;
    }
    const int r = nuif_se3::gn_step(sums, nuif_se3::kGnStep2, S.Rd, S.td, ws);
    if (r == 2) {
        S.failed = 1;
    } else if (r == 1) {
        S.brk[level] = 1;
    } else {
        for (int k = 0; k < 9; ++k) S.R[k] = (float)S.Rd[k];
        for (int k = 0; k < 3; ++k) S.t[k] = (float)S.td[k];
    }
}

// The epoch that follows (level, iter) given the updated state.
__device__ uint32_t next_epoch(const IcpState& S, const IcpJob& J, int level, int iter) {
    if (S.failed || J.single) return kFinished;
    if (!S.brk[level] && iter + 1 < J.iters[level]) return (uint32_t)(level * kEpochIters + iter + 1);
    for (int l = level - 1; l >= 0; --l)
        if (J.iters[l] > 0 && !S.brk[l]) return (uint32_t)(l * kEpochIters);
    return kFinished;
}

// The finish block of an epoch: state into LDS, update, write back, the next
// epoch.  sums (FP64) or sumsf (reference, float) are in LDS.
template <int MODE>
__device__ __attribute__((noinline)) void finish_epoch(const IcpJob& J, IcpState& S, const double* sums, const float* sumsf, int level,
                             int iter) {
    __shared__ double Av[27];
    __shared__ double ws[60];
    __shared__ float ang[3], tinc[3];
    __shared__ int go;
    constexpr int kWords = (int)(sizeof(IcpState) / 4);
    const int tid = threadIdx.x;
    IcpState* st = J.st;
    const uint32_t* src = reinterpret_cast<const uint32_t*>(st);
    uint32_t* lds = reinterpret_cast<uint32_t*>(&S);
    for (int i = tid; i < kWords; i += blockDim.x) lds[i] = src[i];
    __syncthreads();
    if (MODE == kRefOrder) {
        if (tid == 0) go = ref_accept(S, sumsf, level, iter, J.thr[level]);
        __syncthreads();
        if (go) {
            if (tid < 27) Av[tid] = (double)(S.goodf[tid] / S.count);
            __syncthreads();
            if (tid == 0) {
                go = ref_solve(S, Av, level, ang, tinc);
                if (go) ref_compose(S, ang, tinc);
            }
        }
    } else if (tid == 0) {
        comp_update(S, sums, level, J.thr[level], ws);
    }
    __syncthreads();
    uint32_t* dst = reinterpret_cast<uint32_t*>(st);
    for (int i = tid; i < kWords; i += blockDim.x) dst[i] = lds[i];
    if (tid == 0) sync_of(st)->epoch = next_epoch(S, J, level, iter);
}

// ---- the tile launch --------------------------------------------------------------

// LDS of a tile block (256 threads)
struct IcpLdsBase {
    IcpJob single;  // the job of a one-session launch (passed by value)
    IcpJob job;     // the job of the session being processed
    TilePose pose;
    int level;
};
template <int MODE> struct IcpLds;
template <> struct IcpLds<kRefOrder> : IcpLdsBase {
    float xfer[2][64 * kTerms];  // the tree's d = 64 step of two groups
};
// compensated: the threads' FP64 values of up to kXtTerms terms at a time, term-major,
// one double of padding after every 16 threads (conflict-free 16-thread runs)
// product term k (k < 28) is c[a] * c[b] for the k-th pair a <= b < 7 with a < 6, then
// (6, 6) (K/icp.cl:271-289 order)
constexpr int term_a(int k) {
    int a = 0;
    while (a < 6 && k >= 7 - a) k -= 7 - a++;
    return a;
}
constexpr int term_b(int k) {
    int a = 0;
    while (a < 6 && k >= 7 - a) k -= 7 - a++;
    return a + k;
}
static_assert(term_a(6) == 0 && term_b(6) == 6 && term_a(7) == 1 && term_b(7) == 1 && term_a(26) == 5 &&
                  term_b(26) == 6 && term_a(27) == 6 && term_b(27) == 6,
              "term pairs");

template <int I, int N, class F>
__device__ __forceinline__ void static_for(F&& f) {
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        static_for<I + 1, N>(f);
    }
}
constexpr int kXtPasses = 3;                                      // passes of 10, 10 and 9 terms:
constexpr int kXtTerms = (kTerms + kXtPasses - 1) / kXtPasses;  // 22 KB, five blocks per CU
constexpr int kXtRow = kBlock + kBlock / 16;
__device__ __forceinline__ int xt_index(int term, int thread) { return term * kXtRow + thread + (thread >> 4); }
template <> struct IcpLds<kCompensated> : IcpLdsBase {
    double xt[kXtTerms * kXtRow];
};
template <> struct IcpLds<kJoint> : IcpLds<kCompensated> {};

// Tile `tile` of level L.level of job L.job (with L.pose loaded): the 29 terms of its
// pixels and its partial sums.
template <int MODE>
__device__ __forceinline__ void process_tile(IcpLds<MODE>& L, uint32_t tile) {
    const int tid = threadIdx.x;
    const IcpJob& J = L.job;
    const int level = L.level;
    const int npx = level_pixels(J.cam, level);
    const float* __restrict__ vcur = in_global(J.vmap[level]);
    const float* __restrict__ vref = in_global(J.vref);
    const float* __restrict__ nref = in_global(J.nref);

    if constexpr (MODE == kRefOrder) {
        // four 128-pixel groups, two at a time (one per 128 threads)
        const int sub = tid >> 7, l = tid & 127, lane = l & 63, wave = l >> 6;
        const int ngroups = (npx + 127) >> 7;
        float* part = in_global(reinterpret_cast<float*>(J.partials));
        for (int r = 0; r < kTileGroups / 2; ++r) {
            const int g = (int)tile * kTileGroups + 2 * r + sub;
            const int p = g * 128 + l;
            float v[kTerms];
            const f3 vc = p < npx ? ld3(vcur, p) : mk3(kNaN, kNaN, kNaN);
            icp_terms(vc, L.pose, J.cam, vref, nref, J.dist, v);
            // K/icp.cl:162-190 reduce(): step d = 64 through LDS, steps 32..1 in the group's wave 0
            if (wave == 1) {
#pragma unroll
                for (int k = 0; k < kTerms; ++k) L.xfer[sub][lane * kTerms + k] = v[k];
            }
            __syncthreads();
            if (wave == 0) {
#pragma unroll
                for (int k = 0; k < kTerms; ++k) v[k] = wave_tree64(v[k] + L.xfer[sub][lane * kTerms + k]);
                if (lane == 0 && g < ngroups) {
#pragma unroll
                    for (int k = 0; k < kTerms; ++k) part[(size_t)g * kTerms + k] = v[k];
                }
            }
            __syncthreads();
        }
    } else {
        // compensated: the thread's pixels in order, their gathers in flight together
        static_assert(kPxPerThread == 2, "the folded Kahan sums below take two pixels");
        const int p0 = (int)tile * kTilePx + tid;
        float cf[kPxPerThread][7];
        bool ok[kPxPerThread];
        float pc[MODE == kJoint ? kPxPerThread : 1][7];
        bool ph[kPxPerThread] = {};
#pragma unroll
        for (int i = 0; i < kPxPerThread; ++i) {
            const int p = p0 + i * kBlock;
            const f3 vc = p < npx ? ld3(vcur, p) : mk3(kNaN, kNaN, kNaN);
            if constexpr (MODE == kJoint) {
                joint_assoc(vc, L.pose, J.cam, vref, nref, in_global(J.model), in_global(J.icur[level]), p, J.dist,
                            J.psw, J.pdiff, J.pgrad2, cf[i], pc[i], ok[i], ph[i]);
            } else {
                ok[i] = icp_assoc(vc, L.pose, J.cam, vref, nref, J.dist, cf[i]);
            }
        }
        // Kahan sums of the two pixels from zero, folded: after the first pixel the sum is
        // 0 + v0 (never -0) and the compensation exactly +0, so the second step is
        // s = s1 + v1, c = (s - s1) - v1; a rejected pixel adds +0, which changes neither.
        // Each term is formed as it goes to LDS: only the 14 coefficients stay live.
        auto term = [&](auto kc) -> double {
            constexpr int k = decltype(kc)::value;
            if constexpr (k < kTerms - 1) {
                constexpr int a = term_a(k), b = term_b(k);
                float v0 = ok[0] ? cf[0][a] * cf[0][b] : 0.f;
                float v1 = ok[1] ? cf[1][a] * cf[1][b] : 0.f;
                if constexpr (MODE == kJoint) {  // a pixel's value: geometric + photometric product
                    if (ph[0]) v0 = v0 + pc[0][a] * pc[0][b];
                    if (ph[1]) v1 = v1 + pc[MODE == kJoint ? 1 : 0][a] * pc[MODE == kJoint ? 1 : 0][b];
                }
                const float s1 = 0.f + v0;
                const float sm = s1 + v1;
                return (double)sm - (double)((sm - s1) - v1);
            } else {  // the count, summed plainly
                return (double)((0.f + (ok[0] ? 1.f : 0.f)) + (ok[1] ? 1.f : 0.f));
            }
        };
        // across the tile's threads, a pass of up to kXtTerms terms at a time through
        // LDS: 16 threads per term, thread q of a term summing the values of threads
        // 16q .. 16q + 15 in order, then a tree over the 16 (row DPP, lane l takes
        // lane l + d for d = 8, 4, 2, 1)
        double* part = in_global(reinterpret_cast<double*>(J.partials));
        const int q = tid & 15;
        static_for<0, kXtPasses>([&](auto pc) {
            constexpr int t0 = decltype(pc)::value * kXtTerms;
            constexpr int nt = kTerms - t0 < kXtTerms ? kTerms - t0 : kXtTerms;
            static_for<0, nt>(
                [&](auto kc) { L.xt[xt_index(decltype(kc)::value, tid)] = term(std::integral_constant<int, t0 + decltype(kc)::value>{}); });
            __syncthreads();
            const int tl = min(tid >> 4, nt - 1);  // rows past the pass's terms repeat its last
            const double* row = &L.xt[xt_index(tl, 16 * q)];
            double v = row[0];
#pragma unroll
            for (int k = 1; k < 16; ++k) v += row[k];
            v += lane_from_above<8>(v);
            v += lane_from_above<4>(v);
            v += lane_from_above<2>(v);
            v += lane_from_above<1>(v);
            if (q == 0 && (tid >> 4) < nt) part[(size_t)tile * kTerms + t0 + tl] = v;
            __syncthreads();  // xt is rewritten by the next pass or tile
        });
    }
}

// The tiles a session works on now: those of its epoch's level, none when its frame
// is over or (planned launches) when its epoch is not `planned`.
__device__ __forceinline__ uint32_t open_tiles(const IcpJob& J, uint32_t e, uint32_t planned) {
    if (e == kFinished || (planned != kAnyEpoch && e != planned)) return 0u;
    return (uint32_t)level_tiles(J.cam, (int)(e / kEpochIters));
}

// Every open tile of n sessions (jobs, or `single` when jobs is null).  The sessions'
// open tiles are numbered one after another (the prefix of their counts in dynamic
// LDS: n + 1 words) and block b takes the xcd_block(b)-th of gridDim.x equal runs of
// them (the b-th when gridDim.x is not a multiple of the XCD count).
template <int MODE>
__global__ void __launch_bounds__(kBlock)
    icp_tiles_kernel(const IcpJob* __restrict__ jobs, int n, IcpJob single, uint32_t planned, int remap) {
    __shared__ IcpLds<MODE> L;
    extern __shared__ uint32_t pre[];
    const int tid = threadIdx.x;
    if (!jobs) {
        if (tid == 0) L.single = single;
        __syncthreads();
    }
    const IcpJob* table = jobs ? jobs : &L.single;
    // exclusive prefix of the open tiles, 256 sessions at a time
    __shared__ uint32_t carry, wtot[kBlock / 64];
    if (tid == 0) carry = 0u;
    __syncthreads();
    for (int c0 = 0; c0 < n; c0 += kBlock) {
        const int s = c0 + tid;
        uint32_t v = 0u;
        if (s < n) v = open_tiles(table[s], sync_of(in_global(table[s].st))->epoch, planned);
        // inclusive scan in the wave, then over the waves
        const int lane = tid & 63, wv = tid >> 6;
        uint32_t x = v;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t y = __shfl_up(x, d, 64);
            if (lane >= d) x += y;
        }
        if (lane == 63) wtot[wv] = x;
        __syncthreads();
        uint32_t before = carry;
        for (int w = 0; w < wv; ++w) before += wtot[w];
        if (s < n) pre[s] = before + x - v;
        __syncthreads();
        if (tid == kBlock - 1) carry = before + x;
        __syncthreads();
    }
    if (tid == 0) pre[n] = carry;
    __syncthreads();
    const uint32_t total = pre[n];
    // runs of consecutive tiles (one session, one band of its image) share an XCD's L2
    const unsigned lb = remap && gridDim.x % kXcds == 0 ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
    const uint32_t f0 = (uint32_t)((unsigned long long)lb * total / gridDim.x);
    const uint32_t f1 = (uint32_t)((unsigned long long)(lb + 1) * total / gridDim.x);
    int cur = -1;
    for (uint32_t f = f0; f < f1; ++f) {
        // the session of tile f: the last with pre[s] <= f
        int lo = 0, hi = n - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (pre[mid] <= f)
                lo = mid;
            else
                hi = mid - 1;
        }
        if (lo != cur) {
            // the job, word by word; the estimate of this epoch and the previous camera pose
            constexpr int kJobWords = (int)(sizeof(IcpJob) / 4);
            const uint32_t* js = reinterpret_cast<const uint32_t*>(&table[lo]);
            uint32_t* jd = reinterpret_cast<uint32_t*>(&L.job);
            for (int w = tid; w < kJobWords; w += kBlock) jd[w] = js[w];
            const IcpJob& J = table[lo];
            float* dp = reinterpret_cast<float*>(&L.pose);
            IcpState* st = in_global(J.st);
            const NuiCLRigidTransform* pose = in_global(J.pose);
            if (tid < 12) {
                dp[tid] = reinterpret_cast<const float*>(st)[tid];
            } else if (tid < 21) {
                L.pose.Ri[tid - 12] = pose->R_inv[tid - 12];
            } else if (tid < 24) {
                L.pose.tp[tid - 21] = pose->t[tid - 21];
            } else if (tid == 24) {
                L.level = (int)(sync_of(st)->epoch / kEpochIters);
            }
            cur = lo;
            __syncthreads();
        }
        process_tile<MODE>(L, f - pre[lo]);
    }
}

// ---- the finish launch -------------------------------------------------------------

constexpr int kFinSegs = 32;  // compensated: tile segments per term, summed in parallel
static_assert(kTerms * kFinSegs <= kFinBlock, "one thread per term and segment");
struct FinLds {
    IcpJob job;
    uint32_t epoch;
    double seg[kTerms * kFinSegs];         // compensated: segment sums
    double sums[kTerms];
    float sumsf[kTerms];
    float chunk[kMaxChunks * kTerms];      // reference: compute_sums results
    IcpState S;
};

// One block per session: the epoch's partials reduced in a fixed order, the update,
// the next epoch.
template <int MODE>
__global__ void __launch_bounds__(kFinBlock)
    icp_finish_kernel(const IcpJob* __restrict__ jobs, IcpJob single, uint32_t planned) {
    __shared__ FinLds F;
    const int tid = threadIdx.x;
    if (tid == 0) {
        F.job = jobs ? jobs[blockIdx.x] : single;
        F.epoch = sync_of(F.job.st)->epoch;
    }
    __syncthreads();
    const IcpJob& J = F.job;
    const uint32_t e = F.epoch;
    if (e == kFinished || (planned != kAnyEpoch && e != planned)) return;
    const int level = (int)(e / kEpochIters), iter = (int)(e % kEpochIters);
    const int npx = level_pixels(J.cam, level);
    if constexpr (MODE == kRefOrder) {
        // compute_sums' tree over each chunk of 128 group sums (one wave per chunk, its
        // 29 terms together), then the host's sequential float sum over the chunk
        // results (OCL/NuiKinfuOpenCLDepthTracker.cpp:394-402, 428-440)
        const float* part = in_global(reinterpret_cast<const float*>(J.partials));
        const int ngroups = (npx + 127) >> 7, nchunks = (ngroups + 127) >> 7;
        const int lane = tid & 63, wv = tid >> 6;
        for (int q = wv; q < nchunks; q += kFinBlock / 64) {
            const int g0 = (q << 7) + lane, g1 = g0 + 64;
            float x[kTerms];
#pragma unroll
            for (int k = 0; k < kTerms; ++k) {
                const float a = g0 < ngroups ? part[(size_t)g0 * kTerms + k] : 0.f;
                const float b = g1 < ngroups ? part[(size_t)g1 * kTerms + k] : 0.f;
                x[k] = a + b;
            }
#pragma unroll
            for (int k = 0; k < kTerms; ++k) {
                const float sum = wave_tree64(x[k]);
                if (lane == 0) F.chunk[q * kTerms + k] = sum;
            }
        }
        __syncthreads();
        if (tid < kTerms) {
            float sum = 0.f;
            for (int q = 0; q < nchunks; ++q) sum += F.chunk[q * kTerms + tid];
            F.sumsf[tid] = sum;
        }
        __syncthreads();
        finish_epoch<MODE>(J, F.S, nullptr, F.sumsf, level, iter);
    } else {
        const double* part = in_global(reinterpret_cast<const double*>(J.partials));
        const int ntiles = level_tiles(J.cam, level);
        if (tid < kTerms * kFinSegs) {
            // term k, segment j: tiles j, j + kFinSegs, ... in order, 16 loads in flight
            // (past the last tile a load is replaced by +0, which leaves the sum as it is)
            const int k = tid / kFinSegs, j = tid % kFinSegs;
            const double* pk = part + k;
            double s = 0.0;
            for (int q0 = j; q0 < ntiles; q0 += kFinSegs * 16) {
                double v[16];
#pragma unroll
                for (int u = 0; u < 16; ++u) {
                    const int q = q0 + kFinSegs * u;
                    v[u] = q < ntiles ? pk[(size_t)q * kTerms] : 0.0;
                }
#pragma unroll
                for (int u = 0; u < 16; ++u) s += v[u];
            }
            F.seg[k * kFinSegs + j] = s;
        }
        __syncthreads();
        if (tid < kTerms) {
            double s = 0.0;
            for (int j = 0; j < kFinSegs; ++j) s += F.seg[tid * kFinSegs + j];
            F.sums[tid] = s;
        }
        __syncthreads();
        finish_epoch<MODE>(J, F.S, F.sums, nullptr, level, iter);
    }
}

// ---- per-frame begin / end (one block per session) ----------------------------------

// First epoch of a frame: the coarsest level with iterations.
__device__ uint32_t first_epoch(const IcpJob& J, const IcpState* st) {
    // (called after the frame's flags are set; a single epoch is skipped after a break or failure)
    if (J.single) {
        const int level = (int)(J.first_epoch / kEpochIters);
        return (st->failed || st->brk[level]) ? kFinished : (uint32_t)J.first_epoch;
    }
    for (int l = J.levels - 1; l >= 0; --l)
        if (J.iters[l] > 0) return (uint32_t)(l * kEpochIters);
    return kFinished;
}

// Begin a frame: estimate <- the camera pose, flags cleared, the first epoch
// opened.  Compensated: the FP64 pose carries over from the previous frame unless
// the camera pose was changed from outside (reset, hint, volume shift), in which
// case it restarts from that pose.
__global__ void icp_begin_kernel(const IcpJob* __restrict__ jobs, IcpJob single) {
    __shared__ IcpJob J;
    if (threadIdx.x == 0) J = jobs ? jobs[blockIdx.x] : single;
    __syncthreads();
    if (threadIdx.x != 0) return;
    IcpState* st = J.st;
    IcpSync* sy = sync_of(st);
    const NuiCLRigidTransform* pose = J.pose;
    if (!J.single) {
        if (J.mode != kRefOrder) {
            bool same = true;
            for (int k = 0; k < 9; ++k) same = same && (float)st->Rd[k] == pose->R[k];
            for (int k = 0; k < 3; ++k) same = same && (float)st->td[k] == pose->t[k];
            if (!same) {
                for (int k = 0; k < 9; ++k) st->Rd[k] = (double)pose->R[k];
                for (int k = 0; k < 3; ++k) st->td[k] = (double)pose->t[k];
            }
            for (int k = 0; k < 9; ++k) st->R[k] = (float)st->Rd[k];
            for (int k = 0; k < 3; ++k) st->t[k] = (float)st->td[k];
        } else {
            for (int k = 0; k < 9; ++k) st->R[k] = pose->R[k];
            for (int k = 0; k < 3; ++k) st->t[k] = pose->t[k];
        }
        for (int l = 0; l < NUIF_ICP_MAX_LEVELS; ++l) st->brk[l] = st->level_iters[l] = 0;
        st->failed = 0;
        st->iterations = 0;
        st->ok = 0;
        st->timeout = 0;
    }
    sy->epoch = first_epoch(J, st);
}

// src/Foundation/NuiMatrixUtilities.cpp:7-52 rodrigues2(R).norm() for a rotation
__device__ float rodrigues_norm(const float Rm[9]) {
    const double rx = Rm[7] - Rm[5], ry = Rm[2] - Rm[6], rz = Rm[3] - Rm[1];
    const double s = sqrt((rx * rx + ry * ry + rz * rz) * 0.25);
    double c = ((double)Rm[0] + Rm[4] + Rm[8] - 1) * 0.5;
    c = c > 1. ? 1. : c < -1. ? -1. : c;
    const double theta = acos(c);
    if (s < 1e-5 && c > 0) return 0.f;
    return (float)theta;
}

// Finish the frame (one thread per session): tracking outcome, pose, the
// integration gate (VO/NuiKinfuTrackingEngine.cpp:113-127), the history row.
// A session whose epochs did not all run (fewer launched than its schedule needs)
// fails with the incomplete flag set.
__device__ void end_frame(IcpState* st, NuiCLRigidTransform* pose, float gate_threshold, float* history,
                          int capacity, int32_t* outcome = nullptr) {
    if (sync_of(st)->epoch != kFinished) {
        st->failed = 1;
        st->timeout = 1;
    }
    if (outcome) *outcome = st->failed ? 0 : 1;
    const int frame = st->frame;
    st->frame = frame + 1;
    float* row = history ? history + (size_t)(frame % capacity) * NUIF_POSE_ROW : nullptr;
    if (row) {
        const float* R = st->failed ? pose->R : st->R;
        const float* t = st->failed ? pose->t : st->t;
        for (int k = 0; k < 9; ++k) row[k] = R[k];
        for (int k = 0; k < 3; ++k) row[9 + k] = t[k];
        row[12] = st->failed ? 0.f : 1.f;
        row[13] = st->error;
        row[14] = st->count;
        row[15] = 0.f;
    }
    if (st->failed) {  // RunTracking returns false: no pose update, integrate or raycast
        st->ok = 0;
        st->gate = 0;
        return;
    }
    st->ok = 1;
    for (int k = 0; k < 9; ++k) pose->R[k] = st->R[k];
    for (int k = 0; k < 3; ++k) pose->t[k] = st->t[k];
    inverse3f(st->R, pose->R_inv);
    float Rrel[9];
    {
        const float* Ri = pose->R_inv;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                Rrel[3 * i + j] = Ri[3 * i] * st->Rlast[j] + Ri[3 * i + 1] * st->Rlast[3 + j] + Ri[3 * i + 2] * st->Rlast[6 + j];
    }
    const float rnorm = rodrigues_norm(Rrel);
    // global translations (local + the shifting volume's offset; NuiCameraPos::getGlobalTranslation)
    const float d0 = (st->t[0] + st->off[0]) - (st->tlast[0] + st->off[0]);
    const float d1 = (st->t[1] + st->off[1]) - (st->tlast[1] + st->off[1]);
    const float d2 = (st->t[2] + st->off[2]) - (st->tlast[2] + st->off[2]);
    const float tnorm = sqrtf(d0 * d0 + d1 * d1 + d2 * d2);
    const int gate = (rnorm + 1.f * tnorm) / 2 >= gate_threshold;
    st->gate = gate;
    if (row) row[15] = (float)gate;
    if (gate) {
        for (int k = 0; k < 9; ++k) st->Rlast[k] = st->R[k];
        for (int k = 0; k < 3; ++k) st->tlast[k] = st->t[k];
    }
}

__global__ void icp_end_kernel(const IcpJob* __restrict__ jobs, IcpJob single) {
    if (threadIdx.x != 0) return;
    __shared__ IcpJob J;
    J = jobs ? jobs[blockIdx.x] : single;
    end_frame(J.st, J.pose, J.gate_threshold, J.history, J.capacity, J.outcome);
}

// A frame without tracking (the first frame): the pose is kept and the frame
// integrates; estimate <- pose.
__global__ void icp_record_kernel(IcpState* st, NuiCLRigidTransform* pose, float* history, int capacity) {
    for (int k = 0; k < 9; ++k) st->R[k] = pose->R[k];
    for (int k = 0; k < 3; ++k) st->t[k] = pose->t[k];
    for (int l = 0; l < NUIF_ICP_MAX_LEVELS; ++l) st->brk[l] = st->level_iters[l] = 0;
    st->failed = 0;
    st->iterations = 0;
    st->timeout = 0;
    sync_of(st)->epoch = kFinished;
    end_frame(st, pose, -1.0f, history, capacity);
}

__global__ void icp_state_init_kernel(IcpState* st, const NuiCLRigidTransform* pose) {
    if (threadIdx.x != 0) return;
    sync_of(st)->epoch = kFinished;
    for (int k = 0; k < 9; ++k) st->R[k] = st->Rg[k] = st->Rlast[k] = pose->R[k];
    for (int k = 0; k < 3; ++k) st->t[k] = st->tg[k] = st->tlast[k] = pose->t[k];
    for (int k = 0; k < 9; ++k) st->Rd[k] = (double)pose->R[k];
    for (int k = 0; k < 3; ++k) st->td[k] = (double)pose->t[k];
    for (int k = 0; k < 3; ++k) st->off[k] = 0.f;
    for (int l = 0; l < NUIF_ICP_MAX_LEVELS; ++l) st->brk[l] = st->level_iters[l] = 0;
    st->failed = 0;
    st->iterations = 0;
    st->gate = 1;
    st->ok = 1;
    st->timeout = 0;
    st->frame = 0;
    st->lambda = 1.0f;
    st->err_old = 1e20f;
    st->error = 0.f;
    st->count = 0.f;
}

// A shifting volume moved the local frame after this frame's tracking
// (OCL/NuiKinfuOpenCLShiftScene.cpp:353-364): the new local translation becomes the
// camera pose, the estimate (FP64 pose included), the last integration pose (taken
// after the shift, as VO/NuiKinfuTrackingEngine.cpp:126 does) and the frame's history
// row; `off` is the frame's offset for the following integration gates.
__global__ void icp_set_translation_kernel(IcpState* st, NuiCLRigidTransform* pose, float* history, int capacity,
                                           float tx, float ty, float tz, float ox, float oy, float oz) {
    const float t[3] = {tx, ty, tz}, o[3] = {ox, oy, oz};
    for (int k = 0; k < 3; ++k) {
        st->t[k] = st->tlast[k] = pose->t[k] = t[k];
        st->td[k] = (double)t[k];
        st->off[k] = o[k];
    }
    if (history && st->frame > 0) {
        float* row = history + (size_t)((st->frame - 1) % capacity) * NUIF_POSE_ROW;
        for (int k = 0; k < 3; ++k) row[9 + k] = t[k];
    }
}

// ---- the allocation-free variant's scratch ------------------------------------------
// Partial sums of nuif_icp_estimate_static: a file-scope device array, one copy per GPU
// for the life of the process.  Host code never passes `scratch` itself (its name is a
// host-side stand-in, not a device address): the address comes from
// hipGetSymbolAddress, resolved once per device with that device current.
constexpr size_t kStaticScratchBytes = 1u << 20;
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// The current device's address of `scratch` (resolved on first use for the device).
int static_scratch(void** out) {
    int dev = 0;
    NUIF_HIP(hipGetDevice(&dev));
    NUIF_CHECK_ARG(dev >= 0 && dev < kMaxDevices, "device %d beyond the scratch table", dev);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        NUIF_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        NUIF_CHECK_ARG(p != nullptr, "hipGetSymbolAddress returned no address for the ICP scratch");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return NUIF_OK;
}

// ---- host side ---------------------------------------------------------------------

int device_cus() {
    static std::once_flag once;
    static int cus = 256;
    std::call_once(once, [] {
        int dev = 0, n = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && n > 0)
            cus = n;
    });
    return cus;
}

int make_job(const nuif_icp_session& in, IcpJob* j) {
    NUIF_CHECK_ARG(in.state && in.partials && in.ref_vmap && in.ref_nmap && in.pose, "ICP session: null argument");
    const nuif_icp_config& cfg = in.cfg;
    NUIF_CHECK_ARG(cfg.levels >= 1 && cfg.levels <= NUIF_ICP_MAX_LEVELS, "ICP session: bad level count %d",
                   cfg.levels);
    NUIF_CHECK_ARG(cfg.reduction == NUIF_ICP_REDUCE_REFERENCE || cfg.reduction == NUIF_ICP_REDUCE_COMPENSATED ||
                       cfg.reduction == NUIF_ICP_REDUCE_JOINT,
                   "ICP session: unknown reduction %d", cfg.reduction);
    NUIF_CHECK_ARG(!in.history || in.capacity > 0, "ICP session: history capacity must be positive");
    const int W = (int)in.cam.depthImageWidth, H = (int)in.cam.depthImageHeight;
    NUIF_CHECK_ARG(W > 0 && H > 0, "ICP session: empty image");
    NUIF_CHECK_ARG(((size_t)W * H + 127) / 128 <= (size_t)kMaxChunks * 128,
                   "ICP session: %d x %d pixels exceed the limit of %d", W, H, kMaxChunks * 128 * 128);
    IcpJob& J = *j;
    J = IcpJob{};
    J.st = (IcpState*)in.state;
    J.partials = in.partials;
    for (int l = 0; l < NUIF_ICP_MAX_LEVELS; ++l) {
        J.vmap[l] = l < cfg.levels ? in.vmap_pyr[l] : nullptr;
        J.iters[l] = l < cfg.levels ? cfg.iterations[l] : 0;
        NUIF_CHECK_ARG(J.iters[l] >= 0 && J.iters[l] < kEpochIters, "ICP session: %d iterations at level %d",
                       J.iters[l], l);
        NUIF_CHECK_ARG(J.iters[l] == 0 || (J.vmap[l] && (W >> l) > 0 && (H >> l) > 0),
                       "ICP session: level %d has iterations but no vertices", l);
    }
    J.vref = in.ref_vmap;
    J.nref = in.ref_nmap;
    J.pose = in.pose;
    J.history = in.history;
    J.capacity = in.capacity;
    J.outcome = in.outcome;
    J.cam = in.cam;
    const float L = (float)cfg.levels;
    J.dist = cfg.dist_threshold + cfg.dist_threshold / L;  // :300-304
    J.gate_threshold = cfg.gate_threshold;
    int rc = nuif_icp_count_thresholds(W, H, &cfg, J.thr);
    if (rc) return rc;
    J.levels = cfg.levels;
    J.mode = cfg.reduction;
    if (J.mode == kJoint) {
        const nuif_icp_photometric& ph = in.photometric;
        NUIF_CHECK_ARG(ph.ref_model && ph.weight >= 0.f, "ICP session: the joint order needs the photometric inputs");
        NUIF_CHECK_ARG(W % (1 << (cfg.levels - 1)) == 0 && H % (1 << (cfg.levels - 1)) == 0,
                       "ICP session: the joint order needs an image size divisible by 2^(levels - 1)");
        for (int l = 0; l < NUIF_ICP_MAX_LEVELS; ++l) {
            J.icur[l] = l < cfg.levels ? ph.intensity_pyr[l] : nullptr;
            NUIF_CHECK_ARG(J.iters[l] == 0 || J.icur[l], "ICP session: level %d has no intensities", l);
        }
        J.model = reinterpret_cast<const float4*>(ph.ref_model);
        J.psw = sqrtf(ph.weight);
        J.pdiff = ph.max_difference;
        J.pgrad2 = ph.min_gradient * ph.min_gradient;
    }
    return NUIF_OK;
}

// tile blocks the device holds at once (the cap of a batch step's grid)
template <int MODE> int resident_tile_blocks() {
    static std::once_flag once;
    static int blocks = 0;
    std::call_once(once, [] {
        int per_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, icp_tiles_kernel<MODE>, kBlock, 0) != hipSuccess ||
            per_cu <= 0)
            per_cu = 1;
        blocks = per_cu * device_cus();
    });
    return blocks;
}

// Epochs a session runs at most (every iteration of every level).
int max_epochs(const IcpJob& J) {
    int n = 0;
    for (int l = 0; l < NUIF_ICP_MAX_LEVELS; ++l) n += J.iters[l];
    return n;
}

// The most tiles a session can have open at step k: a level is reachable at step k if
// the coarser levels with iterations can take k steps between them, at least one and
// at most all of their iterations each.
uint32_t step_tiles(const IcpJob& J, int k) {
    uint32_t most = 0;
    int least = 0, longest = 0;  // steps the coarser levels take at least / at most
    for (int l = J.levels - 1; l >= 0; --l) {
        if (J.iters[l] <= 0) continue;
        if (k >= least && k < longest + J.iters[l]) most = std::max(most, (uint32_t)level_tiles(J.cam, l));
        least += 1;
        longest += J.iters[l];
    }
    return most;
}

size_t icp_job_bytes() { return sizeof(IcpJob); }

// The jobs of n sessions grouped by order (reference, compensated, joint: one series of
// steps per order); counts[0..1] receive the reference and compensated numbers.
int make_icp_jobs(const nuif_icp_session* sessions, int n, void* host_out, int* counts) {
    IcpJob* jobs = static_cast<IcpJob*>(host_out);
    for (int i = 0; i < n; ++i) {
        const int rc = make_job(sessions[i], &jobs[i]);
        if (rc) return rc;
        for (int j = 0; j < i; ++j)
            NUIF_CHECK_ARG(jobs[j].st != jobs[i].st, "ICP batch: a state appears twice");
    }
    std::stable_sort(jobs, jobs + n, [](const IcpJob& a, const IcpJob& b) { return a.mode < b.mode; });
    counts[0] = (int)std::count_if(jobs, jobs + n, [](const IcpJob& j) { return j.mode == kRefOrder; });
    counts[1] = (int)std::count_if(jobs, jobs + n, [](const IcpJob& j) { return j.mode == kCompensated; });
    return NUIF_OK;
}

// The lockstep steps of n sessions of one order: per step, the tile launch (grid: the
// most tiles the sessions can have open, capped at what the device holds at once)
// and the finish launch.
template <int MODE>
void launch_steps(const IcpJob* dj, const IcpJob* hj, int n, hipStream_t st) {
    int steps = 0;
    for (int s = 0; s < n; ++s) steps = std::max(steps, max_epochs(hj[s]));
    const size_t cap = (size_t)resident_tile_blocks<MODE>();
    const size_t lds = (size_t)(n + 1) * sizeof(uint32_t);
    const int remap = xcd_remap(kXcdIcp) ? 1 : 0;
    for (int k = 0; k < steps; ++k) {
        size_t tiles = 0;
        for (int s = 0; s < n; ++s) tiles += step_tiles(hj[s], k);
        unsigned grid = (unsigned)std::max<size_t>(1, std::min(tiles, cap));
        if (grid >= kXcds) grid -= grid % kXcds;  // whole rounds of the XCDs (see xcd_block)
        icp_tiles_kernel<MODE><<<grid, kBlock, lds, st>>>(dj, n, IcpJob{}, kAnyEpoch, remap);
        icp_finish_kernel<MODE><<<n, kFinBlock, 0, st>>>(dj, IcpJob{}, kAnyEpoch);
    }
}

// begin (all sessions), the steps of each order over its part of the table, end (all
// sessions); dtable: device copy of hjobs
int launch_icp_jobs(const void* dtable, const void* hjobs, int n, const int* counts, hipStream_t st) {
    if (n <= 0) return NUIF_OK;
    NUIF_CHECK_ARG(n <= kMaxBatch, "ICP batch: %d sessions exceed the limit of %d", n, kMaxBatch);
    const IcpJob* dj = static_cast<const IcpJob*>(dtable);
    const IcpJob* hj = static_cast<const IcpJob*>(hjobs);
    icp_begin_kernel<<<n, 64, 0, st>>>(dj, IcpJob{});
    NUIF_LAUNCHED("icp_begin_kernel");
    const int nref = counts[0], ncomp = counts[1], njoint = n - nref - ncomp;
    if (nref > 0) launch_steps<kRefOrder>(dj, hj, nref, st);
    if (ncomp > 0) launch_steps<kCompensated>(dj + nref, hj + nref, ncomp, st);
    if (njoint > 0) launch_steps<kJoint>(dj + nref + ncomp, hj + nref + ncomp, njoint, st);
    NUIF_LAUNCHED("icp_tiles_kernel / icp_finish_kernel");
    icp_end_kernel<<<n, 64, 0, st>>>(dj, IcpJob{});
    NUIF_LAUNCHED("icp_end_kernel");
    return NUIF_OK;
}

// one session: begin, the pair of launches of each planned epoch in the frame's order,
// end (the job as a kernel argument)
int run_single(const IcpJob& J, hipStream_t st, bool end) {
    icp_begin_kernel<<<1, 64, 0, st>>>(nullptr, J);
    NUIF_LAUNCHED("icp_begin_kernel");
    auto launch = [&](int level, int iter) {
        const uint32_t e = (uint32_t)(level * kEpochIters + iter);
        const unsigned grid = (unsigned)level_tiles(J.cam, level);
        const size_t lds = 2 * sizeof(uint32_t);
        const int remap = xcd_remap(kXcdIcp) ? 1 : 0;
        if (J.mode == kRefOrder) {
            icp_tiles_kernel<kRefOrder><<<grid, kBlock, lds, st>>>(nullptr, 1, J, e, remap);
            icp_finish_kernel<kRefOrder><<<1, kFinBlock, 0, st>>>(nullptr, J, e);
        } else if (J.mode == kJoint) {
            icp_tiles_kernel<kJoint><<<grid, kBlock, lds, st>>>(nullptr, 1, J, e, remap);
            icp_finish_kernel<kJoint><<<1, kFinBlock, 0, st>>>(nullptr, J, e);
        } else {
            icp_tiles_kernel<kCompensated><<<grid, kBlock, lds, st>>>(nullptr, 1, J, e, remap);
            icp_finish_kernel<kCompensated><<<1, kFinBlock, 0, st>>>(nullptr, J, e);
        }
    };
    if (J.single) {
        launch((int)(J.first_epoch / kEpochIters), (int)(J.first_epoch % kEpochIters));
    } else {
        for (int level = J.levels - 1; level >= 0; --level)
            for (int it = 0; it < J.iters[level]; ++it) launch(level, it);
    }
    NUIF_LAUNCHED("icp_tiles_kernel / icp_finish_kernel");
    if (end) {
        icp_end_kernel<<<1, 64, 0, st>>>(nullptr, J);
        NUIF_LAUNCHED("icp_end_kernel");
    }
    return NUIF_OK;
}

}  // namespace nuif

using namespace nuif;

extern "C" {

int nuif_icp_set_translation(void* state, NuiCLRigidTransform* pose, float* history, int capacity, const float* t3,
                             const float* offset3, nuif_stream_t s) {
    NUIF_REQUIRE_INIT();
    NUIF_CHECK_ARG(state && pose && t3 && offset3, "nuif_icp_set_translation: null argument");
    NUIF_CHECK_ARG(!history || capacity > 0, "nuif_icp_set_translation: bad history capacity");
    icp_set_translation_kernel<<<1, 1, 0, stream_of(s)>>>((IcpState*)state, pose, history, capacity, t3[0], t3[1],
                                                         t3[2], offset3[0], offset3[1], offset3[2]);
    NUIF_LAUNCHED("icp_set_translation_kernel");
    return NUIF_OK;
}

size_t nuif_icp_state_bytes(void) { return kStateBytes + kSyncBytes; }

size_t nuif_icp_partials_bytes(int width, int height) {
    if (width <= 0 || height <= 0) return 0;
    const size_t npx = (size_t)width * height;
    const size_t compensated = (npx + kTilePx - 1) / kTilePx * kTerms * sizeof(double);
    const size_t groups = (npx + 127) / 128;
    // reference order: the group sums of whole tiles
    const size_t tgroups = (groups + kTileGroups - 1) / kTileGroups * kTileGroups;
    const size_t reference = tgroups * kTerms * sizeof(float);
    return std::max(compensated, reference);
}

const int32_t* nuif_icp_gate_flag(const void* state) {
    return state ? &reinterpret_cast<const IcpState*>(state)->gate : nullptr;
}

const int32_t* nuif_icp_ok_flag(const void* state) {
    return state ? &reinterpret_cast<const IcpState*>(state)->ok : nullptr;
}

int nuif_icp_state_init(void* state, const NuiCLRigidTransform* pose, nuif_stream_t s) {
    NUIF_REQUIRE_INIT();
    NUIF_CHECK_ARG(state && pose, "nuif_icp_state_init: null argument");
    icp_state_init_kernel<<<1, 64, 0, stream_of(s)>>>((IcpState*)state, pose);
    NUIF_LAUNCHED("icp_state_init_kernel");
    return NUIF_OK;
}

int nuif_icp_state_read(const void* state, nuif_icp_result* out, nuif_stream_t s) {
    NUIF_REQUIRE_INIT();
    NUIF_CHECK_ARG(state && out, "nuif_icp_state_read: null argument");
    IcpState h;
    NUIF_HIP(hipMemcpyAsync(&h, state, sizeof(IcpState), hipMemcpyDeviceToHost, stream_of(s)));
    NUIF_HIP(hipStreamSynchronize(stream_of(s)));
    for (int k = 0; k < 9; ++k) out->R[k] = h.R[k];
    for (int k = 0; k < 3; ++k) out->t[k] = h.t[k];
    out->error = h.error;
    out->count = h.count;
    out->status = h.timeout ? 2 : h.failed ? 1 : 0;
    out->iterations = h.iterations;
    out->integrate = h.gate;
    out->frame = h.frame;
    for (int k = 0; k < kTerms; ++k) out->sums[k] = h.sums[k];
    for (int k = 0; k < 9; ++k) out->Rd[k] = h.Rd[k];
    for (int k = 0; k < 3; ++k) out->td[k] = h.td[k];
    for (int l = 0; l < NUIF_ICP_MAX_LEVELS; ++l) out->level_iterations[l] = h.level_iters[l];
    return NUIF_OK;
}

int nuif_icp_begin(void* state, const NuiCLRigidTransform* pose, nuif_stream_t s) {
    NUIF_REQUIRE_INIT();
    NUIF_CHECK_ARG(state && pose, "nuif_icp_begin: null argument");
    // a frame with no iterations: estimate <- pose, flags cleared
    IcpJob J{};
    J.st = (IcpState*)state;
    J.pose = const_cast<NuiCLRigidTransform*>(pose);
    J.levels = 1;
    J.mode = kRefOrder;
    icp_begin_kernel<<<1, 64, 0, stream_of(s)>>>(nullptr, J);
    NUIF_LAUNCHED("icp_begin_kernel");
    return NUIF_OK;
}

int nuif_icp_iteration(void* state, float* partials, const NuiCLCameraParams* cam, const float* vmap_level,
                       int level, const float* ref_vmap, const float* ref_nmap, const NuiCLRigidTransform* prev_pose,
                       float dist_threshold, uint32_t count_threshold, int iter, int reduction, nuif_stream_t s) {
    NUIF_REQUIRE_INIT();
    NUIF_CHECK_ARG(state && partials && cam && vmap_level && ref_vmap && ref_nmap && prev_pose,
                   "nuif_icp_iteration: null argument");
    NUIF_CHECK_ARG(level >= 0 && level < NUIF_ICP_MAX_LEVELS && iter >= 0 && iter < kEpochIters,
                   "nuif_icp_iteration: bad level/iter");
    NUIF_CHECK_ARG(reduction == NUIF_ICP_REDUCE_REFERENCE || reduction == NUIF_ICP_REDUCE_COMPENSATED,
                   "nuif_icp_iteration: unknown reduction %d", reduction);
    const int Wl = (int)cam->depthImageWidth >> level, Hl = (int)cam->depthImageHeight >> level;
    NUIF_CHECK_ARG(Wl > 0 && Hl > 0, "nuif_icp_iteration: level %d too small", level);
    NUIF_CHECK_ARG(((size_t)Wl * Hl + 127) / 128 <= (size_t)kMaxChunks * 128,
                   "nuif_icp_iteration: %d pixels per level exceed the limit %d", Wl * Hl, kMaxChunks * 128 * 128);
    IcpJob J{};
    J.st = (IcpState*)state;
    J.partials = partials;
    J.vmap[level] = vmap_level;
    J.vref = ref_vmap;
    J.nref = ref_nmap;
    J.pose = const_cast<NuiCLRigidTransform*>(prev_pose);
    J.cam = *cam;
    J.dist = dist_threshold;
    J.thr[level] = count_threshold;
    J.iters[level] = iter + 1;
    J.levels = level + 1;
    J.mode = reduction;
    J.single = 1;
    J.first_epoch = level * kEpochIters + iter;
    return run_single(J, stream_of(s), false);
}

int nuif_icp_end(void* state, NuiCLRigidTransform* pose, float gate_threshold, float* history, int capacity,
                 nuif_stream_t s) {
    NUIF_REQUIRE_INIT();
    NUIF_CHECK_ARG(state && pose, "nuif_icp_end: null argument");
    NUIF_CHECK_ARG(!history || capacity > 0, "nuif_icp_end: history capacity must be positive");
    IcpJob J{};
    J.st = (IcpState*)state;
    J.pose = pose;
    J.gate_threshold = gate_threshold;
    J.history = history;
    J.capacity = capacity;
    icp_end_kernel<<<1, 64, 0, stream_of(s)>>>(nullptr, J);
    NUIF_LAUNCHED("icp_end_kernel");
    return NUIF_OK;
}

int nuif_icp_record(void* state, NuiCLRigidTransform* pose, float* history, int capacity, nuif_stream_t s) {
    NUIF_REQUIRE_INIT();
    NUIF_CHECK_ARG(state && pose, "nuif_icp_record: null argument");
    NUIF_CHECK_ARG(!history || capacity > 0, "nuif_icp_record: history capacity must be positive");
    icp_record_kernel<<<1, 1, 0, stream_of(s)>>>((IcpState*)state, pose, history, capacity);
    NUIF_LAUNCHED("icp_record_kernel");
    return NUIF_OK;
}

int nuif_icp_count_thresholds(int W, int H, const nuif_icp_config* cfg, uint32_t* thr) {
    NUIF_CHECK_ARG(cfg && thr && W > 0 && H > 0, "nuif_icp_count_thresholds: bad arguments");
    NUIF_CHECK_ARG(cfg->levels >= 1 && cfg->levels <= NUIF_ICP_MAX_LEVELS, "bad level count %d", cfg->levels);
    // OCL/NuiKinfuOpenCLDepthTracker.cpp:309: (kernelGlobalSize[0] >> 2l) * 0.1 where
    // kernelGlobalSize[0] is left over from the previous level's compute_sums launch
    size_t kgs = (size_t)W * H;
    for (int level = cfg->levels - 1; level >= 0; --level) {
        thr[level] = (uint32_t)((double)(kgs >> (2 * level)) * 0.1);
        if (cfg->iterations[level] > 0) {
            const size_t groups = ((size_t)(W >> level) * (H >> level) + 127) / 128;
            kgs = (groups + 127) / 128 * 128;
        }
    }
    for (int level = cfg->levels; level < NUIF_ICP_MAX_LEVELS; ++level) thr[level] = 0;
    return NUIF_OK;
}

int nuif_icp_estimate(void* state, float* partials, const NuiCLCameraParams* cam, const float* const* vmap_pyr,
                      const float* ref_vmap, const float* ref_nmap, NuiCLRigidTransform* pose,
                      const nuif_icp_config* cfg, float* history, int capacity, nuif_stream_t s) {
    NUIF_REQUIRE_INIT();
    NUIF_CHECK_ARG(state && partials && cam && vmap_pyr && ref_vmap && ref_nmap && pose && cfg,
                   "nuif_icp_estimate: null argument");
    nuif_icp_session in{};
    in.state = state;
    in.partials = partials;
    in.cam = *cam;
    for (int l = 0; l < NUIF_ICP_MAX_LEVELS; ++l) in.vmap_pyr[l] = l < cfg->levels ? vmap_pyr[l] : nullptr;
    in.ref_vmap = ref_vmap;
    in.ref_nmap = ref_nmap;
    in.pose = pose;
    in.cfg = *cfg;
    in.history = history;
    in.capacity = capacity;
    IcpJob J;
    const int rc = make_job(in, &J);
    if (rc) return rc;
    return run_single(J, stream_of(s), true);
}

int nuif_icp_estimate_joint(void* state, float* partials, const NuiCLCameraParams* cam, const float* const* vmap_pyr,
                            const float* ref_vmap, const float* ref_nmap, NuiCLRigidTransform* pose,
                            const nuif_icp_config* cfg, const nuif_icp_photometric* photometric, float* history,
                            int capacity, nuif_stream_t s) {
    NUIF_REQUIRE_INIT();
    NUIF_CHECK_ARG(state && partials && cam && vmap_pyr && ref_vmap && ref_nmap && pose && cfg && photometric,
                   "nuif_icp_estimate_joint: null argument");
    NUIF_CHECK_ARG(cfg->reduction == NUIF_ICP_REDUCE_JOINT, "nuif_icp_estimate_joint: the reduction must be joint");
    nuif_icp_session in{};
    in.state = state;
    in.partials = partials;
    in.cam = *cam;
    for (int l = 0; l < NUIF_ICP_MAX_LEVELS; ++l) in.vmap_pyr[l] = l < cfg->levels ? vmap_pyr[l] : nullptr;
    in.ref_vmap = ref_vmap;
    in.ref_nmap = ref_nmap;
    in.pose = pose;
    in.cfg = *cfg;
    in.history = history;
    in.capacity = capacity;
    in.photometric = *photometric;
    IcpJob J;
    const int rc = make_job(in, &J);
    if (rc) return rc;
    return run_single(J, stream_of(s), true);
}

size_t nuif_icp_static_scratch_bytes(void) { return kStaticScratchBytes; }

int nuif_icp_estimate_static(void* state, const NuiCLCameraParams* cam, const float* const* vmap_pyr,
                             const float* ref_vmap, const float* ref_nmap, NuiCLRigidTransform* pose,
                             const nuif_icp_config* cfg, float* history, int capacity, nuif_stream_t s) {
    NUIF_REQUIRE_INIT();
    NUIF_CHECK_ARG(cam, "nuif_icp_estimate_static: null argument");
    const size_t need = nuif_icp_partials_bytes((int)cam->depthImageWidth, (int)cam->depthImageHeight);
    NUIF_CHECK_ARG(need <= kStaticScratchBytes,
                   "nuif_icp_estimate_static: a %u x %u image needs %zu bytes of partial sums, above the %zu-byte "
                   "static scratch (use nuif_icp_estimate with a workspace)",
                   cam->depthImageWidth, cam->depthImageHeight, need, kStaticScratchBytes);
    void* ptr = nullptr;
    const int rc = static_scratch(&ptr);
    if (rc) return rc;
    return nuif_icp_estimate(state, (float*)ptr, cam, vmap_pyr, ref_vmap, ref_nmap, pose, cfg, history, capacity, s);
}

size_t nuif_icp_batch_table_bytes(int n) { return n > 0 ? ((size_t)n * sizeof(IcpJob) + 255) / 256 * 256 : 0; }

int nuif_icp_estimate_batch(const nuif_icp_session* sessions, int n, void* table, nuif_stream_t s) {
    NUIF_REQUIRE_INIT();
    NUIF_CHECK_ARG(n >= 0 && (n == 0 || (sessions && table)), "nuif_icp_estimate_batch: bad arguments");
    if (n == 0) return NUIF_OK;
    std::vector<IcpJob> jobs((size_t)n);
    int counts[2] = {0, 0};
    int rc = make_icp_jobs(sessions, n, jobs.data(), counts);
    if (rc) return rc;
    hipStream_t st = stream_of(s);
    NUIF_HIP(hipMemcpyAsync(table, jobs.data(), (size_t)n * sizeof(IcpJob), hipMemcpyHostToDevice, st));
    return launch_icp_jobs(table, jobs.data(), n, counts, st);
}

}  // extern "C"
