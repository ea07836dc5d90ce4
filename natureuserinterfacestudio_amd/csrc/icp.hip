// Projective point-to-plane ICP.
//
// The reference runs per iteration: icp_block_kernel (128-wide LDS trees of
// 29 floats), compute_sums, a blocking read-back, and the Levenberg-Marquardt
// bookkeeping + Eigen FP64 6x6 solve on the host
// (OCL/NuiKinfuOpenCLDepthTracker.cpp:240-504).  Here every iteration is ONE
// device launch and the whole pyramid runs without a host round trip:
//
//   icp_ref_kernel   reference order: one 128-item block per 128 pixels and the
//                    reference's tree pairing (item i takes item i + d for
//                    d = 64..1) as one LDS step plus wave shuffles; the last
//                    block of every 128 to finish runs compute_sums' tree over
//                    their partials, and the last of those sums the group
//                    results in group order, as the host does, then solves.
//   icp_comp_kernel  compensated: grid-stride blocks, Kahan FP32 per thread,
//                    FP64 across threads; the last block to finish sums the
//                    block partials in block order in FP64, then solves.
//
// Completion is counted with device-scope atomics after a release fence, and
// the counting block acquires before it reads other blocks' partials.  The
// block that completes a counter re-arms it, so the next launch starts from
// zero without a fill.  Every sum is taken in a fixed order: the result does
// not depend on which block finishes last.
//
// The solve runs on one thread over an LDS copy of the state: the reference's
// accept/reject state machine (count threshold left over from the previous
// level, constant distance threshold, reject restores the last good pose),
// the damped 6x6 system normalised by the count, solved in FP64 (LLT in
// reference order, LDL^T compensated), and the Euler-angle increment.
//
// The tracking state (estimate, last good pose, lambda, break flags, last
// integration pose, integration gate) lives in one device block so that
// integrate/raycast launches can consume the ICP outcome without the host.

#include "det_math.h"
#include "icp_math.h"
#include "nuif_internal.h"

namespace nuif {

constexpr int kTerms = NUIF_ICP_TERMS;
constexpr int kBlock = 256;
constexpr int kMaxBlocks = 256;    // compensated: blocks per launch (grid-stride over the pixels)
constexpr int kMaxGroups2 = 256;   // reference: 128-block groups per launch (W*H <= 2^22)

struct IcpState {
    float R[9], t[3];          // current estimate (Rcurr, tcurr)
    float Rg[9], tg[3];        // last accepted (Rgood, tgood)
    float Rlast[9], tlast[3];  // pose of the last integration
    float lambda, err_old;
    float error, count;
    int32_t brk[NUIF_ICP_MAX_LEVELS];
    int32_t failed, iterations, gate, frame, ok, pad;
    double good[27];       // compensated mode: accepted sums
    double sums[kTerms];   // sums of the last executed iteration (both modes)
    float goodf[27];       // reference mode: goodCores as the reference keeps them (float)
    float pad2;
    float off[3];          // shifting volume: offset of the local frame (metres), 0 otherwise
    float pad3;
};
static_assert(sizeof(IcpState) % 8 == 0, "state is copied as 32-bit words");

// Completion counters, after the state block in the same allocation
struct IcpSync {
    uint32_t arrive[kMaxGroups2];  // reference: finished blocks of each 128-block group
    uint32_t arrive_all;           // finished groups (reference) or blocks (compensated)
    uint32_t pad[63];
};
constexpr size_t kStateBytes = (sizeof(IcpState) + 255) / 256 * 256;
constexpr size_t kSyncBytes = (sizeof(IcpSync) + 255) / 256 * 256;

__device__ __forceinline__ IcpSync* sync_of(IcpState* st) {
    return reinterpret_cast<IcpSync*>(reinterpret_cast<char*>(st) + kStateBytes);
}

__device__ __forceinline__ void clear_sync(IcpSync* sy, int tid, int nthreads) {
    for (int i = tid; i < kMaxGroups2; i += nthreads) sy->arrive[i] = 0u;
    if (tid == 0) sy->arrive_all = 0u;
}

// Reduction orders (nuif_icp_config::reduction)
constexpr int kRefOrder = 0;      // the reference's float trees, bit for bit
constexpr int kCompensated = 1;   // Kahan FP32 per thread + FP64 across blocks

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

// LDS: 29 terms x 8 segments x 33 (padded) floats, reused as doubles later
constexpr int kSeg = 8, kSegLen = 32, kSegPad = 33;
static_assert(kMaxBlocks % kSeg == 0, "the last block reads kMaxBlocks / kSeg partials per segment");

template <int MODE>
__device__ void finish_iteration(IcpState* st, IcpState& S, const double* sums, const float* sumsf, int level,
                                 int iter, uint32_t count_threshold);

// The state and the partials are written and read by different blocks of one
// launch: no __restrict__ on them.
__global__ void __launch_bounds__(kBlock)
    icp_comp_kernel(IcpState* st, const float* __restrict__ vcur, int npx, NuiCLCameraParams cam,
                    const float* __restrict__ vref, const float* __restrict__ nref,
                    const NuiCLRigidTransform* __restrict__ prev, float distThr, int level, int iter,
                    uint32_t count_threshold, double* partials) {
    if (st->failed || st->brk[level]) return;
    __shared__ float lds[kTerms * kSeg * kSegPad];
    __shared__ double part[kTerms * kSeg];
    __shared__ double sums[kTerms];
    __shared__ IcpState S;
    __shared__ int last;

    float R[9], t[3], Ri[9], tp[3];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        R[k] = st->R[k];
        Ri[k] = prev->R_inv[k];
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        t[k] = st->t[k];
        tp[k] = prev->t[k];
    }
    float acc[kTerms], cmp[kTerms];
#pragma unroll
    for (int k = 0; k < kTerms; ++k) acc[k] = cmp[k] = 0.f;

    const int W = (int)cam.depthImageWidth, H = (int)cam.depthImageHeight;
    // kPB pixels of the grid-stride sequence at a time: their gathers are in flight
    // together, their terms are accumulated in the sequence's order
    constexpr int kPB = 4;
    const int stride = gridDim.x * kBlock;
    for (int p0 = blockIdx.x * kBlock + threadIdx.x; p0 < npx; p0 += kPB * stride) {
        f3 pv[kPB], rn[kPB], rv[kPB];
        bool ok[kPB];
#pragma unroll
        for (int j = 0; j < kPB; ++j) {
            const int p = p0 + j * stride;
            const f3 vc = p < npx ? ld3(vcur, p) : mk3(kNaN, kNaN, kNaN);
            // K/icp.cl:243-289
            pv[j] = add3(rot3(R, vc), mk3(t[0], t[1], t[2]));
            const f3 pp = rot3(Ri, sub3(pv[j], mk3(tp[0], tp[1], tp[2])));
            const float px = pp.x * cam.fx / pp.z + cam.cx, py = pp.y * cam.fy / pp.z + cam.cy;
            ok[j] = !nan3(vc) && px >= 0.0f && px < (float)W && py >= 0.0f && py < (float)H;
            rn[j] = ok[j] ? bilinear_holes(nref, px, py, W, H) : mk3(kNaN, kNaN, kNaN);
            rv[j] = ok[j] ? bilinear_holes(vref, px, py, W, H) : mk3(kNaN, kNaN, kNaN);
        }
#pragma unroll
        for (int j = 0; j < kPB; ++j) {
            if (!ok[j] || nan3(rn[j]) || nan3(rv[j])) continue;
            const f3 diff = sub3(rv[j], pv[j]);
            const float dist = sqrtf(dot3(diff, diff));
            if (!(dist < distThr)) continue;
            const f3 row0 = cross3(rn[j], pv[j]);
            const float c[7] = {row0.x, row0.y, row0.z, rn[j].x, rn[j].y, rn[j].z, dot3(rn[j], diff)};
            int s = 0;
#pragma unroll
            for (int i = 0; i < 6; ++i)
#pragma unroll
                for (int q = i; q < 7; ++q) {
                    kahan_add(acc[s], cmp[s], c[i] * c[q]);
                    ++s;
                }
            kahan_add(acc[27], cmp[27], c[6] * c[6]);
            acc[28] += 1.0f;  // counts are exact in FP32 up to 2^24
        }
    }

    // transpose through LDS: term k, thread tid -> lds[k][tid / 32][tid % 32]
    const int tid = threadIdx.x;
#pragma unroll
    for (int k = 0; k < kTerms; ++k) lds[k * kSeg * kSegPad + (tid / kSegLen) * kSegPad + (tid % kSegLen)] = acc[k];
    __syncthreads();
    if (tid < kTerms * kSeg) {
        const int k = tid / kSeg, sg = tid % kSeg;
        const float* src = &lds[k * kSeg * kSegPad + sg * kSegPad];
        float s = 0.f, c = 0.f;
        for (int j = 0; j < kSegLen; ++j) kahan_add(s, c, src[j]);
        part[k * kSeg + sg] = (double)s - (double)c;
    }
    __syncthreads();
    if (tid < kTerms) {
        double s = 0.0;
        for (int j = 0; j < kSeg; ++j) s += part[tid * kSeg + j];
        partials[(size_t)blockIdx.x * kTerms + tid] = s;
    }

    // the last block to finish reduces the block partials and solves
    IcpSync* sy = sync_of(st);
    __syncthreads();  // the partial (threads 0..28) is written
    if (tid == 0) {
        fence_release_agent();  // release this block's partial
        last = atomicAdd(&sy->arrive_all, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (!last) return;
    fence_acquire_agent();  // acquire the other blocks' partials
    const int nblocks = (int)gridDim.x;
    if (tid < kTerms * kSeg) {
        // term k, segment j: blocks j, j + 8, ... in order; the loads of all of them
        // (at most kMaxBlocks / kSeg) are issued together, the FP64 sum keeps the
        // fixed order and adds only blocks that exist
        const int k = tid / kSeg, j = tid % kSeg;
        constexpr int kPer = kMaxBlocks / kSeg;
        double v[kPer];
#pragma unroll
        for (int q = 0; q < kPer; ++q) {
            const int b = j + q * kSeg;
            v[q] = b < nblocks ? partials[(size_t)b * kTerms + k] : 0.0;
        }
        double s = 0.0;
#pragma unroll
        for (int q = 0; q < kPer; ++q)
            if (j + q * kSeg < nblocks) s += v[q];
        part[k * kSeg + j] = s;
    }
    __syncthreads();
    if (tid < kTerms) {
        double s = 0.0;
        for (int j = 0; j < kSeg; ++j) s += part[tid * kSeg + j];
        sums[tid] = s;
    }
    if (tid == 0) sy->arrive_all = 0u;  // re-arm: every block has arrived
    finish_iteration<kCompensated>(st, S, sums, nullptr, level, iter, count_threshold);
}

// One pixel's 29 terms of K/icp.cl:243-289 (zeros when rejected).
__device__ __forceinline__ void icp_terms(const f3 vc, const float* R, const float* t, const float* Ri,
                                          const float* tp, const NuiCLCameraParams& cam,
                                          const float* __restrict__ vref, const float* __restrict__ nref,
                                          float distThr, float out[kTerms]) {
#pragma unroll
    for (int k = 0; k < kTerms; ++k) out[k] = 0.f;
    if (nan3(vc)) return;
    const int W = (int)cam.depthImageWidth, H = (int)cam.depthImageHeight;
    const f3 pv = add3(rot3(R, vc), mk3(t[0], t[1], t[2]));
    const f3 pp = rot3(Ri, sub3(pv, mk3(tp[0], tp[1], tp[2])));
    const float px = pp.x * cam.fx / pp.z + cam.cx, py = pp.y * cam.fy / pp.z + cam.cy;
    if (!(px >= 0.0f && px < (float)W && py >= 0.0f && py < (float)H)) return;
    // both lookups are issued before either is tested (pure reads)
    const f3 rn = bilinear_holes(nref, px, py, W, H);
    const f3 rv = bilinear_holes(vref, px, py, W, H);
    if (nan3(rn) || nan3(rv)) return;
    const f3 diff = sub3(rv, pv);
    const float dist = sqrtf(dot3(diff, diff));
    if (!(dist < distThr)) return;
    const f3 row0 = cross3(rn, pv);
    const float c[7] = {row0.x, row0.y, row0.z, rn.x, rn.y, rn.z, dot3(rn, diff)};
    int s = 0;
#pragma unroll
    for (int i = 0; i < 6; ++i)
#pragma unroll
        for (int j = i; j < 7; ++j) out[s++] = c[i] * c[j];
    out[27] = c[6] * c[6];
    out[28] = 1.0f;
}

// Reference order, one launch per iteration.  Group g (128 items) takes pixels
// 128g..128g+127; a partial last group counts its missing items as zeros
// (K/icp.cl:222-292).  A block holds kRefGroups groups, each reduced with the
// reference's 128-wide tree on its own 128 threads.  The partials buffer holds
// ngroups x 29 group sums followed by ceil(ngroups / 128) x 29 compute_sums
// results; the last block of every 128 groups to finish runs compute_sums' tree
// over them, the last of those sums the results in order and solves.
constexpr int kRefGroups = 4;
static_assert(128 % kRefGroups == 0, "a block's groups lie in one compute_sums chunk");
__global__ void __launch_bounds__(128 * kRefGroups)
    icp_ref_kernel(IcpState* st, const float* __restrict__ vcur, int npx, NuiCLCameraParams cam,
                   const float* __restrict__ vref, const float* __restrict__ nref,
                   const NuiCLRigidTransform* __restrict__ prev, float distThr, int level, int iter,
                   uint32_t count_threshold, float* partials) {
    if (st->failed || st->brk[level]) return;
    __shared__ float xfer[kRefGroups][64 * kTerms];
    __shared__ float sumsf[kTerms];
    __shared__ double sums[kTerms];
    __shared__ IcpState S;
    __shared__ int last, last2;
    float R[9], t[3], Ri[9], tp[3];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        R[k] = st->R[k];
        Ri[k] = prev->R_inv[k];
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        t[k] = st->t[k];
        tp[k] = prev->t[k];
    }
    const int tid = threadIdx.x, sub = tid >> 7, l = tid & 127, lane = l & 63, wave = l >> 6;
    const int ngroups = (npx + 127) >> 7;
    const int g = blockIdx.x * kRefGroups + sub;
    const int p = g * 128 + l;
    float v[kTerms];
    const f3 vc = p < npx ? ld3(vcur, p) : mk3(kNaN, kNaN, kNaN);
    icp_terms(vc, R, t, Ri, tp, cam, vref, nref, distThr, v);

    // K/icp.cl:162-190 reduce(): step d = 64 through LDS, steps 32..1 in the group's wave 0
    if (wave == 1) {
#pragma unroll
        for (int k = 0; k < kTerms; ++k) xfer[sub][lane * kTerms + k] = v[k];
    }
    __syncthreads();
    if (wave == 0) {
#pragma unroll
        for (int k = 0; k < kTerms; ++k) v[k] = wave_tree64(v[k] + xfer[sub][lane * kTerms + k]);
        if (lane == 0 && g < ngroups) {
#pragma unroll
            for (int k = 0; k < kTerms; ++k) partials[(size_t)g * kTerms + k] = v[k];
        }
    }
    __syncthreads();  // the block's group partials are written

    // the last block of each 128-group chunk runs compute_sums' tree over the chunk
    IcpSync* sy = sync_of(st);
    const int chunk = (blockIdx.x * kRefGroups) >> 7;
    if (tid == 0) {
        fence_release_agent();  // release this block's partials
        const int groups_in_chunk = min(128, ngroups - (chunk << 7));
        const uint32_t members = (uint32_t)((groups_in_chunk + kRefGroups - 1) / kRefGroups);
        last = atomicAdd(&sy->arrive[chunk], 1u) == members - 1;
    }
    __syncthreads();
    if (!last) return;
    fence_acquire_agent();  // acquire the chunk's partials
    float* tree = partials + (size_t)ngroups * kTerms;
    const int wid = tid >> 6, nwaves = (128 * kRefGroups) >> 6;
    for (int k = wid; k < kTerms; k += nwaves) {
        const int g0 = (chunk << 7) + (tid & 63), g1 = g0 + 64;
        const float x = g0 < ngroups ? partials[(size_t)g0 * kTerms + k] : 0.f;
        const float y = g1 < ngroups ? partials[(size_t)g1 * kTerms + k] : 0.f;
        const float sum = wave_tree64(x + y);
        if ((tid & 63) == 0) tree[chunk * kTerms + k] = sum;
    }
    __syncthreads();  // the chunk result is written (lanes 0 of the waves)
    const int nchunks = (ngroups + 127) >> 7;
    if (tid == 0) {
        sy->arrive[chunk] = 0u;  // re-arm: every block of the chunk has arrived
        fence_release_agent();  // release the chunk result
        last2 = atomicAdd(&sy->arrive_all, 1u) == (uint32_t)(nchunks - 1);
    }
    __syncthreads();
    if (!last2) return;
    fence_acquire_agent();  // acquire every chunk result

    // the host's sequential float sum over the compute_sums results
    // (OCL/NuiKinfuOpenCLDepthTracker.cpp:394-402, 428-440)
    if (tid < kTerms) {
        float sum = 0.f;
        for (int c = 0; c < nchunks; ++c) sum += tree[c * kTerms + tid];
        sumsf[tid] = sum;
        sums[tid] = (double)sum;
    }
    if (tid == 0) sy->arrive_all = 0u;
    finish_iteration<kRefOrder>(st, S, sums, sumsf, level, iter, count_threshold);
}

// ---- small FP64 linear algebra for the solve ----------------------------------

// A = L D L^T (unit lower L), in place; returns false if a pivot is not positive.
__device__ bool ldlt6(double A[36], double* det) {
    double d = 1.0;
    for (int j = 0; j < 6; ++j) {
        double Dj = A[j * 6 + j];
        for (int k = 0; k < j; ++k) Dj -= A[j * 6 + k] * A[j * 6 + k] * A[k * 6 + k];
        A[j * 6 + j] = Dj;
        d *= Dj;
        if (!(Dj > 0.0)) {
            *det = d;
            return false;
        }
        for (int i = j + 1; i < 6; ++i) {
            double v = A[i * 6 + j];
            for (int k = 0; k < j; ++k) v -= A[i * 6 + k] * A[j * 6 + k] * A[k * 6 + k];
            A[i * 6 + j] = v / Dj;
        }
    }
    *det = d;
    return true;
}

__device__ void ldlt6_solve(const double F[36], const double b[6], double x[6]) {
    double y[6];
    for (int i = 0; i < 6; ++i) {
        double s = b[i];
        for (int k = 0; k < i; ++k) s -= F[i * 6 + k] * y[k];
        y[i] = s;
    }
    for (int i = 0; i < 6; ++i) y[i] /= F[i * 6 + i];
    for (int i = 5; i >= 0; --i) {
        double s = y[i];
        for (int k = i + 1; k < 6; ++k) s -= F[k * 6 + i] * x[k];
        x[i] = s;
    }
}

__device__ void matmul3d(const double a[9], const double b[9], double c[9]) {
    double r[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r[3 * i + j] = a[3 * i] * b[j] + a[3 * i + 1] * b[3 + j] + a[3 * i + 2] * b[6 + j];
    for (int k = 0; k < 9; ++k) c[k] = r[k];
}

// The host loop of OCL/NuiKinfuOpenCLDepthTracker.cpp:402-497 for one iteration,
// given the iteration's sums.  Reference order: float sums and goodCores,
// Eigen-like LLT in double and the float Euler increment.  Compensated order:
// FP64 sums, LDL^T in FP64 and the increment composed in FP64.  The serial steps
// run on one thread over the LDS copy of the state; the 27 normalised entries of
// the system and (compensated order) the six sines and cosines of the increment
// are computed one per lane in between, the same operations on the same operands.

// :402-458: error/count, accept or reject (thread 0).  Returns whether a system
// is solved in this iteration.
template <int MODE>
__device__ bool icp_accept(IcpState& S, const double* sums, const float* sumsf, int level, int iter,
                           uint32_t count_threshold) {
    for (int k = 0; k < kTerms; ++k) S.sums[k] = sums[k];
    const float cnt = (float)sums[28];
    const float errsum = MODE == kRefOrder ? sumsf[27] : (float)sums[27];
    const float err = cnt > 0.f ? sqrtf(errsum) / cnt : 3.402823466e38f;
    S.error = err;
    S.count = cnt;
    S.iterations += 1;
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
        for (int k = 0; k < 27; ++k) {
            S.good[k] = sums[k];
            S.goodf[k] = MODE == kRefOrder ? sumsf[k] : (float)sums[k];
        }
    }
    return true;
}

// Entry s of the system normalised by the current count (one lane each).
template <int MODE>
__device__ __forceinline__ double icp_entry(const IcpState& S, int s) {
    const float cnt = S.count;
    return MODE == kRefOrder ? (double)(S.goodf[s] / cnt) : S.good[s] / (double)cnt;
}

// The damped system, its factorisation and solve, the convergence test (thread 0).
// Returns whether the increment (angles al, be, ga, translation tinc) is applied.
template <int MODE>
__device__ bool icp_solve(IcpState& S, const double* Av, int level, float ang[3], float tinc[3]) {
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
    const bool spd = MODE == kRefOrder ? llt6(A, &det) : ldlt6(A, &det);
    if (!spd || fabs(det) < 1e-15 || isnan(det)) {
        S.failed = 1;
        return false;
    }
    if (MODE == kRefOrder)
        llt6_solve(A, b, x);
    else
        ldlt6_solve(A, b, x);
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
__device__ void icp_compose_ref(IcpState& S, const float ang[3], const float tinc[3]) {
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

// The same composition in FP64; trig = sin, cos of alpha, beta, gamma (as doubles).
__device__ void icp_compose_comp(IcpState& S, const double trig[6], const float tinc[3]) {
    const double sa = trig[0], ca = trig[1], sb = trig[2], cb = trig[3], sg = trig[4], cg = trig[5];
    const double Rz[9] = {cg, -sg, 0, sg, cg, 0, 0, 0, 1};
    const double Ry[9] = {cb, 0, sb, 0, 1, 0, -sb, 0, cb};
    const double Rx[9] = {1, 0, 0, 0, ca, -sa, 0, sa, ca};
    double Rzy[9], Rinc[9];
    matmul3d(Rz, Ry, Rzy);
    matmul3d(Rzy, Rx, Rinc);
    double Ri[9];  // inverse of a rotation: its transpose
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) Ri[3 * i + j] = Rinc[3 * j + i];
    double Rc[9], tc[3];
    for (int k = 0; k < 9; ++k) Rc[k] = S.R[k];
    for (int k = 0; k < 3; ++k) tc[k] = S.t[k];
    double Rn[9];
    matmul3d(Ri, Rc, Rn);
    for (int i = 0; i < 3; ++i)
        S.t[i] = (float)(Ri[3 * i] * tc[0] + Ri[3 * i + 1] * tc[1] + Ri[3 * i + 2] * tc[2] + (double)tinc[i]);
    for (int k = 0; k < 9; ++k) S.R[k] = (float)Rn[k];
}

// Run by the whole block that completes the reduction: copy the state into LDS,
// update it, write it back.  No other block reads the state after it has
// arrived, so the write-back races with nothing.  Every branch between the
// barriers reads a flag written before the previous barrier: it is uniform.
template <int MODE>
__device__ void finish_iteration(IcpState* st, IcpState& S, const double* sums, const float* sumsf, int level,
                                 int iter, uint32_t count_threshold) {
    __shared__ double Av[27], trig[6];
    __shared__ float ang[3], tinc[3];
    __shared__ int go;
    constexpr int kWords = (int)(sizeof(IcpState) / 4);
    const int tid = threadIdx.x;
    const uint32_t* src = reinterpret_cast<const uint32_t*>(st);
    uint32_t* lds = reinterpret_cast<uint32_t*>(&S);
    for (int i = tid; i < kWords; i += blockDim.x) lds[i] = src[i];
    __syncthreads();
    if (tid == 0) go = icp_accept<MODE>(S, sums, sumsf, level, iter, count_threshold);
    __syncthreads();
    if (go) {
        if (tid < 27) Av[tid] = icp_entry<MODE>(S, tid);
        __syncthreads();
        if (tid == 0) go = icp_solve<MODE>(S, Av, level, ang, tinc);
        __syncthreads();
        if (go) {
            if (MODE == kCompensated && tid < 6) {
                const double a = (double)ang[tid >> 1];
                trig[tid] = (tid & 1) ? cos(a) : sin(a);
            }
            __syncthreads();
            if (tid == 0) {
                if (MODE == kRefOrder)
                    icp_compose_ref(S, ang, tinc);
                else
                    icp_compose_comp(S, trig, tinc);
            }
        }
    }
    __syncthreads();
    uint32_t* dst = reinterpret_cast<uint32_t*>(st);
    for (int i = tid; i < kWords; i += blockDim.x) dst[i] = lds[i];
}

__global__ void icp_begin_kernel(IcpState* st, const NuiCLRigidTransform* pose) {
    clear_sync(sync_of(st), threadIdx.x, blockDim.x);
    if (threadIdx.x != 0) return;
    for (int k = 0; k < 9; ++k) st->R[k] = pose->R[k];
    for (int k = 0; k < 3; ++k) st->t[k] = pose->t[k];
    for (int l = 0; l < NUIF_ICP_MAX_LEVELS; ++l) st->brk[l] = 0;
    st->failed = 0;
    st->iterations = 0;
    st->ok = 0;
}

__global__ void icp_state_init_kernel(IcpState* st, const NuiCLRigidTransform* pose) {
    clear_sync(sync_of(st), threadIdx.x, blockDim.x);
    if (threadIdx.x != 0) return;
    for (int k = 0; k < 9; ++k) st->R[k] = st->Rg[k] = st->Rlast[k] = pose->R[k];
    for (int k = 0; k < 3; ++k) st->t[k] = st->tg[k] = st->tlast[k] = pose->t[k];
    for (int k = 0; k < 3; ++k) st->off[k] = 0.f;
    for (int l = 0; l < NUIF_ICP_MAX_LEVELS; ++l) st->brk[l] = 0;
    st->failed = 0;
    st->iterations = 0;
    st->gate = 1;
    st->ok = 1;
    st->frame = 0;
    st->lambda = 1.0f;
    st->err_old = 1e20f;
    st->error = 0.f;
    st->count = 0.f;
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

__global__ void icp_end_kernel(IcpState* st, NuiCLRigidTransform* pose, float gate_threshold, float* history,
                               int capacity) {
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
    // VO/NuiKinfuTrackingEngine.cpp:113-127 integration gate
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

// A shifting volume moved the local frame after this frame's tracking
// (OCL/NuiKinfuOpenCLShiftScene.cpp:353-364): the new local translation becomes the
// camera pose, the last integration pose (taken after the shift, as
// VO/NuiKinfuTrackingEngine.cpp:126 does) and the frame's history row; `off` is the
// frame's offset for the following integration gates.
__global__ void icp_set_translation_kernel(IcpState* st, NuiCLRigidTransform* pose, float* history, int capacity,
                                           float tx, float ty, float tz, float ox, float oy, float oz) {
    const float t[3] = {tx, ty, tz}, o[3] = {ox, oy, oz};
    for (int k = 0; k < 3; ++k) {
        st->t[k] = st->tlast[k] = pose->t[k] = t[k];
        st->off[k] = o[k];
    }
    if (history && st->frame > 0) {
        float* row = history + (size_t)((st->frame - 1) % capacity) * NUIF_POSE_ROW;
        for (int k = 0; k < 3; ++k) row[9 + k] = t[k];
    }
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
    const size_t compensated = (size_t)kMaxBlocks * kTerms * sizeof(double);
    const size_t groups = ((size_t)width * height + 127) / 128;
    const size_t reference = (groups + (groups + 127) / 128) * kTerms * sizeof(float);
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
    out->status = h.failed ? 1 : 0;
    out->iterations = h.iterations;
    out->integrate = h.gate;
    out->frame = h.frame;
    for (int k = 0; k < kTerms; ++k) out->sums[k] = h.sums[k];
    return NUIF_OK;
}

int nuif_icp_begin(void* state, const NuiCLRigidTransform* pose, nuif_stream_t s) {
    NUIF_REQUIRE_INIT();
    NUIF_CHECK_ARG(state && pose, "nuif_icp_begin: null argument");
    icp_begin_kernel<<<1, 64, 0, stream_of(s)>>>((IcpState*)state, pose);
    NUIF_LAUNCHED("icp_begin_kernel");
    return NUIF_OK;
}

int nuif_icp_iteration(void* state, float* partials, const NuiCLCameraParams* cam, const float* vmap_level,
                       int level, const float* ref_vmap, const float* ref_nmap, const NuiCLRigidTransform* prev_pose,
                       float dist_threshold, uint32_t count_threshold, int iter, int reduction, nuif_stream_t s) {
    NUIF_REQUIRE_INIT();
    NUIF_CHECK_ARG(state && partials && cam && vmap_level && ref_vmap && ref_nmap && prev_pose,
                   "nuif_icp_iteration: null argument");
    NUIF_CHECK_ARG(level >= 0 && level < NUIF_ICP_MAX_LEVELS && iter >= 0, "nuif_icp_iteration: bad level/iter");
    NUIF_CHECK_ARG(reduction == NUIF_ICP_REDUCE_REFERENCE || reduction == NUIF_ICP_REDUCE_COMPENSATED,
                   "nuif_icp_iteration: unknown reduction %d", reduction);
    const int Wl = (int)cam->depthImageWidth >> level, Hl = (int)cam->depthImageHeight >> level;
    NUIF_CHECK_ARG(Wl > 0 && Hl > 0, "nuif_icp_iteration: level %d too small", level);
    const int npx = Wl * Hl;
    hipStream_t st = stream_of(s);
    if (reduction == NUIF_ICP_REDUCE_REFERENCE) {
        const int groups = (int)div_up(npx, 128);
        NUIF_CHECK_ARG(div_up(groups, 128) <= kMaxGroups2,
                       "nuif_icp_iteration: %d pixels per level exceed the reference-order limit %d", npx,
                       kMaxGroups2 * 128 * 128);
        const int blocks = (int)div_up((size_t)groups, kRefGroups);
        icp_ref_kernel<<<blocks, 128 * kRefGroups, 0, st>>>((IcpState*)state, vmap_level, npx, *cam, ref_vmap,
                                                            ref_nmap, prev_pose, dist_threshold, level, iter,
                                                            count_threshold, partials);
        NUIF_LAUNCHED("icp_ref_kernel");
    } else {
        const int nblocks = std::min<int>(div_up(npx, kBlock), kMaxBlocks);
        icp_comp_kernel<<<nblocks, kBlock, 0, st>>>((IcpState*)state, vmap_level, npx, *cam, ref_vmap, ref_nmap,
                                                    prev_pose, dist_threshold, level, iter, count_threshold,
                                                    (double*)partials);
        NUIF_LAUNCHED("icp_comp_kernel");
    }
    return NUIF_OK;
}

int nuif_icp_end(void* state, NuiCLRigidTransform* pose, float gate_threshold, float* history, int capacity,
                 nuif_stream_t s) {
    NUIF_REQUIRE_INIT();
    NUIF_CHECK_ARG(state && pose, "nuif_icp_end: null argument");
    NUIF_CHECK_ARG(!history || capacity > 0, "nuif_icp_end: history capacity must be positive");
    icp_end_kernel<<<1, 1, 0, stream_of(s)>>>((IcpState*)state, pose, gate_threshold, history, capacity);
    NUIF_LAUNCHED("icp_end_kernel");
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
    return NUIF_OK;
}

int nuif_icp_estimate(void* state, float* partials, const NuiCLCameraParams* cam, const float* const* vmap_pyr,
                      const float* ref_vmap, const float* ref_nmap, NuiCLRigidTransform* pose,
                      const nuif_icp_config* cfg, float* history, int capacity, nuif_stream_t s) {
    NUIF_REQUIRE_INIT();
    NUIF_CHECK_ARG(state && partials && cam && vmap_pyr && ref_vmap && ref_nmap && pose && cfg,
                   "nuif_icp_estimate: null argument");
    uint32_t thr[NUIF_ICP_MAX_LEVELS];
    int rc = nuif_icp_count_thresholds((int)cam->depthImageWidth, (int)cam->depthImageHeight, cfg, thr);
    if (rc) return rc;
    const float L = (float)cfg->levels;
    const float dist = cfg->dist_threshold + cfg->dist_threshold / L;  // :300-304
    rc = nuif_icp_begin(state, pose, s);
    if (rc) return rc;
    for (int level = cfg->levels - 1; level >= 0; --level) {
        NUIF_CHECK_ARG(vmap_pyr[level], "nuif_icp_estimate: vmap level %d is NULL", level);
        for (int it = 0; it < cfg->iterations[level]; ++it) {
            rc = nuif_icp_iteration(state, partials, cam, vmap_pyr[level], level, ref_vmap, ref_nmap, pose, dist,
                                    thr[level], it, cfg->reduction, s);
            if (rc) return rc;
        }
    }
    return nuif_icp_end(state, pose, cfg->gate_threshold, history, capacity, s);
}

}  // extern "C"
