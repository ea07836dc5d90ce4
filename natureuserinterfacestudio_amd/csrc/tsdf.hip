// Dense TSDF volume: reset, integration and raycast.
//
// Volume layout is the reference's ((y*Nx + x)*Nz + z), z contiguous, so a
// z-run of one (x,y) column is a contiguous byte range.
//
// Integration.  The reference walks one (x,y) column per work-item along z,
// updating the projection incrementally (K/tsdfVolume.cl:74-189); adjacent
// work-items then touch addresses Nz*4 bytes apart, i.e. no coalescing.  Here
// one wavefront owns an 8x8 tile of columns and one z segment of them (64, 128
// or 256 voxels, chosen per launch; the walk state at each segment's start comes
// from a column-state pass), and processes the segment in chunks of 64:
//   phase 1  every lane walks its own column for 64 steps with exactly the
//            reference's incremental arithmetic (bit-identical projections) and
//            parks the new tsdf sample (or "no update") in LDS;
//   phase 2  the wave sweeps the 64 columns of the tile with 16-byte vector
//            read-modify-writes, 1 KiB contiguous per instruction, 16 loads in
//            flight per wave before the first store.
// Voxels that the reference does not update are neither read nor written
// unless they share a 16-byte vector with one that is (then they are written
// back unchanged).  Algorithmic traffic: 2 x bytes/voxel per updated voxel +
// the depth image; the time is set by phase 1 (one dependent depth gather per
// voxel along each column), far above what that traffic costs at HBM rate.
//
// Raycast.  One lane per pixel, 8x8 pixel tiles per wavefront so that
// neighbouring rays share cache lines; the fixed-step march, zero-crossing
// test, trilinear refinement and gradient normal follow K/raycast.cl:128-277
// operation by operation.

#include "batch_jobs.h"
#include "dense_volume.h"
#include "schedule.h"

#include <algorithm>

namespace nuif {

constexpr int kMaxWeight = 128;  // K/tsdfVolume.cl:4

template <class S>
__device__ __forceinline__ typename S::word update_voxel(typename S::word old, float tsdf) {
    const float prev = S::tsdf(old);
    const int w = S::weight(old);
    const float fnew = (prev * (float)w + 1.0f * tsdf) / (float)(w + 1);
    return S::pack(fnew, min(w + 1, kMaxWeight));
}

__device__ __forceinline__ uchar4 update_color(uchar4 vc, uchar4 nc, int maxcw) {
    const float Wrk = 1.f;
    const float nx = ((float)(vc.x * vc.w) + Wrk * (float)nc.x) / ((float)vc.w + Wrk);
    const float ny = ((float)(vc.y * vc.w) + Wrk * (float)nc.y) / ((float)vc.w + Wrk);
    const float nz = ((float)(vc.z * vc.w) + Wrk * (float)nc.z) / ((float)vc.w + Wrk);
    const unsigned char wn = (unsigned char)(vc.w + 1);
    uchar4 r;
    r.x = (unsigned char)(int)fminf(fmaxf(nx, 0.0f), 255.f);
    r.y = (unsigned char)(int)fminf(fmaxf(ny, 0.0f), 255.f);
    r.z = (unsigned char)(int)fminf(fmaxf(nz, 0.0f), 255.f);
    r.w = (unsigned char)min(maxcw, (int)wn);
    return r;
}

// ---- reset ------------------------------------------------------------------

template <class S>
__global__ void reset_kernel(typename S::word* __restrict__ vol, uint32_t* __restrict__ cvol, size_t n) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    const typename S::word w0 = S::init();
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        vol[i] = w0;
        if (cvol) cvol[i] = 0u;
    }
}

// K/tsdfVolume.cl:192-211 initializeVolume over a range with a voxel offset (clearSlice of
// the shifting volume, OCL/NuiKinfuOpenCLShiftScene.cpp:72-117): the voxels stored at
// ((x + w.x) % Nx, (y + w.y) % Ny, (z + w.z) % Nz) for x < rx, y < ry, z < rz.  One thread
// per voxel, z fastest, so a wavefront clears 64 consecutive stored voxels of a column.
template <class S>
__global__ void reset_slice_kernel(typename S::word* __restrict__ vol, uint32_t* __restrict__ cvol, int Nx, int Ny,
                                   int Nz, VoxelWrap w, int rx, int ry, int rz) {
    const size_t n = (size_t)rx * ry * rz;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    const typename S::word w0 = S::init();
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const int z = (int)(i % rz);
        const size_t c = i / rz;
        const int x = (int)(c % rx), y = (int)(c / rx);
        const size_t idx = wrapped_column(x, y, w, Nx, Ny, Nz) + wrap1(z, w.z, Nz);
        vol[idx] = w0;
        if (cvol) cvol[idx] = 0u;
    }
}

// ---- integration ---------------------------------------------------------------

// Column chunk of the integration: 32 voxels keep the LDS transpose buffer at 8.3 KB
// per 64-column wave (5 waves per SIMD; 64 voxels allowed 3 and ran 20 % slower).
constexpr int kZChunk = 32;
static_assert(kMinSegVoxels % kZChunk == 0, "segments hold whole chunks");
constexpr int kLdsPitch = kZChunk + 1;

struct ColumnWalk {  // the reference's per-column incremental state
    float vgz, part, vx, vy, vz, zs, incx, incy;
};

__device__ __forceinline__ ColumnWalk column_setup(int x, int y, int Nx, int Ny, const TsdfParams& p,
                                                   const NuiCLCameraParams& cam, const NuiCLRigidTransform& g) {
    ColumnWalk c;
    const float vgx = ((float)(x - Nx / 2) + 0.5f) * p.cell_size[0] - g.t[0];
    const float vgy = ((float)(y - Ny / 2) + 0.5f) * p.cell_size[1] - g.t[1];
    c.vgz = (0.5f) * p.cell_size[2] - g.t[2];
    c.part = vgx * vgx + vgy * vgy;
    c.vx = (g.R_inv[0] * vgx + g.R_inv[1] * vgy + g.R_inv[2] * c.vgz) * cam.fx;
    c.vy = (g.R_inv[3] * vgx + g.R_inv[4] * vgy + g.R_inv[5] * c.vgz) * cam.fy;
    c.vz = (g.R_inv[6] * vgx + g.R_inv[7] * vgy + g.R_inv[8] * c.vgz);
    c.zs = 0;
    c.incx = g.R_inv[2] * p.cell_size[2] * cam.fx;
    c.incy = g.R_inv[5] * p.cell_size[2] * cam.fy;
    return c;
}

// The pixel that one step of K/tsdfVolume.cl:74-120 reads for the current z, or -1
// where the reference skips the voxel before reading (behind the camera or
// outside the image).
__device__ __forceinline__ int column_pixel(const ColumnWalk& c, const NuiCLCameraParams& cam, float R8) {
    const float inv_z = 1.0f / (c.vz + R8 * c.zs);
    if (inv_z < 0.f) return -1;
    const int u = to_int(roundf(c.vx * inv_z + cam.cx));
    const int v = to_int(roundf(c.vy * inv_z + cam.cy));
    const int W = (int)cam.depthImageWidth, H = (int)cam.depthImageHeight;
    if (!(u >= 0 && v >= 0 && u < W && v < H)) return -1;
    return v * W + u;
}

// column_step's result from its pixel id, the depth read there and the distance
// of the voxel from the camera
__device__ __forceinline__ float column_finish(int id, float Dp, float dist, float trunc, float tr_inv, int* pix) {
    *pix = -1;
    if (id < 0) return kNaN;
    const float sdf = Dp - dist;
    if (isnan(Dp) || !(sdf >= -trunc)) return kNaN;
    if (fabsf(sdf) < trunc) *pix = id;
    return fminf(1.0f, sdf * tr_inv);
}

// One step of K/tsdfVolume.cl:74-120 for the current z; returns the new tsdf
// sample or NaN when the reference skips the voxel; *pix receives the pixel
// for colour integration (|sdf| < trunc) or -1.
__device__ __forceinline__ float column_step(const ColumnWalk& c, const float* __restrict__ depth,
                                             const NuiCLCameraParams& cam, float R8, float trunc, float tr_inv,
                                             int* pix) {
    const int id = column_pixel(c, cam, R8);
    const float Dp = id >= 0 ? depth[id] : 0.f;
    return column_finish(id, Dp, sqrtf(c.vgz * c.vgz + c.part), trunc, tr_inv, pix);
}

__device__ __forceinline__ void column_advance(ColumnWalk& c, float cz) {
    c.vgz += cz;
    c.zs += cz;
    c.vx += c.incx;
    c.vy += c.incy;
}

// Where the next n steps of a column walk can project.  The camera-space point
// of the column is affine in z, hence so are the projection's numerators and its
// depth; with the depth of one sign over the segment the projected coordinates
// are monotonic and the end points bound them.  The margins cover the rounding
// of the serial accumulation (relative 1e-4 of the magnitudes, far above 64 ulps)
// and of the quotients here (hardware reciprocals, 1 ulp); slack_x / slack_y bound
// how far the numerators given may lie from the walk's own (an estimate of a later
// state, see integrate_tile_body).
//   outside: no step can land inside the image (column_step returns NaN for all)
//   rect:    the pixel rectangle [x0, x1] x [y0, y1] holds every pixel a step can read
struct ChunkFootprint {
    bool outside, rect;
    int x0, x1, y0, y1;
};

__device__ __forceinline__ ChunkFootprint chunk_footprint(const ColumnWalk& c, float R8, float cz, int n,
                                                          const NuiCLCameraParams& cam, float slack_x = 0.f,
                                                          float slack_y = 0.f) {
    ChunkFootprint f{false, false, 0, 0, 0, 0};
    const float steps = (float)(n - 1);
    const float z0 = c.vz + R8 * c.zs;
    const float z1 = c.vz + R8 * (c.zs + steps * cz);
    const float ez = 1e-4f * (fabsf(c.vz) + fabsf(R8 * c.zs) + fabsf(R8 * steps * cz)) + 1e-6f;
    if (z0 < -ez && z1 < -ez) {  // behind the camera over the whole chunk
        f.outside = true;
        return f;
    }
    if (!(z0 > ez && z1 > ez)) return f;
    const float W = (float)cam.depthImageWidth, H = (float)cam.depthImageHeight;
    const float x0 = c.vx, x1 = c.vx + steps * c.incx;
    const float y0 = c.vy, y1 = c.vy + steps * c.incy;
    const float r0 = __builtin_amdgcn_rcpf(z0), r1 = __builtin_amdgcn_rcpf(z1), rmin = fmaxf(r0, r1);
    const float u0 = x0 * r0, u1 = x1 * r1, v0 = y0 * r0, v1 = y1 * r1;
    const float mu = 2.f + (1e-4f * (fabsf(x0) + fabsf(steps * c.incx)) + slack_x) * rmin;
    const float mv = 2.f + (1e-4f * (fabsf(y0) + fabsf(steps * c.incy)) + slack_y) * rmin;
    // pixel u is hit iff round(u' + cx) in [0, W-1], i.e. u' + cx in [-0.5, W - 0.5)
    const float ulo = fminf(u0, u1) + cam.cx, uhi = fmaxf(u0, u1) + cam.cx;
    const float vlo = fminf(v0, v1) + cam.cy, vhi = fmaxf(v0, v1) + cam.cy;
    if (uhi < -0.5f - mu || ulo > W - 0.5f + mu || vhi < -0.5f - mv || vlo > H - 0.5f + mv) {
        f.outside = true;
        return f;
    }
    f.rect = true;
    f.x0 = (int)fmaxf(floorf(ulo - mu), 0.f);
    f.x1 = (int)fminf(ceilf(uhi + mu), W - 1.f);
    f.y0 = (int)fmaxf(floorf(vlo - mv), 0.f);
    f.y1 = (int)fminf(ceilf(vhi + mv), H - 1.f);
    return f;
}

// ---- depth bounds: max valid depth over square pixel cells --------------------
// Level L has cells of 2^(L+3) pixels a side, stored level after level; a cell
// without a valid pixel holds -1.  Levels go up to a single cell.

constexpr int kBoundsBase = 3;

__host__ __device__ __forceinline__ int bounds_levels(int W, int H) {
    int L = 0;
    while (((W - 1) >> (L + kBoundsBase)) > 0 || ((H - 1) >> (L + kBoundsBase)) > 0) ++L;
    return L + 1;
}
__host__ __device__ __forceinline__ int bounds_dim(int n, int L) { return ((n - 1) >> (L + kBoundsBase)) + 1; }
__host__ __device__ __forceinline__ size_t bounds_offset(int W, int H, int L) {
    size_t o = 0;
    for (int k = 0; k < L; ++k) o += (size_t)bounds_dim(W, k) * bounds_dim(H, k);
    return o;
}

__device__ __forceinline__ void depth_bounds_base_body(const BoundsJob& J, int b) {
    if (J.gate && *in_global(J.gate) == 0) return;
    const int W = J.W, H = J.H;
    const float* __restrict__ depth = in_global(J.depth);
    const int cw = bounds_dim(W, 0), ch = bounds_dim(H, 0);
    const int i = b * blockDim.x + threadIdx.x;
    if (i >= cw * ch) return;
    const int cx = i % cw, cy = i / cw;
    float m = -1.f;
    for (int y = cy * 8; y < min(H, cy * 8 + 8); ++y)
        for (int x = cx * 8; x < min(W, cx * 8 + 8); ++x) {
            const float d = depth[(size_t)y * W + x];
            if (!isnan(d)) m = fmaxf(m, d);
        }
    in_global(J.out)[i] = m;
}

__device__ __forceinline__ void depth_bounds_levels_body(const BoundsJob& J) {
    if (J.gate && *in_global(J.gate) == 0) return;
    const int W = J.W, H = J.H;
    float* __restrict__ out = in_global(J.out);
    const int levels = bounds_levels(W, H);
    for (int L = 1; L < levels; ++L) {
        const float* src = out + bounds_offset(W, H, L - 1);
        float* dst = out + bounds_offset(W, H, L);
        const int sw = bounds_dim(W, L - 1), sh = bounds_dim(H, L - 1);
        const int dw = bounds_dim(W, L), dh = bounds_dim(H, L);
        for (int i = threadIdx.x; i < dw * dh; i += blockDim.x) {
            const int x = (i % dw) * 2, y = (i / dw) * 2;
            float m = src[y * sw + x];
            if (x + 1 < sw) m = fmaxf(m, src[y * sw + x + 1]);
            if (y + 1 < sh) {
                m = fmaxf(m, src[(y + 1) * sw + x]);
                if (x + 1 < sw) m = fmaxf(m, src[(y + 1) * sw + x + 1]);
            }
            dst[i] = m;
        }
        __syncthreads();
    }
}

__global__ void depth_bounds_base_kernel(BoundsJob J) { depth_bounds_base_body(J, blockIdx.x); }
__global__ void __launch_bounds__(1024) depth_bounds_levels_kernel(BoundsJob J) { depth_bounds_levels_body(J); }

__global__ void depth_bounds_base_batch_kernel(const BoundsJob* __restrict__ jobs, int n) {
    const BoundsJob& J = jobs[job_of(jobs, n, blockIdx.x)];
    depth_bounds_base_body(J, (int)(blockIdx.x - J.block0));
}
// one block per job
__global__ void __launch_bounds__(1024) depth_bounds_levels_batch_kernel(const BoundsJob* __restrict__ jobs) {
    depth_bounds_levels_body(jobs[blockIdx.x]);
}

// Largest valid depth of the pixels in [x0, x1] x [y0, y1] (within the image), or
// an upper bound of it: the max over the at most 2 x 2 cells of the first level
// whose cells cover the rectangle that way.
__device__ __forceinline__ float bounds_max(const float* __restrict__ b, int W, int H, int x0, int x1, int y0, int y1) {
    const int levels = bounds_levels(W, H);
    int L = 0;
    while (L + 1 < levels && (((x1 >> (L + kBoundsBase)) - (x0 >> (L + kBoundsBase))) > 1 ||
                              ((y1 >> (L + kBoundsBase)) - (y0 >> (L + kBoundsBase))) > 1))
        ++L;
    const float* lv = b + bounds_offset(W, H, L);
    const int cw = bounds_dim(W, L);
    const int cx0 = x0 >> (L + kBoundsBase), cx1 = x1 >> (L + kBoundsBase);
    const int cy0 = y0 >> (L + kBoundsBase), cy1 = y1 >> (L + kBoundsBase);
    float m = lv[cy0 * cw + cx0];
    m = fmaxf(m, lv[cy0 * cw + cx1]);
    m = fmaxf(m, lv[cy1 * cw + cx0]);
    m = fmaxf(m, lv[cy1 * cw + cx1]);
    return m;
}

// True when every voxel of the next n steps lies more than trunc behind the
// largest depth its pixels can hold, so column_step skips all of them (sdf =
// D - |v - c| < -trunc).  |v - c|^2 = vgz^2 + part with vgz affine in the step;
// the margin (2e-4 m + 1e-5 relative) covers the serial accumulation of vgz.
__device__ __forceinline__ bool chunk_behind_surface(const ColumnWalk& c, float cz, int n, const ChunkFootprint& f,
                                                     const float* __restrict__ bounds, int W, int H, float trunc) {
    if (!bounds || !f.rect) return false;
    const float a = c.vgz, e = c.vgz + (float)(n - 1) * cz;
    const float lo = fminf(a, e), hi = fmaxf(a, e);
    const float m2 = (lo <= 0.f && hi >= 0.f) ? 0.f : fminf(lo * lo, hi * hi);
    const float dmin = __builtin_amdgcn_sqrtf(m2 + c.part);  // 1 ulp, far inside the margin
    const float dmax = bounds_max(bounds, W, H, f.x0, f.x1, f.y0, f.y1);
    return dmax + trunc < dmin * (1.0f - 1e-5f) - 2e-4f;
}

// The walk state of every column at the start of z segments 1, 2, ...: the
// additions of K/tsdfVolume.cl:74-189 run once here, instead of every segment's
// wavefront replaying them from z = 0.  The projection numerators (vx, vy) are
// per column, float2[(k - 1) * Nx * Ny + tile * 64 + lane] in the order of the
// integration's 8x8-column tiles (one wave reads 512 contiguous bytes); the z
// terms (vgz, zs) are the same for every column and follow as
// float2[(nseg - 1) * Nx * Ny + k - 1], written by column 0.  Nx, Ny are
// multiples of 8 (the tile kernel's condition).
// After the space of the shortest segments follows the z table (ztable_offset).
__device__ __forceinline__ void column_states_body(const TsdfParams& p, const NuiCLCameraParams& cam,
                                                   const NuiCLRigidTransform* __restrict__ pose,
                                                   const int32_t* __restrict__ gate, float2* __restrict__ walk_states,
                                                   int b, int seg) {
    if (gate && *gate == 0) return;
    const int Nx = (int)p.resolution[0], Ny = (int)p.resolution[1], Nz = (int)p.resolution[2];
    const int i = b * blockDim.x + threadIdx.x;
    if (i >= Nx * Ny) return;
    const int tile = i >> 6, lane = i & 63, tx = Nx / 8;
    const int x = (tile % tx) * 8 + (lane & 7), y = (tile / tx) * 8 + (lane >> 3);
    const NuiCLRigidTransform g = *pose;
    const ColumnWalk c0 = column_setup(x, y, Nx, Ny, p, cam, g);
    const float cz = p.cell_size[2];
    const int nseg = (Nz + seg - 1) / seg;
    float vx = c0.vx, vy = c0.vy;
    for (int k = 1; k < nseg; ++k) {
#pragma unroll 16
        for (int j = 0; j < seg; ++j) {
            vx += c0.incx;
            vy += c0.incy;
        }
        walk_states[(size_t)(k - 1) * Nx * Ny + i] = make_float2(vx, vy);
    }
    if (i == 0) {
        float2* zt = walk_states + ztable_offset(Nx, Ny, Nz);
        float tz = c0.vgz, tzs = c0.zs;
        for (int z = 0; z < Nz; ++z) {
            zt[z] = make_float2(tz, tzs);
            tz += cz;
            tzs += cz;
        }
        float vgz = c0.vgz, zs = c0.zs;
        for (int k = 1; k < nseg; ++k) {
#pragma unroll 16
            for (int j = 0; j < seg; ++j) {
                vgz += cz;
                zs += cz;
            }
            walk_states[(size_t)(nseg - 1) * Nx * Ny + k - 1] = make_float2(vgz, zs);
        }
    }
}

__global__ void __launch_bounds__(256)
    column_states_kernel(TsdfParams p, NuiCLCameraParams cam, const NuiCLRigidTransform* __restrict__ pose,
                         const int32_t* __restrict__ gate, float2* __restrict__ walk_states, int seg) {
    column_states_body(p, cam, pose, gate, walk_states, blockIdx.x, seg);
}

// Fast path: Nx, Ny multiples of 8, Nz a multiple of the column chunk.
template <class S, bool COLOR>
__device__ __forceinline__ void integrate_tile_body(const float* __restrict__ depth, const uchar4* __restrict__ bgra,
                                                    typename S::word* __restrict__ vol, uchar4* __restrict__ cvol,
                                                    const TsdfParams& p, const NuiCLCameraParams& cam,
                                                    const NuiCLRigidTransform* __restrict__ pose, int maxcw,
                                                    const int32_t* __restrict__ gate,
                                                    unsigned long long* __restrict__ updated,
                                                    const float* __restrict__ bounds, const SurfaceBlocks& surf,
                                                    VoxelWrap wr, const float2* __restrict__ walk_states, int bx,
                                                    int by, int bz, int gdx, int gdz, int seg) {
    if (gate && *gate == 0) return;
    using word = typename S::word;
    constexpr int VPL = 16 / S::kBytes;      // voxels per lane per 16-byte vector
    constexpr int LPC = kZChunk / VPL;        // lanes covering one column chunk
    constexpr int CPI = 64 / LPC;             // columns per vector instruction
    constexpr int NIT = 64 / CPI;             // instructions per chunk
    constexpr int BATCH = kRmwBatch;  // column chunks of phase 2 whose loads are in flight together
    constexpr int KB = kWalkBatch;    // walk steps whose depth reads are in flight together
    static_assert(kZChunk % KB == 0, "walk batches tile the chunk");
    static_assert(NIT % BATCH == 0 && kZChunk % VPL == 0, "phase 2 covers the 64 columns exactly");
    __shared__ float ts[64 * kLdsPitch];
    __shared__ int cp[COLOR ? 64 * kLdsPitch : 1];

    const int Nx = (int)p.resolution[0], Ny = (int)p.resolution[1], Nz = (int)p.resolution[2];
    const int lane = threadIdx.x;
    const int x = bx * 8 + (lane & 7), y = by * 8 + (lane >> 3);
    const NuiCLRigidTransform g = *pose;
    ColumnWalk c = column_setup(x, y, Nx, Ny, p, cam, g);
    const float trunc = p.tranc_dist, tr_inv = 1.f / p.tranc_dist, cz = p.cell_size[2], R8 = g.R_inv[8];

    const int sub = lane % LPC;  // position inside a column chunk (phase 2)
    unsigned nupd = 0;
    // z segment of this wave (bz): the walk state at its start is the one the
    // same sequence of additions reaches, read from the per-column states of
    // column_states_kernel or replayed here
    const int zbeg = bz * seg, zend = min(Nz, zbeg + seg);
    if (zbeg > 0 && walk_states) {
        const float2 w = walk_states[(size_t)(bz - 1) * Nx * Ny + ((size_t)by * gdx + bx) * 64 + lane];
        const float2 z = walk_states[(size_t)(gdz - 1) * Nx * Ny + bz - 1];
        c.vx = w.x;
        c.vy = w.y;
        c.vgz = z.x;
        c.zs = z.y;
    } else {
#pragma unroll 1
        for (int k = 0; k < zbeg; ++k) column_advance(c, cz);
    }
    // With the z table (walk states given) a chunk that no voxel of the tile can update is
    // tested on an estimate of the columns' numerators at its start, x + k incx (the slack
    // below bounds its distance from the serial sums, (k + 1) 2^-23 of the magnitudes), and
    // skipped without stepping the columns through it: the serial additions run only up to
    // a chunk that is walked (zx: the z where c holds the walk's own state).
    const float2* __restrict__ ztab = walk_states ? walk_states + ztable_offset(Nx, Ny, Nz) : nullptr;
    int zx = zbeg;
    // The estimated state at z (exact when z == zx) and its slack.
    auto estimate = [&](int z, ColumnWalk& a, float& sx, float& sy) {
        a = c;
        sx = sy = 0.f;
        if (ztab && z > zx) {
            const float k = (float)(z - zx);
            a.vx = c.vx + k * c.incx;
            a.vy = c.vy + k * c.incy;
            sx = (k + 1.f) * 1.1920929e-7f * (fabsf(c.vx) + k * fabsf(c.incx));
            sy = (k + 1.f) * 1.1920929e-7f * (fabsf(c.vy) + k * fabsf(c.incy));
            const float2 zt = ztab[z];
            a.vgz = zt.x;
            a.zs = zt.y;
        }
    };
    for (int zc = zbeg; zc < zend; zc += kZChunk) {
        // groups of kCullGroup chunks are tested first, and skipped whole when no voxel of
        // the tile can update (the same test over the longer run)
        if (ztab && (zc - zbeg) % (kCullGroup * kZChunk) == 0 && zend - zc > kZChunk) {
            const int n = min(kCullGroup * kZChunk, zend - zc);
            ColumnWalk a;
            float sx, sy;
            estimate(zc, a, sx, sy);
            const ChunkFootprint fg = chunk_footprint(a, R8, cz, n, cam, sx, sy);
            if (__all(fg.outside || chunk_behind_surface(a, cz, n, fg, bounds, (int)cam.depthImageWidth,
                                                         (int)cam.depthImageHeight, trunc))) {
                zc += n - kZChunk;
                continue;
            }
        }
        // phase 1: serial walk of this lane's column; chunks that cannot project
        // into the image only advance the walk
        ColumnWalk a;
        float sx, sy;
        estimate(zc, a, sx, sy);
        const ChunkFootprint fp = chunk_footprint(a, R8, cz, kZChunk, cam, sx, sy);
        const bool outside = fp.outside || chunk_behind_surface(a, cz, kZChunk, fp, bounds,
                                                                (int)cam.depthImageWidth,
                                                                (int)cam.depthImageHeight, trunc);
        if (__all(outside)) {
            if (!ztab)
                for (int k = 0; k < kZChunk; ++k) column_advance(c, cz);
            continue;
        }
        if (ztab && zc > zx) {  // the walk's own state at zc: the numerators' serial sums
            for (; zx < zc; ++zx) {
                c.vx += c.incx;
                c.vy += c.incy;
            }
            const float2 zt = ztab[zc];
            c.vgz = zt.x;
            c.zs = zt.y;
        }
        zx = zc + kZChunk;  // the walk below steps c through this chunk
        bool any = false;
        if (outside) {
            for (int k = 0; k < kZChunk; ++k) {
                ts[lane * kLdsPitch + k] = kNaN;
                if (COLOR) cp[lane * kLdsPitch + k] = -1;
                column_advance(c, cz);
            }
        } else {
            // KB steps at a time: their pixels first, then their KB depth reads in
            // flight together, then the samples (the same float operations as
            // column_step, in the same order per voxel)
            for (int k0 = 0; k0 < kZChunk; k0 += KB) {
                int id[KB];
                float dist[KB], Dp[KB];
#pragma unroll
                for (int j = 0; j < KB; ++j) {
                    id[j] = column_pixel(c, cam, R8);
                    dist[j] = sqrtf(c.vgz * c.vgz + c.part);
                    column_advance(c, cz);
                }
#pragma unroll
                for (int j = 0; j < KB; ++j) Dp[j] = id[j] >= 0 ? depth[id[j]] : 0.f;
#pragma unroll
                for (int j = 0; j < KB; ++j) {
                    int pix;
                    const float v = column_finish(id[j], Dp[j], dist[j], trunc, tr_inv, &pix);
                    ts[lane * kLdsPitch + k0 + j] = v;
                    if (COLOR) cp[lane * kLdsPitch + k0 + j] = pix;
                    any |= !isnan(v);
                    nupd += !isnan(v);
                }
            }
        }
        if (!__any(any)) continue;  // the whole tile misses the frustum / band in this chunk
        __syncthreads();
        // phase 2: coalesced read-modify-write, CPI columns per instruction
        for (int it0 = 0; it0 < NIT; it0 += BATCH) {
            word buf[BATCH][VPL];
            bool act[BATCH];
#pragma unroll
            for (int j = 0; j < BATCH; ++j) {
                const int col = (it0 + j) * CPI + lane / LPC;
                const float* tv = &ts[col * kLdsPitch + sub * VPL];
                bool a = false;
#pragma unroll
                for (int q = 0; q < VPL; ++q) a |= !isnan(tv[q]);
                act[j] = a;
                if (a) {
                    const int gx = bx * 8 + (col & 7), gy = by * 8 + (col >> 3);
                    const size_t base = wrapped_column(gx, gy, wr, Nx, Ny, Nz) + wrap1(zc + sub * VPL, wr.z, Nz);
                    const uint4 raw = *reinterpret_cast<const uint4*>(vol + base);
                    const word* wv = reinterpret_cast<const word*>(&raw);
#pragma unroll
                    for (int q = 0; q < VPL; ++q) buf[j][q] = wv[q];
                }
            }
#pragma unroll
            for (int j = 0; j < BATCH; ++j) {
                if (!act[j]) continue;
                const int col = (it0 + j) * CPI + lane / LPC;
                const float* tv = &ts[col * kLdsPitch + sub * VPL];
                const int gx = bx * 8 + (col & 7), gy = by * 8 + (col >> 3);
                const int pz = wrap1(zc + sub * VPL, wr.z, Nz);  // stored z of the vector (w.z % VPL == 0)
                const size_t base = wrapped_column(gx, gy, wr, Nx, Ny, Nz) + pz;
                uint4 outv;
                word* ov = reinterpret_cast<word*>(&outv);
                bool neg = false;
#pragma unroll
                for (int q = 0; q < VPL; ++q) {
                    ov[q] = isnan(tv[q]) ? buf[j][q] : update_voxel<S>(buf[j][q], tv[q]);
                    neg |= !isnan(tv[q]) && S::nonpositive(ov[q]);
                }
                *reinterpret_cast<uint4*>(vol + base) = outv;
                // the VPL voxels share one 8-voxel z block (VPL divides 8); the flags
                // describe stored positions
                if (neg && surf.b) surf.mark(wrap1(gx, wr.x, Nx), wrap1(gy, wr.y, Ny), pz);
                if (COLOR) {
                    const int* pv = &cp[col * kLdsPitch + sub * VPL];
#pragma unroll
                    for (int q = 0; q < VPL; ++q) {
                        const int pix = pv[q];
                        if (isnan(tv[q]) || pix < 0) continue;
                        const uchar4 nc = bgra[pix];
                        if (nc.w == 0) continue;
                        cvol[base + q] = update_color(cvol[base + q], nc, maxcw);
                    }
                }
            }
        }
        __syncthreads();
    }
    if (updated) {
        for (int off = 32; off > 0; off >>= 1) nupd += __shfl_xor(nupd, off);
        if (lane == 0 && nupd) atomicAdd(updated, (unsigned long long)nupd);
    }
}

template <class S, bool COLOR>
__global__ void __launch_bounds__(64)
    integrate_tile_kernel(const float* __restrict__ depth, const uchar4* __restrict__ bgra,
                          typename S::word* __restrict__ vol, uchar4* __restrict__ cvol, TsdfParams p,
                          NuiCLCameraParams cam, const NuiCLRigidTransform* __restrict__ pose, int maxcw,
                          const int32_t* __restrict__ gate, unsigned long long* __restrict__ updated,
                          const float* __restrict__ bounds, SurfaceBlocks surf, VoxelWrap wr,
                          const float2* __restrict__ walk_states, int seg) {
    integrate_tile_body<S, COLOR>(depth, bgra, vol, cvol, p, cam, pose, maxcw, gate, updated, bounds, surf, wr,
                                  walk_states, blockIdx.x, blockIdx.y, blockIdx.z, gridDim.x, gridDim.z, seg);
}

// all sessions' tiles of one storage: block -> (session, tile x, tile y, z segment)
template <class S, bool COLOR>
__global__ void __launch_bounds__(64) integrate_tile_batch_kernel(const IntJob* __restrict__ jobs, int n,
                                                                   unsigned blocks, int remap) {
    const unsigned lb = remap ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
    if (lb >= blocks) return;
    const IntJob& J = jobs[job_of(jobs, n, lb)];
    unsigned b = lb - J.block0;
    const int bx = (int)(b % J.gx);
    b /= J.gx;
    const int by = (int)(b % J.gy), bz = (int)(b / J.gy);
    // generic pointers here: with device-memory (global_*) accesses the FP16 variant needs
    // 134 VGPRs instead of 120 (three waves per SIMD instead of four) and runs slower.  The
    // volume and camera parameters are copied once: read through the job table they would
    // be reloaded after every store the walk makes (the stores might alias them)
    const TsdfParams p = J.p;
    const NuiCLCameraParams cam = J.cam;
    integrate_tile_body<S, COLOR>(J.depth, J.bgra, reinterpret_cast<typename S::word*>(J.vol), J.cvol, p, cam, J.pose,
                                  J.maxcw, J.gate, J.updated, J.bounds, J.surf, VoxelWrap{0, 0, 0}, J.walk, bx, by,
                                  bz, (int)J.gx, (int)J.gz, J.seg);
}

// column walk states of all sessions: block -> session by cblock0
__global__ void __launch_bounds__(256) column_states_batch_kernel(const IntJob* __restrict__ jobs, int n) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].cblock0 <= blockIdx.x)
            lo = mid;
        else
            hi = mid - 1;
    }
    const IntJob& J = jobs[lo];
    if (!J.walk) return;
    column_states_body(J.p, J.cam, in_global(J.pose), in_global(J.gate), in_global(J.walk),
                       (int)(blockIdx.x - J.cblock0), J.seg);
}

// Generic path (any resolution): the reference's one-column-per-thread walk.
template <class S, bool COLOR>
__global__ void integrate_column_kernel(const float* __restrict__ depth, const uchar4* __restrict__ bgra,
                                        typename S::word* __restrict__ vol, uchar4* __restrict__ cvol, TsdfParams p,
                                        NuiCLCameraParams cam, const NuiCLRigidTransform* __restrict__ pose,
                                        int maxcw, const int32_t* __restrict__ gate,
                                        unsigned long long* __restrict__ updated, SurfaceBlocks surf,
                                        VoxelWrap wr) {
    if (gate && *gate == 0) return;
    const int Nx = (int)p.resolution[0], Ny = (int)p.resolution[1], Nz = (int)p.resolution[2];
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= Nx || y >= Ny) return;
    const NuiCLRigidTransform g = *pose;
    ColumnWalk c = column_setup(x, y, Nx, Ny, p, cam, g);
    const float trunc = p.tranc_dist, tr_inv = 1.f / p.tranc_dist, cz = p.cell_size[2], R8 = g.R_inv[8];
    const size_t base = wrapped_column(x, y, wr, Nx, Ny, Nz);
    const int px = wrap1(x, wr.x, Nx), py = wrap1(y, wr.y, Ny);
    unsigned nupd = 0;
    for (int z = 0; z < Nz; ++z, column_advance(c, cz)) {
        int pix;
        const float v = column_step(c, depth, cam, R8, trunc, tr_inv, &pix);
        if (isnan(v)) continue;
        ++nupd;
        const int pz = wrap1(z, wr.z, Nz);
        const typename S::word nw = update_voxel<S>(vol[base + pz], v);
        vol[base + pz] = nw;
        if (surf.b && S::nonpositive(nw)) surf.mark(px, py, pz);
        if (COLOR && pix >= 0) {
            const uchar4 nc = bgra[pix];
            if (nc.w != 0) cvol[base + pz] = update_color(cvol[base + pz], nc, maxcw);
        }
    }
    if (updated && nupd) atomicAdd(updated, (unsigned long long)nupd);
}

// ---- raycast -------------------------------------------------------------------

// COUNT: every voxel read is also marked in the bitset `mark` (the footprint
// measurement of nuif_tsdf_raycast_footprint; the raycast kernels do not count)
// I32: the volume has fewer than 2^31 voxels and linear indices use 32-bit arithmetic
template <class S, bool WRAP, bool COUNT = false, bool I32 = false> struct VolView {
    const typename S::word* __restrict__ v;
    int Nx, Ny, Nz;
    VoxelWrap w;
    uint32_t* mark = nullptr;
    __device__ __forceinline__ float sample(long long idx) const {
        if (COUNT) atomicOr(&mark[idx >> 5], 1u << (idx & 31));
        const typename S::word w = v[idx];
        return S::weight(w) > 0 ? S::tsdf(w) : kNaN;
    }
    // stored index of the voxel at logical g (every component inside the volume)
    __device__ __forceinline__ long long index(int gx, int gy, int gz) const {
        if (WRAP) return (long long)wrapped_column(gx, gy, w, Nx, Ny, Nz) + wrap1(gz, w.z, Nz);
        if (I32) return (long long)(unsigned)(((unsigned)gy * (unsigned)Nx + (unsigned)gx) * (unsigned)Nz + (unsigned)gz);
        return ((long long)gy * Nx + gx) * Nz + gz;
    }
};

__device__ __forceinline__ int vol_ny(const TsdfParams& p) { return (int)p.resolution[1]; }

__device__ __forceinline__ void get_voxel(f3 pt, const TsdfParams& p, int g[3]) {
    g[0] = to_int(roundf(pt.x / p.cell_size[0]));
    g[1] = to_int(roundf(pt.y / p.cell_size[1]));
    g[2] = to_int(roundf(pt.z / p.cell_size[2]));
}

// to_int(roundf(x / c)) without the division where its rounding is decided anyway:
// a = x * rc (rc = 1 / c rounded) is within 2^-22 |x / c| of the rounded quotient, so
// when a lies farther than 2^-20 max(|a|, 1) from every half-integer both round to the
// same integer; otherwise (or for NaN / huge values) the quotient is computed.  Away
// from half-integers round-half-to-even (one instruction) and round-half-away agree.
__device__ __forceinline__ int round_div(float x, float c, float rc) {
    const float a = x * rc;
    const float r = rintf(a);
    const float d = fabsf(fabsf(a - r) - 0.5f);
    if (d > 9.5367431640625e-7f * fmaxf(fabsf(a), 1.0f)) return to_int(r);
    return to_int(roundf(x / c));
}
// round_div's rounded quotient as a float (an integer value, or NaN / huge beyond the
// int range where to_int would give INT_MIN): range tests on it decide what to_int's
// result would, without the conversion's saturation
__device__ __forceinline__ float round_div_f(float x, float c, float rc) {
    const float a = x * rc;
    const float r = rintf(a);
    const float d = fabsf(fabsf(a - r) - 0.5f);
    if (d > 9.5367431640625e-7f * fmaxf(fabsf(a), 1.0f)) return r;
    return roundf(x / c);
}
struct CellInv {
    float x, y, z;
};
__device__ __forceinline__ void get_voxel_fast(f3 pt, const TsdfParams& p, const CellInv& ri, int g[3]) {
    g[0] = round_div(pt.x, p.cell_size[0], ri.x);
    g[1] = round_div(pt.y, p.cell_size[1], ri.y);
    g[2] = round_div(pt.z, p.cell_size[2], ri.z);
}

// K/raycast.cl:36-120 interpolateTrilineary, corner weights exactly as written
// (the x and y fractions land on swapped corners relative to the y-major layout).
// In a shifting volume every corner is wrapped on its own; the reference wraps only
// the first corner and steps to the others by the unwrapped strides, which reads the
// wrong voxels (or past the end of the volume) where a cell straddles the wrap.
template <class S, bool WRAP, bool COUNT, bool I32>
__device__ float trilinear(f3 pt, const VolView<S, WRAP, COUNT, I32>& vv, const TsdfParams& p, const CellInv& ri) {
    int g[3];
    get_voxel_fast(pt, p, ri, g);
    if (g[0] <= 0 || g[0] >= vv.Nx - 1) return kNaN;
    if (g[1] <= 0 || g[1] >= vv.Ny - 1) return kNaN;
    if (g[2] <= 0 || g[2] >= vv.Nz - 1) return kNaN;
    const float vx = ((float)g[0] + 0.5f) * p.cell_size[0];
    const float vy = ((float)g[1] + 0.5f) * p.cell_size[1];
    const float vz = ((float)g[2] + 0.5f) * p.cell_size[2];
    g[0] = pt.x < vx ? g[0] - 1 : g[0];
    g[1] = pt.y < vy ? g[1] - 1 : g[1];
    g[2] = pt.z < vz ? g[2] - 1 : g[2];
    const float a = (pt.x - ((float)g[0] + 0.5f) * p.cell_size[0]) / p.cell_size[0];
    const float b = (pt.y - ((float)g[1] + 0.5f) * p.cell_size[1]) / p.cell_size[1];
    const float c = (pt.z - ((float)g[2] + 0.5f) * p.cell_size[2]) / p.cell_size[2];
    // the eight corners (all inside the volume) are read together; any unobserved
    // corner makes the result NaN, as the reference's early returns do
    float f000, f001, f011, f010, f110, f111, f101, f100;
    if (WRAP) {
        f000 = vv.sample(vv.index(g[0], g[1], g[2]));
        f001 = vv.sample(vv.index(g[0], g[1], g[2] + 1));
        f011 = vv.sample(vv.index(g[0] + 1, g[1], g[2] + 1));
        f010 = vv.sample(vv.index(g[0] + 1, g[1], g[2]));
        f110 = vv.sample(vv.index(g[0] + 1, g[1] + 1, g[2]));
        f111 = vv.sample(vv.index(g[0] + 1, g[1] + 1, g[2] + 1));
        f101 = vv.sample(vv.index(g[0], g[1] + 1, g[2] + 1));
        f100 = vv.sample(vv.index(g[0], g[1] + 1, g[2]));
    } else {
        const long long sx = vv.Nz, sy = (long long)vv.Nx * vv.Nz;
        const long long idx = vv.index(g[0], g[1], g[2]);
        f000 = vv.sample(idx);
        f001 = vv.sample(idx + 1);
        f011 = vv.sample(idx + sx + 1);
        f010 = vv.sample(idx + sx);
        f110 = vv.sample(idx + sx + sy);
        f111 = vv.sample(idx + sx + sy + 1);
        f101 = vv.sample(idx + sy + 1);
        f100 = vv.sample(idx + sy);
    }
    if (isnan(f000) || isnan(f001) || isnan(f011) || isnan(f010) || isnan(f110) || isnan(f111) || isnan(f101) ||
        isnan(f100))
        return kNaN;
    return f000 * (1 - a) * (1 - b) * (1 - c) + f001 * (1 - a) * (1 - b) * c + f010 * (1 - a) * b * (1 - c) +
           f011 * (1 - a) * b * c + f100 * a * (1 - b) * (1 - c) + f101 * a * (1 - b) * c +
           f110 * a * b * (1 - c) + f111 * a * b * c;
}

__device__ __forceinline__ float min_time(const float* vmax, f3 o, f3 d) {
    const float tx = fabs((double)d.x) < 1e-14 ? 0.f : ((d.x > 0 ? 0.f : vmax[0]) - o.x) / d.x;
    const float ty = fabs((double)d.y) < 1e-14 ? 0.f : ((d.y > 0 ? 0.f : vmax[1]) - o.y) / d.y;
    const float tz = fabs((double)d.z) < 1e-14 ? 0.f : ((d.z > 0 ? 0.f : vmax[2]) - o.z) / d.z;
    return fmaxf(fmaxf(tx, ty), tz);
}
__device__ __forceinline__ float max_time(const float* vmax, f3 o, f3 d) {
    const float tx = fabs((double)d.x) < 1e-14 ? 1e14f : ((d.x > 0 ? vmax[0] : 0.f) - o.x) / d.x;
    const float ty = fabs((double)d.y) < 1e-14 ? 1e14f : ((d.y > 0 ? vmax[1] : 0.f) - o.y) / d.y;
    const float tz = fabs((double)d.z) < 1e-14 ? 1e14f : ((d.z > 0 ? vmax[2] : 0.f) - o.z) / d.z;
    return fminf(fminf(tx, ty), tz);
}

// Ray parameters u with rs + dir * u inside the superblock of voxel g shrunk by
// kSpanMargin voxels on every side (voxel i spans [(i - 1/2) c, (i + 1/2) c)):
// [*u_in, *u_out].  The margin (0.05 voxel, 0.3 mm at 5.9 mm voxels) is some 60
// times the rounding of the sample positions (< 1e-6 m inside a 3 m volume), of
// these bounds (reciprocal directions, < 5e-6 m) and of the voxel rounding, so every
// sample with u in the span lies in the superblock.
constexpr float kSpanMargin = 0.05f;
__device__ __forceinline__ void superblock_span(const int g[3], const int N[3], const float c[3], f3 rs, f3 inv_d,
                                                float* u_in, float* u_out) {
    const float r[3] = {rs.x, rs.y, rs.z}, id[3] = {inv_d.x, inv_d.y, inv_d.z};
    float lo_t = -3.402823466e38f, hi_t = 3.402823466e38f;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int x0 = g[a] & ~(kSurfSuper - 1), x1 = min(x0 + kSurfSuper - 1, N[a] - 1);
        const float lo = ((float)x0 - 0.5f + kSpanMargin) * c[a], hi = ((float)x1 + 0.5f - kSpanMargin) * c[a];
        if (id[a] > 0.f) {
            lo_t = fmaxf(lo_t, (lo - r[a]) * id[a]);
            hi_t = fminf(hi_t, (hi - r[a]) * id[a]);
        } else if (id[a] < 0.f) {
            lo_t = fmaxf(lo_t, (hi - r[a]) * id[a]);
            hi_t = fminf(hi_t, (lo - r[a]) * id[a]);
        } else if (!(r[a] >= lo && r[a] <= hi)) {
            hi_t = -3.402823466e38f;
        }
    }
    *u_in = lo_t;
    *u_out = hi_t;
}

// The time after the march's steps t <- fl(t + step) that stay inside [t, s_out] and
// below max_t, in one operation.  With t in the binade [2^e, E), E = 2^(e+1), of ulp
// u = 2^(e-23), every sum t + step that stays at or below E - u rounds to a multiple of u
// with the same offset: fl(t + step) = t + d, d = fl(t + step) - t (exact), unless step
// lies exactly half way between two multiples of u (then the rounding to even alternates
// and no jump is made).  So m steps give exactly t + m d while t + m d <= E - 2u; m is
// taken one short of the float estimate and the result checked, so a rounding of the
// estimate only ever shortens the jump.
__device__ __forceinline__ float march_jump(float t, float step, float s_out, float max_t) {
    const unsigned ex = __float_as_uint(t) & 0x7f800000u;
    if (!(t > 0.f) || ex < (24u << 23) || ex >= (254u << 23)) return t;
    const float d = (t + step) - t;
    const float u = __uint_as_float(ex - (23u << 23));
    const float top = __uint_as_float(ex + (1u << 23)) - 2.f * u;
    if (fabsf(step - d) == 0.5f * u) return t;
    const float lim = fminf(fminf(s_out, top), max_t);
    const float m = floorf((lim - t) * __builtin_amdgcn_rcpf(d)) - 1.f;
    if (!(m >= 1.f)) return t;
    const float tm = t + m * d;
    return (tm <= s_out && tm <= top && tm < max_t) ? tm : t;
}

// superblock flags kept in LDS when the volume has at most this many superblocks
constexpr int kSuperLds = 4096;

template <class S, bool WRAP, bool COUNT = false, bool I32 = false>
__device__ __forceinline__ void raycast_body(const typename S::word* __restrict__ vol, const uchar4* __restrict__ cvol,
                                             const TsdfParams& p, const NuiCLCameraParams& cam,
                                             const NuiCLRigidTransform* __restrict__ pose, float* __restrict__ vmap,
                                             float* __restrict__ nmap, uchar4* __restrict__ cmap,
                                             const int32_t* __restrict__ enable, const SurfaceBlocks& surf,
                                             VoxelWrap wr, int bx, int by, uint32_t* mark = nullptr) {
    if (enable && *enable == 0) return;
    // the superblock flags, read by every lane at every superblock it enters
    __shared__ __align__(16) uint8_t sflags[kSuperLds];
    const int nsup = surf.b ? (int)(surf.nsx * surf.nsz * super_dim(vol_ny(p))) : 0;
    const bool sup_lds = nsup > 0 && nsup <= kSuperLds;
    if (sup_lds) {
        for (int i = threadIdx.x * 16; i < nsup; i += 64 * 16) {
            if (i + 16 <= nsup && ((surf.nb + i) & 15) == 0) {
                *reinterpret_cast<uint4*>(&sflags[i]) = *reinterpret_cast<const uint4*>(surf.b + surf.nb + i);
            } else {
                for (int j = i; j < min(i + 16, nsup); ++j) sflags[j] = surf.b[surf.nb + j];
            }
        }
        __syncthreads();
    }
    const int W = (int)cam.depthImageWidth, H = (int)cam.depthImageHeight;
    const int x = bx * 8 + (threadIdx.x & 7), y = by * 8 + (threadIdx.x >> 3);
    if (x >= W || y >= H) return;
    const size_t id = (size_t)y * W + x;
    const f3 NaN3 = mk3(kNaN, kNaN, kNaN);
    VolView<S, WRAP, COUNT, I32> vv{vol, (int)p.resolution[0], (int)p.resolution[1], (int)p.resolution[2], wr, mark};
    const NuiCLRigidTransform g = *pose;
    const f3 off = mk3(p.dimension[0] / 2.0f, p.dimension[1] / 2.0f, 0.f);
    const f3 rs = add3(mk3(g.t[0], g.t[1], g.t[2]), off);
    f3 rn = mk3(((float)x - cam.cx) * cam.fx_inv, ((float)y - cam.cy) * cam.fy_inv, 1.f);
    rn = add3(add3(rot3(g.R, rn), mk3(g.t[0], g.t[1], g.t[2])), off);
    const f3 dir = normalize3(sub3(rn, rs));
    const CellInv ri{1.0f / p.cell_size[0], 1.0f / p.cell_size[1], 1.0f / p.cell_size[2]};
    const f3 Nm1 = mk3(p.resolution[0] - 1.f, p.resolution[1] - 1.f, p.resolution[2] - 1.f);
    // reciprocal directions for the superblock spans; a component below 1e-30 counts as
    // zero (it moves the ray by less than 1e-28 m over the whole march)
    const f3 inv_dir = mk3(fabsf(dir.x) >= 1e-30f ? 1.0f / dir.x : 0.f, fabsf(dir.y) >= 1e-30f ? 1.0f / dir.y : 0.f,
                           fabsf(dir.z) >= 1e-30f ? 1.0f / dir.z : 0.f);
    float tcur = min_time(p.dimension, rs, dir);
    const float texit = max_time(p.dimension, rs, dir);
    tcur = fmaxf(tcur, 0.f);
    f3 vout = NaN3, nout = NaN3;
    uchar4 cout = make_uchar4(0, 0, 0, 0);
    if (tcur < texit) {
        int gv[3];
        get_voxel(add3(rs, mul3(dir, tcur)), p, gv);
        gv[0] = max(0, min(gv[0], vv.Nx - 1));
        gv[1] = max(0, min(gv[1], vv.Ny - 1));
        gv[2] = max(0, min(gv[2], vv.Nz - 1));
        long long idx = vv.index(gv[0], gv[1], gv[2]);
        float tsdf = vv.sample(idx);
        const float max_t = 3.f * (p.dimension[0] + p.dimension[1] + p.dimension[2]);
        const float step = p.tranc_dist * 0.8f;
        // The march of K/raycast.cl:193-276, staged kRayBatch samples at a time: the
        // sample positions of the next steps do not depend on the samples, so their
        // loads are issued together and then consumed in order with the reference's
        // exit tests (loads past the exit are in bounds and simply unused).
        // Surface blocks: a sample in a block without tsdf <= 0 (value > 0 or NaN)
        // whose neighbouring samples are too can trigger neither exit test, at its
        // own step or at the next one, whatever its exact value; such a sample is
        // not read and stands in as +1.
        bool marching = true, crossing = false, prev_safe = false;
        const int Nv[3] = {vv.Nx, vv.Ny, vv.Nz};
        while (marching) {
            // Inside clear superblocks no sample can trigger an exit test when the
            // previous sample is not negative: only the times advance (the same float
            // additions), and the last crossed sample is read for the next test.  A
            // sample inside the current superblock's one-voxel-shrunk box is known to be
            // in it; any other sample is located exactly and continues the crossing
            // only if its superblock is clear too.
            if (surf.b && !(tsdf < 0.f)) {
                float tn = tcur, s_in = 1.f, s_out = 0.f;  // empty span
                unsigned cur = ~0u;                       // superblock of the span
                while (tn < max_t) {
                    const float un = tn + step;
                    if (!(un >= s_in && un <= s_out)) {
                        int gq[3];
                        get_voxel_fast(add3(rs, mul3(dir, un)), p, ri, gq);
                        if (!(gq[0] >= 0 && gq[1] >= 0 && gq[2] >= 0 && gq[0] < vv.Nx && gq[1] < vv.Ny &&
                              gq[2] < vv.Nz))
                            break;
                        const unsigned si = surf.sindex32(gq[0], gq[1], gq[2]);
                        if (si != cur) {
                            if ((sup_lds ? sflags[si - surf.nb] : surf.b[si]) != 0) break;
                            superblock_span(gq, Nv, p.cell_size, rs, inv_dir, &s_in, &s_out);
                            cur = si;
                        }
                    }
                    tn = un;
                    // further steps inside the clear span only advance the time
                    if (tn >= s_in && tn <= s_out) tn = march_jump(tn, step, s_out, max_t);
                }
                if (tn != tcur) {
                    int gs[3];
                    get_voxel_fast(add3(rs, mul3(dir, tn)), p, ri, gs);
                    tsdf = vv.sample(vv.index(gs[0], gs[1], gs[2]));
                    tcur = tn;
                    prev_safe = true;
                }
            }
            float tb[kRayBatch], sv[kRayBatch];
            long long ib[kRayBatch];
            bool inb[kRayBatch], safe[kRayBatch + 1];
            float tn = tcur;
#pragma unroll
            for (int k = 0; k < kRayBatch; ++k) {
                tb[k] = tn;
                const f3 q = add3(rs, mul3(dir, tn + step));
                const float rx = round_div_f(q.x, p.cell_size[0], ri.x), ry = round_div_f(q.y, p.cell_size[1], ri.y),
                            rz = round_div_f(q.z, p.cell_size[2], ri.z);
                inb[k] = rx >= 0.f && ry >= 0.f && rz >= 0.f && rx <= Nm1.x && ry <= Nm1.y && rz <= Nm1.z;
                gv[0] = (int)rx;
                gv[1] = (int)ry;
                gv[2] = (int)rz;
                ib[k] = inb[k] ? vv.index(gv[0], gv[1], gv[2]) : 0;
                safe[k] = surf.b && inb[k] && surf.b[surf.index32(gv[0], gv[1], gv[2])] == 0;
                tn += step;  // the reference's `time_curr += time_step`
            }
            safe[kRayBatch] = false;  // the first sample of the next batch
            if (surf.b) {
                get_voxel_fast(add3(rs, mul3(dir, tn + step)), p, ri, gv);
                safe[kRayBatch] = gv[0] >= 0 && gv[1] >= 0 && gv[2] >= 0 && gv[0] < vv.Nx && gv[1] < vv.Ny &&
                                  gv[2] < vv.Nz && surf.b[surf.index32(gv[0], gv[1], gv[2])] == 0;
            }
#pragma unroll
            for (int k = 0; k < kRayBatch; ++k) {
                const bool need = !safe[k] || !(k ? safe[k - 1] : prev_safe) || !safe[k + 1];
                sv[k] = inb[k] ? (need ? vv.sample(ib[k]) : 1.0f) : kNaN;
            }
            for (int k = 0; k < kRayBatch && marching; ++k) {
                tcur = tb[k];
                if (!(tcur < max_t) || !inb[k]) {  // loop guard / checkInds: leave the march
                    marching = false;
                    break;
                }
                const float tprev = tsdf;
                idx = ib[k];
                tsdf = sv[k];
                if (isnan(tsdf) || isnan(tprev)) continue;
                if (tprev < 0.f && tsdf > 0.f) marching = false;  // leaving a surface from behind
                else if (tprev >= 0.f && tsdf <= 0.f) {           // zero crossing at tcur
                    marching = false;
                    crossing = true;
                }
            }
            if (marching) {
                tcur = tn;
                prev_safe = safe[kRayBatch - 1];
            }
        }
        if (crossing) {
            do {
                // both interpolations are evaluated (pure reads) so that their loads overlap
                const float Ftdt = trilinear<S, WRAP, COUNT, I32>(add3(rs, mul3(dir, tcur + step)), vv, p, ri);
                const float Ft = trilinear<S, WRAP, COUNT, I32>(add3(rs, mul3(dir, tcur)), vv, p, ri);
                if (isnan(Ftdt) || isnan(Ft)) break;
                const float Ts = tcur - step * Ft / (Ftdt - Ft);
                const f3 vf = add3(rs, mul3(dir, Ts));
                vout = sub3(vf, off);
                if (cvol && cmap) cout = cvol[idx];
                int gg[3];
                get_voxel_fast(add3(rs, mul3(dir, tcur)), p, ri, gg);
                if (gg[0] > 1 && gg[1] > 1 && gg[2] > 1 && gg[0] < vv.Nx - 2 && gg[1] < vv.Ny - 2 && gg[2] < vv.Nz - 2) {
                    f3 n, q;
                    q = vf; q.x += p.cell_size[0]; const float fx1 = trilinear<S, WRAP, COUNT, I32>(q, vv, p, ri);
                    q = vf; q.x -= p.cell_size[0]; const float fx2 = trilinear<S, WRAP, COUNT, I32>(q, vv, p, ri);
                    n.x = fx1 - fx2;
                    q = vf; q.y += p.cell_size[1]; const float fy1 = trilinear<S, WRAP, COUNT, I32>(q, vv, p, ri);
                    q = vf; q.y -= p.cell_size[1]; const float fy2 = trilinear<S, WRAP, COUNT, I32>(q, vv, p, ri);
                    n.y = fy1 - fy2;
                    q = vf; q.z += p.cell_size[2]; const float fz1 = trilinear<S, WRAP, COUNT, I32>(q, vv, p, ri);
                    q = vf; q.z -= p.cell_size[2]; const float fz2 = trilinear<S, WRAP, COUNT, I32>(q, vv, p, ri);
                    n.z = fz1 - fz2;
                    nout = normalize3(n);
                }
            } while (false);
        }
    }
    st3(vmap, id, vout);
    st3(nmap, id, nout);
    if (cmap) cmap[id] = cout;
}

// linear voxel indices fit 31 bits (I32 launches)
inline bool index32(const TsdfParams& p) {
    return (double)p.resolution[0] * (double)p.resolution[1] * (double)p.resolution[2] < 2147483648.0;
}

template <class S, bool WRAP, bool I32>
__global__ void __launch_bounds__(64)
    raycast_kernel(const typename S::word* __restrict__ vol, const uchar4* __restrict__ cvol, TsdfParams p,
                   NuiCLCameraParams cam, const NuiCLRigidTransform* __restrict__ pose, float* __restrict__ vmap,
                   float* __restrict__ nmap, uchar4* __restrict__ cmap, const int32_t* __restrict__ enable,
                   SurfaceBlocks surf, VoxelWrap wr) {
    raycast_body<S, WRAP, false, I32>(vol, cvol, p, cam, pose, vmap, nmap, cmap, enable, surf, wr, blockIdx.x,
                                      blockIdx.y);
}

// the raycast of raycast_kernel (no wrap) marking every voxel it reads
template <class S>
__global__ void __launch_bounds__(64)
    raycast_count_kernel(const typename S::word* __restrict__ vol, TsdfParams p, NuiCLCameraParams cam,
                         const NuiCLRigidTransform* __restrict__ pose, float* __restrict__ vmap,
                         float* __restrict__ nmap, SurfaceBlocks surf, uint32_t* __restrict__ mark) {
    raycast_body<S, false, true>(vol, nullptr, p, cam, pose, vmap, nmap, nullptr, nullptr, surf, VoxelWrap{0, 0, 0},
                                 blockIdx.x, blockIdx.y, mark);
}

__global__ void popcount_kernel(const uint32_t* __restrict__ bits, size_t n, unsigned long long* __restrict__ out) {
    unsigned long long c = 0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        c += __popc(bits[i]);
    for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off);
    if ((threadIdx.x & 63) == 0 && c) atomicAdd(out, c);
}

// all sessions' 8 x 8 pixel tiles of one storage (volumes without wrap)
template <class S, bool I32>
__global__ void __launch_bounds__(64) raycast_batch_kernel(const RayJob* __restrict__ jobs, int n, unsigned blocks,
                                                            int remap) {
    const unsigned lb = remap ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
    if (lb >= blocks) return;
    const RayJob& J = jobs[job_of(jobs, n, lb)];
    const unsigned b = lb - J.block0;
    raycast_body<S, false, false, I32>(in_global(reinterpret_cast<const typename S::word*>(J.vol)), in_global(J.cvol),
                                       J.p, J.cam, in_global(J.pose), in_global(J.vmap), in_global(J.nmap),
                                       in_global(J.cmap), in_global(J.enable), in_global(J.surf), VoxelWrap{0, 0, 0},
                                       (int)(b % J.gx), (int)(b / J.gx));
}

// ---- host launchers ------------------------------------------------------------

template <class S>
int launch_integrate(const float* depth, const uint8_t* bgra, void* volume, uint8_t* color_volume,
                     const TsdfParams& p, const NuiCLCameraParams& cam, const NuiCLRigidTransform* pose, int maxcw,
                     const int32_t* gate, unsigned long long* updated, const float* bounds, uint8_t* surface,
                     VoxelWrap wr, float* walk_states, hipStream_t st) {
    const int Nx = (int)p.resolution[0], Ny = (int)p.resolution[1], Nz = (int)p.resolution[2];
    const SurfaceBlocks surf = surface_blocks_of(surface, Nx, Ny, Nz);
    const bool color = bgra && color_volume;
    auto* vol = reinterpret_cast<typename S::word*>(volume);
    auto* cv = reinterpret_cast<uchar4*>(color_volume);
    auto* bg = reinterpret_cast<const uchar4*>(bgra);
    // the tile kernel's 16-byte vectors stay aligned and unsplit when w.z is a multiple of
    // the voxels per vector
    if (Nx % 8 == 0 && Ny % 8 == 0 && Nz % kZChunk == 0 && wr.z % (16 / S::kBytes) == 0) {
        const int seg = segment_voxels(1);
        dim3 grid(Nx / 8, Ny / 8, div_up(Nz, seg));
        float2* ws = reinterpret_cast<float2*>(walk_states);
        if (ws) {
            column_states_kernel<<<column_state_blocks(true, Nx, Ny, grid.z), 256, 0, st>>>(p, cam, pose, gate, ws,
                                                                                            seg);
            NUIF_LAUNCHED("column_states_kernel");
        }
        if (color)
            integrate_tile_kernel<S, true><<<grid, 64, 0, st>>>(depth, bg, vol, cv, p, cam, pose, maxcw, gate, updated,
                                                                 bounds, surf, wr, ws, seg);
        else
            integrate_tile_kernel<S, false><<<grid, 64, 0, st>>>(depth, bg, vol, cv, p, cam, pose, maxcw, gate, updated,
                                                                  bounds, surf, wr, ws, seg);
        NUIF_LAUNCHED("integrate_tile_kernel");
    } else {
        dim3 block(16, 16), grid(div_up(Nx, 16), div_up(Ny, 16));
        if (color)
            integrate_column_kernel<S, true><<<grid, block, 0, st>>>(depth, bg, vol, cv, p, cam, pose, maxcw, gate, updated,
                                                                     surf, wr);
        else
            integrate_column_kernel<S, false><<<grid, block, 0, st>>>(depth, bg, vol, cv, p, cam, pose, maxcw, gate,
                                                                      updated, surf, wr);
        NUIF_LAUNCHED("integrate_column_kernel");
    }
    return NUIF_OK;
}

template <class S>
int launch_raycast(const void* volume, const uint8_t* color_volume, const TsdfParams& p,
                   const NuiCLCameraParams& cam, const NuiCLRigidTransform* pose, float* vmap, float* nmap,
                   uint8_t* cmap, const int32_t* enable, const uint8_t* surface, VoxelWrap wr, hipStream_t st) {
    const int W = (int)cam.depthImageWidth, H = (int)cam.depthImageHeight;
    dim3 grid(div_up(W, 8), div_up(H, 8));
    const auto* vol = reinterpret_cast<const typename S::word*>(volume);
    const auto* cv = reinterpret_cast<const uchar4*>(color_volume);
    auto* cm = reinterpret_cast<uchar4*>(cmap);
    if (wr.zero()) {
        const SurfaceBlocks surf = surface_blocks_of(const_cast<uint8_t*>(surface), (int)p.resolution[0],
                                                     (int)p.resolution[1], (int)p.resolution[2]);
        if (index32(p))
            raycast_kernel<S, false, true><<<grid, 64, 0, st>>>(vol, cv, p, cam, pose, vmap, nmap, cm, enable, surf, wr);
        else
            raycast_kernel<S, false, false><<<grid, 64, 0, st>>>(vol, cv, p, cam, pose, vmap, nmap, cm, enable, surf, wr);
    } else {
        // the surface-block skip works on logical positions: not used in a wrapped volume
        const SurfaceBlocks none = surface_blocks_of(nullptr, (int)p.resolution[0], (int)p.resolution[1],
                                                     (int)p.resolution[2]);
        raycast_kernel<S, true, false><<<grid, 64, 0, st>>>(vol, cv, p, cam, pose, vmap, nmap, cm, enable, none, wr);
    }
    NUIF_LAUNCHED("raycast_kernel");
    return NUIF_OK;
}

template <class S>
int launch_raycast_count(const void* volume, const TsdfParams& p, const NuiCLCameraParams& cam,
                         const NuiCLRigidTransform* pose, float* vmap, float* nmap, const uint8_t* surface,
                         uint32_t* bits, hipStream_t st) {
    const int W = (int)cam.depthImageWidth, H = (int)cam.depthImageHeight;
    const SurfaceBlocks surf = surface_blocks_of(const_cast<uint8_t*>(surface), (int)p.resolution[0],
                                                 (int)p.resolution[1], (int)p.resolution[2]);
    raycast_count_kernel<S><<<dim3(div_up(W, 8), div_up(H, 8)), 64, 0, st>>>(
        reinterpret_cast<const typename S::word*>(volume), p, cam, pose, vmap, nmap, surf, bits);
    NUIF_LAUNCHED("raycast_count_kernel");
    return NUIF_OK;
}

int launch_bounds_batch(const BoundsJob* jobs, const BoundsJob* hjobs, int n, hipStream_t st) {
    if (n <= 0) return NUIF_OK;
    const BoundsJob& last = hjobs[n - 1];
    const unsigned blocks = last.block0 + div_up((size_t)bounds_dim(last.W, 0) * bounds_dim(last.H, 0), 256);
    depth_bounds_base_batch_kernel<<<blocks, 256, 0, st>>>(jobs, n);
    NUIF_LAUNCHED("depth_bounds_base_batch_kernel");
    depth_bounds_levels_batch_kernel<<<n, 1024, 0, st>>>(jobs);
    NUIF_LAUNCHED("depth_bounds_levels_batch_kernel");
    return NUIF_OK;
}

int launch_integrate_batch(const IntJob* jobs, const IntJob* hjobs, int n, int storage, bool color, hipStream_t st) {
    if (n <= 0) return NUIF_OK;
    const IntJob& last = hjobs[n - 1];
    const unsigned cblocks = last.cblock0 + column_state_blocks(last.walk != nullptr, (int)last.p.resolution[0],
                                                                (int)last.p.resolution[1], last.gz);
    if (cblocks > 0) {
        column_states_batch_kernel<<<cblocks, 256, 0, st>>>(jobs, n);
        NUIF_LAUNCHED("column_states_batch_kernel");
    }
    const unsigned blocks = last.block0 + last.gx * last.gy * last.gz;
    const int remap = xcd_remap(kXcdIntegrate) ? 1 : 0;
    switch (storage) {
        case NUIF_TSDF_S16:
            if (color) integrate_tile_batch_kernel<StS16, true><<<xcd_grid(blocks), 64, 0, st>>>(jobs, n, blocks, remap);
            else integrate_tile_batch_kernel<StS16, false><<<xcd_grid(blocks), 64, 0, st>>>(jobs, n, blocks, remap);
            break;
        case NUIF_TSDF_F16:
            if (color) integrate_tile_batch_kernel<StF16, true><<<xcd_grid(blocks), 64, 0, st>>>(jobs, n, blocks, remap);
            else integrate_tile_batch_kernel<StF16, false><<<xcd_grid(blocks), 64, 0, st>>>(jobs, n, blocks, remap);
            break;
        default:
            if (color) integrate_tile_batch_kernel<StF32, true><<<xcd_grid(blocks), 64, 0, st>>>(jobs, n, blocks, remap);
            else integrate_tile_batch_kernel<StF32, false><<<xcd_grid(blocks), 64, 0, st>>>(jobs, n, blocks, remap);
            break;
    }
    NUIF_LAUNCHED("integrate_tile_batch_kernel");
    return NUIF_OK;
}

int launch_raycast_batch(const RayJob* jobs, const RayJob* hjobs, int n, int storage, hipStream_t st) {
    if (n <= 0) return NUIF_OK;
    const RayJob& last = hjobs[n - 1];
    const unsigned blocks = last.block0 + last.gx * div_up(last.cam.depthImageHeight, 8);
    const int remap = xcd_remap(kXcdRaycast) ? 1 : 0;
    const bool i32 = std::all_of(hjobs, hjobs + n, [](const RayJob& j) { return index32(j.p); });
    const unsigned g = xcd_grid(blocks);
    switch (storage * 2 + (i32 ? 1 : 0)) {
        case NUIF_TSDF_S16 * 2: raycast_batch_kernel<StS16, false><<<g, 64, 0, st>>>(jobs, n, blocks, remap); break;
        case NUIF_TSDF_S16 * 2 + 1: raycast_batch_kernel<StS16, true><<<g, 64, 0, st>>>(jobs, n, blocks, remap); break;
        case NUIF_TSDF_F16 * 2: raycast_batch_kernel<StF16, false><<<g, 64, 0, st>>>(jobs, n, blocks, remap); break;
        case NUIF_TSDF_F16 * 2 + 1: raycast_batch_kernel<StF16, true><<<g, 64, 0, st>>>(jobs, n, blocks, remap); break;
        case NUIF_TSDF_F32 * 2: raycast_batch_kernel<StF32, false><<<g, 64, 0, st>>>(jobs, n, blocks, remap); break;
        default: raycast_batch_kernel<StF32, true><<<g, 64, 0, st>>>(jobs, n, blocks, remap); break;
    }
    NUIF_LAUNCHED("raycast_batch_kernel");
    return NUIF_OK;
}

static int check_params(const TsdfParams* p) {
    NUIF_CHECK_ARG(p != nullptr, "null TsdfParams");
    for (int k = 0; k < 3; ++k) {
        const float r = p->resolution[k];
        NUIF_CHECK_ARG(r >= 1.f && r <= 4096.f && r == (float)(int)r, "bad resolution %g", (double)r);
        NUIF_CHECK_ARG(p->cell_size[k] > 0.f && p->dimension[k] > 0.f, "bad cell size / dimension");
    }
    NUIF_CHECK_ARG(p->tranc_dist > 0.f, "bad truncation distance");
    return NUIF_OK;
}

}  // namespace nuif

using namespace nuif;

extern "C" {

size_t nuif_tsdf_voxel_bytes(int storage) {
    switch (storage) {
        case NUIF_TSDF_S16: return 4;
        case NUIF_TSDF_F16: return 4;
        case NUIF_TSDF_F32: return 8;
        default: return 0;
    }
}

int nuif_tsdf_params(int nx, int ny, int nz, float dx, float dy, float dz, float trunc, TsdfParams* out) {
    NUIF_CHECK_ARG(out && nx > 0 && ny > 0 && nz > 0 && dx > 0 && dy > 0 && dz > 0, "nuif_tsdf_params: bad arguments");
    out->resolution[0] = (float)nx;
    out->resolution[1] = (float)ny;
    out->resolution[2] = (float)nz;
    out->dimension[0] = dx;
    out->dimension[1] = dy;
    out->dimension[2] = dz;
    out->cell_size[0] = dx / (float)nx;
    out->cell_size[1] = dy / (float)ny;
    out->cell_size[2] = dz / (float)nz;
    const float c = std::max(out->cell_size[0], std::max(out->cell_size[1], out->cell_size[2]));
    out->tranc_dist = std::max(trunc, 2.1f * c);  // OCL/NuiKinfuOpenCLScene.cpp:41
    return NUIF_OK;
}

int nuif_tsdf_reset(void* volume, uint8_t* color_volume, const TsdfParams* params, int storage, nuif_stream_t s) {
    NUIF_REQUIRE_INIT();
    int rc = check_params(params);
    if (rc) return rc;
    NUIF_CHECK_ARG(volume != nullptr, "nuif_tsdf_reset: null volume");
    const size_t n = (size_t)params->resolution[0] * (size_t)params->resolution[1] * (size_t)params->resolution[2];
    const unsigned blocks = std::min<unsigned>(div_up(n, 256), 65536u);
    hipStream_t st = stream_of(s);
    uint32_t* cv = reinterpret_cast<uint32_t*>(color_volume);
    switch (storage) {
        case NUIF_TSDF_S16: reset_kernel<StS16><<<blocks, 256, 0, st>>>((uint32_t*)volume, cv, n); break;
        case NUIF_TSDF_F16: reset_kernel<StF16><<<blocks, 256, 0, st>>>((uint32_t*)volume, cv, n); break;
        case NUIF_TSDF_F32: reset_kernel<StF32><<<blocks, 256, 0, st>>>((uint2*)volume, cv, n); break;
        default: set_error("nuif_tsdf_reset: unknown storage %d", storage); return NUIF_ERR_INVALID;
    }
    NUIF_LAUNCHED("reset_kernel");
    return NUIF_OK;
}

// host wrap (any integers) -> the reduced VoxelWrap of a volume
static VoxelWrap wrap_of(const int32_t* w, const TsdfParams& p) {
    VoxelWrap r{0, 0, 0};
    if (!w) return r;
    int* o[3] = {&r.x, &r.y, &r.z};
    for (int k = 0; k < 3; ++k) {
        const long long n = (long long)p.resolution[k];
        *o[k] = (int)((((long long)w[k] % n) + n) % n);
    }
    return r;
}

size_t nuif_tsdf_walk_states_bytes(const TsdfParams* params) {
    if (!params || check_params(params)) return 0;
    const int Nx = (int)params->resolution[0], Ny = (int)params->resolution[1], Nz = (int)params->resolution[2];
    return (ztable_offset(Nx, Ny, Nz) + (size_t)Nz) * sizeof(float2);  // column states, then the z table
}

int nuif_tsdf_integrate_wrapped(const float* depth, const uint8_t* bgra, void* volume, uint8_t* color_volume,
                                const TsdfParams* params, const NuiCLCameraParams* cam,
                                const NuiCLRigidTransform* pose, int storage, int max_color_weight,
                                const int32_t* wrap, const int32_t* gate, unsigned long long* updated,
                                const float* depth_bounds, uint8_t* surface_blocks, float* walk_states,
                                nuif_stream_t s) {
    NUIF_REQUIRE_INIT();
    int rc = check_params(params);
    if (rc) return rc;
    NUIF_CHECK_ARG(depth && volume && cam && pose, "nuif_tsdf_integrate: null argument");
    NUIF_CHECK_ARG(cam->depthImageWidth > 0 && cam->depthImageHeight > 0, "nuif_tsdf_integrate: empty image");
    hipStream_t st = stream_of(s);
    const VoxelWrap wr = wrap_of(wrap, *params);
    switch (storage) {
        case NUIF_TSDF_S16:
            return launch_integrate<StS16>(depth, bgra, volume, color_volume, *params, *cam, pose, max_color_weight,
                                          gate, updated, depth_bounds, surface_blocks, wr, walk_states, st);
        case NUIF_TSDF_F16:
            return launch_integrate<StF16>(depth, bgra, volume, color_volume, *params, *cam, pose, max_color_weight,
                                          gate, updated, depth_bounds, surface_blocks, wr, walk_states, st);
        case NUIF_TSDF_F32:
            return launch_integrate<StF32>(depth, bgra, volume, color_volume, *params, *cam, pose, max_color_weight,
                                          gate, updated, depth_bounds, surface_blocks, wr, walk_states, st);
        default: set_error("nuif_tsdf_integrate: unknown storage %d", storage); return NUIF_ERR_INVALID;
    }
}

int nuif_tsdf_integrate(const float* depth, const uint8_t* bgra, void* volume, uint8_t* color_volume,
                        const TsdfParams* params, const NuiCLCameraParams* cam, const NuiCLRigidTransform* pose,
                        int storage, int max_color_weight, const int32_t* gate, unsigned long long* updated,
                        const float* depth_bounds, uint8_t* surface_blocks, nuif_stream_t s) {
    return nuif_tsdf_integrate_wrapped(depth, bgra, volume, color_volume, params, cam, pose, storage,
                                       max_color_weight, nullptr, gate, updated, depth_bounds, surface_blocks, nullptr,
                                       s);
}

int nuif_tsdf_reset_slice(void* volume, uint8_t* color_volume, const TsdfParams* params, int storage,
                          const int32_t* wrap, const int32_t* range, nuif_stream_t s) {
    NUIF_REQUIRE_INIT();
    int rc = check_params(params);
    if (rc) return rc;
    NUIF_CHECK_ARG(volume && wrap && range, "nuif_tsdf_reset_slice: null argument");
    const int Nx = (int)params->resolution[0], Ny = (int)params->resolution[1], Nz = (int)params->resolution[2];
    NUIF_CHECK_ARG(range[0] >= 0 && range[0] <= Nx && range[1] >= 0 && range[1] <= Ny && range[2] >= 0 &&
                       range[2] <= Nz,
                   "nuif_tsdf_reset_slice: range outside the volume");
    const size_t n = (size_t)range[0] * range[1] * range[2];
    if (n == 0) return NUIF_OK;
    const VoxelWrap w = wrap_of(wrap, *params);
    const unsigned blocks = std::min<unsigned>(div_up(n, 256), 65536u);
    hipStream_t st = stream_of(s);
    uint32_t* cv = reinterpret_cast<uint32_t*>(color_volume);
    switch (storage) {
        case NUIF_TSDF_S16:
            reset_slice_kernel<StS16><<<blocks, 256, 0, st>>>((uint32_t*)volume, cv, Nx, Ny, Nz, w, range[0], range[1],
                                                               range[2]);
            break;
        case NUIF_TSDF_F16:
            reset_slice_kernel<StF16><<<blocks, 256, 0, st>>>((uint32_t*)volume, cv, Nx, Ny, Nz, w, range[0], range[1],
                                                               range[2]);
            break;
        case NUIF_TSDF_F32:
            reset_slice_kernel<StF32><<<blocks, 256, 0, st>>>((uint2*)volume, cv, Nx, Ny, Nz, w, range[0], range[1],
                                                               range[2]);
            break;
        default: set_error("nuif_tsdf_reset_slice: unknown storage %d", storage); return NUIF_ERR_INVALID;
    }
    NUIF_LAUNCHED("reset_slice_kernel");
    return NUIF_OK;
}

size_t nuif_tsdf_depth_bounds_bytes(int width, int height) {
    if (width <= 0 || height <= 0) return 0;
    return bounds_offset(width, height, bounds_levels(width, height)) * sizeof(float);
}

int nuif_tsdf_depth_bounds(const float* depth, int width, int height, float* bounds, const int32_t* gate,
                           nuif_stream_t s) {
    NUIF_REQUIRE_INIT();
    NUIF_CHECK_ARG(depth && bounds && width > 0 && height > 0, "nuif_tsdf_depth_bounds: bad argument");
    hipStream_t st = stream_of(s);
    const int cells = bounds_dim(width, 0) * bounds_dim(height, 0);
    const BoundsJob bj{depth, width, height, bounds, gate, 0u};
    depth_bounds_base_kernel<<<div_up(cells, 256), 256, 0, st>>>(bj);
    NUIF_LAUNCHED("depth_bounds_base_kernel");
    depth_bounds_levels_kernel<<<1, 1024, 0, st>>>(bj);
    NUIF_LAUNCHED("depth_bounds_levels_kernel");
    return NUIF_OK;
}

int nuif_tsdf_raycast_wrapped(const void* volume, const uint8_t* color_volume, const TsdfParams* params,
                              const NuiCLCameraParams* cam, const NuiCLRigidTransform* pose, float* vmap, float* nmap,
                              uint8_t* cmap, int storage, const int32_t* wrap, const int32_t* enable,
                              const uint8_t* surface_blocks, nuif_stream_t s) {
    NUIF_REQUIRE_INIT();
    int rc = check_params(params);
    if (rc) return rc;
    NUIF_CHECK_ARG(volume && cam && pose && vmap && nmap, "nuif_tsdf_raycast: null argument");
    NUIF_CHECK_ARG(cam->depthImageWidth > 0 && cam->depthImageHeight > 0, "nuif_tsdf_raycast: empty image");
    hipStream_t st = stream_of(s);
    const VoxelWrap wr = wrap_of(wrap, *params);
    switch (storage) {
        case NUIF_TSDF_S16: return launch_raycast<StS16>(volume, color_volume, *params, *cam, pose, vmap, nmap, cmap, enable,
                                                            surface_blocks, wr, st);
        case NUIF_TSDF_F16: return launch_raycast<StF16>(volume, color_volume, *params, *cam, pose, vmap, nmap, cmap, enable,
                                                            surface_blocks, wr, st);
        case NUIF_TSDF_F32: return launch_raycast<StF32>(volume, color_volume, *params, *cam, pose, vmap, nmap, cmap, enable,
                                                            surface_blocks, wr, st);
        default: set_error("nuif_tsdf_raycast: unknown storage %d", storage); return NUIF_ERR_INVALID;
    }
}

int nuif_schedule_segment(int sessions) { return sessions > 0 ? segment_voxels(sessions) : NUIF_ERR_INVALID; }

int nuif_schedule(int32_t* out4) {
    NUIF_CHECK_ARG(out4, "nuif_schedule: null argument");
    out4[0] = kRayBatch;
    out4[1] = kWalkBatch;
    out4[2] = kRmwBatch;
    out4[3] = segment_voxels(1);
    return NUIF_OK;
}

size_t nuif_tsdf_footprint_bytes(const TsdfParams* params) {
    if (!params) return 0;
    const size_t n = (size_t)params->resolution[0] * params->resolution[1] * params->resolution[2];
    return ((n + 31) / 32 * sizeof(uint32_t) + 255) / 256 * 256 + 256;  // the bitset, then the count
}

long long nuif_tsdf_raycast_footprint(const void* volume, const TsdfParams* params, const NuiCLCameraParams* cam,
                                      const NuiCLRigidTransform* pose, int storage, const uint8_t* surface_blocks,
                                      float* vmap, float* nmap, void* workspace, nuif_stream_t s) {
    NUIF_REQUIRE_INIT();
    int rc = check_params(params);
    if (rc) return rc;
    NUIF_CHECK_ARG(volume && cam && pose && vmap && nmap && workspace, "nuif_tsdf_raycast_footprint: null argument");
    NUIF_CHECK_ARG(cam->depthImageWidth > 0 && cam->depthImageHeight > 0, "nuif_tsdf_raycast_footprint: empty image");
    hipStream_t st = stream_of(s);
    const size_t n = (size_t)params->resolution[0] * params->resolution[1] * params->resolution[2];
    const size_t words = (n + 31) / 32;
    uint32_t* bits = static_cast<uint32_t*>(workspace);
    auto* count = reinterpret_cast<unsigned long long*>(static_cast<char*>(workspace) + (words * 4 + 255) / 256 * 256);
    NUIF_HIP(hipMemsetAsync(workspace, 0, nuif_tsdf_footprint_bytes(params), st));
    switch (storage) {
        case NUIF_TSDF_S16: rc = launch_raycast_count<StS16>(volume, *params, *cam, pose, vmap, nmap, surface_blocks, bits, st); break;
        case NUIF_TSDF_F16: rc = launch_raycast_count<StF16>(volume, *params, *cam, pose, vmap, nmap, surface_blocks, bits, st); break;
        case NUIF_TSDF_F32: rc = launch_raycast_count<StF32>(volume, *params, *cam, pose, vmap, nmap, surface_blocks, bits, st); break;
        default: set_error("nuif_tsdf_raycast_footprint: unknown storage %d", storage); return NUIF_ERR_INVALID;
    }
    if (rc) return rc;
    popcount_kernel<<<(unsigned)std::min<size_t>(div_up(words, 256), 4096), 256, 0, st>>>(bits, words, count);
    NUIF_LAUNCHED("popcount_kernel");
    unsigned long long h = 0;
    NUIF_HIP(hipMemcpyAsync(&h, count, sizeof(h), hipMemcpyDeviceToHost, st));
    NUIF_HIP(hipStreamSynchronize(st));
    return (long long)h;
}

int nuif_tsdf_raycast_gated(const void* volume, const uint8_t* color_volume, const TsdfParams* params,
                            const NuiCLCameraParams* cam, const NuiCLRigidTransform* pose, float* vmap, float* nmap,
                            uint8_t* cmap, int storage, const int32_t* enable, const uint8_t* surface_blocks,
                            nuif_stream_t s) {
    return nuif_tsdf_raycast_wrapped(volume, color_volume, params, cam, pose, vmap, nmap, cmap, storage, nullptr,
                                     enable, surface_blocks, s);
}

int nuif_tsdf_raycast(const void* volume, const uint8_t* color_volume, const TsdfParams* params,
                      const NuiCLCameraParams* cam, const NuiCLRigidTransform* pose, float* vmap, float* nmap,
                      uint8_t* cmap, int storage, nuif_stream_t s) {
    return nuif_tsdf_raycast_gated(volume, color_volume, params, cam, pose, vmap, nmap, cmap, storage, nullptr,
                                   nullptr, s);
}

size_t nuif_tsdf_surface_blocks_bytes(const TsdfParams* params) {
    if (!params || check_params(params)) return 0;
    const int nx = (int)params->resolution[0], ny = (int)params->resolution[1], nz = (int)params->resolution[2];
    return (size_t)surf_dim(nx) * surf_dim(ny) * surf_dim(nz) + (size_t)super_dim(nx) * super_dim(ny) * super_dim(nz);
}

int nuif_tsdf_surface_blocks_clear(uint8_t* surface_blocks, const TsdfParams* params, nuif_stream_t s) {
    NUIF_REQUIRE_INIT();
    int rc = check_params(params);
    if (rc) return rc;
    NUIF_CHECK_ARG(surface_blocks != nullptr, "nuif_tsdf_surface_blocks_clear: null buffer");
    return nuif_fill_u8(surface_blocks, 0, nuif_tsdf_surface_blocks_bytes(params), s);
}

}  // extern "C"
