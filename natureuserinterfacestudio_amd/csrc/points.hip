// Point cloud of the dense volume: K/extract.cl:4-78 fetchVolumeKernel driven by
// OCL/NuiKinfuOpenCLScene.cpp:436-570 FetchVolume (range_z = resolution z).
//
// A voxel is a point where its tsdf is exactly 0 (the voxel centre) or where the
// tsdf changes sign from the voxel below it in z (the linear zero crossing
// between the two); unobserved voxels (weight <= 0) read as NaN and never pair.
//
// The reference runs one work-item per (x,y) column with a local size of 1x1,
// walking z serially, and appends through a global atomic counter, so its point
// order changes from run to run and adjacent work-items read addresses Nz*4
// bytes apart.  Here one wavefront walks one column 64 voxels per step (one
// 256-byte load per step; the predecessor comes from the neighbouring lane):
//   pass 1  point count per column;
//   scan    exclusive prefix sum over the columns (device-wide, deterministic);
//   pass 2  each lane writes its point at column offset + in-column prefix.
// The points come out in (y, x, z) order, the order of the reference's serial
// walk.  The total stays in device memory; the cap (the reference's 5,000,000,
// K/gpu_def.h:3) only limits what is written.
//
// With the surface-block flags of the integration, a 64-voxel step whose voxels
// and predecessor all lie in clear blocks is skipped without reading it: a point
// needs a stored tsdf <= 0 at z (a zero or a crossing into the negative) or at
// z-1 (a crossing out of it), and clear blocks hold none.
//
// The same kernel is fetchShiftVolumeKernel (K/extract.cl:80-190) of the shifting
// volume: the slab that a shift evicts, or the whole volume placed at its offset.

#include "dense_volume.h"

namespace nuif {

constexpr int kPointColumns = 4;  // columns (wavefronts) per 256-thread block

template <class S>
__device__ __forceinline__ float observed_tsdf(typename S::word w) {
    return S::observed(w) ? S::tsdf(w) : __builtin_nanf("");
}

// The extraction of K/extract.cl:80-190 fetchShiftVolumeKernel, of which the whole-volume
// fetchVolumeKernel is the case w = 0, o = 0, r = N: the work-item grid is rx x ry columns,
// column (x, y) is the stored column ((x + w.x) % Nx, (y + w.y) % Ny), its walk runs over
// voxel_z = w.z .. w.z + rz - 1 and reads stored z (voxel_z + w.z) % Nz (the reference adds
// w.z twice), and the points are placed at voxel coordinates shifted by the volume offset o
// as the reference computes it (the global position of the stored voxel).
struct SliceArgs {
    VoxelWrap w;     // as passed, 0 <= w <= N
    int ox, oy, oz;  // the volume's voxel offset
    int rx, ry, rz;  // work-item grid and walk length
};

// K/extract.cl:104-123 (and :139-148 for z): the multiple of N added to coordinate v
__device__ __forceinline__ int slice_offset(int v, int o, int n) {
    int b = (o / n) * n;  // C division: toward zero
    if (o > 0) {
        if (v < o % n) b += n;
    } else {
        if (v > n + o % n) b -= n;
    }
    return b;
}

template <class S, bool EMIT>
__global__ void __launch_bounds__(256)
    points_dense_kernel(const typename S::word* __restrict__ vol, const uchar4* __restrict__ cvol, TsdfParams p,
                        SliceArgs a, const uint8_t* __restrict__ surface, uint32_t* __restrict__ column_counts,
                        float* __restrict__ pts, float* __restrict__ cols, uint32_t capacity) {
    const int Nx = (int)p.resolution[0], Ny = (int)p.resolution[1], Nz = (int)p.resolution[2];
    const int lane = threadIdx.x & 63;
    const int x = blockIdx.x * kPointColumns + (threadIdx.x >> 6), y = blockIdx.y;
    if (x >= a.rx) return;  // whole wavefront
    const size_t cid = (size_t)y * a.rx + x;
    const size_t col = wrapped_column(x, y, a.w, Nx, Ny, Nz);
    // the surface flags are stored-position flags: used only for the unwrapped walk (w = 0)
    const SurfaceBlocks sb = surface_blocks_of(const_cast<uint8_t*>(surface), Nx, Ny, Nz);
    const int gx = x + slice_offset(x, a.ox, Nx), gy = y + slice_offset(y, a.oy, Ny);
    const float px = ((float)gx + 0.5f - (float)(Nx >> 1)) * p.cell_size[0];
    const float py = (-(float)gy - 0.5f + (float)(Ny >> 1)) * p.cell_size[1];
    uint32_t base = EMIT ? column_counts[cid] : 0u;
    for (int k0 = 0; k0 < a.rz; k0 += 64) {
        if (surface) {
            // blocks of voxels k0-1 .. k0+63 of this column (at most 9)
            const int zlo = max(k0 - 1, 0), zhi = min(k0 + 63, a.rz - 1);
            const int nblk = (zhi >> 3) - (zlo >> 3) + 1;
            const bool marked = lane < nblk && sb.b[sb.index(x, y, zlo + 8 * lane)] != 0;
            if (!__any(marked)) continue;
        }
        const int k = k0 + lane;
        const int vz = a.w.z + k;  // the reference's voxel_z
        const float t = k < a.rz ? observed_tsdf<S>(vol[col + (vz + a.w.z) % Nz]) : __builtin_nanf("");
        float tp = __shfl_up(t, 1, 64);
        if (lane == 0) tp = k0 > 0 ? observed_tsdf<S>(vol[col + (vz - 1 + a.w.z) % Nz]) : __builtin_nanf("");
        const bool zero = (0.f == t);
        const bool cross = (tp < 0.f && t > 0.f) || (tp > 0.f && t < 0.f);
        const bool emit = zero || cross;
        const unsigned long long m = __ballot(emit);
        if (EMIT && emit) {
            const uint32_t o = base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
            if (o < capacity) {
                const int gz = vz + slice_offset(vz, a.oz, Nz);
                float pz;
                if (zero) {
                    pz = ((float)gz + 0.5f) * p.cell_size[2];
                } else {
                    const float tsdf_inv = 1.f / fabsf(t);
                    const float tsdf_prev_inv = 1.f / fabsf(tp);
                    const float dp = (tsdf_inv * (float)gz + tsdf_prev_inv * (float)(gz - 1)) / (tsdf_inv + tsdf_prev_inv);
                    pz = (dp + 0.5f) * p.cell_size[2];
                }
                st3(pts, o, mk3(px, py, pz));
                if (cols) {
                    float4 c = make_float4(0.f, 0.f, 0.f, 1.f);
                    if (cvol) {
                        const uchar4 q = cvol[col + (vz + a.w.z) % Nz];
                        // the two branches read the colour bytes in opposite orders (K/extract.cl:41-46, 62-67)
                        c = zero ? make_float4((float)q.x / 255.f, (float)q.y / 255.f, (float)q.z / 255.f, 1.f)
                                 : make_float4((float)q.z / 255.f, (float)q.y / 255.f, (float)q.x / 255.f, 1.f);
                    }
                    reinterpret_cast<float4*>(cols)[o] = c;
                }
            }
        }
        base += (uint32_t)__popcll(m);
    }
    if (!EMIT && lane == 0) column_counts[cid] = base;
}

template <class S>
static void launch_points(bool emit, const void* vol, const uchar4* cv, const TsdfParams& p, const SliceArgs& a,
                          const uint8_t* surface, uint32_t* counts, float* pts, float* cols, uint32_t capacity,
                          hipStream_t st) {
    const dim3 grid((unsigned)div_up(a.rx, kPointColumns), (unsigned)a.ry);
    const auto* v = static_cast<const typename S::word*>(vol);
    if (emit)
        points_dense_kernel<S, true><<<grid, 256, 0, st>>>(v, cv, p, a, surface, counts, pts, cols, capacity);
    else
        points_dense_kernel<S, false><<<grid, 256, 0, st>>>(v, cv, p, a, surface, counts, pts, cols, capacity);
}

static int points_run(const void* volume, const uint8_t* color_volume, const TsdfParams* p, int storage,
                      const SliceArgs& a, const uint8_t* surface, float* points, float* colors, uint32_t capacity,
                      uint32_t* total, void* scratch, nuif_stream_t s) {
    const size_t ncol = (size_t)a.rx * a.ry;
    uint32_t* counts = static_cast<uint32_t*>(scratch);
    void* scan_scratch = counts + ncol + 64;
    hipStream_t st = stream_of(s);
    const uchar4* cv = reinterpret_cast<const uchar4*>(color_volume);
    for (int pass = 0; pass < 2; ++pass) {
        const bool emit = pass == 1;
        if (emit && capacity == 0) break;
        switch (storage) {
            case NUIF_TSDF_S16:
                launch_points<StS16>(emit, volume, cv, *p, a, surface, counts, points, colors, capacity, st);
                break;
            case NUIF_TSDF_F16:
                launch_points<StF16>(emit, volume, cv, *p, a, surface, counts, points, colors, capacity, st);
                break;
            default:
                launch_points<StF32>(emit, volume, cv, *p, a, surface, counts, points, colors, capacity, st);
                break;
        }
        NUIF_LAUNCHED("points_dense_kernel");
        if (!emit) {
            const int rc = launch_scan(counts, false, counts, ncol, total, scan_scratch, st);
            if (rc) return rc;
        }
    }
    return NUIF_OK;
}

}  // namespace nuif

using namespace nuif;

extern "C" {

size_t nuif_points_dense_scratch_bytes(const TsdfParams* p) {
    if (!p) return 0;
    const size_t cols = (size_t)p->resolution[0] * (size_t)p->resolution[1];
    return (cols + 64) * sizeof(uint32_t) + nuif_scan_scratch_bytes(cols);
}

int nuif_points_dense(const void* volume, const uint8_t* color_volume, const TsdfParams* p, int storage,
                      const uint8_t* surface_blocks, float* points, float* colors, uint32_t capacity,
                      uint32_t* total, void* scratch, nuif_stream_t s) {
    NUIF_REQUIRE_INIT();
    NUIF_CHECK_ARG(volume && p && total && scratch, "nuif_points_dense: null argument");
    NUIF_CHECK_ARG(capacity == 0 || points, "nuif_points_dense: point buffer missing");
    for (int k = 0; k < 3; ++k)
        NUIF_CHECK_ARG(p->resolution[k] >= 1.f && p->resolution[k] <= 65536.f, "nuif_points_dense: bad resolution");
    NUIF_CHECK_ARG(storage == NUIF_TSDF_S16 || storage == NUIF_TSDF_F16 || storage == NUIF_TSDF_F32,
                   "nuif_points_dense: unknown storage %d", storage);
    const SliceArgs a{{0, 0, 0}, 0, 0, 0, (int)p->resolution[0], (int)p->resolution[1], (int)p->resolution[2]};
    return points_run(volume, color_volume, p, storage, a, surface_blocks, points, colors, capacity, total, scratch,
                      s);
}

int nuif_points_dense_slice(const void* volume, const uint8_t* color_volume, const TsdfParams* p, int storage,
                            const int32_t* wrap, const int32_t* offset, const int32_t* range, float* points,
                            float* colors, uint32_t capacity, uint32_t* total, void* scratch, nuif_stream_t s) {
    NUIF_REQUIRE_INIT();
    NUIF_CHECK_ARG(volume && p && wrap && offset && range && total && scratch, "nuif_points_dense_slice: null argument");
    NUIF_CHECK_ARG(capacity == 0 || points, "nuif_points_dense_slice: point buffer missing");
    NUIF_CHECK_ARG(storage == NUIF_TSDF_S16 || storage == NUIF_TSDF_F16 || storage == NUIF_TSDF_F32,
                   "nuif_points_dense_slice: unknown storage %d", storage);
    int N[3];
    for (int k = 0; k < 3; ++k) {
        NUIF_CHECK_ARG(p->resolution[k] >= 1.f && p->resolution[k] <= 65536.f, "nuif_points_dense_slice: bad resolution");
        N[k] = (int)p->resolution[k];
        NUIF_CHECK_ARG(wrap[k] >= 0 && wrap[k] <= N[k], "nuif_points_dense_slice: wrap outside [0, N]");
        NUIF_CHECK_ARG(range[k] >= 0 && range[k] <= N[k], "nuif_points_dense_slice: range outside [0, N]");
    }
    const SliceArgs a{{wrap[0], wrap[1], wrap[2]}, offset[0], offset[1], offset[2], range[0], range[1], range[2]};
    if ((size_t)a.rx * a.ry == 0) {
        NUIF_CHECK_ARG(capacity == 0 || points, "nuif_points_dense_slice: point buffer missing");
        return nuif_fill_u32(total, 0u, 1, s);
    }
    return points_run(volume, color_volume, p, storage, a, nullptr, points, colors, capacity, total, scratch, s);
}

}  // extern "C"
