// Dense TSDF volume access shared by the integration, the raycast and the
// point extraction: voxel storage policies and the surface-block flags.
#pragma once

#include <hip/hip_fp16.h>

#include "nuif_internal.h"

namespace nuif {

// ---- voxel storage policies ----------------------------------------------------

// Reference short2: int16 tsdf * 32766 (round toward zero), int16 weight (K/utils.cl:69-84).
struct StS16 {
    using word = uint32_t;
    static constexpr int kBytes = 4;
    __device__ static word init() { return 0x00007fffu; }  // (SHORT_NAN 32767, 0)
    __device__ static float tsdf(word w) { return (float)(int16_t)(w & 0xffffu) / 32766.0f; }
    __device__ static bool nonpositive(word w) { return (int16_t)(w & 0xffffu) <= 0; }  // tsdf(w) <= 0
    __device__ static int weight(word w) { return (int16_t)(w >> 16); }
    __device__ static bool observed(word w) { return (int16_t)(w >> 16) > 0; }  // weight > 0
    __device__ static word pack(float f, int wt) {
        const int16_t q = (int16_t)(int)(f * 32766.0f);
        return (uint32_t)(uint16_t)q | ((uint32_t)(uint16_t)wt << 16);
    }
};
// FP16 tsdf (round to nearest even) + uint16 weight.
struct StF16 {
    using word = uint32_t;
    static constexpr int kBytes = 4;
    __device__ static word init() { return (uint32_t)__half_as_ushort(__float2half_rn(1.0f)); }
    __device__ static float tsdf(word w) { return __half2float(__ushort_as_half((unsigned short)(w & 0xffffu))); }
    // tsdf(w) <= 0: sign bit set or +0 (no NaN is ever stored)
    __device__ static bool nonpositive(word w) { return (w & 0x8000u) || (w & 0x7fffu) == 0; }
    __device__ static int weight(word w) { return (int)(w >> 16); }
    __device__ static bool observed(word w) { return (w >> 16) != 0u; }
    __device__ static word pack(float f, int wt) {
        return (uint32_t)__half_as_ushort(__float2half_rn(f)) | ((uint32_t)(uint16_t)wt << 16);
    }
};
// FP32 tsdf + FP32 weight.
struct StF32 {
    using word = uint2;
    static constexpr int kBytes = 8;
    __device__ static word init() { return make_uint2(__float_as_uint(1.0f), 0u); }
    __device__ static float tsdf(word w) { return __uint_as_float(w.x); }
    __device__ static bool nonpositive(word w) { return !(__uint_as_float(w.x) > 0.f); }
    __device__ static int weight(word w) { return (int)__uint_as_float(w.y); }
    __device__ static bool observed(word w) { return __uint_as_float(w.y) > 0.f; }
    __device__ static word pack(float f, int wt) { return make_uint2(__float_as_uint(f), __float_as_uint((float)wt)); }
};

// Surface blocks: one byte per 8 x 8 x 8 voxel block (same y-major, z-fastest order
// as the voxels), set by the integration when it stores a tsdf <= 0 and never
// cleared, so a zero byte guarantees that every voxel of the block reads positive
// or unobserved.  A second level holds one byte per 32 x 32 x 32 superblock (the OR
// of its blocks).  The raycast reads the voxels of clear blocks only where its exit
// tests can need their values, and crosses the inside of clear superblocks without
// locating its samples.
constexpr int kSurfBlock = 8, kSurfSuper = 32;
struct SurfaceBlocks {
    uint8_t* b;     // blocks, then superblocks at b + nb
    int nbx, nbz;   // blocks along x and z
    size_t nb;      // block count
    int nsx, nsz;   // superblocks along x and z
    __device__ __forceinline__ size_t index(int x, int y, int z) const {
        return ((size_t)(y >> 3) * nbx + (x >> 3)) * nbz + (z >> 3);
    }
    __device__ __forceinline__ size_t sindex(int x, int y, int z) const {
        return nb + ((size_t)(y >> 5) * nsx + (x >> 5)) * nsz + (z >> 5);
    }
    // the same in 32-bit arithmetic (at most 512^3 blocks: 4096 voxels a side)
    __device__ __forceinline__ unsigned index32(int x, int y, int z) const {
        return ((unsigned)(y >> 3) * (unsigned)nbx + (unsigned)(x >> 3)) * (unsigned)nbz + (unsigned)(z >> 3);
    }
    __device__ __forceinline__ unsigned sindex32(int x, int y, int z) const {
        return (unsigned)nb + ((unsigned)(y >> 5) * (unsigned)nsx + (unsigned)(x >> 5)) * (unsigned)nsz +
               (unsigned)(z >> 5);
    }
    __device__ __forceinline__ void mark(int x, int y, int z) const {
        b[index(x, y, z)] = 1;
        b[sindex(x, y, z)] = 1;
    }
};
__host__ __device__ __forceinline__ int surf_dim(int n) { return (n + kSurfBlock - 1) / kSurfBlock; }
__host__ __device__ __forceinline__ int super_dim(int n) { return (n + kSurfSuper - 1) / kSurfSuper; }
__host__ __device__ __forceinline__ SurfaceBlocks surface_blocks_of(uint8_t* b, int Nx, int Ny, int Nz) {
    return SurfaceBlocks{b, surf_dim(Nx), surf_dim(Nz), (size_t)surf_dim(Nx) * surf_dim(Ny) * surf_dim(Nz),
                         super_dim(Nx), super_dim(Nz)};
}

// Cyclic offset of a shifting volume (K/tsdfVolume.cl:71,104, K/raycast.cl:62,184): the
// voxel at logical (x, y, z) is stored at ((x + w.x) % Nx, (y + w.y) % Ny, (z + w.z) % Nz),
// with 0 <= w < N on every axis (the reference's getVoxelWrap reduced modulo N).
struct VoxelWrap {
    int x, y, z;
    __host__ __device__ __forceinline__ bool zero() const { return (x | y | z) == 0; }
};
__device__ __forceinline__ int wrap1(int g, int w, int n) {  // (g + w) % n for 0 <= g, w < n
    const int s = g + w;
    return s >= n ? s - n : s;
}
// first voxel of the stored column of logical column (x, y)
__device__ __forceinline__ size_t wrapped_column(int x, int y, const VoxelWrap& w, int Nx, int Ny, int Nz) {
    return ((size_t)wrap1(y, w.y, Ny) * Nx + wrap1(x, w.x, Nx)) * Nz;
}

}  // namespace nuif
