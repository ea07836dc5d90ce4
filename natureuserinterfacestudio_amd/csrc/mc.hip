// Dense marching cubes: K/marchingCube.cl:100-256 marchingCubeKernel
// driven by OCL/NuiKinfuOpenCLScene.cpp:617-786 Volume2CLMesh.
//
// The reference runs one work-item per (x,y) column with a local size of 1x1
// and appends triangles through a global atomic counter, so the vertex order
// changes from run to run.  Here the cubes are processed in
// 256-cube tiles (4 columns x 64 z, z fastest across the wavefront so that the
// eight corner loads of a wavefront are contiguous 256-byte runs):
//   pass 1  vertex count per tile (sum of the case table's counts);
//   scan    exclusive prefix sum of the tile counts (device-wide, deterministic);
//   pass 2  each cube writes its triangles at tile offset + in-tile prefix.
// Same triangles, same count as the reference, in a fixed order.  The total
// stays in device memory; the vertex cap (the reference's 5,000,000,
// K/gpu_def.h:3) only limits what is written.

#include "block_scan.h"
#include "mc_tables.h"
#include "nuif_internal.h"

#include <hip/hip_fp16.h>

namespace nuif {

__constant__ McTable kMc = mc_decode();

// raw voxel access for the three storages: (weight > 0, tsdf < 0) and tsdf
struct McVoxel {
    float tsdf;
    bool valid;
};

template <int STORAGE>
__device__ __forceinline__ McVoxel mc_load(const void* vol, size_t i) {
    McVoxel v;
    if (STORAGE == NUIF_TSDF_S16) {
        const uint32_t w = static_cast<const uint32_t*>(vol)[i];
        v.valid = (int16_t)(w >> 16) > 0;
        v.tsdf = (float)(int16_t)(w & 0xffffu) / 32766.0f;
    } else if (STORAGE == NUIF_TSDF_F16) {
        const uint32_t w = static_cast<const uint32_t*>(vol)[i];
        v.valid = (w >> 16) > 0;
        v.tsdf = __half2float(__ushort_as_half((unsigned short)(w & 0xffffu)));
    } else {
        const uint2 w = static_cast<const uint2*>(vol)[i];
        v.valid = __uint_as_float(w.y) > 0.f;
        v.tsdf = __uint_as_float(w.x);
    }
    return v;
}

struct Cube {
    size_t id[8];
    float f[8];
    int cubeindex;
    int nverts;
};

// corner ids and case of the cube whose corner 0 is voxel (x,y,z); nverts = 0 when
// any corner has no weight (the reference's `continue`s) or the case is empty
template <int STORAGE>
__device__ __forceinline__ Cube mc_cube(const void* vol, int x, int y, int z, int Nx, int Ny, int Nz) {
    Cube c;
    c.nverts = 0;
    c.cubeindex = 0;
    if (x >= Nx - 1 || y >= Ny - 1 || z >= Nz - 1) return c;
    const size_t id0 = ((size_t)y * Nx + x) * Nz + z;
    const size_t id1 = id0 + Nz;
    const size_t id3 = id0 + (size_t)Nx * Nz;
    const size_t id2 = id3 + Nz;
    c.id[0] = id0; c.id[1] = id1; c.id[2] = id2; c.id[3] = id3;
    c.id[4] = id0 + 1; c.id[5] = id1 + 1; c.id[6] = id2 + 1; c.id[7] = id3 + 1;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const McVoxel v = mc_load<STORAGE>(vol, c.id[k]);
        if (!v.valid) return c;
        c.f[k] = v.tsdf;
    }
    int ci = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) ci += (c.f[k] < 0 ? 1 : 0) << k;
    c.cubeindex = ci;
    c.nverts = kMc.nverts[ci];
    return c;
}

__device__ __forceinline__ unsigned tile_id() {
    return (blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
}

template <int STORAGE>
__global__ void __launch_bounds__(256) mc_count_kernel(const void* __restrict__ vol, TsdfParams p,
                                                       uint32_t* __restrict__ tile_counts) {
    __shared__ uint32_t lds[256 / 64 + 1];
    const int Nx = (int)p.resolution[0], Ny = (int)p.resolution[1], Nz = (int)p.resolution[2];
    const int x = blockIdx.x * 4 + (threadIdx.x >> 6), y = blockIdx.y, z = blockIdx.z * 64 + (threadIdx.x & 63);
    const Cube c = mc_cube<STORAGE>(vol, x, y, z, Nx, Ny, Nz);
    uint32_t total;
    (void)block_exclusive_scan<256>((uint32_t)c.nverts, lds, &total);
    if (threadIdx.x == 0) tile_counts[tile_id()] = total;
}

// K/marchingCube.cl:81-98 getNodeCoo (y flipped), vertex_interp, color_interp
__device__ __forceinline__ f3 node_coo(int x, int y, int z, const TsdfParams& p, float hx, float hy) {
    return mk3(((float)x + 0.5f - hx) * p.cell_size[0], (-(float)y - 0.5f + hy) * p.cell_size[1],
               ((float)z + 0.5f) * p.cell_size[2]);
}

template <int STORAGE>
__global__ void __launch_bounds__(256)
    mc_emit_kernel(const void* __restrict__ vol, const uchar4* __restrict__ cvol, TsdfParams p,
                   const uint32_t* __restrict__ tile_offsets, float* __restrict__ verts, float* __restrict__ colors,
                   uint32_t capacity) {
    __shared__ uint32_t lds[256 / 64 + 1];
    const int Nx = (int)p.resolution[0], Ny = (int)p.resolution[1], Nz = (int)p.resolution[2];
    const int x = blockIdx.x * 4 + (threadIdx.x >> 6), y = blockIdx.y, z = blockIdx.z * 64 + (threadIdx.x & 63);
    const Cube c = mc_cube<STORAGE>(vol, x, y, z, Nx, Ny, Nz);
    uint32_t total;
    const uint32_t local = block_exclusive_scan<256>((uint32_t)c.nverts, lds, &total);
    if (c.nverts == 0) return;
    const uint32_t base = tile_offsets[tile_id()] + local;
    const float hx = (float)(Nx >> 1), hy = (float)(Ny >> 1);
    const int co[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
    for (int k = 0; k < c.nverts; ++k) {
        const uint32_t o = base + k;
        if (o >= capacity) break;
        const int e = kMc.tri[c.cubeindex][k];
        const int a = kMc.edge[e][0], b = kMc.edge[e][1];
        const f3 p0 = node_coo(x + co[a][0], y + co[a][1], z + co[a][2], p, hx, hy);
        const f3 p1 = node_coo(x + co[b][0], y + co[b][1], z + co[b][2], p, hx, hy);
        const float f0 = c.f[a], f1 = c.f[b];
        float t = (0.f - f0) / (f1 - f0 + 1e-15f);
        t = fminf(fmaxf(t, 0.0f), 1.0f);
        st3(verts, o, add3(p0, mul3(sub3(p1, p0), t)));
        if (colors) {
            float4 col = make_float4(0.f, 0.f, 0.f, 1.f);
            if (cvol) {
                const uchar4 cc = fabsf(f0) < fabsf(f1) ? cvol[c.id[a]] : cvol[c.id[b]];
                col = make_float4((float)cc.z / 255.f, (float)cc.y / 255.f, (float)cc.x / 255.f, 1.f);
            }
            reinterpret_cast<float4*>(colors)[o] = col;
        }
    }
}

struct McGrid {
    dim3 grid;
    size_t tiles;
};

static McGrid mc_grid(const TsdfParams& p) {
    const int Nx = (int)p.resolution[0], Ny = (int)p.resolution[1], Nz = (int)p.resolution[2];
    McGrid g;
    g.grid = dim3((unsigned)div_up(std::max(Nx - 1, 1), 4), (unsigned)std::max(Ny - 1, 1),
                  (unsigned)div_up(std::max(Nz - 1, 1), 64));
    g.tiles = (size_t)g.grid.x * g.grid.y * g.grid.z;
    return g;
}

}  // namespace nuif

using namespace nuif;

extern "C" {

size_t nuif_mc_scratch_bytes(const TsdfParams* p) {
    if (!p) return 0;
    const McGrid g = mc_grid(*p);
    return (g.tiles + 64) * sizeof(uint32_t) + nuif_scan_scratch_bytes(g.tiles);
}

int nuif_mc_dense(const void* volume, const uint8_t* color_volume, const TsdfParams* p, int storage, float* vertices,
                  float* colors, uint32_t capacity, uint32_t* total, void* scratch, nuif_stream_t s) {
    NUIF_REQUIRE_INIT();
    NUIF_CHECK_ARG(volume && p && total && scratch, "nuif_mc_dense: null argument");
    NUIF_CHECK_ARG(capacity == 0 || vertices, "nuif_mc_dense: vertex buffer missing");
    for (int k = 0; k < 3; ++k) NUIF_CHECK_ARG(p->resolution[k] >= 2.f, "nuif_mc_dense: resolution below 2");
    const McGrid g = mc_grid(*p);
    uint32_t* counts = static_cast<uint32_t*>(scratch);
    void* scan_scratch = counts + g.tiles + 64;
    hipStream_t st = stream_of(s);
    const uchar4* cv = reinterpret_cast<const uchar4*>(color_volume);
    switch (storage) {
        case NUIF_TSDF_S16: mc_count_kernel<NUIF_TSDF_S16><<<g.grid, 256, 0, st>>>(volume, *p, counts); break;
        case NUIF_TSDF_F16: mc_count_kernel<NUIF_TSDF_F16><<<g.grid, 256, 0, st>>>(volume, *p, counts); break;
        case NUIF_TSDF_F32: mc_count_kernel<NUIF_TSDF_F32><<<g.grid, 256, 0, st>>>(volume, *p, counts); break;
        default: set_error("nuif_mc_dense: unknown storage %d", storage); return NUIF_ERR_INVALID;
    }
    NUIF_LAUNCHED("mc_count_kernel");
    int rc = launch_scan(counts, false, counts, g.tiles, total, scan_scratch, st);
    if (rc) return rc;
    switch (storage) {
        case NUIF_TSDF_S16:
            mc_emit_kernel<NUIF_TSDF_S16><<<g.grid, 256, 0, st>>>(volume, cv, *p, counts, vertices, colors, capacity);
            break;
        case NUIF_TSDF_F16:
            mc_emit_kernel<NUIF_TSDF_F16><<<g.grid, 256, 0, st>>>(volume, cv, *p, counts, vertices, colors, capacity);
            break;
        default:
            mc_emit_kernel<NUIF_TSDF_F32><<<g.grid, 256, 0, st>>>(volume, cv, *p, counts, vertices, colors, capacity);
            break;
    }
    NUIF_LAUNCHED("mc_emit_kernel");
    return NUIF_OK;
}

}  // extern "C"
