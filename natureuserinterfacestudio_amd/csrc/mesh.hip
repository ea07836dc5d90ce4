// Indexed mesh of a dense volume: OCL/NuiKinfuOpenCLScene.cpp:788-984
// Volume2Mesh, which reads the whole volume back and runs marching cubes on the
// host, keying each vertex by the ordered voxel pair of its cube edge (an edge
// walked in opposite directions by two cubes gives two vertices) and numbering
// vertices in first-use order of its (y, x, z) cube traversal.
//
// Here the same mesh is built on the device, in the same order:
//   count   one wavefront per (y, x) cube column, 64 z per step: vertex count
//   scan    column offsets in (y, x) order
//   emit    triangle corners in traversal order: 64-bit edge key + position
//   dedup   stable radix sort of (key, corner index) (scan.hip); runs of equal
//           keys are numbered by a scan of their heads, and each corner learns the
//           first corner of its run; first uses are numbered by a scan in corner
//           order, which is the reference's first-use order
// Points and indices equal the host mesher's bit for bit.

#include "mc_tables.h"
#include "nuif_internal.h"

#include <hip/hip_fp16.h>

namespace nuif {
namespace {

__constant__ McTable kMcMesh = mc_decode();

template <int STORAGE>
__device__ __forceinline__ bool voxel_at(const void* vol, size_t i, float* tsdf) {
    if (STORAGE == NUIF_TSDF_S16) {
        const uint32_t w = static_cast<const uint32_t*>(vol)[i];
        *tsdf = (float)(int16_t)(w & 0xffffu) / 32766.0f;
        return (int16_t)(w >> 16) > 0;
    } else if (STORAGE == NUIF_TSDF_F16) {
        const uint32_t w = static_cast<const uint32_t*>(vol)[i];
        *tsdf = __half2float(__ushort_as_half((unsigned short)(w & 0xffffu)));
        return (w >> 16) > 0;
    } else {
        const uint2 w = static_cast<const uint2*>(vol)[i];
        *tsdf = __uint_as_float(w.x);
        return __uint_as_float(w.y) > 0.f;
    }
}

struct MeshCube {
    uint32_t id[8];
    float f[8];
    int ci, nv;
};

template <int STORAGE>
__device__ __forceinline__ MeshCube mesh_cube(const void* vol, int x, int y, int z, int Nx, int Nz) {
    MeshCube c;
    c.nv = 0;
    c.ci = 0;
    const uint32_t id0 = ((uint32_t)y * Nx + x) * Nz + z;
    const uint32_t id1 = id0 + Nz, id3 = id0 + (uint32_t)Nx * Nz, id2 = id3 + Nz;
    const uint32_t ids[8] = {id0, id1, id2, id3, id0 + 1, id1 + 1, id2 + 1, id3 + 1};
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        c.id[k] = ids[k];
        if (!voxel_at<STORAGE>(vol, ids[k], &c.f[k])) return c;
    }
    int ci = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) ci += (c.f[k] < 0 ? 1 : 0) << k;
    c.ci = ci;
    c.nv = kMcMesh.nverts[ci];
    return c;
}

__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// one wave per (y, x) column of cubes; 4 columns per 256-thread block
template <int STORAGE>
__global__ void __launch_bounds__(256)
    mesh_count_kernel(const void* __restrict__ vol, int Nx, int Ny, int Nz, uint32_t* __restrict__ counts) {
    const int col = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int ncols = (Nx - 1) * (Ny - 1);
    if (col >= ncols) return;  // whole wave
    const int y = col / (Nx - 1), x = col % (Nx - 1);
    uint32_t n = 0;
    for (int z0 = 0; z0 < Nz - 1; z0 += 64) {
        const int z = z0 + lane;
        if (z < Nz - 1) n += (uint32_t)mesh_cube<STORAGE>(vol, x, y, z, Nx, Nz).nv;
    }
    n = wave_sum_u32(n);
    if (lane == 0) counts[col] = n;
}

// OCL/NuiKinfuOpenCLScene.cpp:84-91 getNodeCoo (y flipped, all axes centred)
__device__ __forceinline__ f3 node_cpu(int x, int y, int z, const TsdfParams& p) {
    return mk3(((float)x + 0.5f - p.resolution[0] / 2.0f) * p.cell_size[0],
               ((float)(-y) - 0.5f + p.resolution[1] / 2.0f) * p.cell_size[1],
               ((float)z + 0.5f - p.resolution[2] / 2.0f) * p.cell_size[2]);
}

template <int STORAGE>
__global__ void __launch_bounds__(256)
    mesh_emit_kernel(const void* __restrict__ vol, TsdfParams p, const uint32_t* __restrict__ offsets,
                     unsigned long long* __restrict__ keys, float* __restrict__ pos) {
    const int Nx = (int)p.resolution[0], Ny = (int)p.resolution[1], Nz = (int)p.resolution[2];
    const int col = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (col >= (Nx - 1) * (Ny - 1)) return;
    const int y = col / (Nx - 1), x = col % (Nx - 1);
    const int co[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
    uint32_t base = offsets[col];
    for (int z0 = 0; z0 < Nz - 1; z0 += 64) {
        const int z = z0 + lane;
        MeshCube c;
        c.nv = 0;
        if (z < Nz - 1) c = mesh_cube<STORAGE>(vol, x, y, z, Nx, Nz);
        const uint32_t nv = (uint32_t)c.nv;
        uint32_t incl = nv;
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t t = __shfl_up(incl, off, 64);
            if (lane >= off) incl += t;
        }
        const uint32_t mine = base + incl - nv;
        base += __shfl(incl, 63, 64);
        for (uint32_t k = 0; k < nv; ++k) {
            const int e = kMcMesh.tri[c.ci][k];
            const int a = kMcMesh.edge[e][0], b = kMcMesh.edge[e][1];
            keys[mine + k] = ((unsigned long long)c.id[a] << 32) + (unsigned long long)c.id[b];
            // OCL/NuiKinfuOpenCLScene.cpp:776-786 vertex_interp
            const f3 p0 = node_cpu(x + co[a][0], y + co[a][1], z + co[a][2], p);
            const f3 p1 = node_cpu(x + co[b][0], y + co[b][1], z + co[b][2], p);
            const float f0 = c.f[a], f1 = c.f[b];
            float t = (0.f - f0) / (f1 - f0 + 1e-15f);
            t = fmaxf(fminf(t, 1.0f), 0.0f);
            st3(pos, mine + k, add3(p0, mul3(sub3(p1, p0), t)));
        }
    }
}

__global__ void mesh_iota_kernel(uint32_t* __restrict__ v, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = i;
}

// head[j] = 1 at the first position of each run of equal sorted keys
__global__ void mesh_heads_kernel(const unsigned long long* __restrict__ sk, uint32_t* __restrict__ head, uint32_t n) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) head[j] = (j == 0 || sk[j] != sk[j - 1]) ? 1u : 0u;
}

// run r = excl[j] (heads before j) starts at a head; the sort is stable, so the run's
// first element is the key's first corner
__global__ void mesh_runfirst_kernel(const uint32_t* __restrict__ sv, const uint32_t* __restrict__ head,
                                     const uint32_t* __restrict__ excl, uint32_t* __restrict__ runfirst, uint32_t n) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n && head[j]) runfirst[excl[j]] = sv[j];
}

__global__ void mesh_first_kernel(const uint32_t* __restrict__ sv, const uint32_t* __restrict__ head,
                                  const uint32_t* __restrict__ excl, const uint32_t* __restrict__ runfirst,
                                  uint32_t* __restrict__ first, uint32_t n) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) first[sv[j]] = runfirst[excl[j] + head[j] - 1u];
}

__global__ void mesh_isfirst_kernel(const uint32_t* __restrict__ first, uint32_t* __restrict__ flag, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) flag[i] = first[i] == i ? 1u : 0u;
}

__global__ void mesh_output_kernel(const uint32_t* __restrict__ first, const uint32_t* __restrict__ vid,
                                   const float* __restrict__ pos, float* __restrict__ points,
                                   uint32_t* __restrict__ indices, uint32_t n, uint32_t* __restrict__ npoints) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t f = first[i];
    indices[i] = vid[f];
    if (f == i) st3(points, vid[i], ld3(pos, i));
    if (i == n - 1) *npoints = vid[i] + (f == i ? 1u : 0u);
}

struct MeshScratch {
    uint32_t* counts;
    void* scan;
    unsigned long long *keys, *keys2;
    uint32_t *vals, *vals2, *head, *excl, *runfirst, *first, *flag, *vid, *npoints;
    float* pos;
    void* sort;
    void* scan2;
    size_t total;
};

size_t a256(size_t v) { return (v + 255) / 256 * 256; }

MeshScratch mesh_scratch(void* base, const TsdfParams& p, uint32_t max_vertices) {
    const size_t ncols = (size_t)((int)p.resolution[0] - 1) * ((int)p.resolution[1] - 1);
    const size_t m = std::max<size_t>(max_vertices, 1);
    MeshScratch s{};
    char* b = static_cast<char*>(base);
    size_t o = 0;
    auto take = [&](size_t bytes) {
        char* r = b ? b + o : nullptr;
        o = a256(o + bytes);
        return r;
    };
    s.counts = (uint32_t*)take(ncols * 4);
    s.scan = take(nuif_scan_scratch_bytes(ncols));
    s.keys = (unsigned long long*)take(m * 8);
    s.keys2 = (unsigned long long*)take(m * 8);
    s.vals = (uint32_t*)take(m * 4);
    s.vals2 = (uint32_t*)take(m * 4);
    s.head = (uint32_t*)take(m * 4);
    s.excl = (uint32_t*)take(m * 4);
    s.runfirst = (uint32_t*)take(m * 4);
    s.first = (uint32_t*)take(m * 4);
    s.flag = (uint32_t*)take(m * 4);
    s.vid = (uint32_t*)take(m * 4);
    s.npoints = (uint32_t*)take(256);
    s.pos = (float*)take(m * 12);
    s.sort = take(sort_scratch_bytes((uint32_t)m));
    s.scan2 = take(nuif_scan_scratch_bytes(m));
    s.total = o;
    return s;
}

}  // namespace
}  // namespace nuif

using namespace nuif;

extern "C" {

size_t nuif_mesh_dense_scratch_bytes(const TsdfParams* p, uint32_t max_vertices) {
    if (!p || p->resolution[0] < 2.f || p->resolution[1] < 2.f || p->resolution[2] < 2.f) return 0;
    return mesh_scratch(nullptr, *p, max_vertices).total;
}

int nuif_mesh_dense(const void* volume, const TsdfParams* p, int storage, uint32_t max_vertices, float* points,
                    uint32_t* indices, uint32_t* n_points, uint32_t* n_indices, void* scratch, nuif_stream_t s) {
    NUIF_REQUIRE_INIT();
    NUIF_CHECK_ARG(volume && p && points && indices && n_points && n_indices && scratch,
                   "nuif_mesh_dense: null argument");
    for (int k = 0; k < 3; ++k) NUIF_CHECK_ARG(p->resolution[k] >= 2.f, "nuif_mesh_dense: resolution below 2");
    NUIF_CHECK_ARG((double)p->resolution[0] * p->resolution[1] * p->resolution[2] < 4294967296.0,
                   "nuif_mesh_dense: more than 2^32 voxels");
    NUIF_CHECK_ARG(storage == NUIF_TSDF_S16 || storage == NUIF_TSDF_F16 || storage == NUIF_TSDF_F32,
                   "nuif_mesh_dense: unknown storage %d", storage);
    const int Nx = (int)p->resolution[0], Ny = (int)p->resolution[1], Nz = (int)p->resolution[2];
    const size_t ncols = (size_t)(Nx - 1) * (Ny - 1);
    MeshScratch sc = mesh_scratch(scratch, *p, max_vertices);
    hipStream_t st = stream_of(s);
    const unsigned cgrid = div_up(ncols, 4);
    switch (storage) {
        case NUIF_TSDF_S16: mesh_count_kernel<NUIF_TSDF_S16><<<cgrid, 256, 0, st>>>(volume, Nx, Ny, Nz, sc.counts); break;
        case NUIF_TSDF_F16: mesh_count_kernel<NUIF_TSDF_F16><<<cgrid, 256, 0, st>>>(volume, Nx, Ny, Nz, sc.counts); break;
        default: mesh_count_kernel<NUIF_TSDF_F32><<<cgrid, 256, 0, st>>>(volume, Nx, Ny, Nz, sc.counts); break;
    }
    NUIF_LAUNCHED("mesh_count_kernel");
    int rc = launch_scan(sc.counts, false, sc.counts, ncols, sc.npoints, sc.scan, st);
    if (rc) return rc;
    uint32_t n = 0;
    NUIF_HIP(hipMemcpyAsync(&n, sc.npoints, 4, hipMemcpyDeviceToHost, st));
    NUIF_HIP(hipStreamSynchronize(st));
    *n_indices = n;
    *n_points = 0;
    NUIF_CHECK_ARG(n <= max_vertices, "nuif_mesh_dense: %u triangle corners exceed the capacity %u", n, max_vertices);
    if (n == 0) return NUIF_OK;
    switch (storage) {
        case NUIF_TSDF_S16: mesh_emit_kernel<NUIF_TSDF_S16><<<cgrid, 256, 0, st>>>(volume, *p, sc.counts, sc.keys, sc.pos); break;
        case NUIF_TSDF_F16: mesh_emit_kernel<NUIF_TSDF_F16><<<cgrid, 256, 0, st>>>(volume, *p, sc.counts, sc.keys, sc.pos); break;
        default: mesh_emit_kernel<NUIF_TSDF_F32><<<cgrid, 256, 0, st>>>(volume, *p, sc.counts, sc.keys, sc.pos); break;
    }
    NUIF_LAUNCHED("mesh_emit_kernel");
    const unsigned g = div_up(n, 256);
    mesh_iota_kernel<<<g, 256, 0, st>>>(sc.vals, n);
    NUIF_LAUNCHED("mesh_iota_kernel");
    // keys are (voxel a) << 32 | (voxel b) with voxel ids below Nx * Ny * Nz
    int vbits = 1;
    while (vbits < 32 && ((size_t)1 << vbits) < (size_t)Nx * Ny * Nz) ++vbits;
    bool alt = false;
    rc = launch_sort_pairs_u64(sc.keys, sc.keys2, sc.vals, sc.vals2, n, 32 + vbits, sc.sort, st, &alt);
    if (rc) return rc;
    const unsigned long long* sk = alt ? sc.keys2 : sc.keys;
    const uint32_t* sv = alt ? sc.vals2 : sc.vals;
    mesh_heads_kernel<<<g, 256, 0, st>>>(sk, sc.head, n);
    NUIF_LAUNCHED("mesh_heads_kernel");
    rc = launch_scan(sc.head, false, sc.excl, n, nullptr, sc.scan2, st);
    if (rc) return rc;
    mesh_runfirst_kernel<<<g, 256, 0, st>>>(sv, sc.head, sc.excl, sc.runfirst, n);
    NUIF_LAUNCHED("mesh_runfirst_kernel");
    mesh_first_kernel<<<g, 256, 0, st>>>(sv, sc.head, sc.excl, sc.runfirst, sc.first, n);
    NUIF_LAUNCHED("mesh_first_kernel");
    mesh_isfirst_kernel<<<g, 256, 0, st>>>(sc.first, sc.flag, n);
    NUIF_LAUNCHED("mesh_isfirst_kernel");
    rc = launch_scan(sc.flag, false, sc.vid, n, nullptr, sc.scan2, st);
    if (rc) return rc;
    mesh_output_kernel<<<g, 256, 0, st>>>(sc.first, sc.vid, sc.pos, points, indices, n, sc.npoints);
    NUIF_LAUNCHED("mesh_output_kernel");
    uint32_t np = 0;
    NUIF_HIP(hipMemcpyAsync(&np, sc.npoints, 4, hipMemcpyDeviceToHost, st));
    NUIF_HIP(hipStreamSynchronize(st));
    *n_points = np;
    return NUIF_OK;
}

}  // extern "C"
