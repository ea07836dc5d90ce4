// Voxel-hashing scene: block allocation, visible list, integration, ray ranges,
// raycast and marching cubes (OCL/NuiKinfuOpenCLHashScene.cpp with the kernels
// of K/KinfuHashScene.cl, K/KinfuHashRaycast.cl and K/KinfuHashVisualization.cl).
//
// Table: 2^b buckets of 4 entries {64-bit block key, brick index}; a full bucket
// overflows into the next buckets (bucketed linear probing).  Insertion is one
// 64-bit compare-and-swap on the key, the winner takes the next brick of a
// bump allocator.  Every block that any pixel's truncation band touches is
// allocated in that frame (the reference allocates at most one block per
// bucket per frame and its winner is racy), and the visible list is every
// flagged entry, compacted in entry order by the device-wide scan (the
// reference scans only the first 2^20 entries).
//
// Bricks: 8^3 voxels, voxel index x + 8y + 64z, one 32-bit word per voxel --
// int16 sdf * 32767 (round toward zero) with the weight in the high half (the
// reference's NuiKinfuVoxel values), or FP16 sdf with a uint16 weight -- and an
// optional uchar4 colour per voxel.  The pool starts at sdf 0, weight 0.
//
// Kernels walk the visible list with a fixed grid and read its length from the
// scene header, so a frame needs no host round trip and can be captured in a
// graph.  One wave owns one brick: 64 lanes cover an 8x8 (x, y) slice and loop
// over z, so every brick access is a coalesced 256-byte row.
//
// Float arithmetic follows the reference kernels operation by operation
// (compiled without contraction).  Out-of-range float -> int conversions
// saturate and NaN converts to 0, as the GPU's conversion instructions do.

#include <hip/hip_fp16.h>

#include "mc_tables.h"
#include "nuif_internal.h"

namespace nuif {
namespace {

constexpr int kSB = 8;        // SDF_BLOCK_SIZE (K/hashing_gpu_def.h:9)
constexpr int kSB3 = 512;     // SDF_BLOCK_SIZE3
constexpr int kSlots = 4;     // entries per bucket
constexpr int kMaxProbe = 64; // buckets an insertion or lookup visits at most
constexpr int kWalkGrid = 2048;

struct alignas(16) HashEntry {
    unsigned long long key;  // 0: empty; else bit 48 | x:16 | y:16 | z:16
    int ptr;                 // brick index, -1 when the pool was exhausted
    uint32_t pad;
};

struct HashHeader {
    uint32_t nbricks;   // bricks in use (depth of the free-brick stack)
    uint32_t nvisible;  // length of the visible list
    uint32_t failed;    // touched blocks without an entry or a brick
    uint32_t nentries;  // occupied entries
    uint32_t pad[60];
};
static_assert(sizeof(HashHeader) == 256, "header is 256 bytes");

__constant__ McTable kMcHash = mc_decode();

struct Layout {
    size_t header, entries, flags, offsets, scan, visible, bricks, colors, swap, freelist, total;
    size_t nentries;
};

size_t align256(size_t v) { return (v + 255) / 256 * 256; }

Layout layout_of(const nuif_hash_params& p) {
    Layout L{};
    L.nentries = (size_t)kSlots << p.buckets_log2;
    size_t o = 0;
    L.header = o;
    o += 256;
    L.entries = o;
    o = align256(o + L.nentries * sizeof(HashEntry));
    L.flags = o;
    o = align256(o + L.nentries);
    L.offsets = o;
    o = align256(o + L.nentries * 4);
    L.scan = o;
    o = align256(o + nuif_scan_scratch_bytes(L.nentries));
    L.visible = o;
    o = align256(o + L.nentries * 4);
    L.bricks = o;
    o = align256(o + (size_t)p.bricks * kSB3 * 4);
    L.colors = o;
    if (p.color) o = align256(o + (size_t)p.bricks * kSB3 * 4);
    L.swap = o;  // swap state per entry (host swapping)
    o = align256(o + L.nentries);
    L.freelist = o;  // free-brick stack: entries [nbricks, bricks) are free
    o = align256(o + (size_t)p.bricks * 4);
    L.total = o;
    return L;
}

struct Scene {
    HashHeader* hdr;
    HashEntry* entries;
    uint8_t* flags;
    uint32_t* offsets;
    void* scan;
    int32_t* visible;
    uint32_t* bricks;
    uchar4* colors;
    uint8_t* swap;       // 0: nothing to do, 1: wants its stored data, 2: synchronised
    int32_t* freelist;   // brick ids; [nbricks, capacity) are free
    uint32_t mask;  // bucket mask
};

Scene scene_of(void* base, const nuif_hash_params& p) {
    const Layout L = layout_of(p);
    char* b = static_cast<char*>(base);
    Scene s;
    s.hdr = reinterpret_cast<HashHeader*>(b + L.header);
    s.entries = reinterpret_cast<HashEntry*>(b + L.entries);
    s.flags = reinterpret_cast<uint8_t*>(b + L.flags);
    s.offsets = reinterpret_cast<uint32_t*>(b + L.offsets);
    s.scan = b + L.scan;
    s.visible = reinterpret_cast<int32_t*>(b + L.visible);
    s.bricks = reinterpret_cast<uint32_t*>(b + L.bricks);
    s.colors = p.color ? reinterpret_cast<uchar4*>(b + L.colors) : nullptr;
    s.swap = reinterpret_cast<uint8_t*>(b + L.swap);
    s.freelist = reinterpret_cast<int32_t*>(b + L.freelist);
    s.mask = (1u << p.buckets_log2) - 1u;
    return s;
}

// ---- voxel words -------------------------------------------------------------

struct VS16 {  // the reference's short sdf (x32767, round toward zero) + weight
    __device__ static float sdf(uint32_t w) { return (float)(int16_t)(w & 0xffffu) / 32767.0f; }
    __device__ static int weight(uint32_t w) { return (int)(w >> 16); }
    // K/KinfuHashVisualization.cl:25 abs(sdf) < 32767 * truncScale on the short sdf
    __device__ static bool in_band(uint32_t w, float ts) { return (float)abs((int)(int16_t)(w & 0xffffu)) < 32767 * ts; }
};
struct VF16 {  // FP16 sdf (round to nearest even) + uint16 weight
    __device__ static float sdf(uint32_t w) { return __half2float(__ushort_as_half((unsigned short)(w & 0xffffu))); }
    __device__ static int weight(uint32_t w) { return (int)(w >> 16); }
    __device__ static bool in_band(uint32_t w, float ts) { return fabsf(sdf(w)) < ts; }
};

__device__ __forceinline__ int cvt_sat(float v) {
    if (v != v) return 0;
    if (v >= 2147483648.0f) return INT32_MAX;
    if (v <= -2147483648.0f) return INT32_MIN;
    return (int)v;
}

__device__ __forceinline__ uint32_t pack_s16(float f, int wt) {
    return (uint32_t)(uint16_t)(int16_t)cvt_sat(f * 32767.0f) | ((uint32_t)wt << 16);
}
__device__ __forceinline__ uint32_t pack_f16(float f, int wt) {
    return (uint32_t)__half_as_ushort(__float2half_rn(f)) | ((uint32_t)wt << 16);
}

// ---- keys, buckets, lookups ----------------------------------------------------

__device__ __forceinline__ unsigned long long block_key(int x, int y, int z) {
    return (1ull << 48) | ((unsigned long long)(uint16_t)(int16_t)x << 32) |
           ((unsigned long long)(uint16_t)(int16_t)y << 16) | (unsigned long long)(uint16_t)(int16_t)z;
}
__device__ __forceinline__ void key_pos(unsigned long long k, int* p) {
    p[0] = (int16_t)(uint16_t)(k >> 32);
    p[1] = (int16_t)(uint16_t)(k >> 16);
    p[2] = (int16_t)(uint16_t)k;
}
// K/KinfuHashUtils.cl:9-16 computeHashIndex, masked to the table's bucket count
__device__ __forceinline__ uint32_t bucket_of(int x, int y, int z, uint32_t mask) {
    return (((uint32_t)(int16_t)x * 73856093u) ^ ((uint32_t)(int16_t)y * 19349669u) ^
            ((uint32_t)(int16_t)z * 83492791u)) & mask;
}

// A brick from the free-brick stack, -1 when the pool is exhausted (the count is
// restored, so it never exceeds the capacity).
__device__ __forceinline__ int pop_brick(const Scene& s, uint32_t capacity) {
    const uint32_t k = atomicAdd(&s.hdr->nbricks, 1u);
    if (k < capacity) return s.freelist[k];
    atomicSub(&s.hdr->nbricks, 1u);
    return -1;
}

// Entry of a block, inserting it (and taking a brick) when absent; -1 when the
// probe limit is reached.
__device__ int find_or_insert(const Scene& s, int bx, int by, int bz, uint32_t capacity) {
    const unsigned long long key = block_key(bx, by, bz);
    uint32_t b = bucket_of(bx, by, bz, s.mask);
    for (int probe = 0; probe < kMaxProbe; ++probe, b = (b + 1) & s.mask) {
        for (int k = 0; k < kSlots; ++k) {
            const int id = (int)(b * kSlots + k);
            HashEntry* e = s.entries + id;
            // keys never change once set: a stale read can only show an empty
            // slot, and the compare-and-swap below then reports the real key
            const unsigned long long cur = __atomic_load_n(&e->key, __ATOMIC_RELAXED);
            if (cur == key) return id;
            if (cur != 0ull) continue;
            const unsigned long long old = atomicCAS(&e->key, 0ull, key);
            if (old == 0ull) {
                const int v = pop_brick(s, capacity);
                if (v < 0) atomicAdd(&s.hdr->failed, 1u);
                e->ptr = v;
                atomicAdd(&s.hdr->nentries, 1u);
                return id;
            }
            if (old == key) return id;
        }
    }
    atomicAdd(&s.hdr->failed, 1u);
    return -1;
}

// Brick index of a block, -1 when absent or without a brick.
__device__ __forceinline__ int lookup_ptr(const Scene& s, int bx, int by, int bz) {
    const unsigned long long key = block_key(bx, by, bz);
    uint32_t b = bucket_of(bx, by, bz, s.mask);
    for (int probe = 0; probe < kMaxProbe; ++probe, b = (b + 1) & s.mask) {
        const HashEntry* e = s.entries + (size_t)b * kSlots;
#pragma unroll
        for (int k = 0; k < kSlots; ++k) {
            const unsigned long long cur = e[k].key;
            if (cur == key) return e[k].ptr;
            if (cur == 0ull) return -1;
        }
    }
    return -1;
}

// One-block cache of the raycast and the mesher (K/KinfuHashRaycast.cl:78-111)
struct BrickCache {
    int bx = INT32_MIN, by = 0, bz = 0, ptr = -1;
};

template <class V>
__device__ __forceinline__ bool read_voxel(const Scene& s, BrickCache& c, int x, int y, int z, float* sdf,
                                           uint32_t* vidx = nullptr) {
    const int bx = x >> 3, by = y >> 3, bz = z >> 3;  // floor division by 8
    if (bx != c.bx || by != c.by || bz != c.bz) {
        c.bx = bx;
        c.by = by;
        c.bz = bz;
        c.ptr = lookup_ptr(s, bx, by, bz);
    }
    if (c.ptr < 0) {
        *sdf = 1.0f;
        return false;
    }
    const uint32_t i = (uint32_t)c.ptr * kSB3 + (uint32_t)((x - bx * 8) + (y - by * 8) * 8 + (z - bz * 8) * 64);
    *sdf = V::sdf(s.bricks[i]);
    if (vidx) *vidx = i;
    return true;
}

// ---- camera helpers (K/cameraUtils.cl) -----------------------------------------

__device__ __forceinline__ f3 xform(const NuiCLRigidTransform& g, f3 v) {
    return add3(rot3(g.R, v), mk3(g.t[0], g.t[1], g.t[2]));
}
__device__ __forceinline__ f3 xform_inv(const NuiCLRigidTransform& g, f3 v) {
    return rot3(g.R_inv, sub3(v, mk3(g.t[0], g.t[1], g.t[2])));
}
__device__ __forceinline__ f3 depth_to_skeleton(const NuiCLCameraParams& cam, int ux, int uy, float d) {
    const float x = ((float)ux - cam.cx) / cam.fx;
    const float y = ((float)uy - cam.cy) / cam.fy;
    return mk3(d * x, d * y, d);
}
__device__ __forceinline__ float flen(f3 v) { return sqrtf(v.x * v.x + v.y * v.y + v.z * v.z); }

// ---- allocation and visible list (K/KinfuHashScene.cl:52-285) ------------------

__global__ void __launch_bounds__(256)
    hash_touch_kernel(Scene s, const float* __restrict__ depth, NuiCLCameraParams cam,
                      const NuiCLRigidTransform* __restrict__ pose, float trunc_scale, float inv_block,
                      uint32_t capacity, const int32_t* __restrict__ gate) {
    if (gate && *gate == 0) return;
    const int W = (int)cam.depthImageWidth, H = (int)cam.depthImageHeight;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= W * H) return;
    const float Dp = depth[i];
    if (isnan(Dp)) return;
    const NuiCLRigidTransform g = *pose;
    const f3 pc = depth_to_skeleton(cam, i % W, i / W, Dp);
    const float norm = trunc_scale / flen(pc);
    f3 rmin = mul3(xform(g, mul3(pc, 1.0f - norm)), inv_block);
    const f3 rmax = mul3(xform(g, mul3(pc, 1.0f + norm)), inv_block);
    f3 dir = sub3(rmax, rmin);
    const float rn = flen(dir);
    const int steps = cvt_sat(ceilf(2.0f * rn));
    const float div = (float)(steps - 1);
    dir = mk3(dir.x / div, dir.y / div, dir.z / div);
    for (int k = 0; k < steps; ++k) {
        const int bx = (int16_t)cvt_sat(floorf(rmin.x));
        const int by = (int16_t)cvt_sat(floorf(rmin.y));
        const int bz = (int16_t)cvt_sat(floorf(rmin.z));
        const int id = find_or_insert(s, bx, by, bz, capacity);
        if (id >= 0) s.flags[id] = 1;
        rmin = add3(rmin, dir);
    }
}

// K/KinfuHashScene.cl:197-212 checkPointVisibility
__device__ __forceinline__ bool point_visible(const NuiCLCameraParams& cam, const NuiCLRigidTransform& g, f3 p) {
    const f3 q = xform_inv(g, p);
    if (nan3(q)) return false;
    const float sx = q.x * cam.fx / q.z + cam.cx, sy = q.y * cam.fy / q.z + cam.cy;
    const int u = cvt_sat(sx + 0.5f), v = cvt_sat(sy + 0.5f);
    return u >= 0 && u < (int)cam.depthImageWidth && v >= 0 && v < (int)cam.depthImageHeight;
}

// K/KinfuHashScene.cl:214-260 checkBlockVisibility: the eight corners in the
// reference's order, reached by the same float additions
__device__ bool block_visible(const NuiCLCameraParams& cam, const NuiCLRigidTransform& g, const int* p, float vs) {
    const float f = (float)kSB * vs;
    f3 pt = mk3((float)p[0] * f, (float)p[1] * f, (float)p[2] * f);
    if (point_visible(cam, g, pt)) return true;
    pt.z += f;
    if (point_visible(cam, g, pt)) return true;
    pt.y += f;
    if (point_visible(cam, g, pt)) return true;
    pt.x += f;
    if (point_visible(cam, g, pt)) return true;
    pt.z -= f;
    if (point_visible(cam, g, pt)) return true;
    pt.y -= f;
    if (point_visible(cam, g, pt)) return true;
    pt.x -= f;
    pt.y += f;
    if (point_visible(cam, g, pt)) return true;
    pt.x += f;
    pt.y -= f;
    pt.z += f;
    return point_visible(cam, g, pt);
}

// Entries of the previous visible list that no band touched this frame stay
// visible when a corner projects into the image (K/KinfuHashScene.cl:262-285).
__global__ void __launch_bounds__(256)
    hash_recheck_kernel(Scene s, NuiCLCameraParams cam, const NuiCLRigidTransform* __restrict__ pose, float vs,
                        const int32_t* __restrict__ gate) {
    if (gate && *gate == 0) return;
    const uint32_t n = s.hdr->nvisible;
    const NuiCLRigidTransform g = *pose;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int id = s.visible[i];
        if (s.flags[id]) continue;
        int p[3];
        key_pos(s.entries[id].key, p);
        if (block_visible(cam, g, p, vs)) s.flags[id] = 1;
    }
}

__global__ void hash_compact_kernel(Scene s, size_t n, const int32_t* __restrict__ gate) {
    if (gate && *gate == 0) return;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && s.flags[i]) s.visible[s.offsets[i]] = (int32_t)i;
}

// The scan writes the new list length into the header; when the gate is off the
// length must survive, so the scan goes to a side slot and this copies it.
__global__ void hash_commit_count_kernel(Scene s, const uint32_t* __restrict__ total,
                                         const int32_t* __restrict__ gate) {
    if (gate && *gate == 0) return;
    s.hdr->nvisible = *total;
}

// ---- integration (K/KinfuHashScene.cl:387-470) -----------------------------------

template <class V, bool F16, bool COLOR>
__global__ void __launch_bounds__(64)
    hash_integrate_kernel(Scene s, const float* __restrict__ depth, const uchar4* __restrict__ bgra,
                          NuiCLCameraParams cam, const NuiCLRigidTransform* __restrict__ pose, float vs,
                          float trunc_scale, int wmax, const int32_t* __restrict__ gate,
                          unsigned long long* __restrict__ updated) {
    if (gate && *gate == 0) return;
    const uint32_t n = s.hdr->nvisible;
    const NuiCLRigidTransform g = *pose;
    const int lane = threadIdx.x, lx = lane & 7, ly = lane >> 3;
    const int W = (int)cam.depthImageWidth, H = (int)cam.depthImageHeight;
    unsigned nupd = 0;
    for (uint32_t i = blockIdx.x; i < n; i += gridDim.x) {
        const HashEntry e = s.entries[s.visible[i]];
        if (e.ptr < 0) continue;
        int p[3];
        key_pos(e.key, p);
        for (int lz = 0; lz < kSB; ++lz) {
            const uint32_t vi = (uint32_t)e.ptr * kSB3 + (uint32_t)(lane + 64 * lz);
            const uint32_t word = s.bricks[vi];
            const int w0 = V::weight(word);
            if (w0 == wmax) continue;
            const f3 pm = mul3(mk3((float)(p[0] * kSB + lx), (float)(p[1] * kSB + ly), (float)(p[2] * kSB + lz)), vs);
            const f3 pc = xform_inv(g, pm);
            if (pc.z <= 0) continue;
            const float sx = pc.x * cam.fx / pc.z + cam.cx, sy = pc.y * cam.fy / pc.z + cam.cy;
            const int u = cvt_sat(sx + 0.5f), v = cvt_sat(sy + 0.5f);
            if (u < 0 || u >= W || v < 0 || v >= H) continue;
            const int sid = v * W + u;
            const float d = depth[sid];
            if (isnan(d)) continue;
            const float eta = d - pc.z;
            const float sdf = eta / trunc_scale;
            if (eta < -trunc_scale) continue;
            const float oldF = V::sdf(word);
            const float oldW = (float)w0;
            float newF = fminf(1.0f, sdf);
            const float newW = 1.0f;
            const float cw = oldW + newW;
            if (COLOR && eta < trunc_scale && fabsf(sdf) < 0.25f) {
                const uchar4 col = bgra[sid];
                if (col.w == 255) {
                    uchar4 c = s.colors[vi];
                    const float r = (oldW * (float)c.x + newW * (float)col.z) / cw;
                    const float gg = (oldW * (float)c.y + newW * (float)col.y) / cw;
                    const float b = (oldW * (float)c.z + newW * (float)col.x) / cw;
                    c.x = (unsigned char)cvt_sat(fminf(fmaxf(r, 0.0f), 255.0f));
                    c.y = (unsigned char)cvt_sat(fminf(fmaxf(gg, 0.0f), 255.0f));
                    c.z = (unsigned char)cvt_sat(fminf(fmaxf(b, 0.0f), 255.0f));
                    s.colors[vi] = c;
                }
            }
            newF = oldW * oldF + newW * newF;
            newF /= cw;
            const int wt = min(cvt_sat(cw), wmax);
            s.bricks[vi] = F16 ? pack_f16(newF, wt) : pack_s16(newF, wt);
            ++nupd;
        }
    }
    if (updated) {
        for (int off = 32; off > 0; off >>= 1) nupd += __shfl_xor(nupd, off);
        if (lane == 0 && nupd) atomicAdd(updated, (unsigned long long)nupd);
    }
}

// ---- ray ranges (K/KinfuHashRaycast.cl:4-76) -------------------------------------

// (999999.9, 0.05) everywhere (OCL/NuiKinfuOpenCLAcceleratedFeedbackFrame.cpp:45-49)
__global__ void hash_range_init_kernel(float2* __restrict__ range, int n, const int32_t* __restrict__ enable) {
    if (enable && *enable == 0) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) range[i] = make_float2(999999.9f, 0.05f);
}

// One wave per visible block: lanes 0-7 project the corners, the wave splats the
// bounding rectangle.  Near and far are positive, so their float bits order as
// integers and the splat is an exact atomic min / max.
__global__ void __launch_bounds__(64)
    hash_range_kernel(Scene s, float2* __restrict__ range, NuiCLCameraParams cam,
                      const NuiCLRigidTransform* __restrict__ pose, float vs, const int32_t* __restrict__ enable) {
    if (enable && *enable == 0) return;
    const uint32_t n = s.hdr->nvisible;
    const NuiCLRigidTransform g = *pose;
    const int lane = threadIdx.x;
    const int W = (int)cam.depthImageWidth, H = (int)cam.depthImageHeight;
    for (uint32_t i = blockIdx.x; i < n; i += gridDim.x) {
        const HashEntry e = s.entries[s.visible[i]];
        if (e.ptr < 0) continue;
        int p[3];
        key_pos(e.key, p);
        int fx = W, cxm = 0, fy = H, cym = 0;
        float z = 999999.9f;
        bool valid = false;
        if (lane < 8) {
            const int tx = p[0] + ((lane & 1) ? 1 : 0), ty = p[1] + ((lane & 2) ? 1 : 0), tz = p[2] + ((lane & 4) ? 1 : 0);
            const f3 wpt = mk3((float)tx * (float)kSB * vs, (float)ty * (float)kSB * vs, (float)tz * (float)kSB * vs);
            const f3 q = xform_inv(g, wpt);
            valid = !nan3(q) && !(q.z < 1e-6f);
            if (valid) {
                const float sx = q.x * cam.fx / q.z + cam.cx, sy = q.y * cam.fy / q.z + cam.cy;
                fx = cvt_sat(floorf(sx));
                cxm = cvt_sat(ceilf(sx));
                fy = cvt_sat(floorf(sy));
                cym = cvt_sat(ceilf(sy));
                z = q.z;
            }
        }
        // the far bound is the depth of the last valid corner (K/KinfuHashRaycast.cl:49)
        const unsigned long long vmask = __ballot(valid);
        float zfar = 0.05f;
        if (vmask) zfar = __shfl(z, 63 - __clzll(vmask), 64);
        for (int off = 4; off > 0; off >>= 1) {
            fx = min(fx, __shfl_xor(fx, off, 64));
            fy = min(fy, __shfl_xor(fy, off, 64));
            cxm = max(cxm, __shfl_xor(cxm, off, 64));
            cym = max(cym, __shfl_xor(cym, off, 64));
            z = fminf(z, __shfl_xor(z, off, 64));
        }
        // the reference starts the rectangle at upperLeft = (W, H), lowerRight = (0, 0)
        // (K/KinfuHashRaycast.cl:24-25): those start values take part in the min / max
        fx = min(__shfl(fx, 0, 64), W);
        fy = min(__shfl(fy, 0, 64), H);
        cxm = max(__shfl(cxm, 0, 64), 0);
        cym = max(__shfl(cym, 0, 64), 0);
        float znear = __shfl(z, 0, 64);
        if (fx < 0) fx = 0;
        if (fy < 0) fy = 0;
        if (cxm >= W) cxm = W - 1;
        if (cym >= H) cym = H - 1;
        if (znear < 0.05f) znear = 0.05f;
        if (cxm < fx || cym < fy) continue;
        const int rw = cxm - fx + 1;
        const int cnt = rw * (cym - fy + 1);
        for (int k = lane; k < cnt; k += 64) {
            float2* r = range + (size_t)(fy + k / rw) * W + (fx + k % rw);
            atomicMin(reinterpret_cast<int*>(&r->x), __float_as_int(znear));
            atomicMax(reinterpret_cast<int*>(&r->y), __float_as_int(zfar));
        }
    }
}

// ---- raycast (K/KinfuHashRaycast.cl:155-448 castRay) -----------------------------

template <class V>
__device__ float interp_sdf(const Scene& s, BrickCache& c, f3 p) {
    const int x = cvt_sat(floorf(p.x)), y = cvt_sat(floorf(p.y)), z = cvt_sat(floorf(p.z));
    const f3 co = mk3(p.x - (float)x, p.y - (float)y, p.z - (float)z);
    float v1, v2;
    read_voxel<V>(s, c, x, y, z, &v1);
    read_voxel<V>(s, c, x + 1, y, z, &v2);
    float res1 = (1.0f - co.x) * v1 + co.x * v2;
    read_voxel<V>(s, c, x, y + 1, z, &v1);
    read_voxel<V>(s, c, x + 1, y + 1, z, &v2);
    res1 = (1.0f - co.y) * res1 + co.y * ((1.0f - co.x) * v1 + co.x * v2);
    read_voxel<V>(s, c, x, y, z + 1, &v1);
    read_voxel<V>(s, c, x + 1, y, z + 1, &v2);
    float res2 = (1.0f - co.x) * v1 + co.x * v2;
    read_voxel<V>(s, c, x, y + 1, z + 1, &v1);
    read_voxel<V>(s, c, x + 1, y + 1, z + 1, &v2);
    res2 = (1.0f - co.y) * res2 + co.y * ((1.0f - co.x) * v1 + co.x * v2);
    return (1.0f - co.z) * res1 + co.z * res2;
}

template <class V>
__device__ __forceinline__ float rd(const Scene& s, BrickCache& c, int x, int y, int z) {
    float v;
    read_voxel<V>(s, c, x, y, z, &v);
    return v;
}

// castRay of pixel (x, y): vertex (metres, world), normal (nmap may be null) and the
// nearest corner's colour (cmap may be null)
template <class V>
__device__ void cast_ray(const Scene& s, const float2* __restrict__ range, const NuiCLCameraParams& cam,
                         const NuiCLRigidTransform& g, float vs, float truncation, int x, int y,
                         float* __restrict__ vmap, float* __restrict__ nmap, uchar4* __restrict__ cmap) {
    const size_t pid = (size_t)y * cam.depthImageWidth + x;
    const float ivs = 1.0f / vs;
    const float step_scale = truncation * ivs * 0.6f;
    const float2 mm = range[pid];
    f3 ps = depth_to_skeleton(cam, x, y, mm.x);
    float total = flen(ps) * ivs;
    ps = mul3(xform(g, ps), ivs);
    f3 pe = depth_to_skeleton(cam, x, y, mm.y);
    const float total_max = flen(pe) * ivs;
    pe = mul3(xform(g, pe), ivs);
    const f3 dir = normalize3(sub3(pe, ps));
    f3 pt = ps;
    float sdf = 1.0f, step;
    BrickCache cache;
    while (total < total_max) {
        const int vx = cvt_sat(roundf(pt.x)), vy = cvt_sat(roundf(pt.y)), vz = cvt_sat(roundf(pt.z));
        float v;
        if (!read_voxel<V>(s, cache, vx, vy, vz, &v)) {
            sdf = 1.0f;  // missing block: no surface here
            step = (float)kSB;
        } else {
            sdf = v;
            if (sdf <= 0.1f && sdf >= -0.5f) sdf = interp_sdf<V>(s, cache, pt);
            if (sdf <= 0.0f) break;
            step = fmaxf(sdf * step_scale, 1.0f);
        }
        pt = add3(pt, mul3(dir, step));
        total += step;
    }
    if (!(sdf <= 0.0f)) {
        st3(vmap, pid, mk3(kNaN, kNaN, kNaN));
        if (nmap) st3(nmap, pid, mk3(kNaN, kNaN, kNaN));
        if (cmap) cmap[pid] = make_uchar4(0, 0, 0, 0);
        return;
    }
    step = sdf * step_scale;
    pt = add3(pt, mul3(dir, step));
    const int vx = cvt_sat(floorf(pt.x)), vy = cvt_sat(floorf(pt.y)), vz = cvt_sat(floorf(pt.z));
    const f3 co = mk3(pt.x - (float)vx, pt.y - (float)vy, pt.z - (float)vz);
    const f3 nc = mk3(1.0f - co.x, 1.0f - co.y, 1.0f - co.z);
    float val[8];
    uchar4 col = make_uchar4(0, 0, 0, 0);
    float nearest = 1.0e20f;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int ox = k & 1, oy = (k >> 1) & 1, oz = k >> 2;
        uint32_t vi;
        const bool found = read_voxel<V>(s, cache, vx + ox, vy + oy, vz + oz, &val[k], &vi);
        if (cmap && found) {
            const float dist = flen(mk3(ox ? nc.x : co.x, oy ? nc.y : co.y, oz ? nc.z : co.z));
            if (nearest > dist) {
                nearest = dist;
                if (s.colors) {
                    const uchar4 c = s.colors[vi];
                    col = make_uchar4(c.x, c.y, c.z, 255);
                } else {
                    col = make_uchar4(0, 0, 0, 255);
                }
            }
        }
    }
    const float fx_ = val[0], fy_ = val[1], fz_ = val[2], fw_ = val[3];
    const float bx_ = val[4], by_ = val[5], bz_ = val[6], bw_ = val[7];
    sdf = fx_ * nc.x * nc.y * nc.z + fy_ * co.x * nc.y * nc.z + fz_ * nc.x * co.y * nc.z + fw_ * co.x * co.y * nc.z +
          bx_ * nc.x * nc.y * co.z + by_ * co.x * nc.y * co.z + bz_ * nc.x * co.y * co.z + bw_ * co.x * co.y * co.z;
    step = sdf * step_scale;
    pt = add3(pt, mul3(dir, step));
    st3(vmap, pid, mul3(pt, vs));
    if (nmap) {
        f3 ret;
        float p1, p2, v1, t0, t1, t2, t3;
        p1 = fx_ * nc.y * nc.z + fz_ * co.y * nc.z + bx_ * nc.y * co.z + bz_ * co.y * co.z;
        t0 = rd<V>(s, cache, vx - 1, vy, vz), t1 = rd<V>(s, cache, vx - 1, vy + 1, vz);
        t2 = rd<V>(s, cache, vx - 1, vy, vz + 1), t3 = rd<V>(s, cache, vx - 1, vy + 1, vz + 1);
        p2 = t0 * nc.y * nc.z + t1 * co.y * nc.z + t2 * nc.y * co.z + t3 * co.y * co.z;
        v1 = p1 * co.x + p2 * nc.x;
        p1 = fy_ * nc.y * nc.z + fw_ * co.y * nc.z + by_ * nc.y * co.z + bw_ * co.y * co.z;
        t0 = rd<V>(s, cache, vx + 2, vy, vz), t1 = rd<V>(s, cache, vx + 2, vy + 1, vz);
        t2 = rd<V>(s, cache, vx + 2, vy, vz + 1), t3 = rd<V>(s, cache, vx + 2, vy + 1, vz + 1);
        p2 = t0 * nc.y * nc.z + t1 * co.y * nc.z + t2 * nc.y * co.z + t3 * co.y * co.z;
        ret.x = p1 * nc.x + p2 * co.x - v1;
        p1 = fx_ * nc.x * nc.z + fy_ * co.x * nc.z + bx_ * nc.x * co.z + by_ * co.x * co.z;
        t0 = rd<V>(s, cache, vx, vy - 1, vz), t1 = rd<V>(s, cache, vx + 1, vy - 1, vz);
        t2 = rd<V>(s, cache, vx, vy - 1, vz + 1), t3 = rd<V>(s, cache, vx + 1, vy - 1, vz + 1);
        p2 = t0 * nc.x * nc.z + t1 * co.x * nc.z + t2 * nc.x * co.z + t3 * co.x * co.z;
        v1 = p1 * co.y + p2 * nc.y;
        p1 = fz_ * nc.x * nc.z + fw_ * co.x * nc.z + bz_ * nc.x * co.z + bw_ * co.x * co.z;
        t0 = rd<V>(s, cache, vx, vy + 2, vz), t1 = rd<V>(s, cache, vx + 1, vy + 2, vz);
        t2 = rd<V>(s, cache, vx, vy + 2, vz + 1), t3 = rd<V>(s, cache, vx + 1, vy + 2, vz + 1);
        p2 = t0 * nc.x * nc.z + t1 * co.x * nc.z + t2 * nc.x * co.z + t3 * co.x * co.z;
        ret.y = p1 * nc.y + p2 * co.y - v1;
        p1 = fx_ * nc.x * nc.y + fy_ * co.x * nc.y + fz_ * nc.x * co.y + fw_ * co.x * co.y;
        t0 = rd<V>(s, cache, vx, vy, vz - 1), t1 = rd<V>(s, cache, vx + 1, vy, vz - 1);
        t2 = rd<V>(s, cache, vx, vy + 1, vz - 1), t3 = rd<V>(s, cache, vx + 1, vy + 1, vz - 1);
        p2 = t0 * nc.x * nc.y + t1 * co.x * nc.y + t2 * nc.x * co.y + t3 * co.x * co.y;
        v1 = p1 * co.z + p2 * nc.z;
        p1 = bx_ * nc.x * nc.y + by_ * co.x * nc.y + bz_ * nc.x * co.y + bw_ * co.x * co.y;
        t0 = rd<V>(s, cache, vx, vy, vz + 2), t1 = rd<V>(s, cache, vx + 1, vy, vz + 2);
        t2 = rd<V>(s, cache, vx, vy + 1, vz + 2), t3 = rd<V>(s, cache, vx + 1, vy + 1, vz + 2);
        p2 = t0 * nc.x * nc.y + t1 * co.x * nc.y + t2 * nc.x * co.y + t3 * co.x * co.y;
        ret.z = p1 * nc.z + p2 * co.z - v1;
        st3(nmap, pid, normalize3(ret));
    }
    if (cmap) cmap[pid] = col;
}

// K/KinfuHashRaycast.cl:450-471 raycast_kernel: castRay at every pixel, 8x8 pixel tiles
// per wave so that neighbouring rays share bricks
template <class V>
__global__ void __launch_bounds__(64)
    hash_raycast_kernel(Scene s, const float2* __restrict__ range, NuiCLCameraParams cam,
                        const NuiCLRigidTransform* __restrict__ pose, float vs, float truncation,
                        float* __restrict__ vmap, float* __restrict__ nmap, uchar4* __restrict__ cmap,
                        const int32_t* __restrict__ enable) {
    if (enable && *enable == 0) return;
    const int W = (int)cam.depthImageWidth, H = (int)cam.depthImageHeight;
    const int x = blockIdx.x * 8 + (threadIdx.x & 7), y = blockIdx.y * 8 + (threadIdx.x >> 3);
    if (x >= W || y >= H) return;
    cast_ray<V>(s, range, cam, *pose, vs, truncation, x, y, vmap, nmap, cmap);
}

// ---- forward render (OCL/NuiKinfuOpenCLHashScene.cpp:991-1012 ForwardRender) -------
//
// The reference moves each point of the previous vertex map to the pixel it projects
// to, then casts rays only where no point landed and the ray range is not empty.  Here
// the projection is a z-buffer: every point does one 64-bit atomicMin of
// (ordered z << 32 | source pixel) on its target pixel, so the nearest point wins
// (smallest source index on ties) whatever the order of the writes, which the
// reference leaves to a race.  The pixels to raycast are appended per wave (ballot +
// one atomic) to a list that a grid-stride castRay kernel consumes; each pixel's result
// depends on that pixel only, so the list order does not matter.

constexpr unsigned long long kNoPoint = ~0ull;

// float -> uint32 key that orders as the float does (negative values included)
__device__ __forceinline__ uint32_t ordered_z(float z) {
    const uint32_t b = __float_as_uint(z);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
// float -> uint32 as the GPU converts: round toward zero, saturating, NaN -> 0
__device__ __forceinline__ uint32_t cvt_u32_sat(float v) {
    if (!(v > 0.0f)) return 0u;
    if (v >= 4294967296.0f) return 0xffffffffu;
    return (uint32_t)v;
}

__global__ void hash_fwd_clear_kernel(unsigned long long* __restrict__ keys, int n, uint32_t* __restrict__ count,
                                      const int32_t* __restrict__ enable) {
    if (enable && *enable == 0) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) *count = 0;
    if (i < n) keys[i] = kNoPoint;
}

// K/KinfuHashRaycast.cl:473-506 forwardProject_kernel.  reference_transform: the
// reference's projection (camera->world applied to the world point, scaled by the voxel
// size; uint conversion of p + 0.5); otherwise world->camera, z > 0 and p + 0.5 >= 0.
__global__ void __launch_bounds__(256)
    hash_fwd_splat_kernel(const float* __restrict__ vmap, NuiCLCameraParams cam,
                          const NuiCLRigidTransform* __restrict__ pose, float vs, int reference_transform,
                          unsigned long long* __restrict__ keys, const int32_t* __restrict__ enable) {
    if (enable && *enable == 0) return;
    const int W = (int)cam.depthImageWidth, H = (int)cam.depthImageHeight;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= W * H) return;
    const f3 v = ld3(vmap, i);
    if (nan3(v)) return;
    const NuiCLRigidTransform g = *pose;
    f3 q;
    uint32_t ux, uy;
    if (reference_transform) {
        q = mul3(xform(g, v), vs);
        ux = cvt_u32_sat(q.x * cam.fx / q.z + cam.cx + 0.5f);
        uy = cvt_u32_sat(q.y * cam.fy / q.z + cam.cy + 0.5f);
    } else {
        q = xform_inv(g, v);
        if (!(q.z > 0.0f)) return;
        const float sx = q.x * cam.fx / q.z + cam.cx + 0.5f, sy = q.y * cam.fy / q.z + cam.cy + 0.5f;
        if (!(sx >= 0.0f && sy >= 0.0f)) return;
        ux = cvt_u32_sat(sx);
        uy = cvt_u32_sat(sy);
    }
    if (ux >= (uint32_t)W || uy >= (uint32_t)H) return;
    atomicMin(&keys[(size_t)uy * W + ux], ((unsigned long long)ordered_z(q.z) << 32) | (uint32_t)i);
}

// The forwarded point of each pixel (NaN where none landed) into fwd, and its colour
// (the pixel's own where none landed) into the low half of the pixel's key
__global__ void __launch_bounds__(256)
    hash_fwd_gather_kernel(const float* __restrict__ vmap, const uchar4* __restrict__ cmap, float* __restrict__ fwd,
                           unsigned long long* __restrict__ keys, int n, const int32_t* __restrict__ enable) {
    if (enable && *enable == 0) return;
    const int d = blockIdx.x * 256 + threadIdx.x;
    if (d >= n) return;
    const unsigned long long k = keys[d];
    const bool hit = k != kNoPoint;
    const uint32_t src = hit ? (uint32_t)k : (uint32_t)d;
    st3(fwd, d, hit ? ld3(vmap, src) : mk3(kNaN, kNaN, kNaN));
    if (cmap) {
        const uchar4 c = cmap[src];
        keys[d] = (unsigned long long)((uint32_t)c.x | ((uint32_t)c.y << 8) | ((uint32_t)c.z << 16) |
                                       ((uint32_t)c.w << 24));
    }
}

// K/KinfuHashRaycast.cl:508-529 findMissingPoints_kernel: keep the forwarded point, or
// list the pixel for a ray when none landed and its ray range is not empty.  8x8 tiles
// per wave keep the listed pixels of a wave close together.
__global__ void __launch_bounds__(64)
    hash_fwd_missing_kernel(const float* __restrict__ fwd, const float2* __restrict__ range, float* __restrict__ vmap,
                            uchar4* __restrict__ cmap, const unsigned long long* __restrict__ keys, int W, int H,
                            uint32_t* __restrict__ list, uint32_t* __restrict__ count,
                            const int32_t* __restrict__ enable) {
    if (enable && *enable == 0) return;
    const int x = blockIdx.x * 8 + (threadIdx.x & 7), y = blockIdx.y * 8 + (threadIdx.x >> 3);
    const size_t d = (size_t)y * W + x;
    bool needed = false;
    if (x < W && y < H) {
        const f3 f = ld3(fwd, d);
        const float2 mm = range[d];
        needed = nan3(f) && mm.x < mm.y;
        if (!needed) st3(vmap, d, f);
        if (cmap) {
            const uint32_t c = (uint32_t)keys[d];
            cmap[d] = make_uchar4(c & 255u, (c >> 8) & 255u, (c >> 16) & 255u, c >> 24);
        }
    }
    const unsigned long long mask = __ballot(needed);
    const int lane = threadIdx.x & 63;
    uint32_t base = 0;
    if (lane == 0 && mask) base = atomicAdd(count, (uint32_t)__popcll(mask));
    base = __shfl(base, 0, 64);
    if (needed) list[base + __popcll(mask & ((1ull << lane) - 1ull))] = (uint32_t)d;
}

// K/KinfuHashRaycast.cl:531-559 genericRaycastMissingPoints_kernel: castRay (no normal)
// at the listed pixels
template <class V>
__global__ void __launch_bounds__(64)
    hash_fwd_raycast_kernel(Scene s, const float2* __restrict__ range, NuiCLCameraParams cam,
                            const NuiCLRigidTransform* __restrict__ pose, float vs, float truncation,
                            float* __restrict__ vmap, uchar4* __restrict__ cmap, const uint32_t* __restrict__ list,
                            const uint32_t* __restrict__ count, const int32_t* __restrict__ enable) {
    if (enable && *enable == 0) return;
    const uint32_t n = *count, W = cam.depthImageWidth;
    const NuiCLRigidTransform g = *pose;
    for (uint32_t k = blockIdx.x * 64 + threadIdx.x; k < n; k += gridDim.x * 64) {
        const uint32_t d = list[k];
        cast_ray<V>(s, range, cam, g, vs, truncation, (int)(d % W), (int)(d / W), vmap, nullptr, cmap);
    }
}

// OCL/NuiKinfuOpenCLHashScene.cpp:943-989 Vertex2Normal (threshold 0.3 m)
__global__ void hash_fwd_normals_kernel(const float* __restrict__ vmap, float* __restrict__ nmap, int W, int H,
                                        const int32_t* __restrict__ enable) {
    if (enable && *enable == 0) return;
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= W || y >= H) return;
    st3(nmap, (size_t)y * W + x, normal_simple_at(vmap, x, y, W, H, 0.3f));
}

// ---- marching cubes (K/KinfuHashVisualization.cl:88-271) -------------------------

// Cube (x, y, z) of voxel coordinates: its corner sdf values, false when a corner's
// block is missing (voxel weights are not consulted, as in the reference).
template <class V>
__device__ bool cube_at(const Scene& s, BrickCache& c, int x, int y, int z, float val[8], uint32_t vidx[8]) {
    const int ox[8] = {0, 1, 1, 0, 0, 1, 1, 0}, oy[8] = {0, 0, 1, 1, 0, 0, 1, 1}, oz[8] = {0, 0, 0, 0, 1, 1, 1, 1};
#pragma unroll
    for (int k = 0; k < 8; ++k)
        if (!read_voxel<V>(s, c, x + ox[k], y + oy[k], z + oz[k], &val[k], &vidx[k])) return false;
    return true;
}

__device__ __forceinline__ int cube_index(const float val[8]) {
    int ci = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k)
        if (val[k] < 0) ci |= 1 << k;
    return ci;
}

template <class V>
__global__ void __launch_bounds__(64) hash_mc_count_kernel(Scene s, uint32_t* __restrict__ counts) {
    const uint32_t n = s.hdr->nvisible;
    const int lane = threadIdx.x, lx = lane & 7, ly = lane >> 3;
    for (uint32_t i = blockIdx.x; i < n; i += gridDim.x) {
        const HashEntry e = s.entries[s.visible[i]];
        uint32_t cnt = 0;
        if (e.ptr >= 0) {
            int p[3];
            key_pos(e.key, p);
            BrickCache c;
            for (int lz = 0; lz < kSB; ++lz) {
                float val[8];
                uint32_t vi[8];
                if (cube_at<V>(s, c, p[0] * kSB + lx, p[1] * kSB + ly, p[2] * kSB + lz, val, vi))
                    cnt += kMcHash.nverts[cube_index(val)];
            }
        }
        for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off);
        if (lane == 0) counts[i] = cnt;
    }
}

// K/KinfuHashVisualization.cl:78-86 sdfInterp
__device__ __forceinline__ f3 sdf_interp(f3 p1, f3 p2, float v1, float v2) {
    if (fabsf(0.0f - v1) < 0.00001f) return p1;
    if (fabsf(0.0f - v2) < 0.00001f) return p2;
    if (fabsf(v1 - v2) < 0.00001f) return p1;
    return add3(p1, mul3(sub3(p2, p1), (0.0f - v1) / (v2 - v1)));
}

template <class V>
__global__ void __launch_bounds__(64)
    hash_mc_emit_kernel(Scene s, const uint32_t* __restrict__ offsets, float vs, float* __restrict__ verts,
                        float* __restrict__ colors, uint32_t capacity) {
    const uint32_t n = s.hdr->nvisible;
    const int lane = threadIdx.x, lx = lane & 7, ly = lane >> 3;
    const int ox[8] = {0, 1, 1, 0, 0, 1, 1, 0}, oy[8] = {0, 0, 1, 1, 0, 0, 1, 1}, oz[8] = {0, 0, 0, 0, 1, 1, 1, 1};
    for (uint32_t i = blockIdx.x; i < n; i += gridDim.x) {
        const HashEntry e = s.entries[s.visible[i]];
        if (e.ptr < 0) continue;
        int p[3];
        key_pos(e.key, p);
        BrickCache c;
        uint32_t base = offsets[i];
        for (int lz = 0; lz < kSB; ++lz) {
            const int gx = p[0] * kSB + lx, gy = p[1] * kSB + ly, gz = p[2] * kSB + lz;
            float val[8];
            uint32_t vi[8];
            const bool ok = cube_at<V>(s, c, gx, gy, gz, val, vi);
            const int ci = ok ? cube_index(val) : 0;
            const uint32_t nv = ok ? kMcHash.nverts[ci] : 0u;
            // voxel order inside the brick: x fastest, then y, then z
            uint32_t incl = nv;
            for (int off = 1; off < 64; off <<= 1) {
                const uint32_t t = __shfl_up(incl, off, 64);
                if (lane >= off) incl += t;
            }
            const uint32_t mine = base + incl - nv;
            base += __shfl(incl, 63, 64);
            for (uint32_t k = 0; k < nv; ++k) {
                const uint32_t o = mine + k;
                if (o >= capacity) break;
                const int ed = kMcHash.tri[ci][k];
                const int a = kMcHash.edge[ed][0], b = kMcHash.edge[ed][1];
                const f3 pa = mk3((float)(gx + ox[a]), (float)(gy + oy[a]), (float)(gz + oz[a]));
                const f3 pb = mk3((float)(gx + ox[b]), (float)(gy + oy[b]), (float)(gz + oz[b]));
                st3(verts, o, mul3(sdf_interp(pa, pb, val[a], val[b]), vs));
                if (colors) {
                    f3 ca = mk3(0.f, 0.f, 0.f), cb = ca;
                    if (s.colors) {
                        const uchar4 qa = s.colors[vi[a]], qb = s.colors[vi[b]];
                        ca = mk3((float)qa.x / 255.f, (float)qa.y / 255.f, (float)qa.z / 255.f);
                        cb = mk3((float)qb.x / 255.f, (float)qb.y / 255.f, (float)qb.z / 255.f);
                    }
                    const f3 cc = sdf_interp(ca, cb, val[a], val[b]);
                    float* q = colors + 4 * (size_t)o;
                    q[0] = cc.x;
                    q[1] = cc.y;
                    q[2] = cc.z;
                    q[3] = 1.0f;
                }
            }
        }
    }
}

// K/KinfuHashVisualization.cl:5-52 fetchHashScene_kernel: every observed voxel of a
// visible block inside the band becomes a point at its centre.  The reference
// appends through an atomic counter; here one wavefront takes one visible block
// (8 coalesced 256-byte loads), pass 1 counts its points, the scan turns the
// counts into offsets and pass 2 writes them in visible-list order with the voxel
// index (x + 8y + 64z) ascending inside a block.
template <class V, bool EMIT>
__global__ void __launch_bounds__(64)
    hash_points_kernel(Scene s, float vs, float ts, uint32_t* __restrict__ counts, float* __restrict__ pts,
                       float* __restrict__ cols, uint32_t capacity) {
    const uint32_t n = s.hdr->nvisible;
    const int lane = threadIdx.x;
    const unsigned long long below = (1ull << lane) - 1ull;
    for (uint32_t i = blockIdx.x; i < n; i += gridDim.x) {
        const int32_t id = s.visible[i];
        const HashEntry e = id >= 0 ? s.entries[id] : HashEntry{0ull, -1, 0u};
        uint32_t base = EMIT ? counts[i] : 0u;
        if (e.ptr >= 0) {
            int p[3];
            key_pos(e.key, p);
            for (int lz = 0; lz < kSB; ++lz) {
                const uint32_t vi = (uint32_t)e.ptr * kSB3 + (uint32_t)(lane + 64 * lz);
                const uint32_t w = s.bricks[vi];
                const bool keep = V::weight(w) > 0 && V::in_band(w, ts);
                const unsigned long long m = __ballot(keep);
                if (EMIT && keep) {
                    const uint32_t o = base + (uint32_t)__popcll(m & below);
                    if (o < capacity) {
                        const float pz = ((float)(p[2] * kSB) + ((float)lz + 0.5f)) * vs;
                        const float py = ((float)(p[1] * kSB) + ((float)(lane >> 3) + 0.5f)) * vs;
                        const float px = ((float)(p[0] * kSB) + ((float)(lane & 7) + 0.5f)) * vs;
                        st3(pts, o, mk3(px, py, pz));
                        if (cols) {
                            float4 c = make_float4(0.f, 0.f, 0.f, 1.f);
                            if (s.colors) {
                                const uchar4 q = s.colors[vi];
                                c = make_float4((float)q.x / 255.f, (float)q.y / 255.f, (float)q.z / 255.f, 1.f);
                            }
                            reinterpret_cast<float4*>(cols)[o] = c;
                        }
                    }
                }
                base += (uint32_t)__popcll(m);
            }
        }
        if (!EMIT && lane == 0) counts[i] = base;
    }
}

// ---- host swapping (OCL/NuiKinfuOpenCLHashGlobalCache.cpp, K/KinfuHashSwapping.cl) ----

// K/KinfuHashScene.cl:383-384 and K/KinfuHashSwapping.cl:6-21 reAllocateSwappedOutVoxelBlocks:
// a visible entry not yet synchronised asks for its stored data (state 1); a visible
// entry without a brick (swapped out, or left without one by a full pool) takes one
// from the free-brick stack (bricks come back zeroed).
__global__ void hash_swap_mark_kernel(Scene s, uint32_t capacity, const int32_t* __restrict__ gate) {
    if (gate && *gate == 0) return;
    const uint32_t n = s.hdr->nvisible;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int id = s.visible[i];
        if (s.swap[id] != 2) s.swap[id] = 1;
        if (s.entries[id].ptr == -1) {
            const int v = pop_brick(s, capacity);
            if (v >= 0) s.entries[id].ptr = v;
        }
    }
}

// List flags: swap-in = visible entries in state 1 holding a brick (visible-list
// order); swap-out = entries in state 2 holding a brick that this frame does not see
// (K/KinfuHashSwapping.cl:120-133 buildSwapOutList_kernel; its test ptr > 0 keeps
// whichever block holds brick 0 resident, an accident of allocation order, so here
// every brick can go).
__global__ void hash_swap_in_compact_kernel(Scene s, const uint32_t* __restrict__ flag,
                                            const uint32_t* __restrict__ pos, uint32_t n, int32_t* __restrict__ ids) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        if (flag[i]) ids[pos[i]] = s.visible[i];
}
__global__ void hash_swap_out_flags_kernel(Scene s, uint32_t* __restrict__ flag, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        flag[i] = (s.swap[i] == 2 && s.entries[i].ptr >= 0 && s.flags[i] == 0) ? 1u : 0u;
}
__global__ void hash_swap_out_compact_kernel(const uint32_t* __restrict__ flag, const uint32_t* __restrict__ pos,
                                             size_t n, int32_t* __restrict__ ids) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        if (flag[i]) ids[pos[i]] = (int32_t)i;
}

// the visible list's length as a kernel argument needs the host; the flags kernel
// reads it from the header instead and flags the rest of the capacity 0
__global__ void hash_swap_in_flags_all_kernel(Scene s, uint32_t* __restrict__ flag, size_t cap) {
    const uint32_t n = s.hdr->nvisible;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < cap; i += (size_t)gridDim.x * blockDim.x) {
        uint32_t f = 0u;
        if (i < n) {
            const int id = s.visible[i];
            f = (s.swap[id] == 1 && s.entries[id].ptr >= 0) ? 1u : 0u;
        }
        flag[i] = f;
    }
}

// float -> voxel word of the two storages, NaN -> 0 (a block merged with nothing at a
// voxel neither side observed)
template <class V> __device__ __forceinline__ uint32_t sdf_bits(float f);
template <> __device__ __forceinline__ uint32_t sdf_bits<VS16>(float f) {
    return (uint32_t)(uint16_t)(int16_t)cvt_sat(f * 32767.0f);
}
template <> __device__ __forceinline__ uint32_t sdf_bits<VF16>(float f) {
    return f != f ? 0u : (uint32_t)__half_as_ushort(__float2half_rn(f));
}

// K/KinfuHashSwapping.cl:57-118 integrateOldIntoActiveData_kernel: the stored block
// (zeros when nothing was stored) and the active brick, weight-averaged into the
// brick; the entry is synchronised (state 2).  One wavefront per listed entry.
template <class V>
__global__ void __launch_bounds__(64)
    hash_swap_in_kernel(Scene s, const int32_t* __restrict__ ids, uint32_t n, const uint32_t* __restrict__ words,
                        const uchar4* __restrict__ cols, int wmax) {
    const int lane = threadIdx.x;
    for (uint32_t i = blockIdx.x; i < n; i += gridDim.x) {
        const int id = ids[i];
        const int ptr = s.entries[id].ptr;
        for (int k = 0; k < 8; ++k) {
            const int li = lane + 64 * k;
            const size_t vi = (size_t)ptr * kSB3 + li;
            const uint32_t old = words[(size_t)i * kSB3 + li];
            const uint32_t cur = s.bricks[vi];
            const float oldF = V::sdf(old), oldW = (float)V::weight(old);
            const float newF = V::sdf(cur), newW = (float)V::weight(cur);
            const float combinedW = oldW + newW;
            float f = oldW * oldF + newW * newF;
            f /= combinedW;
            // the reference's convert_uchar of a weight sum above 255 is undefined; it saturates here
            const int w = min(cvt_sat(combinedW), wmax);
            s.bricks[vi] = sdf_bits<V>(f) | ((uint32_t)w << 16);
            if (s.colors) {
                const uchar4 oc = cols ? cols[(size_t)i * kSB3 + li] : make_uchar4(0, 0, 0, 0);
                uchar4 nc = s.colors[vi];
                const float r = (oldW * (float)oc.x + newW * (float)nc.x) / combinedW;
                const float g = (oldW * (float)oc.y + newW * (float)nc.y) / combinedW;
                const float b = (oldW * (float)oc.z + newW * (float)nc.z) / combinedW;
                nc.x = (unsigned char)cvt_sat(fminf(fmaxf(r, 0.0f), 255.f));
                nc.y = (unsigned char)cvt_sat(fminf(fmaxf(g, 0.0f), 255.f));
                nc.z = (unsigned char)cvt_sat(fminf(fmaxf(b, 0.0f), 255.f));
                s.colors[vi] = nc;
            }
        }
        if (lane == 0) s.swap[id] = 2;
    }
}

// K/KinfuHashSwapping.cl:135-180 moveActiveDataToTransferBuffer_kernel + cleanVoxelMemory_kernel:
// the brick goes to the transfer buffer and back to the free-brick stack zeroed; the
// entry keeps its key without a brick (ptr -1, state 0).
__global__ void __launch_bounds__(64)
    hash_swap_out_kernel(Scene s, const int32_t* __restrict__ ids, uint32_t n, uint32_t* __restrict__ words,
                         uchar4* __restrict__ cols) {
    const int lane = threadIdx.x;
    for (uint32_t i = blockIdx.x; i < n; i += gridDim.x) {
        const int id = ids[i];
        const int ptr = s.entries[id].ptr;
        for (int k = 0; k < 8; ++k) {
            const int li = lane + 64 * k;
            const size_t vi = (size_t)ptr * kSB3 + li;
            words[(size_t)i * kSB3 + li] = s.bricks[vi];
            s.bricks[vi] = 0u;
            if (s.colors) {
                if (cols) cols[(size_t)i * kSB3 + li] = s.colors[vi];
                s.colors[vi] = make_uchar4(0, 0, 0, 0);
            }
        }
        if (lane == 0) {
            const uint32_t top = atomicSub(&s.hdr->nbricks, 1u) - 1u;
            s.freelist[top] = ptr;
            s.entries[id].ptr = -1;
            s.swap[id] = 0;
        }
    }
}

__global__ void hash_reset_kernel(Scene s, size_t nentries, size_t nwords, size_t nbricks) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    const size_t i0 = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i0 == 0) {
        s.hdr->nbricks = 0;
        s.hdr->nvisible = 0;
        s.hdr->failed = 0;
        s.hdr->nentries = 0;
    }
    for (size_t i = i0; i < nentries; i += stride) {
        s.entries[i].key = 0ull;
        s.entries[i].ptr = -1;
        s.entries[i].pad = 0u;
        s.swap[i] = 0;
    }
    for (size_t i = i0; i < nbricks; i += stride) s.freelist[i] = (int32_t)i;
    for (size_t i = i0; i < nwords; i += stride) {
        s.bricks[i] = 0u;
        if (s.colors) s.colors[i] = make_uchar4(0, 0, 0, 0);
    }
}

int check_params(const nuif_hash_params* p) {
    NUIF_CHECK_ARG(p != nullptr, "hash: null parameters");
    NUIF_CHECK_ARG(p->voxel_size > 0.f && p->trunc_scale > 0.f && p->truncation > 0.f,
                   "hash: voxel size and truncations must be positive");
    NUIF_CHECK_ARG(p->weight_max >= 1 && p->weight_max <= 255, "hash: weight_max %d outside 1..255", p->weight_max);
    NUIF_CHECK_ARG(p->buckets_log2 >= 4 && p->buckets_log2 <= 24, "hash: buckets_log2 %d outside 4..24",
                   p->buckets_log2);
    NUIF_CHECK_ARG(p->bricks >= 1 && (size_t)p->bricks * kSB3 < (1ull << 32), "hash: brick count %d", p->bricks);
    NUIF_CHECK_ARG(p->storage == NUIF_TSDF_S16 || p->storage == NUIF_TSDF_F16,
                   "hash: storage must be NUIF_TSDF_S16 or NUIF_TSDF_F16");
    return NUIF_OK;
}

}  // namespace
}  // namespace nuif

using namespace nuif;

extern "C" {

size_t nuif_hash_scene_bytes(const nuif_hash_params* p) {
    if (!p || check_params(p)) return 0;
    return layout_of(*p).total;
}

int nuif_hash_layout(const nuif_hash_params* p, size_t* offsets8) {
    int rc = check_params(p);
    if (rc) return rc;
    NUIF_CHECK_ARG(offsets8 != nullptr, "nuif_hash_layout: null output");
    const Layout L = layout_of(*p);
    offsets8[0] = L.header;
    offsets8[1] = L.entries;
    offsets8[2] = L.flags;
    offsets8[3] = L.visible;
    offsets8[4] = L.bricks;
    offsets8[5] = p->color ? L.colors : 0;
    offsets8[6] = L.nentries;
    offsets8[7] = L.total;
    return NUIF_OK;
}

int nuif_hash_reset(void* scene, const nuif_hash_params* p, nuif_stream_t s) {
    NUIF_REQUIRE_INIT();
    int rc = check_params(p);
    if (rc) return rc;
    NUIF_CHECK_ARG(scene != nullptr, "nuif_hash_reset: null scene");
    const Layout L = layout_of(*p);
    hash_reset_kernel<<<4096, 256, 0, stream_of(s)>>>(scene_of(scene, *p), L.nentries, (size_t)p->bricks * kSB3,
                                                       (size_t)p->bricks);
    NUIF_LAUNCHED("hash_reset_kernel");
    return NUIF_OK;
}

int nuif_hash_allocate(void* scene, const nuif_hash_params* p, const float* depth, const NuiCLCameraParams* cam,
                       const NuiCLRigidTransform* pose, const int32_t* gate, nuif_stream_t s) {
    NUIF_REQUIRE_INIT();
    int rc = check_params(p);
    if (rc) return rc;
    NUIF_CHECK_ARG(scene && depth && cam && pose, "nuif_hash_allocate: null argument");
    const int n = (int)(cam->depthImageWidth * cam->depthImageHeight);
    NUIF_CHECK_ARG(n > 0, "nuif_hash_allocate: empty image");
    const Layout L = layout_of(*p);
    Scene sc = scene_of(scene, *p);
    hipStream_t st = stream_of(s);
    // OCL/NuiKinfuOpenCLHashScene.cpp:187-191 ResetAllocType, :146-185 ResetVisibleEntrys
    rc = nuif_fill_u8(sc.flags, 0, L.nentries, s);
    if (rc) return rc;
    // host float, as OCL/NuiKinfuOpenCLHashScene.cpp:205 computes it
    const float inv_block = 1.0f / (p->voxel_size * (float)kSB);
    hash_touch_kernel<<<div_up(n, 256), 256, 0, st>>>(sc, depth, *cam, pose, p->trunc_scale, inv_block,
                                                       (uint32_t)p->bricks, gate);
    NUIF_LAUNCHED("hash_touch_kernel");
    hash_recheck_kernel<<<kWalkGrid / 4, 256, 0, st>>>(sc, *cam, pose, p->voxel_size, gate);
    NUIF_LAUNCHED("hash_recheck_kernel");
    // the visible list: exclusive scan of the flags, compaction in entry order;
    // the new length goes to the header only when the gate is open
    uint32_t* total = &sc.hdr->pad[0];
    rc = launch_scan(sc.flags, true, sc.offsets, L.nentries, total, sc.scan, st);
    if (rc) return rc;
    hash_compact_kernel<<<div_up(L.nentries, 256), 256, 0, st>>>(sc, L.nentries, gate);
    NUIF_LAUNCHED("hash_compact_kernel");
    hash_commit_count_kernel<<<1, 1, 0, st>>>(sc, total, gate);
    NUIF_LAUNCHED("hash_commit_count_kernel");
    return NUIF_OK;
}

int nuif_hash_integrate(void* scene, const nuif_hash_params* p, const float* depth, const uint8_t* bgra,
                        const NuiCLCameraParams* cam, const NuiCLRigidTransform* pose, const int32_t* gate,
                        unsigned long long* updated, nuif_stream_t s) {
    NUIF_REQUIRE_INIT();
    int rc = check_params(p);
    if (rc) return rc;
    NUIF_CHECK_ARG(scene && depth && cam && pose, "nuif_hash_integrate: null argument");
    Scene sc = scene_of(scene, *p);
    hipStream_t st = stream_of(s);
    const bool color = p->color && bgra;
    const uchar4* c4 = reinterpret_cast<const uchar4*>(bgra);
#define NUIF_HASH_INT(V, F, C)                                                                                       \
    hash_integrate_kernel<V, F, C><<<kWalkGrid, 64, 0, st>>>(sc, depth, c4, *cam, pose, p->voxel_size,            \
                                                             p->trunc_scale, p->weight_max, gate, updated)
    if (p->storage == NUIF_TSDF_S16) {
        if (color)
            NUIF_HASH_INT(VS16, false, true);
        else
            NUIF_HASH_INT(VS16, false, false);
    } else {
        if (color)
            NUIF_HASH_INT(VF16, true, true);
        else
            NUIF_HASH_INT(VF16, true, false);
    }
#undef NUIF_HASH_INT
    NUIF_LAUNCHED("hash_integrate_kernel");
    return NUIF_OK;
}

}  // extern "C"

namespace {
// resetExpectedRange + CreateExpectedDepths (OCL/NuiKinfuOpenCLHashScene.cpp:1030-1040)
int launch_ranges(const Scene& sc, const nuif_hash_params* p, const NuiCLCameraParams* cam,
                  const NuiCLRigidTransform* pose, float2* r2, const int32_t* enable, hipStream_t st) {
    const int W = (int)cam->depthImageWidth, H = (int)cam->depthImageHeight;
    hash_range_init_kernel<<<div_up((size_t)W * H, 256), 256, 0, st>>>(r2, W * H, enable);
    NUIF_LAUNCHED("hash_range_init_kernel");
    hash_range_kernel<<<kWalkGrid, 64, 0, st>>>(sc, r2, *cam, pose, p->voxel_size, enable);
    NUIF_LAUNCHED("hash_range_kernel");
    return NUIF_OK;
}

struct FwdWork {
    unsigned long long* keys;
    uint32_t* list;
    uint32_t* count;
    size_t bytes;
};
FwdWork fwd_work(void* base, size_t n) {
    FwdWork w;
    char* b = static_cast<char*>(base);
    const size_t keys = align256(n * 8), list = align256(n * 4);
    w.keys = reinterpret_cast<unsigned long long*>(b);
    w.list = reinterpret_cast<uint32_t*>(b + keys);
    w.count = reinterpret_cast<uint32_t*>(b + keys + list);
    w.bytes = keys + list + 256;
    return w;
}
}  // namespace

extern "C" {

int nuif_hash_raycast(void* scene, const nuif_hash_params* p, const NuiCLCameraParams* cam,
                      const NuiCLRigidTransform* pose, float* range2, float* vmap, float* nmap, uint8_t* cmap,
                      const int32_t* enable, nuif_stream_t s) {
    NUIF_REQUIRE_INIT();
    int rc = check_params(p);
    if (rc) return rc;
    NUIF_CHECK_ARG(scene && cam && pose && range2 && vmap, "nuif_hash_raycast: null argument");
    const int W = (int)cam->depthImageWidth, H = (int)cam->depthImageHeight;
    NUIF_CHECK_ARG(W > 0 && H > 0, "nuif_hash_raycast: empty image");
    Scene sc = scene_of(scene, *p);
    hipStream_t st = stream_of(s);
    float2* r2 = reinterpret_cast<float2*>(range2);
    rc = launch_ranges(sc, p, cam, pose, r2, enable, st);
    if (rc) return rc;
    const dim3 grid(div_up(W, 8), div_up(H, 8));
    uchar4* c4 = reinterpret_cast<uchar4*>(cmap);
    if (p->storage == NUIF_TSDF_S16)
        hash_raycast_kernel<VS16><<<grid, 64, 0, st>>>(sc, r2, *cam, pose, p->voxel_size, p->truncation, vmap, nmap,
                                                        c4, enable);
    else
        hash_raycast_kernel<VF16><<<grid, 64, 0, st>>>(sc, r2, *cam, pose, p->voxel_size, p->truncation, vmap, nmap,
                                                        c4, enable);
    NUIF_LAUNCHED("hash_raycast_kernel");
    return NUIF_OK;
}

size_t nuif_hash_forward_render_bytes(int width, int height) {
    if (width <= 0 || height <= 0) return 0;
    return fwd_work(nullptr, (size_t)width * height).bytes;
}

int nuif_hash_forward_render(void* scene, const nuif_hash_params* p, const NuiCLCameraParams* cam,
                             const NuiCLRigidTransform* pose, float* range2, float* vmap, float* nmap, uint8_t* cmap,
                             void* work, int reference_transform, const int32_t* enable, nuif_stream_t s) {
    NUIF_REQUIRE_INIT();
    int rc = check_params(p);
    if (rc) return rc;
    NUIF_CHECK_ARG(scene && cam && pose && range2 && vmap && nmap && work,
                   "nuif_hash_forward_render: null argument");
    NUIF_CHECK_ARG(vmap != nmap, "nuif_hash_forward_render: vmap and nmap alias");
    const int W = (int)cam->depthImageWidth, H = (int)cam->depthImageHeight;
    NUIF_CHECK_ARG(W > 0 && H > 0, "nuif_hash_forward_render: empty image");
    const int n = W * H;
    Scene sc = scene_of(scene, *p);
    hipStream_t st = stream_of(s);
    float2* r2 = reinterpret_cast<float2*>(range2);
    uchar4* c4 = reinterpret_cast<uchar4*>(cmap);
    const FwdWork w = fwd_work(work, (size_t)n);
    rc = launch_ranges(sc, p, cam, pose, r2, enable, st);
    if (rc) return rc;
    hash_fwd_clear_kernel<<<div_up((size_t)n, 256), 256, 0, st>>>(w.keys, n, w.count, enable);
    NUIF_LAUNCHED("hash_fwd_clear_kernel");
    hash_fwd_splat_kernel<<<div_up((size_t)n, 256), 256, 0, st>>>(vmap, *cam, pose, p->voxel_size,
                                                                  reference_transform, w.keys, enable);
    NUIF_LAUNCHED("hash_fwd_splat_kernel");
    // the forwarded points go to nmap, which the normals overwrite at the end
    hash_fwd_gather_kernel<<<div_up((size_t)n, 256), 256, 0, st>>>(vmap, c4, nmap, w.keys, n, enable);
    NUIF_LAUNCHED("hash_fwd_gather_kernel");
    const dim3 tiles(div_up(W, 8), div_up(H, 8));
    hash_fwd_missing_kernel<<<tiles, 64, 0, st>>>(nmap, r2, vmap, c4, w.keys, W, H, w.list, w.count, enable);
    NUIF_LAUNCHED("hash_fwd_missing_kernel");
    const int rgrid = (int)std::min<size_t>(div_up((size_t)n, 64), 2048);
    if (p->storage == NUIF_TSDF_S16)
        hash_fwd_raycast_kernel<VS16><<<rgrid, 64, 0, st>>>(sc, r2, *cam, pose, p->voxel_size, p->truncation, vmap,
                                                             c4, w.list, w.count, enable);
    else
        hash_fwd_raycast_kernel<VF16><<<rgrid, 64, 0, st>>>(sc, r2, *cam, pose, p->voxel_size, p->truncation, vmap,
                                                             c4, w.list, w.count, enable);
    NUIF_LAUNCHED("hash_fwd_raycast_kernel");
    const dim3 block(16, 16), grid(div_up(W, 16), div_up(H, 16));
    hash_fwd_normals_kernel<<<grid, block, 0, st>>>(vmap, nmap, W, H, enable);
    NUIF_LAUNCHED("hash_fwd_normals_kernel");
    return NUIF_OK;
}

size_t nuif_hash_mesh_scratch_bytes(const nuif_hash_params* p) {
    if (!p || check_params(p)) return 0;
    const size_t n = layout_of(*p).nentries;
    return align256(n * 4) * 2 + align256(nuif_scan_scratch_bytes(n)) + 256;
}

int nuif_hash_mesh(void* scene, const nuif_hash_params* p, float* vertices, float* colors, uint32_t capacity,
                   uint32_t* total, void* scratch, nuif_stream_t s) {
    NUIF_REQUIRE_INIT();
    int rc = check_params(p);
    if (rc) return rc;
    NUIF_CHECK_ARG(scene && total && scratch, "nuif_hash_mesh: null argument");
    NUIF_CHECK_ARG(capacity == 0 || vertices, "nuif_hash_mesh: vertex buffer missing");
    const size_t n = layout_of(*p).nentries;
    Scene sc = scene_of(scene, *p);
    hipStream_t st = stream_of(s);
    char* b = static_cast<char*>(scratch);
    uint32_t* counts = reinterpret_cast<uint32_t*>(b);
    uint32_t* offsets = reinterpret_cast<uint32_t*>(b + align256(n * 4));
    void* scan = b + 2 * align256(n * 4);
    // counts of the whole visible-list capacity are scanned; entries past the
    // list length are zeroed first so the scan's total is the vertex count
    rc = nuif_fill_u32(counts, 0u, n, s);
    if (rc) return rc;
    if (p->storage == NUIF_TSDF_S16)
        hash_mc_count_kernel<VS16><<<kWalkGrid, 64, 0, st>>>(sc, counts);
    else
        hash_mc_count_kernel<VF16><<<kWalkGrid, 64, 0, st>>>(sc, counts);
    NUIF_LAUNCHED("hash_mc_count_kernel");
    rc = launch_scan(counts, false, offsets, n, total, scan, st);
    if (rc) return rc;
    if (capacity > 0) {
        if (p->storage == NUIF_TSDF_S16)
            hash_mc_emit_kernel<VS16><<<kWalkGrid, 64, 0, st>>>(sc, offsets, p->voxel_size, vertices, colors, capacity);
        else
            hash_mc_emit_kernel<VF16><<<kWalkGrid, 64, 0, st>>>(sc, offsets, p->voxel_size, vertices, colors, capacity);
        NUIF_LAUNCHED("hash_mc_emit_kernel");
    }
    return NUIF_OK;
}

size_t nuif_hash_points_scratch_bytes(const nuif_hash_params* p) { return nuif_hash_mesh_scratch_bytes(p); }

int nuif_hash_points(void* scene, const nuif_hash_params* p, float* points, float* colors, uint32_t capacity,
                     uint32_t* total, void* scratch, nuif_stream_t s) {
    NUIF_REQUIRE_INIT();
    int rc = check_params(p);
    if (rc) return rc;
    NUIF_CHECK_ARG(scene && total && scratch, "nuif_hash_points: null argument");
    NUIF_CHECK_ARG(capacity == 0 || points, "nuif_hash_points: point buffer missing");
    const size_t n = layout_of(*p).nentries;
    Scene sc = scene_of(scene, *p);
    hipStream_t st = stream_of(s);
    char* b = static_cast<char*>(scratch);
    uint32_t* counts = reinterpret_cast<uint32_t*>(b);
    uint32_t* offsets = reinterpret_cast<uint32_t*>(b + align256(n * 4));
    void* scan = b + 2 * align256(n * 4);
    // counts past the visible list's length stay zero, so the scan's total is the point count
    rc = nuif_fill_u32(counts, 0u, n, s);
    if (rc) return rc;
    if (p->storage == NUIF_TSDF_S16)
        hash_points_kernel<VS16, false><<<kWalkGrid, 64, 0, st>>>(sc, p->voxel_size, p->trunc_scale, counts, nullptr,
                                                                   nullptr, 0u);
    else
        hash_points_kernel<VF16, false><<<kWalkGrid, 64, 0, st>>>(sc, p->voxel_size, p->trunc_scale, counts, nullptr,
                                                                   nullptr, 0u);
    NUIF_LAUNCHED("hash_points_kernel");
    rc = launch_scan(counts, false, offsets, n, total, scan, st);
    if (rc) return rc;
    if (capacity > 0) {
        if (p->storage == NUIF_TSDF_S16)
            hash_points_kernel<VS16, true><<<kWalkGrid, 64, 0, st>>>(sc, p->voxel_size, p->trunc_scale, offsets,
                                                                      points, colors, capacity);
        else
            hash_points_kernel<VF16, true><<<kWalkGrid, 64, 0, st>>>(sc, p->voxel_size, p->trunc_scale, offsets,
                                                                      points, colors, capacity);
        NUIF_LAUNCHED("hash_points_kernel");
    }
    return NUIF_OK;
}

int nuif_hash_swap_layout(const nuif_hash_params* p, size_t* offsets2) {
    int rc = check_params(p);
    if (rc) return rc;
    NUIF_CHECK_ARG(offsets2 != nullptr, "nuif_hash_swap_layout: null output");
    const Layout L = layout_of(*p);
    offsets2[0] = L.swap;
    offsets2[1] = L.freelist;
    return NUIF_OK;
}

int nuif_hash_swap_prepare(void* scene, const nuif_hash_params* p, const int32_t* gate, nuif_stream_t s) {
    NUIF_REQUIRE_INIT();
    int rc = check_params(p);
    if (rc) return rc;
    NUIF_CHECK_ARG(scene != nullptr, "nuif_hash_swap_prepare: null scene");
    hash_swap_mark_kernel<<<kWalkGrid / 4, 256, 0, stream_of(s)>>>(scene_of(scene, *p), (uint32_t)p->bricks, gate);
    NUIF_LAUNCHED("hash_swap_mark_kernel");
    return NUIF_OK;
}

size_t nuif_hash_swap_scratch_bytes(const nuif_hash_params* p) {
    if (!p || check_params(p)) return 0;
    const size_t n = layout_of(*p).nentries;
    return 2 * align256(n * 4) + align256(nuif_scan_scratch_bytes(n)) + 256;
}

int nuif_hash_swap_list(void* scene, const nuif_hash_params* p, int which, int32_t* ids, uint32_t* count,
                        void* scratch, nuif_stream_t s) {
    NUIF_REQUIRE_INIT();
    int rc = check_params(p);
    if (rc) return rc;
    NUIF_CHECK_ARG(scene && ids && count && scratch, "nuif_hash_swap_list: null argument");
    NUIF_CHECK_ARG(which == NUIF_HASH_SWAP_IN || which == NUIF_HASH_SWAP_OUT, "nuif_hash_swap_list: bad list %d",
                   which);
    const size_t n = layout_of(*p).nentries;
    Scene sc = scene_of(scene, *p);
    hipStream_t st = stream_of(s);
    char* b = static_cast<char*>(scratch);
    uint32_t* flag = reinterpret_cast<uint32_t*>(b);
    uint32_t* pos = reinterpret_cast<uint32_t*>(b + align256(n * 4));
    void* scan = b + 2 * align256(n * 4);
    const unsigned grid = (unsigned)std::min<size_t>(div_up(n, 256), 4096);
    if (which == NUIF_HASH_SWAP_IN)
        hash_swap_in_flags_all_kernel<<<grid, 256, 0, st>>>(sc, flag, n);
    else
        hash_swap_out_flags_kernel<<<grid, 256, 0, st>>>(sc, flag, n);
    NUIF_LAUNCHED("hash_swap_flags_kernel");
    rc = launch_scan(flag, false, pos, n, count, scan, st);
    if (rc) return rc;
    if (which == NUIF_HASH_SWAP_IN)
        hash_swap_in_compact_kernel<<<grid, 256, 0, st>>>(sc, flag, pos, (uint32_t)n, ids);
    else
        hash_swap_out_compact_kernel<<<grid, 256, 0, st>>>(flag, pos, n, ids);
    NUIF_LAUNCHED("hash_swap_compact_kernel");
    return NUIF_OK;
}

int nuif_hash_swap_in(void* scene, const nuif_hash_params* p, const int32_t* ids, uint32_t n, const uint32_t* words,
                      const uint8_t* colors, nuif_stream_t s) {
    NUIF_REQUIRE_INIT();
    int rc = check_params(p);
    if (rc) return rc;
    if (n == 0) return NUIF_OK;
    NUIF_CHECK_ARG(scene && ids && words, "nuif_hash_swap_in: null argument");
    Scene sc = scene_of(scene, *p);
    const unsigned grid = std::min<unsigned>(n, (unsigned)kWalkGrid);
    const auto* c4 = reinterpret_cast<const uchar4*>(colors);
    if (p->storage == NUIF_TSDF_S16)
        hash_swap_in_kernel<VS16><<<grid, 64, 0, stream_of(s)>>>(sc, ids, n, words, c4, p->weight_max);
    else
        hash_swap_in_kernel<VF16><<<grid, 64, 0, stream_of(s)>>>(sc, ids, n, words, c4, p->weight_max);
    NUIF_LAUNCHED("hash_swap_in_kernel");
    return NUIF_OK;
}

int nuif_hash_swap_out(void* scene, const nuif_hash_params* p, const int32_t* ids, uint32_t n, uint32_t* words,
                       uint8_t* colors, nuif_stream_t s) {
    NUIF_REQUIRE_INIT();
    int rc = check_params(p);
    if (rc) return rc;
    if (n == 0) return NUIF_OK;
    NUIF_CHECK_ARG(scene && ids && words, "nuif_hash_swap_out: null argument");
    Scene sc = scene_of(scene, *p);
    const unsigned grid = std::min<unsigned>(n, (unsigned)kWalkGrid);
    hash_swap_out_kernel<<<grid, 64, 0, stream_of(s)>>>(sc, ids, n, words, reinterpret_cast<uchar4*>(colors));
    NUIF_LAUNCHED("hash_swap_out_kernel");
    return NUIF_OK;
}

int nuif_hash_stats(const void* scene, uint32_t* out4, nuif_stream_t s) {
    NUIF_REQUIRE_INIT();
    NUIF_CHECK_ARG(scene && out4, "nuif_hash_stats: null argument");
    HashHeader h;
    NUIF_HIP(hipMemcpyAsync(&h, scene, sizeof(h), hipMemcpyDeviceToHost, stream_of(s)));
    NUIF_HIP(hipStreamSynchronize(stream_of(s)));
    out4[0] = h.nbricks;
    out4[1] = h.nvisible;
    out4[2] = h.failed;
    out4[3] = h.nentries;
    return NUIF_OK;
}

}  // extern "C"
