// TEST INFRASTRUCTURE ONLY (see oracle_common.h).  CPU restatement of the
// reference's voxel-hashing scene (OCL/NuiKinfuOpenCLHashScene.cpp with the
// kernels of K/KinfuHashScene.cl, K/KinfuHashRaycast.cl and
// K/KinfuHashVisualization.cl) in float arithmetic that follows the kernels
// operation by operation, plus the per-frame orchestration of RunTracking with
// a hash scene.
//
// The reference's racy or undefined behaviour is replaced by the intended
// deterministic result:
//   * every block touched by a pixel's truncation band is allocated in the
//     frame that touches it (the reference allocates one block per bucket per
//     frame and the winner is racy; its mask 0x14ffff also leaves most buckets
//     unreachable) -- the allocated set is the union of the touched sets;
//   * every visible entry enters the visible list (the reference scans only
//     the first 2^20 entries);
//   * the voxel pool starts at sdf 0, weight 0, colour 0 (the reset kernel's
//     intent; it indexes gidx+i instead of gidx*512+i);
//   * the ray-range splat is an exact min / max (the reference's is a
//     non-atomic read-modify-write);
//   * a ray sample in a missing block reads sdf +1 (the reference reads the
//     pool at index -1 first);
//   * out-of-range float -> int conversions saturate and NaN converts to 0,
//     as the GPU's conversion instructions do (undefined in OpenCL C).
// Kept as the reference computes them: the far bound of the ray range is the
// running near bound's max with the corner depth (K/KinfuHashRaycast.cl:49),
// the ray step scale uses the truncation (0.06) while sdf values are
// normalised by truncScale (0.02), the integration pixel is (int)(p + 0.5)
// (round toward zero), and marching cubes ignores voxel weights.
// Vendor-approximate functions (fast_length, normalize) are evaluated
// correctly rounded: parity at their last bit is unpinned.

#include "kinfu_core.h"

#include <algorithm>
#include <unordered_map>

namespace hashor {
using namespace oracle;

constexpr int kSB = 8;     // SDF_BLOCK_SIZE (K/hashing_gpu_def.h:9)
constexpr int kSB3 = 512;  // SDF_BLOCK_SIZE3

inline uint64_t pack_key(int x, int y, int z) {
    return (1ull << 48) | ((uint64_t)(uint16_t)(int16_t)x << 32) | ((uint64_t)(uint16_t)(int16_t)y << 16) |
           (uint64_t)(uint16_t)(int16_t)z;
}
inline void unpack_key(uint64_t k, int* p) {
    p[0] = (int16_t)(uint16_t)(k >> 32);
    p[1] = (int16_t)(uint16_t)(k >> 16);
    p[2] = (int16_t)(uint16_t)k;
}
// key order used for every exported list: (x, y, z) lexicographic on signed values
inline bool key_less(uint64_t a, uint64_t b) {
    int pa[3], pb[3];
    unpack_key(a, pa);
    unpack_key(b, pb);
    if (pa[0] != pb[0]) return pa[0] < pb[0];
    if (pa[1] != pb[1]) return pa[1] < pb[1];
    return pa[2] < pb[2];
}

// float -> int as the GPU converts: round toward zero, saturating, NaN -> 0
inline int cvt_sat(float v) {
    if (v != v) return 0;
    if (v >= 2147483648.0f) return INT32_MAX;
    if (v <= -2147483648.0f) return INT32_MIN;
    return (int)v;
}

inline float flen(V3<float> v) { return std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z); }

// K/cameraUtils.cl:9-17 cameraToKinectScreen
inline void to_screen(const Cam<float>& cam, V3<float> p, float* sx, float* sy) {
    *sx = p.x * cam.fx / p.z + cam.cx;
    *sy = p.y * cam.fy / p.z + cam.cy;
}
// K/cameraUtils.cl:43-50 kinectDepthToSkeleton
inline V3<float> depth_to_skeleton(const Cam<float>& cam, int ux, int uy, float d) {
    const float x = ((float)ux - cam.cx) / cam.fx;
    const float y = ((float)uy - cam.cy) / cam.fy;
    return mk(d * x, d * y, d);
}

struct Brick {
    int16_t sdf[kSB3];
    uint8_t w[kSB3];
    uint8_t rgb[kSB3][3];
};

struct HashCfg {
    float vs;          // m_virtualVoxelSize
    float trunc_scale; // m_truncScale (sdf normalisation, band half-width)
    float truncation;  // m_truncation (raycast step scale)
    int wmax;          // m_integrationWeightMax
    int capacity;      // bricks
};

struct HashScene {
    HashCfg c;
    std::unordered_map<uint64_t, int> index;  // block key -> brick id (-1: swapped out, no brick)
    std::vector<Brick> bricks;
    // host swapping (OCL/NuiKinfuOpenCLHashGlobalCache.cpp): swap state per key (0 / 1 / 2),
    // the host store of swapped-out blocks, recycled brick ids
    bool swapping = false;
    std::unordered_map<uint64_t, uint8_t> swap_state;
    std::unordered_map<uint64_t, Brick> store;
    std::vector<int> free_ids;
    std::vector<uint64_t> visible;  // visible list of the last integration (sorted by key)
    std::vector<float> range;       // ray range image of the last raycast (near, far)

    explicit HashScene(const HashCfg& cfg) : c(cfg) {}

    // K/KinfuHashScene.cl:52-136: blocks touched by one pixel's truncation band
    void touched_pixel(const Cam<float>& cam, const Rigid<float>& g, int x, int y, float Dp, float inv_block,
                       std::vector<uint64_t>& out) const {
        if (std::isnan(Dp)) return;
        const V3<float> pc = depth_to_skeleton(cam, x, y, Dp);
        const float norm = c.trunc_scale / flen(pc);
        V3<float> rmin = transform(g, pc * (1.0f - norm)) * inv_block;
        const V3<float> rmax = transform(g, pc * (1.0f + norm)) * inv_block;
        V3<float> dir = rmax - rmin;
        const float rn = flen(dir);
        const int steps = cvt_sat(std::ceil(2.0f * rn));
        const float div = (float)(steps - 1);
        dir = mk(dir.x / div, dir.y / div, dir.z / div);
        for (int i = 0; i < steps; ++i) {
            out.push_back(pack_key((int)(int16_t)cvt_sat(std::floor(rmin.x)), (int)(int16_t)cvt_sat(std::floor(rmin.y)),
                                   (int)(int16_t)cvt_sat(std::floor(rmin.z))));
            rmin = rmin + dir;
        }
    }

    // sorted unique set of blocks touched by a frame
    std::vector<uint64_t> touched(const float* depth, const Cam<float>& cam, const Rigid<float>& g) const {
        // OCL/NuiKinfuOpenCLHashScene.cpp:205: 1 / (voxel size * block size) in float
        const float inv_block = 1.0f / (c.vs * (float)kSB);
        std::vector<uint64_t> all;
        for (int y = 0; y < cam.H; ++y)
            for (int x = 0; x < cam.W; ++x) touched_pixel(cam, g, x, y, depth[(size_t)y * cam.W + x], inv_block, all);
        std::sort(all.begin(), all.end());
        all.erase(std::unique(all.begin(), all.end()), all.end());
        std::sort(all.begin(), all.end(), key_less);
        return all;
    }

    // K/KinfuHashScene.cl:197-260 checkPointVisibility / checkBlockVisibility
    static bool point_visible(const Cam<float>& cam, const Rigid<float>& g, V3<float> p) {
        const V3<float> q = transform_inverse(g, p);
        if (isnan3(q)) return false;
        float sx, sy;
        to_screen(cam, q, &sx, &sy);
        const int u = cvt_sat(sx + 0.5f), v = cvt_sat(sy + 0.5f);
        return u >= 0 && u < cam.W && v >= 0 && v < cam.H;
    }
    bool block_visible(const Cam<float>& cam, const Rigid<float>& g, uint64_t key) const {
        int p[3];
        unpack_key(key, p);
        const float f = (float)kSB * c.vs;
        V3<float> pt = mk((float)p[0] * f, (float)p[1] * f, (float)p[2] * f);
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

    // OCL/NuiKinfuOpenCLHashScene.cpp:404-424 AllocateSceneFromDepth: allocation
    // of the touched blocks, then the visible list = touched blocks + previously
    // visible blocks that still pass the corner test.  Returns the touched count.
    size_t allocate_and_select(const float* depth, const Cam<float>& cam, const Rigid<float>& g) {
        const std::vector<uint64_t> t = touched(depth, cam, g);
        for (uint64_t k : t) {
            if (index.count(k)) continue;
            const int b = take_brick();
            if (b < 0) continue;
            index.emplace(k, b);
        }
        std::vector<uint64_t> vis;
        for (uint64_t k : t)
            if (index.count(k)) vis.push_back(k);
        std::vector<uint64_t> tset(t);
        std::sort(tset.begin(), tset.end());
        for (uint64_t k : visible)
            if (!std::binary_search(tset.begin(), tset.end(), k) && block_visible(cam, g, k)) vis.push_back(k);
        std::sort(vis.begin(), vis.end(), key_less);
        visible.swap(vis);
        return t.size();
    }

    // a zeroed brick: a recycled one, else a new one while the pool has room; -1 when full
    int take_brick() {
        if (!free_ids.empty()) {
            const int b = free_ids.back();
            free_ids.pop_back();
            return b;
        }
        if ((int)bricks.size() >= c.capacity) return -1;
        bricks.emplace_back();
        std::memset(&bricks.back(), 0, sizeof(Brick));
        return (int)bricks.size() - 1;
    }

    // K/KinfuHashScene.cl:383-384: visible blocks not yet synchronised want their stored
    // data; K/KinfuHashSwapping.cl:6-21: visible blocks without a brick take one
    void swap_prepare() {
        for (uint64_t k : visible) {
            uint8_t& st = swap_state[k];
            if (st != 2) st = 1;
            int& b = index.at(k);
            if (b == -1) b = take_brick();
        }
    }

    // K/KinfuHashSwapping.cl:57-118 integrateOldIntoActiveData_kernel: the stored block
    // (zeros when none) weight-averaged into the brick of every visible block in state 1
    void swap_in() {
        for (uint64_t k : visible) {
            uint8_t& st = swap_state[k];
            const int id = index.at(k);
            if (st != 1 || id < 0) continue;
            Brick old;
            std::memset(&old, 0, sizeof(old));
            auto it = store.find(k);
            if (it != store.end()) old = it->second;
            Brick& cur = bricks[id];
            for (int v = 0; v < kSB3; ++v) {
                const float oldF = (float)old.sdf[v] / 32767.0f, oldW = (float)old.w[v];
                const float newF = (float)cur.sdf[v] / 32767.0f, newW = (float)cur.w[v];
                const float combinedW = oldW + newW;
                float f = oldW * oldF + newW * newF;
                f /= combinedW;
                cur.sdf[v] = (int16_t)cvt_sat(f * 32767.0f);
                // convert_uchar of a weight sum above 255 is undefined in the reference: saturated
                cur.w[v] = (uint8_t)std::min(cvt_sat(combinedW), c.wmax);
                for (int ch = 0; ch < 3; ++ch) {
                    const float x = (oldW * (float)old.rgb[v][ch] + newW * (float)cur.rgb[v][ch]) / combinedW;
                    cur.rgb[v][ch] = (uint8_t)cvt_sat(std::min(std::max(x, 0.0f), 255.f));
                }
            }
            st = 2;
        }
    }

    // K/KinfuHashSwapping.cl:120-180 buildSwapOutList / moveActiveDataToTransferBuffer /
    // cleanVoxelMemory: synchronised blocks with a brick that this frame does not see go to
    // the store; their bricks are zeroed and recycled (every brick; the reference keeps
    // the holder of brick 0)
    void swap_out() {
        std::vector<uint64_t> vis(visible);
        std::sort(vis.begin(), vis.end());
        for (auto& kv : index) {
            uint8_t& st = swap_state[kv.first];
            if (st != 2 || kv.second < 0 || std::binary_search(vis.begin(), vis.end(), kv.first)) continue;
            store[kv.first] = bricks[kv.second];
            std::memset(&bricks[kv.second], 0, sizeof(Brick));
            free_ids.push_back(kv.second);
            kv.second = -1;
            st = 0;
        }
    }

    // OCL/NuiKinfuOpenCLHashScene.cpp:533-583 integrateVolume: allocation and visible list,
    // integration, and with swapping the global cache's swap-in and swap-out (:576-580)
    long long integrate_frame(const float* depth, const uint8_t* bgra, const Cam<float>& cam, const Rigid<float>& g) {
        allocate_and_select(depth, cam, g);
        if (swapping) swap_prepare();
        const long long n = integrate(depth, bgra, cam, g);
        if (swapping) {
            swap_in();
            swap_out();
        }
        return n;
    }

    // K/KinfuHashScene.cl:387-470 integrateIntoScene_kernel over the visible list
    long long integrate(const float* depth, const uint8_t* bgra, const Cam<float>& cam, const Rigid<float>& g) {
        long long nupd = 0;
        const int n = (int)visible.size();
#pragma omp parallel for reduction(+ : nupd) schedule(dynamic, 16)
        for (int i = 0; i < n; ++i) {
            const uint64_t key = visible[i];
            const int bid = index.at(key);
            if (bid < 0) continue;  // no brick (pool full)
            Brick& b = bricks[bid];
            int p[3];
            unpack_key(key, p);
            for (int lid = 0; lid < kSB3; ++lid) {
                if (b.w[lid] == c.wmax) continue;
                const int lx = lid % 8, ly = (lid % 64) / 8, lz = lid / 64;
                const V3<float> pm = mk((float)(p[0] * kSB + lx), (float)(p[1] * kSB + ly), (float)(p[2] * kSB + lz)) * c.vs;
                const V3<float> pc = transform_inverse(g, pm);
                if (pc.z <= 0) continue;
                float sx, sy;
                to_screen(cam, pc, &sx, &sy);
                const int u = cvt_sat(sx + 0.5f), v = cvt_sat(sy + 0.5f);
                if (u < 0 || u >= cam.W || v < 0 || v >= cam.H) continue;
                const int sid = v * cam.W + u;
                const float d = depth[sid];
                if (std::isnan(d)) continue;
                const float eta = d - pc.z;
                const float sdf = eta / c.trunc_scale;
                if (eta < -c.trunc_scale) continue;
                const float oldF = (float)b.sdf[lid] / 32767.0f;
                const float oldW = (float)b.w[lid];
                float newF = std::min(1.0f, sdf);
                const float newW = 1.0f;
                const float cw = oldW + newW;
                if (bgra && eta < c.trunc_scale && std::fabs(sdf) < 0.25f) {
                    const uint8_t* col = bgra + 4 * (size_t)sid;
                    if (col[3] == 255) {
                        // colour[0] <- s2 (R of BGRA), [1] <- s1, [2] <- s0
                        const uint8_t src[3] = {col[2], col[1], col[0]};
                        for (int ch = 0; ch < 3; ++ch) {
                            const float nc = (oldW * (float)b.rgb[lid][ch] + newW * (float)src[ch]) / cw;
                            b.rgb[lid][ch] = (uint8_t)cvt_sat(std::min(std::max(nc, 0.0f), 255.0f));
                        }
                    }
                }
                newF = oldW * oldF + newW * newF;
                newF /= cw;
                b.sdf[lid] = (int16_t)cvt_sat(newF * 32767.0f);
                b.w[lid] = (uint8_t)std::min(cvt_sat(cw), c.wmax);
                ++nupd;
            }
        }
        return nupd;
    }

    // ---- lookups (K/KinfuHashRaycast.cl:78-124)
    const Brick* brick_of(int bx, int by, int bz) const {
        auto it = index.find(pack_key(bx, by, bz));
        return it == index.end() || it->second < 0 ? nullptr : &bricks[it->second];
    }
    static int floor_div8(int v) { return v < 0 ? (v - 7) / 8 : v / 8; }  // virtualPosToVoxelBlock
    // sdf of a voxel (world voxel units), +1 when its block is missing
    bool voxel(int x, int y, int z, float* sdf, const uint8_t** rgb = nullptr) const {
        const int bx = floor_div8(x), by = floor_div8(y), bz = floor_div8(z);
        const Brick* b = brick_of(bx, by, bz);
        if (!b) {
            *sdf = 1.0f;
            return false;
        }
        const int lid = (x - bx * 8) + (y - by * 8) * 8 + (z - bz * 8) * 64;
        *sdf = (float)b->sdf[lid] / 32767.0f;
        if (rgb) *rgb = b->rgb[lid];
        return true;
    }
    float read_sdf(int x, int y, int z) const {
        float s;
        voxel(x, y, z, &s);
        return s;
    }
    // K/KinfuHashRaycast.cl:125-153 readInterpolatedVoxelSDF
    float interp_sdf(V3<float> p) const {
        const int x = cvt_sat(std::floor(p.x)), y = cvt_sat(std::floor(p.y)), z = cvt_sat(std::floor(p.z));
        const V3<float> co = mk(p.x - (float)x, p.y - (float)y, p.z - (float)z);
        float v1 = read_sdf(x, y, z), v2 = read_sdf(x + 1, y, z);
        float res1 = (1.0f - co.x) * v1 + co.x * v2;
        v1 = read_sdf(x, y + 1, z);
        v2 = read_sdf(x + 1, y + 1, z);
        res1 = (1.0f - co.y) * res1 + co.y * ((1.0f - co.x) * v1 + co.x * v2);
        v1 = read_sdf(x, y, z + 1);
        v2 = read_sdf(x + 1, y, z + 1);
        float res2 = (1.0f - co.x) * v1 + co.x * v2;
        v1 = read_sdf(x, y + 1, z + 1);
        v2 = read_sdf(x + 1, y + 1, z + 1);
        res2 = (1.0f - co.y) * res2 + co.y * ((1.0f - co.x) * v1 + co.x * v2);
        return (1.0f - co.z) * res1 + co.z * res2;
    }

    // K/KinfuHashRaycast.cl:4-76 project_minmax_depths_kernel over the visible list,
    // splatted with exact min / max into an image initialised to (999999.9, 0.05)
    // (OCL/NuiKinfuOpenCLAcceleratedFeedbackFrame.cpp:45-49)
    void ray_ranges(const Cam<float>& cam, const Rigid<float>& g) {
        range.assign((size_t)cam.W * cam.H * 2, 0.f);
        for (size_t i = 0; i < range.size(); i += 2) {
            range[i] = 999999.9f;
            range[i + 1] = 0.05f;
        }
        for (uint64_t key : visible) {
            int p[3];
            unpack_key(key, p);
            int ulx = cam.W, uly = cam.H, lrx = 0, lry = 0;
            float zx = 999999.9f, zy = 0.05f;
            for (int corner = 0; corner < 8; ++corner) {
                const int tx = p[0] + ((corner & 1) ? 1 : 0);
                const int ty = p[1] + ((corner & 2) ? 1 : 0);
                const int tz = p[2] + ((corner & 4) ? 1 : 0);
                const V3<float> w = mk((float)tx * (float)kSB * c.vs, (float)ty * (float)kSB * c.vs,
                                       (float)tz * (float)kSB * c.vs);
                const V3<float> q = transform_inverse(g, w);
                if (isnan3(q) || q.z < 1e-6f) continue;
                float sx, sy;
                to_screen(cam, q, &sx, &sy);
                ulx = std::min(ulx, cvt_sat(std::floor(sx)));
                lrx = std::max(lrx, cvt_sat(std::ceil(sx)));
                uly = std::min(uly, cvt_sat(std::floor(sy)));
                lry = std::max(lry, cvt_sat(std::ceil(sy)));
                zx = std::fmin(zx, q.z);
                zy = std::fmax(zx, q.z);  // the reference's far bound
            }
            if (ulx < 0) ulx = 0;
            if (uly < 0) uly = 0;
            if (lrx >= cam.W) lrx = cam.W - 1;
            if (lry >= cam.H) lry = cam.H - 1;
            if (zx < 0.05f) zx = 0.05f;
            for (int by = uly; by <= lry; ++by)
                for (int bx = ulx; bx <= lrx; ++bx) {
                    float* r = &range[2 * ((size_t)by * cam.W + bx)];
                    r[0] = std::fmin(r[0], zx);
                    r[1] = std::fmax(r[1], zy);
                }
        }
    }

    // K/KinfuHashRaycast.cl:462-470 raycast_kernel: castRay at every pixel
    void raycast(const Cam<float>& cam, const Rigid<float>& g, float* vmap, float* nmap, uint8_t* cmap) const {
#pragma omp parallel for schedule(dynamic, 4)
        for (int y = 0; y < cam.H; ++y)
            for (int x = 0; x < cam.W; ++x) cast_ray(cam, g, x, y, vmap, nmap, cmap);
    }

    // K/KinfuHashRaycast.cl:155-448 castRay: world vertex / normal / colour of pixel (x, y)
    void cast_ray(const Cam<float>& cam, const Rigid<float>& g, int x, int y, float* vmap, float* nmap,
                  uint8_t* cmap) const {
        const float ivs = 1.0f / c.vs;
        const float step_scale = c.truncation * ivs * 0.6f;
        {
            {
                const size_t pid = (size_t)y * cam.W + x;
                const float* mm = &range[2 * pid];
                V3<float> ps = depth_to_skeleton(cam, x, y, mm[0]);
                float total = flen(ps) * ivs;
                ps = transform(g, ps) * ivs;
                V3<float> pe = depth_to_skeleton(cam, x, y, mm[1]);
                const float total_max = flen(pe) * ivs;
                pe = transform(g, pe) * ivs;
                const V3<float> dir = normalize3(pe - ps);
                V3<float> pt = ps;
                float sdf = 1.0f, step;
                while (total < total_max) {
                    const int vx = cvt_sat(cl_round(pt.x)), vy = cvt_sat(cl_round(pt.y)), vz = cvt_sat(cl_round(pt.z));
                    float s;
                    if (!voxel(vx, vy, vz, &s)) {
                        sdf = 1.0f;
                        step = (float)kSB;
                    } else {
                        sdf = s;
                        if (sdf <= 0.1f && sdf >= -0.5f) sdf = interp_sdf(pt);
                        if (sdf <= 0.0f) break;
                        step = std::max(sdf * step_scale, 1.0f);
                    }
                    pt = pt + dir * step;
                    total += step;
                }
                if (!(sdf <= 0.0f)) {
                    store3(vmap, pid, mk(qnan<float>(), qnan<float>(), qnan<float>()));
                    if (nmap) store3(nmap, pid, mk(qnan<float>(), qnan<float>(), qnan<float>()));
                    if (cmap) std::memset(cmap + 4 * pid, 0, 4);
                    return;
                }
                step = sdf * step_scale;
                pt = pt + dir * step;
                const int vx = cvt_sat(std::floor(pt.x)), vy = cvt_sat(std::floor(pt.y)), vz = cvt_sat(std::floor(pt.z));
                const V3<float> co = mk(pt.x - (float)vx, pt.y - (float)vy, pt.z - (float)vz);
                const V3<float> nc = mk(1.0f - co.x, 1.0f - co.y, 1.0f - co.z);
                // corners front (z) x,y: (0,0) (1,0) (0,1) (1,1); back the same at z+1
                static const int cx8[8] = {0, 1, 0, 1, 0, 1, 0, 1}, cy8[8] = {0, 0, 1, 1, 0, 0, 1, 1},
                                 cz8[8] = {0, 0, 0, 0, 1, 1, 1, 1};
                float val[8];
                uint8_t col[4] = {0, 0, 0, 0};
                float nearest = 1.0e20f;
                for (int k = 0; k < 8; ++k) {
                    const uint8_t* rgb = nullptr;
                    const bool found = voxel(vx + cx8[k], vy + cy8[k], vz + cz8[k], &val[k], &rgb);
                    const V3<float> dv = mk(cx8[k] ? nc.x : co.x, cy8[k] ? nc.y : co.y, cz8[k] ? nc.z : co.z);
                    const float dist = flen(dv);
                    if (cmap && found && (k == 0 || nearest > dist)) {
                        nearest = dist;
                        col[0] = rgb[0];
                        col[1] = rgb[1];
                        col[2] = rgb[2];
                        col[3] = 255;
                    }
                }
                const float fx_ = val[0], fy_ = val[1], fz_ = val[2], fw_ = val[3];
                const float bx_ = val[4], by_ = val[5], bz_ = val[6], bw_ = val[7];
                sdf = fx_ * nc.x * nc.y * nc.z + fy_ * co.x * nc.y * nc.z + fz_ * nc.x * co.y * nc.z +
                      fw_ * co.x * co.y * nc.z + bx_ * nc.x * nc.y * co.z + by_ * co.x * nc.y * co.z +
                      bz_ * nc.x * co.y * co.z + bw_ * co.x * co.y * co.z;
                step = sdf * step_scale;
                pt = pt + dir * step;
                store3(vmap, pid, pt * c.vs);
                if (nmap) {
                    V3<float> ret;
                    float p1, p2, v1, t0, t1, t2, t3;
                    // gradient x
                    p1 = fx_ * nc.y * nc.z + fz_ * co.y * nc.z + bx_ * nc.y * co.z + bz_ * co.y * co.z;
                    t0 = read_sdf(vx - 1, vy, vz), t1 = read_sdf(vx - 1, vy + 1, vz);
                    t2 = read_sdf(vx - 1, vy, vz + 1), t3 = read_sdf(vx - 1, vy + 1, vz + 1);
                    p2 = t0 * nc.y * nc.z + t1 * co.y * nc.z + t2 * nc.y * co.z + t3 * co.y * co.z;
                    v1 = p1 * co.x + p2 * nc.x;
                    p1 = fy_ * nc.y * nc.z + fw_ * co.y * nc.z + by_ * nc.y * co.z + bw_ * co.y * co.z;
                    t0 = read_sdf(vx + 2, vy, vz), t1 = read_sdf(vx + 2, vy + 1, vz);
                    t2 = read_sdf(vx + 2, vy, vz + 1), t3 = read_sdf(vx + 2, vy + 1, vz + 1);
                    p2 = t0 * nc.y * nc.z + t1 * co.y * nc.z + t2 * nc.y * co.z + t3 * co.y * co.z;
                    ret.x = p1 * nc.x + p2 * co.x - v1;
                    // gradient y
                    p1 = fx_ * nc.x * nc.z + fy_ * co.x * nc.z + bx_ * nc.x * co.z + by_ * co.x * co.z;
                    t0 = read_sdf(vx, vy - 1, vz), t1 = read_sdf(vx + 1, vy - 1, vz);
                    t2 = read_sdf(vx, vy - 1, vz + 1), t3 = read_sdf(vx + 1, vy - 1, vz + 1);
                    p2 = t0 * nc.x * nc.z + t1 * co.x * nc.z + t2 * nc.x * co.z + t3 * co.x * co.z;
                    v1 = p1 * co.y + p2 * nc.y;
                    p1 = fz_ * nc.x * nc.z + fw_ * co.x * nc.z + bz_ * nc.x * co.z + bw_ * co.x * co.z;
                    t0 = read_sdf(vx, vy + 2, vz), t1 = read_sdf(vx + 1, vy + 2, vz);
                    t2 = read_sdf(vx, vy + 2, vz + 1), t3 = read_sdf(vx + 1, vy + 2, vz + 1);
                    p2 = t0 * nc.x * nc.z + t1 * co.x * nc.z + t2 * nc.x * co.z + t3 * co.x * co.z;
                    ret.y = p1 * nc.y + p2 * co.y - v1;
                    // gradient z
                    p1 = fx_ * nc.x * nc.y + fy_ * co.x * nc.y + fz_ * nc.x * co.y + fw_ * co.x * co.y;
                    t0 = read_sdf(vx, vy, vz - 1), t1 = read_sdf(vx + 1, vy, vz - 1);
                    t2 = read_sdf(vx, vy + 1, vz - 1), t3 = read_sdf(vx + 1, vy + 1, vz - 1);
                    p2 = t0 * nc.x * nc.y + t1 * co.x * nc.y + t2 * nc.x * co.y + t3 * co.x * co.y;
                    v1 = p1 * co.z + p2 * nc.z;
                    p1 = bx_ * nc.x * nc.y + by_ * co.x * nc.y + bz_ * nc.x * co.y + bw_ * co.x * co.y;
                    t0 = read_sdf(vx, vy, vz + 2), t1 = read_sdf(vx + 1, vy, vz + 2);
                    t2 = read_sdf(vx, vy + 1, vz + 2), t3 = read_sdf(vx + 1, vy + 1, vz + 2);
                    p2 = t0 * nc.x * nc.y + t1 * co.x * nc.y + t2 * nc.x * co.y + t3 * co.x * co.y;
                    ret.z = p1 * nc.z + p2 * co.z - v1;
                    store3(nmap, pid, normalize3(ret));
                }
                if (cmap) std::memcpy(cmap + 4 * pid, col, 4);
            }
        }
    }

    // OCL/NuiKinfuOpenCLHashScene.cpp:991-1012 ForwardRender after the ray ranges:
    //  * forwardProject_kernel (K/KinfuHashRaycast.cl:473-506) moves every valid point of
    //    the previous vertex map (and its colour) to the pixel it projects to under the
    //    current pose;
    //  * findMissingPoints_kernel (:508-529) keeps the forwarded point where one landed and
    //    flags the pixel for a ray where none did and its ray range is not empty;
    //  * genericRaycastMissingPoints_kernel (:531-559) runs castRay (no normal) on the
    //    flagged pixels;
    //  * Vertex2Normal (:943-989) recomputes the normals from the vertex map (threshold 0.3).
    // reference_transform selects the reference's projection (K/KinfuHashRaycast.cl:473-506): camera->world applied
    // to world points, scaled by the voxel size, pixel = (uint)(p + 0.5) with negative and
    // NaN values converting to 0 (the saturating float->uint conversion); otherwise the
    // world->camera transform, z > 0 and p + 0.5 >= 0.  Where several points land on one
    // pixel (a race in the reference) the nearest (smallest z, then smallest source index)
    // wins.  Returns the number of flagged pixels.
    static uint32_t zkey(float z) {
        uint32_t b;
        std::memcpy(&b, &z, 4);
        return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
    }
    static uint32_t cvt_u32_sat(float v) {
        if (!(v > 0.0f)) return 0u;
        if (v >= 4294967296.0f) return 0xffffffffu;
        return (uint32_t)v;
    }
    size_t forward_render(const Cam<float>& cam, const Rigid<float>& g, float* vmap, float* nmap, uint8_t* cmap,
                          bool reference_transform) const {
        const size_t n = (size_t)cam.W * cam.H;
        std::vector<uint64_t> keys(n, ~0ull);
        for (size_t i = 0; i < n; ++i) {
            const V3<float> v = load3(vmap, i);
            if (isnan3(v)) continue;
            V3<float> q;
            float sx, sy;
            uint32_t ux, uy;
            if (reference_transform) {
                q = transform(g, v) * c.vs;
                to_screen(cam, q, &sx, &sy);
                ux = cvt_u32_sat(sx + 0.5f);
                uy = cvt_u32_sat(sy + 0.5f);
            } else {
                q = transform_inverse(g, v);
                if (!(q.z > 0.0f)) continue;
                to_screen(cam, q, &sx, &sy);
                sx += 0.5f;
                sy += 0.5f;
                if (!(sx >= 0.0f && sy >= 0.0f)) continue;
                ux = cvt_u32_sat(sx);
                uy = cvt_u32_sat(sy);
            }
            if (ux >= (uint32_t)cam.W || uy >= (uint32_t)cam.H) continue;
            const size_t dst = (size_t)uy * cam.W + ux;
            keys[dst] = std::min(keys[dst], ((uint64_t)zkey(q.z) << 32) | (uint64_t)i);
        }
        std::vector<float> fwd(3 * n);
        std::vector<uint8_t> col(cmap ? 4 * n : 0);
        for (size_t d = 0; d < n; ++d) {
            const bool hit = keys[d] != ~0ull;
            const size_t src = hit ? (size_t)(uint32_t)keys[d] : d;
            store3(fwd.data(), d, hit ? load3(vmap, src) : mk(qnan<float>(), qnan<float>(), qnan<float>()));
            if (cmap) std::memcpy(&col[4 * d], cmap + 4 * src, 4);
        }
        std::vector<size_t> needed;
        for (size_t d = 0; d < n; ++d) {
            const V3<float> f = load3(fwd.data(), d);
            if (isnan3(f) && range[2 * d] < range[2 * d + 1])
                needed.push_back(d);
            else
                store3(vmap, d, f);
            if (cmap) std::memcpy(cmap + 4 * d, &col[4 * d], 4);
        }
#pragma omp parallel for schedule(dynamic, 64)
        for (long long k = 0; k < (long long)needed.size(); ++k)
            cast_ray(cam, g, (int)(needed[k] % cam.W), (int)(needed[k] / cam.W), vmap, nullptr, cmap);
        normals_simple<float>(vmap, nmap, cam.W, cam.H, 0.3f);
        return needed.size();
    }

    // K/KinfuHashVisualization.cl:88-271 meshHashScene_kernel over the visible list
    // (sorted by key), voxels in z, y, x order; positions in metres.
    size_t mesh(const int* edge_table, const int* tri_table, float* out, size_t cap) const {
        static const int cofs[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                       {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
        static const int edges[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {4, 5}, {5, 6},
                                         {6, 7}, {7, 4}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};
        size_t nv = 0;
        for (uint64_t key : visible) {
            int p[3];
            unpack_key(key, p);
            for (int z = 0; z < kSB; ++z)
                for (int y = 0; y < kSB; ++y)
                    for (int x = 0; x < kSB; ++x) {
                        const int gx = p[0] * kSB + x, gy = p[1] * kSB + y, gz = p[2] * kSB + z;
                        float s[8];
                        V3<float> pts[8];
                        bool ok = true;
                        for (int k = 0; k < 8 && ok; ++k) {
                            const int vx = gx + cofs[k][0], vy = gy + cofs[k][1], vz = gz + cofs[k][2];
                            ok = voxel(vx, vy, vz, &s[k]);
                            pts[k] = mk((float)vx, (float)vy, (float)vz);
                        }
                        if (!ok) continue;
                        int ci = 0;
                        for (int k = 0; k < 8; ++k)
                            if (s[k] < 0) ci |= 1 << k;
                        const int em = edge_table[ci];
                        if (em == 0) continue;
                        V3<float> vl[12];
                        for (int e = 0; e < 12; ++e) {
                            if (!(em & (1 << e))) continue;
                            const int a = edges[e][0], b = edges[e][1];
                            const float va = s[a], vb = s[b];
                            if (std::fabs(0.0f - va) < 0.00001f)
                                vl[e] = pts[a];
                            else if (std::fabs(0.0f - vb) < 0.00001f)
                                vl[e] = pts[b];
                            else if (std::fabs(va - vb) < 0.00001f)
                                vl[e] = pts[a];
                            else
                                vl[e] = pts[a] + (pts[b] - pts[a]) * ((0.0f - va) / (vb - va));
                        }
                        for (int i = 0; i < 16; i += 3) {
                            const int* t = tri_table + ci * 16 + i;
                            if (t[0] < 0 || t[0] >= 12) break;
                            for (int k = 0; k < 3; ++k) {
                                if (out && nv < cap) store3(out, nv, vl[t[k]] * c.vs);
                                ++nv;
                            }
                        }
                    }
        }
        return nv;
    }
};

// ---------------------------------------------------------------- engine

struct HashEngineCfg {
    int W, H;
    float cam[6];
    int filter_radius;
    float sigma_space, sigma_depth, depth_threshold, dist_threshold;
    int levels;
    int iters[8];
    float gate;
    HashCfg hc;
    int forward;  // 0: castRay; 1: forward render; 2: forward render with the reference's transform
};

// src/SLAM/VisualOdometry/NuiKinfuTrackingEngine.cpp:68-142 with the hash scene:
// the scene integrates the frame's unscaled passing-filter depth and its raycast
// (castRay) produces the world vertex / normal maps of the next ICP.
struct HashEngine {
    HashEngineCfg c;
    Cam<float> cam;
    HashScene scene;
    std::vector<float> g, depth, vref, nref;
    std::vector<std::vector<float>> dpyr, vpyr;
    float R[9], t[3], Rlast[9], tlast[3];
    int frames = 0;
    IcpResult last{1, 0, 0, 0};
    long long nupd_last = 0;

    explicit HashEngine(const HashEngineCfg& cfg) : c(cfg), scene(cfg.hc) {
        cam.fx = c.cam[0];
        cam.fy = c.cam[1];
        cam.fx_inv = 1.0f / cam.fx;
        cam.fy_inv = 1.0f / cam.fy;
        cam.cx = c.cam[2];
        cam.cy = c.cam[3];
        cam.W = c.W;
        cam.H = c.H;
        cam.dmin = c.cam[4];
        cam.dmax = c.cam[5];
        g.resize(2 * c.filter_radius + 1);
        gaussian_table<float>(c.filter_radius, c.sigma_space, g.data());
        const size_t n = (size_t)c.W * c.H;
        depth.resize(n);
        vref.resize(3 * n);
        nref.resize(3 * n);
        dpyr.resize(c.levels);
        vpyr.resize(c.levels);
        for (int l = 0; l < c.levels; ++l) {
            const size_t nl = (size_t)(c.W >> l) * (c.H >> l);
            dpyr[l].resize(nl);
            vpyr[l].resize(3 * nl);
        }
        const float I[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
        std::memcpy(R, I, sizeof(I));
        std::memcpy(Rlast, I, sizeof(I));
        t[0] = t[1] = t[2] = tlast[0] = tlast[1] = tlast[2] = 0;
    }

    void integrate(const uint8_t* bgra, const Rigid<float>& pose) {
        nupd_last = scene.integrate_frame(depth.data(), bgra, cam, pose);
    }

    int run(const uint16_t* raw, const uint8_t* bgra) {
        const int W = c.W, H = c.H;
        const size_t n = (size_t)W * H;
        passing_filter<float>(raw, depth.data(), n, cam.dmin, cam.dmax);
        bilateral<float>(depth.data(), dpyr[0].data(), W, H, g.data(), c.filter_radius, c.sigma_depth,
                         c.depth_threshold);
        depth2vertex<float>(dpyr[0].data(), vpyr[0].data(), W, H, cam, 1);
        const Rigid<float> pose = make_rigid<float>(R, t);
        nupd_last = 0;
        if (frames == 0) {
            std::vector<float> nm(3 * n);
            normals_simple<float>(vpyr[0].data(), nm.data(), W, H, c.depth_threshold);
            transform_maps<float>(vpyr[0].data(), nm.data(), vref.data(), nref.data(), n, pose);
            integrate(bgra, pose);
        } else {
            for (int l = 1; l < c.levels; ++l) {
                half_sample<float>(dpyr[l - 1].data(), dpyr[l].data(), W >> l, H >> l, 1, c.depth_threshold);
                depth2vertex<float>(dpyr[l].data(), vpyr[l].data(), W >> l, H >> l, cam, 1 << l);
            }
            IcpConfig ic;
            ic.levels = c.levels;
            for (int l = 0; l < 8; ++l) ic.iters[l] = c.iters[l];
            ic.dist_threshold = c.dist_threshold;
            std::vector<const float*> vp(c.levels);
            for (int l = 0; l < c.levels; ++l) vp[l] = vpyr[l].data();
            float Rn[9], tn[3];
            std::memcpy(Rn, R, sizeof(Rn));
            std::memcpy(tn, t, sizeof(tn));
            last = icp<float>(cam, ic, vp.data(), vref.data(), nref.data(), pose, Rn, tn);
            if (!last.ok) return 0;
            std::memcpy(R, Rn, sizeof(Rn));
            std::memcpy(t, tn, sizeof(tn));
            const Rigid<float> np = make_rigid<float>(R, t);
            float Rinv[9], Rrel[9];
            inverse3(R, Rinv);
            matmul3(Rinv, Rlast, Rrel);
            const float rnorm = rodrigues_norm<float>(Rrel);
            const float d0 = t[0] - tlast[0], d1 = t[1] - tlast[1], d2 = t[2] - tlast[2];
            const float tnorm = std::sqrt(d0 * d0 + d1 * d1 + d2 * d2);
            if ((rnorm + tnorm) / 2.0f >= c.gate) {
                integrate(bgra, np);
                std::memcpy(Rlast, R, sizeof(Rlast));
                std::memcpy(tlast, t, sizeof(tlast));
            }
            scene.ray_ranges(cam, np);
            if (c.forward)
                scene.forward_render(cam, np, vref.data(), nref.data(), nullptr, c.forward == 2);
            else
                scene.raycast(cam, np, vref.data(), nref.data(), nullptr);
        }
        frames++;
        return 1;
    }
};

}  // namespace hashor

using namespace hashor;

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
int export_keys(const std::vector<uint64_t>& keys, int16_t* out, long long cap) {
    for (size_t i = 0; i < keys.size() && (long long)i < cap && out; ++i) {
        int p[3];
        unpack_key(keys[i], p);
        out[3 * i] = (int16_t)p[0];
        out[3 * i + 1] = (int16_t)p[1];
        out[3 * i + 2] = (int16_t)p[2];
    }
    return (int)keys.size();
}
}  // namespace

// Sorted unique blocks touched by one frame's truncation bands (pose: R9 + t3).
// Returns the count; keys (x, y, z int16 triples) up to cap.
OR_EXPORT int or_hash_touched(const float* depth, int W, int H, const float* cam6, const float* R9, const float* t3,
                              float vs, float trunc_scale, int16_t* keys, long long cap) {
    HashCfg hc{vs, trunc_scale, 0.06f, 255, 0};
    HashScene s(hc);
    return export_keys(s.touched(depth, cam_of(cam6, W, H), make_rigid<float>(R9, t3)), keys, cap);
}

// The reference's bucket hash (K/KinfuHashUtils.cl:9-16) with a given mask.
OR_EXPORT unsigned or_hash_index(int x, int y, int z, unsigned mask) {
    const uint32_t p0 = 73856093u, p1 = 19349669u, p2 = 83492791u;
    return (((uint32_t)(int16_t)x * p0) ^ ((uint32_t)(int16_t)y * p1) ^ ((uint32_t)(int16_t)z * p2)) & mask;
}

OR_EXPORT void* or_hash_engine_create(int W, int H, const float* cam6, int filter_radius, float sigma_space,
                                      float sigma_depth, float depth_threshold, float dist_threshold, int levels,
                                      const int* iters, float gate, float vs, float trunc_scale, float truncation,
                                      int wmax, int capacity, int forward) {
    HashEngineCfg c;
    c.W = W;
    c.H = H;
    std::memcpy(c.cam, cam6, sizeof(c.cam));
    c.filter_radius = filter_radius;
    c.sigma_space = sigma_space;
    c.sigma_depth = sigma_depth;
    c.depth_threshold = depth_threshold;
    c.dist_threshold = dist_threshold;
    c.levels = levels;
    for (int l = 0; l < 8; ++l) c.iters[l] = l < levels ? iters[l] : 0;
    c.gate = gate;
    c.hc = HashCfg{vs, trunc_scale, truncation, wmax, capacity};
    c.forward = forward;
    return new HashEngine(c);
}

OR_EXPORT void or_hash_engine_destroy(void* h) { delete (HashEngine*)h; }

// One frame; pose12 out (R, t as double); info3: ICP error, count, voxels updated.
OR_EXPORT int or_hash_engine_run(void* hp, const uint16_t* raw, const uint8_t* bgra, double* pose12, double* info3) {
    HashEngine* e = (HashEngine*)hp;
    const int ok = e->run(raw, bgra);
    if (pose12) {
        for (int k = 0; k < 9; ++k) pose12[k] = e->R[k];
        for (int k = 0; k < 3; ++k) pose12[9 + k] = e->t[k];
    }
    if (info3) {
        info3[0] = e->last.error;
        info3[1] = e->last.count;
        info3[2] = (double)e->nupd_last;
    }
    return ok;
}

// Scene of a hash engine (for the scene-level calls below).
OR_EXPORT void* or_hash_engine_scene(void* hp) { return &((HashEngine*)hp)->scene; }

// Reference maps of the next ICP (world vertices / normals).
OR_EXPORT void or_hash_engine_refmaps(void* hp, float* vmap, float* nmap) {
    HashEngine* e = (HashEngine*)hp;
    if (vmap) std::memcpy(vmap, e->vref.data(), e->vref.size() * sizeof(float));
    if (nmap) std::memcpy(nmap, e->nref.data(), e->nref.size() * sizeof(float));
}

// ---- scene-level calls (explicit poses) ----

OR_EXPORT void* or_hash_scene_create(float vs, float trunc_scale, float truncation, int wmax, int capacity) {
    return new HashScene(HashCfg{vs, trunc_scale, truncation, wmax, capacity});
}
OR_EXPORT void or_hash_scene_destroy(void* sp) { delete (HashScene*)sp; }

// AllocateSceneFromDepth + IntegrateIntoScene for one frame; returns voxels updated.
OR_EXPORT long long or_hash_scene_integrate(void* sp, const float* depth, const uint8_t* bgra, int W, int H,
                                            const float* cam6, const float* R9, const float* t3) {
    HashScene* s = (HashScene*)sp;
    const Cam<float> cam = cam_of(cam6, W, H);
    const Rigid<float> g = make_rigid<float>(R9, t3);
    return s->integrate_frame(depth, bgra, cam, g);
}

OR_EXPORT void or_hash_scene_set_swapping(void* sp, int on) { ((HashScene*)sp)->swapping = on != 0; }
OR_EXPORT void or_hash_engine_set_swapping(void* hp, int on) { ((HashEngine*)hp)->scene.swapping = on != 0; }

// Swap states of the blocks with an entry (sorted by key) and the stored blocks (sorted by key).
OR_EXPORT int or_hash_scene_swap_states(void* sp, int16_t* keys, uint8_t* states, long long cap) {
    HashScene* s = (HashScene*)sp;
    std::vector<uint64_t> ks;
    for (auto& kv : s->index) ks.push_back(kv.first);
    std::sort(ks.begin(), ks.end(), key_less);
    for (size_t i = 0; i < ks.size() && (long long)i < cap; ++i) {
        auto it = s->swap_state.find(ks[i]);
        if (states) states[i] = it == s->swap_state.end() ? 0 : it->second;
    }
    return export_keys(ks, keys, cap);
}
OR_EXPORT int or_hash_scene_store(void* sp, int16_t* keys, int16_t* sdf, uint8_t* w, uint8_t* rgb, long long cap) {
    HashScene* s = (HashScene*)sp;
    std::vector<uint64_t> ks;
    for (auto& kv : s->store) ks.push_back(kv.first);
    std::sort(ks.begin(), ks.end(), key_less);
    for (size_t i = 0; i < ks.size() && (long long)i < cap; ++i) {
        const Brick& b = s->store.at(ks[i]);
        if (sdf) std::memcpy(sdf + i * kSB3, b.sdf, sizeof(b.sdf));
        if (w) std::memcpy(w + i * kSB3, b.w, sizeof(b.w));
        if (rgb) std::memcpy(rgb + i * kSB3 * 3, b.rgb, sizeof(b.rgb));
    }
    return export_keys(ks, keys, cap);
}

// Ray ranges + castRay at a pose: vmap / nmap (W*H*3), range2 (W*H*2), cmap (W*H*4 or NULL).
OR_EXPORT void or_hash_scene_raycast(void* sp, int W, int H, const float* cam6, const float* R9, const float* t3,
                                     float* vmap, float* nmap, float* range2, uint8_t* cmap) {
    HashScene* s = (HashScene*)sp;
    const Cam<float> cam = cam_of(cam6, W, H);
    const Rigid<float> g = make_rigid<float>(R9, t3);
    s->ray_ranges(cam, g);
    s->raycast(cam, g, vmap, nmap, cmap);
    if (range2) std::memcpy(range2, s->range.data(), s->range.size() * sizeof(float));
}

// Ray ranges + forward render at a pose.  vmap holds the previous world vertex map on
// entry and the rendered one on return; nmap (W*H*3) receives the normals, range2 the
// ray ranges, cmap (W*H*4 or NULL) is forwarded in place.  Returns the pixels raycast.
OR_EXPORT long long or_hash_scene_forward_render(void* sp, int W, int H, const float* cam6, const float* R9,
                                                 const float* t3, float* vmap, float* nmap, float* range2,
                                                 uint8_t* cmap, int reference_transform) {
    HashScene* s = (HashScene*)sp;
    const Cam<float> cam = cam_of(cam6, W, H);
    const Rigid<float> g = make_rigid<float>(R9, t3);
    s->ray_ranges(cam, g);
    const size_t n = s->forward_render(cam, g, vmap, nmap, cmap, reference_transform != 0);
    if (range2) std::memcpy(range2, s->range.data(), s->range.size() * sizeof(float));
    return (long long)n;
}

// Allocated blocks sorted by key: keys (int16 x3), sdf (int16 x512), weight (uint8 x512)
// and colour (uint8 x512x3) per block, each optional; returns the block count.
OR_EXPORT int or_hash_scene_blocks(void* sp, int16_t* keys, int16_t* sdf, uint8_t* w, uint8_t* rgb, long long cap) {
    HashScene* s = (HashScene*)sp;
    std::vector<uint64_t> ks;
    ks.reserve(s->index.size());
    for (auto& kv : s->index)
        if (kv.second >= 0) ks.push_back(kv.first);  // blocks holding a brick
    std::sort(ks.begin(), ks.end(), key_less);
    for (size_t i = 0; i < ks.size() && (long long)i < cap; ++i) {
        const Brick& b = s->bricks[s->index.at(ks[i])];
        if (sdf) std::memcpy(sdf + i * kSB3, b.sdf, sizeof(b.sdf));
        if (w) std::memcpy(w + i * kSB3, b.w, sizeof(b.w));
        if (rgb) std::memcpy(rgb + i * kSB3 * 3, b.rgb, sizeof(b.rgb));
    }
    return export_keys(ks, keys, cap);
}

// Visible list of the last integration, sorted by key.
OR_EXPORT int or_hash_scene_visible(void* sp, int16_t* keys, long long cap) {
    return export_keys(((HashScene*)sp)->visible, keys, cap);
}

// Marching cubes over the visible list; returns the vertex count (3 per triangle).
OR_EXPORT long long or_hash_scene_mesh(void* sp, const int* edge_table, const int* tri_table, float* verts,
                                       long long cap) {
    return (long long)((HashScene*)sp)->mesh(edge_table, tri_table, verts, (size_t)cap);
}
