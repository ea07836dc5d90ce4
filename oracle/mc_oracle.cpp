// TEST INFRASTRUCTURE ONLY (see oracle_common.h).  CPU restatement of the
// reference's dense marching cubes: the GPU kernel's definition
// (K/marchingCube.cl:100-256) and the CPU mesher with vertex de-duplication
// (OCL/NuiKinfuOpenCLScene.cpp:788-984).  The case table is the standard one
// shared with the product (its content is data, pinned by tests against the
// reference's table and the table invariants).

#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <vector>

#include "../natureuserinterfacestudio_amd/csrc/mc_tables.h"

#define OR_EXPORT extern "C" __attribute__((visibility("default")))

namespace {

constexpr McTable kTab = mc_decode();
const int kCo[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};

struct P3 {
    float x, y, z;
};

// K/marchingCube.cl:81-84 getNodeCoo (GPU mesher: y flipped, z not centred)
P3 node_gpu(int x, int y, int z, float cell, float hx, float hy) {
    return {((float)x + 0.5f - hx) * cell, (-(float)y - 0.5f + hy) * cell, ((float)z + 0.5f) * cell};
}
// OCL/NuiKinfuOpenCLScene.cpp:82-89 getNodeCoo (CPU mesher: y flipped, z centred)
P3 node_cpu(int x, int y, int z, float cell, int N) {
    return {((float)x + 0.5f - (float)N / 2.0f) * cell, ((float)(-y) - 0.5f + (float)N / 2.0f) * cell,
            ((float)z + 0.5f - (float)N / 2.0f) * cell};
}

P3 interp(P3 p0, P3 p1, float f0, float f1) {
    float t = (0.f - f0) / (f1 - f0 + 1e-15f);
    t = t < 0.f ? 0.f : (t > 1.f ? 1.f : t);
    return {p0.x + (p1.x - p0.x) * t, p0.y + (p1.y - p0.y) * t, p0.z + (p1.z - p0.z) * t};
}

// The eight corners of cube (x,y,z); false if any weight is zero.
bool cube(const int16_t* vol, int N, int x, int y, int z, size_t id[8], int16_t f[8]) {
    const size_t id0 = ((size_t)y * N + x) * N + z;
    const size_t id1 = id0 + N, id3 = id0 + (size_t)N * N, id2 = id3 + N;
    const size_t ids[8] = {id0, id1, id2, id3, id0 + 1, id1 + 1, id2 + 1, id3 + 1};
    for (int k = 0; k < 8; ++k) {
        id[k] = ids[k];
        if (vol[2 * ids[k] + 1] <= 0) return false;
        f[k] = vol[2 * ids[k]];
    }
    return true;
}

int case_index(const int16_t f[8]) {
    int c = 0;
    for (int k = 0; k < 8; ++k) c += (f[k] < 0 ? 1 : 0) << k;
    return c;
}

}  // namespace

// Vertex count of marchingCubeKernel over a short2 volume (3 per triangle).
OR_EXPORT long long or_mc_count_s16(const int16_t* vol, int N) {
    long long total = 0;
#pragma omp parallel for reduction(+ : total) schedule(dynamic, 4)
    for (int y = 0; y < N - 1; ++y)
        for (int x = 0; x < N - 1; ++x)
            for (int z = 0; z < N - 1; ++z) {
                size_t id[8];
                int16_t f[8];
                if (!cube(vol, N, x, y, z, id, f)) continue;
                total += kTab.nverts[case_index(f)];
            }
    return total;
}

// Triangles of marchingCubeKernel in (y, x, z) cube order: vertices (float3)
// and colours (float4 RGBA of the nearest corner); returns the total count.
OR_EXPORT long long or_mc_emit_s16(const int16_t* vol, const uint8_t* cvol, int N, float dim, float* verts,
                                   float* colors, long long cap) {
    const float cell = dim / (float)N;
    const float hx = (float)(N >> 1), hy = (float)(N >> 1);
    long long n = 0;
    for (int y = 0; y < N - 1; ++y)
        for (int x = 0; x < N - 1; ++x)
            for (int z = 0; z < N - 1; ++z) {
                size_t id[8];
                int16_t fi[8];
                if (!cube(vol, N, x, y, z, id, fi)) continue;
                const int ci = case_index(fi);
                float f[8];
                for (int k = 0; k < 8; ++k) f[k] = (float)fi[k] / 32766.0f;
                for (int k = 0; k < kTab.nverts[ci]; ++k, ++n) {
                    if (n >= cap) continue;
                    const int e = kTab.tri[ci][k];
                    const int a = kMcEdgeCorners[e][0], b = kMcEdgeCorners[e][1];
                    const P3 p = interp(node_gpu(x + kCo[a][0], y + kCo[a][1], z + kCo[a][2], cell, hx, hy),
                                        node_gpu(x + kCo[b][0], y + kCo[b][1], z + kCo[b][2], cell, hx, hy), f[a],
                                        f[b]);
                    verts[3 * n] = p.x;
                    verts[3 * n + 1] = p.y;
                    verts[3 * n + 2] = p.z;
                    if (colors) {
                        const uint8_t* c = cvol ? cvol + 4 * (std::fabs(f[a]) < std::fabs(f[b]) ? id[a] : id[b]) : nullptr;
                        colors[4 * n] = c ? (float)c[2] / 255.f : 0.f;
                        colors[4 * n + 1] = c ? (float)c[1] / 255.f : 0.f;
                        colors[4 * n + 2] = c ? (float)c[0] / 255.f : 0.f;
                        colors[4 * n + 3] = 1.f;
                    }
                }
            }
    return n;
}

// OCL/NuiKinfuOpenCLScene.cpp:797-984 Volume2Mesh: vertices keyed by the ordered
// corner pair of the cube edge (so an edge walked in opposite directions by two
// cubes yields two vertices), indices in traversal order.  Returns the point
// count; *ntri receives the index count.
OR_EXPORT long long or_mc_mesh_dedup_s16(const int16_t* vol, int N, float dim, float* points, long long max_points,
                                         int32_t* indices, long long max_indices, long long* nidx) {
    const float cell = dim / (float)N;
    std::map<long long, int> index_map;
    long long np = 0, ni = 0;
    for (int y = 0; y < N - 1; ++y)
        for (int x = 0; x < N - 1; ++x)
            for (int z = 0; z < N - 1; ++z) {
                size_t id[8];
                int16_t fi[8];
                if (!cube(vol, N, x, y, z, id, fi)) continue;
                const int ci = case_index(fi);
                const int nv = kTab.nverts[ci];
                for (int i = 0; i < nv; i += 3) {
                    for (int j = 0; j < 3; ++j) {
                        const int e = kTab.tri[ci][i + j];
                        const int a = kMcEdgeCorners[e][0], b = kMcEdgeCorners[e][1];
                        const long long key = ((long long)id[a] << 32) + (long long)id[b];
                        auto it = index_map.find(key);
                        int vi;
                        if (it == index_map.end()) {
                            vi = (int)np;
                            index_map.emplace(key, vi);
                            const float f0 = (float)fi[a] / 32766.0f, f1 = (float)fi[b] / 32766.0f;
                            const P3 p = interp(node_cpu(x + kCo[a][0], y + kCo[a][1], z + kCo[a][2], cell, N),
                                                node_cpu(x + kCo[b][0], y + kCo[b][1], z + kCo[b][2], cell, N), f0, f1);
                            if (np < max_points) {
                                points[3 * np] = p.x;
                                points[3 * np + 1] = p.y;
                                points[3 * np + 2] = p.z;
                            }
                            ++np;
                        } else {
                            vi = it->second;
                        }
                        if (ni < max_indices) indices[ni] = vi;
                        ++ni;
                    }
                }
            }
    *nidx = ni;
    return np;
}

// Decoded table for the table tests: tri (256 x 16, -1 padded) and counts.
OR_EXPORT void or_mc_table(int8_t* tri, uint8_t* nverts) {
    std::memcpy(tri, kTab.tri, sizeof(kTab.tri));
    std::memcpy(nverts, kTab.nverts, sizeof(kTab.nverts));
}
