// TEST INFRASTRUCTURE ONLY (see oracle_common.h).  CPU restatement of the
// reference's point-cloud extraction of both scenes:
//   dense  K/extract.cl:4-78 fetchVolumeKernel, launched by
//          OCL/NuiKinfuOpenCLScene.cpp:436-570 FetchVolume over (resolution x,
//          resolution y) columns with range_z = resolution z;
//   hash   K/KinfuHashVisualization.cl:5-52 fetchHashScene_kernel, launched by
//          OCL/NuiKinfuOpenCLHashScene.cpp:1066-1190 Volume2CLVertices over the
//          visible list with truncScale = m_truncScale;
//   shift  K/extract.cl:80-190 fetchShiftVolumeKernel, the evicted slab of the
//          shifting volume (OCL/NuiKinfuOpenCLShiftScene.cpp:119-245).
// The reference appends through an atomic counter (its order is a race); the
// oracle writes in the order the device extraction defines: columns in (y, x)
// order with z ascending, and visible-list order with the voxel index ascending.
// Points past `cap` are counted but not written (MAX_OUTPUT_VERTEX_SIZE,
// K/gpu_def.h:3).

#include <cmath>
#include <cstdint>
#include <cstring>

#define OR_EXPORT extern "C" __attribute__((visibility("default")))

namespace {

// IEEE binary16 -> binary32 (exact for every half, NaN payload kept)
float half_to_float(uint16_t h) {
    const uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
    const uint32_t e = (h >> 10) & 0x1fu, m = h & 0x3ffu;
    uint32_t bits;
    if (e == 0) {
        if (m == 0) {
            bits = sign;
        } else {  // subnormal: m * 2^-24
            const float v = (float)m * 0x1p-24f;
            std::memcpy(&bits, &v, 4);
            bits |= sign;
        }
    } else if (e == 31) {
        bits = sign | 0x7f800000u | (m << 13);
    } else {
        bits = sign | ((e + 112u) << 23) | (m << 13);
    }
    float f;
    std::memcpy(&f, &bits, 4);
    return f;
}

// tsdf and "weight > 0" of voxel i for the three storages (0: short2 with
// tsdf * 32766, K/utils.cl:76-84; 1: FP16 tsdf + uint16 weight; 2: FP32 pair)
struct Vox {
    float t;
    bool valid;
};
Vox dense_voxel(const void* vol, int storage, size_t i) {
    Vox v;
    if (storage == 0) {
        const int16_t* s = static_cast<const int16_t*>(vol);
        v.valid = s[2 * i + 1] > 0;
        v.t = (float)s[2 * i] / (float)32766;
    } else if (storage == 1) {
        const uint16_t* s = static_cast<const uint16_t*>(vol);
        v.valid = s[2 * i + 1] > 0;
        v.t = half_to_float(s[2 * i]);
    } else {
        const float* s = static_cast<const float*>(vol);
        v.valid = s[2 * i + 1] > 0.f;
        v.t = s[2 * i];
    }
    return v;
}

void put(float* pts, float* cols, long long o, long long cap, const float p[3], const float c[4]) {
    if (o >= cap) return;
    if (pts) for (int k = 0; k < 3; ++k) pts[3 * o + k] = p[k];
    if (cols) for (int k = 0; k < 4; ++k) cols[4 * o + k] = c[k];
}

}  // namespace

// fetchVolumeKernel over an Nx x Ny x Nz volume in ((y*Nx + x)*Nz + z) layout;
// cvol: uchar4 per voxel or NULL.  Returns the point count (before the cap).
OR_EXPORT long long or_points_dense(const void* vol, const uint8_t* cvol, int storage, int Nx, int Ny, int Nz,
                                    const float* cell, float* pts, float* cols, long long cap) {
    const float half_x = (float)(Nx >> 1), half_y = (float)(Ny >> 1);
    long long n = 0;
    for (int y = 0; y < Ny; ++y) {
        for (int x = 0; x < Nx; ++x) {
            const size_t col = ((size_t)y * Nx + x) * Nz;
            float tsdf = NAN;
            for (int z = 0; z < Nz; ++z) {
                const float tsdf_prev = tsdf;
                const size_t idx = col + z;
                const Vox v = dense_voxel(vol, storage, idx);
                if (!v.valid) {
                    tsdf = NAN;
                    continue;
                }
                tsdf = v.t;
                const float px = ((float)x + 0.5f - half_x) * cell[0];
                const float py = (-(float)y - 0.5f + half_y) * cell[1];
                if (0.f == tsdf) {
                    const float p[3] = {px, py, ((float)z + 0.5f) * cell[2]};
                    float c[4] = {0.f, 0.f, 0.f, 1.f};
                    if (cvol) {
                        const uint8_t* q = cvol + 4 * idx;
                        c[0] = (float)q[0] / 255.f;
                        c[1] = (float)q[1] / 255.f;
                        c[2] = (float)q[2] / 255.f;
                    }
                    put(pts, cols, n++, cap, p, c);
                } else if ((tsdf_prev < 0.f && tsdf > 0.f) || (tsdf_prev > 0.f && tsdf < 0.f)) {
                    const float tsdf_inv = 1.f / std::fabs(tsdf);
                    const float tsdf_prev_inv = 1.f / std::fabs(tsdf_prev);
                    const float dp = (tsdf_inv * (float)z + tsdf_prev_inv * (float)(z - 1)) / (tsdf_inv + tsdf_prev_inv);
                    const float p[3] = {px, py, (dp + 0.5f) * cell[2]};
                    float c[4] = {0.f, 0.f, 0.f, 1.f};
                    if (cvol) {  // channels swapped relative to the exact-zero branch
                        const uint8_t* q = cvol + 4 * idx;
                        c[0] = (float)q[2] / 255.f;
                        c[1] = (float)q[1] / 255.f;
                        c[2] = (float)q[0] / 255.f;
                    }
                    put(pts, cols, n++, cap, p, c);
                }
            }
        }
    }
    return n;
}

// fetchHashScene_kernel over `nvis` visible blocks: pos (int32 x3 per block,
// block coordinates), ptr (brick index, < 0: none), bricks (uint32 per voxel:
// sdf in the low 16 bits -- storage 0: int16 sdf * 32767, 1: FP16 sdf -- and the
// weight in the high 16), colours (uchar4 per voxel or NULL).  A voxel is a
// point when its weight is > 0 and it lies inside the band: |short sdf| <
// 32767 * trunc_scale (storage 0), |sdf| < trunc_scale (storage 1).
OR_EXPORT long long or_points_hash(int nvis, const int32_t* pos, const int32_t* ptr, const uint32_t* bricks,
                                   const uint8_t* colors, int storage, float voxel_size, float trunc_scale,
                                   float* pts, float* cols, long long cap) {
    long long n = 0;
    const float band = 32767 * trunc_scale;
    for (int b = 0; b < nvis; ++b) {
        if (ptr[b] < 0) continue;
        for (int i = 0; i < 512; ++i) {
            const size_t vi = (size_t)ptr[b] * 512 + i;
            const uint32_t w = bricks[vi];
            const int weight = (int)(w >> 16);
            bool in_band;
            if (storage == 0)
                in_band = (float)std::abs((int)(int16_t)(w & 0xffffu)) < band;
            else
                in_band = std::fabs(half_to_float((uint16_t)(w & 0xffffu))) < trunc_scale;
            if (!(weight > 0 && in_band)) continue;
            float p[3];
            p[2] = ((float)(pos[3 * b + 2] * 8) + ((float)(i / 64) + 0.5f)) * voxel_size;
            p[1] = ((float)(pos[3 * b + 1] * 8) + ((float)((i % 64) / 8) + 0.5f)) * voxel_size;
            p[0] = ((float)(pos[3 * b + 0] * 8) + ((float)((i % 64) % 8) + 0.5f)) * voxel_size;
            float c[4] = {0.f, 0.f, 0.f, 1.f};
            if (colors) {
                const uint8_t* q = colors + 4 * vi;
                c[0] = (float)q[0] / 255.f;
                c[1] = (float)q[1] / 255.f;
                c[2] = (float)q[2] / 255.f;
            }
            put(pts, cols, n++, cap, p, c);
        }
    }
    return n;
}

// K/extract.cl:80-190 fetchShiftVolumeKernel over range[0] x range[1] work-items with
// range_z = range[2], voxelWrap = wrap (>= 0) and voxelOffset = offset, restated as
// written: the walk runs voxel_z = wrap.z .. wrap.z + range_z - 1 and reads stored z
// (voxel_z + wrap.z) % Nz (wrap.z added twice), and the position of a voxel adds the
// multiple of N that the reference derives from the offset and the work-item index.
OR_EXPORT long long or_points_dense_slice(const void* vol, const uint8_t* cvol, int storage, int Nx, int Ny, int Nz,
                                          const float* cell, const int* wrap, const int* offset, const int* range,
                                          float* pts, float* cols, long long cap) {
    const float half_x = (float)(Nx >> 1), half_y = (float)(Ny >> 1);
    long long n = 0;
    for (int voxel_y = 0; voxel_y < range[1]; ++voxel_y) {
        for (int voxel_x = 0; voxel_x < range[0]; ++voxel_x) {
            const size_t voxel_id = ((size_t)((voxel_y + wrap[1]) % Ny) * Nx + (size_t)((voxel_x + wrap[0]) % Nx)) * Nz;
            int ob[3] = {(offset[0] / Nx) * Nx, (offset[1] / Ny) * Ny, (offset[2] / Nz) * Nz};
            if (offset[0] > 0) {
                if (voxel_x < offset[0] % Nx) ob[0] += Nx;
            } else {
                if (voxel_x > Nx + offset[0] % Nx) ob[0] -= Nx;
            }
            if (offset[1] > 0) {
                if (voxel_y < offset[1] % Ny) ob[1] += Ny;
            } else {
                if (voxel_y > Ny + offset[1] % Ny) ob[1] -= Ny;
            }
            const int range_z = range[2] + wrap[2];
            float tsdf = NAN;
            for (int voxel_z = wrap[2]; voxel_z < range_z; ++voxel_z) {
                const float tsdf_prev = tsdf;
                const size_t idx = voxel_id + (size_t)((voxel_z + wrap[2]) % Nz);
                const Vox v = dense_voxel(vol, storage, idx);
                if (!v.valid) {
                    tsdf = NAN;
                    continue;
                }
                int oz = ob[2];
                if (offset[2] > 0) {
                    if (voxel_z < offset[2] % Nz) oz += Nz;
                } else {
                    if (voxel_z > Nz + offset[2] % Nz) oz -= Nz;
                }
                tsdf = v.t;
                const float px = ((float)(voxel_x + ob[0]) + 0.5f - half_x) * cell[0];
                const float py = (-(float)(voxel_y + ob[1]) - 0.5f + half_y) * cell[1];
                if (0.f == tsdf) {
                    const float p[3] = {px, py, ((float)(voxel_z + oz) + 0.5f) * cell[2]};
                    float c[4] = {0.f, 0.f, 0.f, 1.f};
                    if (cvol) {
                        const uint8_t* q = cvol + 4 * idx;
                        c[0] = (float)q[0] / 255.f;
                        c[1] = (float)q[1] / 255.f;
                        c[2] = (float)q[2] / 255.f;
                    }
                    put(pts, cols, n++, cap, p, c);
                } else if ((tsdf_prev < 0.f && tsdf > 0.f) || (tsdf_prev > 0.f && tsdf < 0.f)) {
                    const float tsdf_inv = 1.f / std::fabs(tsdf);
                    const float tsdf_prev_inv = 1.f / std::fabs(tsdf_prev);
                    const float dp = (tsdf_inv * (float)(voxel_z + oz) + tsdf_prev_inv * (float)(voxel_z + oz - 1)) /
                                     (tsdf_inv + tsdf_prev_inv);
                    const float p[3] = {px, py, (dp + 0.5f) * cell[2]};
                    float c[4] = {0.f, 0.f, 0.f, 1.f};
                    if (cvol) {
                        const uint8_t* q = cvol + 4 * idx;
                        c[0] = (float)q[2] / 255.f;
                        c[1] = (float)q[1] / 255.f;
                        c[2] = (float)q[0] / 255.f;
                    }
                    put(pts, cols, n++, cap, p, c);
                }
            }
        }
    }
    return n;
}

OR_EXPORT float or_half_to_float(uint16_t h) { return half_to_float(h); }
