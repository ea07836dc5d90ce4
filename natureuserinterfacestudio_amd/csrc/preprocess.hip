// Frame preprocessing kernels: passing filter, bilateral
// filter, depth pyramid, vertex and normal maps, map transform, depth scaling.
//
// Layout: images are row-major W*H float (depth) or packed float3 AoS
// (vertex/normal maps, 12 bytes per pixel) exactly as the reference stores
// them, so the ICP and raycast kernels read the same buffers.  All of these
// are bandwidth-trivial (a few MB per frame); the fused preprocessing kernel
// exists to cut launches and to read the raw depth once.

#include "batch_jobs.h"
#include "det_math.h"
#include "nuif_internal.h"

namespace nuif {

constexpr int kTileW = 32, kTileH = 8, kMaxR = 7;

// K/filters.cl:9-23
__device__ __forceinline__ float passing(uint16_t raw, float dmin, float dmax) {
    float d = (float)raw * 0.001f;
    if (!(d > dmin && d < dmax)) d = kNaN;
    return d;
}

__global__ void passing_filter_kernel(const uint16_t* __restrict__ raw, float* __restrict__ out, size_t n,
                                      float dmin, float dmax) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = passing(raw[i], dmin, dmax);
}

// K/bilateralFilter.cl:68-119 for the pixel (x,y) of an LDS tile whose origin
// is (ox - R, oy - R).  Summation order: rows, then columns (as the reference).
__device__ __forceinline__ float bilateral_px(const float* tile, int tw, int x, int y, int ox, int oy, int W,
                                              int H, const Gauss& g, int r, float sigma_depth, float thr,
                                              const double* exp_tab) {
    const float c = tile[(y - oy + kMaxR) * tw + (x - ox + kMaxR)];
    if (isnan(c)) return kNaN;
    const int x0 = max(x - r, 0), x1 = min(x + r + 1, W);
    const int y0 = max(y - r, 0), y1 = min(y + r + 1, H);
    float sd = 0.f, sw = 0.f;
    for (int cy = y0; cy < y1; ++cy) {
        const float* row = tile + (cy - oy + kMaxR) * tw;
        for (int cx = x0; cx < x1; ++cx) {
            const float n = row[cx - ox + kMaxR];
            if (isnan(n)) continue;
            const float diff = fabsf(c - n);
            if (n > 0.0f && diff < thr) {
                const float d2 = diff * diff;
                const float w = g.g[cx + r - x] * g.g[cy + r - y] * nuif_det::expf_det_t(-(d2 * sigma_depth), exp_tab);
                sd += n * w;
                sw += w;
            }
        }
    }
    return sd / sw;
}

// The exp table of nuif_det::expf_det into LDS (every lane of the block reads it at
// data-dependent indices); the caller synchronises.
__device__ __forceinline__ void load_exp_table(double* t) {
    const int tid = threadIdx.y * blockDim.x + threadIdx.x;
    if (tid < 64) t[tid] = nuif_det::kExp2J64[tid];
}

// Loads the block's tile (with a kMaxR halo) into LDS.  FROM_RAW applies the
// passing filter to uint16 input on the fly.
template <bool FROM_RAW>
__device__ __forceinline__ void load_tile(float* tile, int tw, int th, const void* src, int W, int H, int ox, int oy,
                                          float dmin, float dmax) {
    const int tid = threadIdx.y * blockDim.x + threadIdx.x;
    const int nthr = blockDim.x * blockDim.y;
    for (int k = tid; k < tw * th; k += nthr) {
        const int ty = k / tw, tx = k % tw;
        const int gx = ox - kMaxR + tx, gy = oy - kMaxR + ty;
        float v = kNaN;
        if (gx >= 0 && gy >= 0 && gx < W && gy < H) {
            const size_t gi = (size_t)gy * W + gx;
            if (FROM_RAW)
                v = passing(static_cast<const uint16_t*>(src)[gi], dmin, dmax);
            else
                v = static_cast<const float*>(src)[gi];
        }
        tile[k] = v;
    }
}

__global__ void __launch_bounds__(256) bilateral_kernel(const float* __restrict__ in, float* __restrict__ out, int W,
                                                        int H, Gauss g, int r, float sigma_depth, float thr) {
    __shared__ float tile[(kTileW + 2 * kMaxR) * (kTileH + 2 * kMaxR)];
    __shared__ double etab[64];
    const int tw = kTileW + 2 * kMaxR, th = kTileH + 2 * kMaxR;
    const int ox = blockIdx.x * kTileW, oy = blockIdx.y * kTileH;
    load_tile<false>(tile, tw, th, in, W, H, ox, oy, 0.f, 0.f);
    load_exp_table(etab);
    __syncthreads();
    const int x = ox + threadIdx.x, y = oy + threadIdx.y;
    if (x < W && y < H)
        out[(size_t)y * W + x] = bilateral_px(tile, tw, x, y, ox, oy, W, H, g, r, sigma_depth, thr, etab);
}

// Fused passing filter + bilateral + level-0 vertices + ray-length depth: reads the raw
// depth once.  Body of one 32 x 8 tile (bx, by) of job J.
__device__ __forceinline__ void preprocess_l0_body(const PreJob& J, int bx, int by) {
    __shared__ float tile[(kTileW + 2 * kMaxR) * (kTileH + 2 * kMaxR)];
    __shared__ double etab[64];
    const int tw = kTileW + 2 * kMaxR, th = kTileH + 2 * kMaxR;
    const int W = J.W, H = J.H;
    const NuiCLCameraParams cam = J.cam;
    const int ox = bx * kTileW, oy = by * kTileH;
    load_tile<true>(tile, tw, th, in_global(J.raw), W, H, ox, oy, cam.sensorDepthWorldMin, cam.sensorDepthWorldMax);
    load_exp_table(etab);
    __syncthreads();
    const int x = ox + threadIdx.x, y = oy + threadIdx.y;
    if (x >= W || y >= H) return;
    const size_t i = (size_t)y * W + x;
    float d = tile[(y - oy + kMaxR) * tw + (x - ox + kMaxR)];
    if (J.scale && !isnan(d)) {
        // K/tsdfVolume.cl:8-29 scaleDepthsKernel
        const float xl = ((float)x - cam.cx) * cam.fx_inv;
        const float yl = ((float)y - cam.cy) * cam.fy_inv;
        const float lambda = sqrtf(xl * xl + yl * yl + 1.f);
        d = d * lambda;
    }
    in_global(J.depth_out)[i] = d;
    const float f = bilateral_px(tile, tw, x, y, ox, oy, W, H, J.g, J.r, J.sigma_depth, J.thr, etab);
    in_global(J.depth_f)[i] = f;
    // K/transformMaps.cl:18-42 with div = 1
    f3 v = mk3(kNaN, kNaN, kNaN);
    if (!isnan(f)) v = mk3(f * (((float)x - cam.cx) * cam.fx_inv), f * (((float)y - cam.cy) * cam.fy_inv), f * 1.0f);
    st3(in_global(J.vmap), i, v);
}

__global__ void __launch_bounds__(256) preprocess_l0_kernel(PreJob J) { preprocess_l0_body(J, blockIdx.x, blockIdx.y); }

__global__ void __launch_bounds__(256) preprocess_l0_batch_kernel(const PreJob* __restrict__ jobs, int n) {
    const PreJob& J = jobs[job_of(jobs, n, blockIdx.x)];
    const unsigned b = blockIdx.x - J.block0;
    preprocess_l0_body(J, (int)(b % J.gx), (int)(b / J.gx));
}

// K/transformMaps.cl:18-42 depth2vertex_kernel
__device__ __forceinline__ f3 vertex_at(float d, int x, int y, const NuiCLCameraParams& cam, int div) {
    if (isnan(d)) return mk3(kNaN, kNaN, kNaN);
    const float fxi = (float)div * cam.fx_inv, fyi = (float)div * cam.fy_inv;
    const float cx = cam.cx / (float)div, cy = cam.cy / (float)div;
    return mk3(d * (((float)x - cx) * fxi), d * (((float)y - cy) * fyi), d * 1.0f);
}

__global__ void depth2vertex_kernel(const float* __restrict__ depth, float* __restrict__ vmap, int W, int H,
                                    NuiCLCameraParams cam, int div) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= W || y >= H) return;
    const size_t i = (size_t)y * W + x;
    st3(vmap, i, vertex_at(depth[i], x, y, cam, div));
}

// K/transformMaps.cl:44-88 half_sample_kernel (+ optional depth2vertex of the result),
// one 16 x 16 block (bx, by) of job J
__device__ __forceinline__ void half_sample_body(const HalfJob& J, int bx, int by) {
    const int Wd = J.Wd, Hd = J.Hd, r = J.r;
    const int x = bx * 16 + threadIdx.x, y = by * 16 + threadIdx.y;
    if (x >= Wd || y >= Hd) return;
    const float* __restrict__ src = in_global(J.src);
    const float thr = J.thr;
    const int Ws = Wd << 1, Hs = Hd << 1;
    const int sx = x << 1, sy = y << 1;
    const float c = src[(size_t)sy * Ws + sx];
    const int x0 = max(sx - r, 0), x1 = min(sx + r + 1, Ws);
    const int y0 = max(sy - r, 0), y1 = min(sy + r + 1, Hs);
    float s = 0.0f;
    int cnt = 0;
    for (int cy = y0; cy < y1; ++cy)
        for (int cx = x0; cx < x1; ++cx) {
            const float n = src[(size_t)cy * Ws + cx];
            if (!isnan(n) && (isnan(c) || fabsf(c - n) < thr)) {
                s += n;
                cnt += 1;
            }
        }
    const float o = cnt > 0 ? s / (float)cnt : kNaN;
    const size_t i = (size_t)y * Wd + x;
    in_global(J.dst)[i] = o;
    if (J.vmap) st3(in_global(J.vmap), i, vertex_at(o, x, y, J.cam, J.div));
}

__global__ void half_sample_kernel(HalfJob J) { half_sample_body(J, blockIdx.x, blockIdx.y); }

__global__ void half_sample_batch_kernel(const HalfJob* __restrict__ jobs, int n) {
    const HalfJob& J = jobs[job_of(jobs, n, blockIdx.x)];
    const unsigned b = blockIdx.x - J.block0;
    half_sample_body(J, (int)(b % J.gx), (int)(b / J.gx));
}

// K/normalEstimation.cl:770-822 estimate_normals_simple_kernel
__global__ void normals_simple_kernel(const float* __restrict__ vmap, float* __restrict__ nmap, int W, int H,
                                      float thr) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= W || y >= H) return;
    st3(nmap, (size_t)y * W + x, normal_simple_at(vmap, x, y, W, H, thr));
}

// K/transformMaps.cl:91-127 transform_maps_kernel (in-place safe per pixel)
__global__ void transform_maps_kernel(const float* vmap, const float* nmap, float* vdst, float* ndst, int n,
                                      const NuiCLRigidTransform* __restrict__ pose) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const f3 NaN3 = mk3(kNaN, kNaN, kNaN);
    const f3 v = ld3(vmap, i);
    if (nan3(v)) {
        st3(vdst, i, NaN3);
        st3(ndst, i, NaN3);
        return;
    }
    const f3 t = mk3(pose->t[0], pose->t[1], pose->t[2]);
    st3(vdst, i, add3(rot3(pose->R, v), t));
    const f3 nn = ld3(nmap, i);
    st3(ndst, i, nan3(nn) ? NaN3 : rot3(pose->R, nn));
}

// K/tsdfVolume.cl:8-29 scaleDepthsKernel
__global__ void scale_depths_kernel(float* depth, int W, int H, NuiCLCameraParams cam) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= W || y >= H) return;
    const size_t i = (size_t)y * W + x;
    const float d = depth[i];
    if (isnan(d)) return;
    const float xl = ((float)x - cam.cx) * cam.fx_inv;
    const float yl = ((float)y - cam.cy) * cam.fy_inv;
    const float lambda = sqrtf(xl * xl + yl * yl + 1.f);
    depth[i] = d * lambda;
}

int make_gauss(const float* gaussian, int radius, Gauss* g) {
    NUIF_CHECK_ARG(gaussian != nullptr, "gaussian table is NULL");
    NUIF_CHECK_ARG(radius >= 0 && radius <= kMaxR, "bilateral radius %d outside [0, %d]", radius, kMaxR);
    for (int i = 0; i < 15; ++i) g->g[i] = i < 2 * radius + 1 ? gaussian[i] : 0.f;
    return NUIF_OK;
}

int launch_preprocess_batch(const PreJob* jobs, const PreJob* hjobs, int n, hipStream_t st) {
    if (n <= 0) return NUIF_OK;
    const unsigned blocks = hjobs[n - 1].block0 + hjobs[n - 1].gx * div_up(hjobs[n - 1].H, kTileH);
    preprocess_l0_batch_kernel<<<blocks, dim3(kTileW, kTileH), 0, st>>>(jobs, n);
    NUIF_LAUNCHED("preprocess_l0_batch_kernel");
    return NUIF_OK;
}

int launch_half_batch(const HalfJob* jobs, const HalfJob* hjobs, int n, hipStream_t st) {
    if (n <= 0) return NUIF_OK;
    const unsigned blocks = hjobs[n - 1].block0 + hjobs[n - 1].gx * div_up(hjobs[n - 1].Hd, 16);
    half_sample_batch_kernel<<<blocks, dim3(16, 16), 0, st>>>(jobs, n);
    NUIF_LAUNCHED("half_sample_batch_kernel");
    return NUIF_OK;
}

// OCL/NuiKinfuOpenCLFrame.cpp:405-448: the colour of depth pixel i through the
// depth-to-colour map, read at map[lut[i]] (map[i] without a lookup table)
__global__ void map_color_kernel(const float2* __restrict__ map, const uint32_t* __restrict__ lut, uint32_t n,
                                 const uchar4* __restrict__ image, int iw, int ih, uchar4* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t j = lut ? lut[i] : i;
    uchar4 c = make_uchar4(0, 0, 0, 0);
    if (j < n) {
        const float2 q = map[j];
        const int cx = (int)(q.x + 0.5f), cy = (int)(q.y + 0.5f);  // truncation toward zero, as written
        if (cx >= 0 && cx < iw && cy >= 0 && cy < ih) {
            c = image[(size_t)cy * iw + cx];
            c.w = 255;
        }
    }
    out[i] = c;
}

}  // namespace nuif

using namespace nuif;

extern "C" {

int nuif_map_depth_to_color(const float* map_xy, const uint32_t* lut, uint32_t n, const uint8_t* image, int iw, int ih,
                            uint8_t* bgra_out, nuif_stream_t s) {
    NUIF_REQUIRE_INIT();
    NUIF_CHECK_ARG(map_xy && image && bgra_out, "nuif_map_depth_to_color: null argument");
    NUIF_CHECK_ARG(iw > 0 && ih > 0, "nuif_map_depth_to_color: empty colour image");
    if (n == 0) return NUIF_OK;
    map_color_kernel<<<div_up(n, 256), 256, 0, stream_of(s)>>>(reinterpret_cast<const float2*>(map_xy), lut, n,
                                                                reinterpret_cast<const uchar4*>(image), iw, ih,
                                                                reinterpret_cast<uchar4*>(bgra_out));
    NUIF_LAUNCHED("map_color_kernel");
    return NUIF_OK;
}

int nuif_depth_passing_filter(const uint16_t* raw, float* depth, size_t n, float dmin, float dmax, nuif_stream_t s) {
    NUIF_REQUIRE_INIT();
    if (n == 0) return NUIF_OK;
    NUIF_CHECK_ARG(raw && depth, "nuif_depth_passing_filter: null buffer");
    passing_filter_kernel<<<div_up(n, 256), 256, 0, stream_of(s)>>>(raw, depth, n, dmin, dmax);
    NUIF_LAUNCHED("passing_filter_kernel");
    return NUIF_OK;
}

int nuif_bilateral_filter_depth(const float* in, float* out, int W, int H, const float* gaussian, int radius,
                                float sigma_depth2_inv_half, float depth_threshold, nuif_stream_t s) {
    NUIF_REQUIRE_INIT();
    NUIF_CHECK_ARG(in && out && W > 0 && H > 0, "nuif_bilateral_filter_depth: bad arguments");
    NUIF_CHECK_ARG(in != out, "nuif_bilateral_filter_depth: in-place filtering is not supported");
    Gauss g;
    int rc = make_gauss(gaussian, radius, &g);
    if (rc) return rc;
    dim3 grid(div_up(W, kTileW), div_up(H, kTileH)), block(kTileW, kTileH);
    bilateral_kernel<<<grid, block, 0, stream_of(s)>>>(in, out, W, H, g, radius, sigma_depth2_inv_half,
                                                       depth_threshold);
    NUIF_LAUNCHED("bilateral_kernel");
    return NUIF_OK;
}

int nuif_half_sample(const float* src, float* dst, int Wd, int Hd, int radius, float depth_threshold,
                     nuif_stream_t s) {
    NUIF_REQUIRE_INIT();
    NUIF_CHECK_ARG(src && dst && Wd > 0 && Hd > 0 && radius >= 0, "nuif_half_sample: bad arguments");
    NuiCLCameraParams cam{};
    dim3 block(16, 16), grid(div_up(Wd, 16), div_up(Hd, 16));
    const HalfJob hj{src, dst, nullptr, Wd, Hd, radius, depth_threshold, cam, 1, 0u, 0u};
    half_sample_kernel<<<grid, block, 0, stream_of(s)>>>(hj);
    NUIF_LAUNCHED("half_sample_kernel");
    return NUIF_OK;
}

int nuif_depth2vertex(const float* depth, float* vmap, int W, int H, const NuiCLCameraParams* cam, int div,
                      nuif_stream_t s) {
    NUIF_REQUIRE_INIT();
    NUIF_CHECK_ARG(depth && vmap && cam && W > 0 && H > 0 && div > 0, "nuif_depth2vertex: bad arguments");
    dim3 block(16, 16), grid(div_up(W, 16), div_up(H, 16));
    depth2vertex_kernel<<<grid, block, 0, stream_of(s)>>>(depth, vmap, W, H, *cam, div);
    NUIF_LAUNCHED("depth2vertex_kernel");
    return NUIF_OK;
}

int nuif_estimate_normals_simple(const float* vmap, float* nmap, int W, int H, float depth_threshold,
                                 nuif_stream_t s) {
    NUIF_REQUIRE_INIT();
    NUIF_CHECK_ARG(vmap && nmap && W > 0 && H > 0 && vmap != nmap, "nuif_estimate_normals_simple: bad arguments");
    dim3 block(16, 16), grid(div_up(W, 16), div_up(H, 16));
    normals_simple_kernel<<<grid, block, 0, stream_of(s)>>>(vmap, nmap, W, H, depth_threshold);
    NUIF_LAUNCHED("normals_simple_kernel");
    return NUIF_OK;
}

int nuif_transform_maps(const float* vmap, const float* nmap, float* vdst, float* ndst, int W, int H,
                        const NuiCLRigidTransform* pose, nuif_stream_t s) {
    NUIF_REQUIRE_INIT();
    NUIF_CHECK_ARG(vmap && nmap && vdst && ndst && pose && W > 0 && H > 0, "nuif_transform_maps: bad arguments");
    const int n = W * H;
    transform_maps_kernel<<<div_up(n, 256), 256, 0, stream_of(s)>>>(vmap, nmap, vdst, ndst, n, pose);
    NUIF_LAUNCHED("transform_maps_kernel");
    return NUIF_OK;
}

int nuif_scale_depths(float* depth, int W, int H, const NuiCLCameraParams* cam, nuif_stream_t s) {
    NUIF_REQUIRE_INIT();
    NUIF_CHECK_ARG(depth && cam && W > 0 && H > 0, "nuif_scale_depths: bad arguments");
    dim3 block(16, 16), grid(div_up(W, 16), div_up(H, 16));
    scale_depths_kernel<<<grid, block, 0, stream_of(s)>>>(depth, W, H, *cam);
    NUIF_LAUNCHED("scale_depths_kernel");
    return NUIF_OK;
}

int nuif_frame_preprocess(const uint16_t* raw, float* depth_out, float* const* depth_pyr, float* const* vmap_pyr,
                          int levels, int W, int H, const NuiCLCameraParams* cam, const float* gaussian, int radius,
                          float sigma_depth2_inv_half, float depth_threshold, int scale_for_tsdf, nuif_stream_t s) {
    NUIF_REQUIRE_INIT();
    NUIF_CHECK_ARG(raw && depth_out && depth_pyr && vmap_pyr && cam, "nuif_frame_preprocess: null argument");
    NUIF_CHECK_ARG(W > 0 && H > 0 && levels >= 1 && levels <= NUIF_ICP_MAX_LEVELS,
                   "nuif_frame_preprocess: bad shape %dx%d / levels %d", W, H, levels);
    for (int l = 0; l < levels; ++l)
        NUIF_CHECK_ARG(depth_pyr[l] && vmap_pyr[l], "nuif_frame_preprocess: level %d buffer is NULL", l);
    NUIF_CHECK_ARG((W >> (levels - 1)) > 0 && (H >> (levels - 1)) > 0, "nuif_frame_preprocess: too many levels");
    Gauss g;
    int rc = make_gauss(gaussian, radius, &g);
    if (rc) return rc;
    hipStream_t st = stream_of(s);
    dim3 grid(div_up(W, kTileW), div_up(H, kTileH)), block(kTileW, kTileH);
    const PreJob pj{raw, depth_out, depth_pyr[0], vmap_pyr[0], W, H, *cam, g, radius, sigma_depth2_inv_half,
                    depth_threshold, scale_for_tsdf, 0u, 0u};
    preprocess_l0_kernel<<<grid, block, 0, st>>>(pj);
    NUIF_LAUNCHED("preprocess_l0_kernel");
    for (int l = 1; l < levels; ++l) {
        const int Wd = W >> l, Hd = H >> l;
        dim3 b2(16, 16), g2(div_up(Wd, 16), div_up(Hd, 16));
        // OCL/NuiKinfuOpenCLDepthTracker.cpp:158 subSampleRadius = 1
        const HalfJob hj{depth_pyr[l - 1], depth_pyr[l], vmap_pyr[l], Wd, Hd, 1, depth_threshold, *cam, 1 << l, 0u, 0u};
        half_sample_kernel<<<g2, b2, 0, st>>>(hj);
        NUIF_LAUNCHED("half_sample_kernel");
    }
    return NUIF_OK;
}

}  // extern "C"
