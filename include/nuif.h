/*
 * nuif.h - C ABI of the MI355X-native KinectFusion / voxel-hashing fusion path.
 *
 * This library replaces the reference's OpenCL runtime layer
 * (src/OpenCLUtilities: NuiOpenCLGlobal, NuiOpenCLKernelManager,
 * NuiGPUMemManager, NuiOpenCLFoundationUtils) and its OpenCL C kernels
 * (the .cl files of src/Kernels) for the fusion path of src/SLAM.  Every entry point is
 * plain C: device pointers, host pointers, sizes and a stream handle.  All
 * functions return NUIF_OK (0) or a negative status; nuif_last_error() gives
 * the message of the calling thread's last failure.  Launch entry points are
 * asynchronous on the given stream (NULL selects the library's default
 * stream of the current device).
 *
 * Citations: K/ = src/Kernels/, OCL/ = src/SLAM/VisualOdometry/DeviceSpecific/OpenCL/,
 * CLU/ = src/OpenCLUtilities/.
 */
#ifndef NUIF_H
#define NUIF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NUIF_OK 0
#define NUIF_ERR_INVALID -1   /* bad argument / shape */
#define NUIF_ERR_HIP -2       /* HIP runtime error */
#define NUIF_ERR_OOM -3       /* device allocation failed */
#define NUIF_ERR_NODEV -4     /* no GPU / not initialised */
#define NUIF_ERR_TRACKING -5  /* ICP failure (singular normal equations) */

/* ---- shared structures (layouts fixed by K/gpu_def.h:5-34) ---------------- */

/* K/gpu_def.h:13-27 NuiCLCameraParams (40 bytes) */
typedef struct NuiCLCameraParams {
    float fx, fy, fx_inv, fy_inv, cx, cy;
    uint32_t depthImageWidth, depthImageHeight;
    float sensorDepthWorldMin, sensorDepthWorldMax;
} NuiCLCameraParams;

/* K/gpu_def.h:29-34 NuiCLRigidTransform (84 bytes): world = R * cam + t, row-major */
typedef struct NuiCLRigidTransform {
    float R[9];
    float R_inv[9];
    float t[3];
} NuiCLRigidTransform;

/* K/gpu_def.h:5-11 TsdfParams (40 bytes) */
typedef struct TsdfParams {
    float resolution[3];
    float dimension[3];
    float cell_size[3];
    float tranc_dist;
} TsdfParams;

/* Dense TSDF voxel storage (4 or 8 bytes per voxel, layout ((y*Nx + x)*Nz + z)). */
#define NUIF_TSDF_S16 0   /* reference short2: int16 tsdf*32766 (RTZ), int16 weight (K/utils.cl:69-84) */
#define NUIF_TSDF_F16 1   /* FP16 tsdf + uint16 weight, FP32 arithmetic, round-to-nearest */
#define NUIF_TSDF_F32 2   /* FP32 tsdf + FP32 weight */

typedef struct nuif_stream_s* nuif_stream_t;   /* a hipStream_t */
typedef struct nuif_event_s* nuif_event_t;     /* a hipEvent_t */
typedef struct nuif_graph_s* nuif_graph_t;     /* an instantiated hipGraphExec_t */

/* ---- runtime: replaces CLU/NuiOpenCLGlobal.cpp:75-220 (context + queue) ---- */
int nuif_init(int device);                          /* select + initialise a GPU */
int nuif_device_count(int* count);
int nuif_device_info(char* name, int name_len, int* cu_count, size_t* hbm_bytes, int* clock_khz);
const char* nuif_last_error(void);
const char* nuif_version(void);

/* ---- memory: replaces CLU/NuiGPUMemManager (CreateBufferCL / ReleaseMemObjectCL) */
int nuif_malloc(void** dptr, size_t bytes);         /* 256-byte aligned, zero-filled */
int nuif_free(void* dptr);
int nuif_mem_in_use(size_t* bytes);                 /* sum of live nuif_malloc bytes (admission) */
int nuif_host_alloc(void** hptr, size_t bytes);     /* pinned host memory */
int nuif_host_free(void* hptr);
int nuif_memcpy_h2d(void* dst, const void* src, size_t bytes, nuif_stream_t s);
int nuif_memcpy_d2h(void* dst, const void* src, size_t bytes, nuif_stream_t s);
int nuif_memcpy_d2d(void* dst, const void* src, size_t bytes, nuif_stream_t s);
/* byte fill by kernel (safe inside graph capture); CLU/NuiOpenCLFoundationUtils.cpp:8-150 */
int nuif_fill_u8(void* dst, uint8_t value, size_t bytes, nuif_stream_t s);
int nuif_fill_u32(void* dst, uint32_t value, size_t count, nuif_stream_t s);

/* ---- streams, events, graphs --------------------------------------------- */
int nuif_stream_create(nuif_stream_t* s);
int nuif_stream_destroy(nuif_stream_t s);
int nuif_stream_synchronize(nuif_stream_t s);
int nuif_device_synchronize(void);
int nuif_event_create(nuif_event_t* e);
int nuif_event_destroy(nuif_event_t e);
int nuif_event_record(nuif_event_t e, nuif_stream_t s);
int nuif_event_synchronize(nuif_event_t e);
int nuif_event_elapsed_ms(nuif_event_t start, nuif_event_t stop, float* ms);
int nuif_stream_wait_event(nuif_stream_t s, nuif_event_t e);
int nuif_graph_begin(nuif_stream_t s);                       /* start capture on s */
int nuif_graph_end(nuif_stream_t s, nuif_graph_t* g);        /* end capture + instantiate */
int nuif_graph_launch(nuif_graph_t g, nuif_stream_t s);
int nuif_graph_destroy(nuif_graph_t g);

/* ---- frame preprocessing (one entry per reference kernel) ------------------ */

/* OCL/NuiKinfuOpenCLFrame.cpp:405-448 UpdateColorBuffers(ColorSpacePoint*, UINT*, UINT,
 * const NuiColorImage&): the BGRA colour of each of the n depth pixels through the
 * depth-to-colour map.  For pixel i, j = lut ? lut[i] : i; when j < n and
 * ((int)(x_j + 0.5f), (int)(y_j + 0.5f)) lies in the iw x ih colour image, the output is that
 * pixel with alpha 255, else 0.  map_xy: n (x, y) float pairs in colour-image pixels (the
 * Kinect SDK's ColorSpacePoint), lut: n uint32 or NULL, image: iw*ih BGRA; device memory. */
int nuif_map_depth_to_color(const float* map_xy, const uint32_t* lut, uint32_t n, const uint8_t* image, int iw,
                            int ih, uint8_t* bgra_out, nuif_stream_t s);

/* K/filters.cl:9-23 depth_passing_filter_kernel: depth[i] = raw[i]*0.001 m, NaN unless dmin < d < dmax */
int nuif_depth_passing_filter(const uint16_t* raw, float* depth, size_t n, float dmin, float dmax,
                              nuif_stream_t s);

/* K/bilateralFilter.cl:68-119 bilateral_filter_depth_kernel; gaussian: 2*radius+1 host floats
 * (OCL/NuiKinfuOpenCLFrame.cpp:165-189), radius <= 7 */
int nuif_bilateral_filter_depth(const float* in, float* out, int width, int height, const float* gaussian,
                                int radius, float sigma_depth2_inv_half, float depth_threshold, nuif_stream_t s);

/* K/transformMaps.cl:44-88 half_sample_kernel; dst is (dst_width x dst_height), src 2x that */
int nuif_half_sample(const float* src, float* dst, int dst_width, int dst_height, int radius,
                     float depth_threshold, nuif_stream_t s);

/* K/transformMaps.cl:18-42 depth2vertex_kernel; vmap packed float3 AoS; div = 2^level */
int nuif_depth2vertex(const float* depth, float* vmap, int width, int height, const NuiCLCameraParams* cam,
                      int div, nuif_stream_t s);

/* K/normalEstimation.cl:770-822 estimate_normals_simple_kernel */
int nuif_estimate_normals_simple(const float* vmap, float* nmap, int width, int height, float depth_threshold,
                                 nuif_stream_t s);

/* K/transformMaps.cl:91-127 transform_maps_kernel; pose: device NuiCLRigidTransform */
int nuif_transform_maps(const float* vmap, const float* nmap, float* vmap_dst, float* nmap_dst, int width,
                        int height, const NuiCLRigidTransform* pose, nuif_stream_t s);

/* K/tsdfVolume.cl:8-29 scaleDepthsKernel (in place) */
int nuif_scale_depths(float* depth, int width, int height, const NuiCLCameraParams* cam, nuif_stream_t s);

/* Fused frame preprocessing: OCL/NuiKinfuOpenCLFrame.cpp:314-367 UpdateVertexBuffers
 * (passing filter + bilateral + level-0 vertices) and OCL/NuiKinfuOpenCLDepthTracker.cpp:140-236
 * (pyramid + vertices per level) in as few launches as the data flow allows.
 *   depth_out      : passing-filter depth (W*H); if scale_for_tsdf, already scaled to ray
 *                    length (K/tsdfVolume.cl:8-29) for the dense scene's integration
 *   depth_pyr[l]   : filtered depth per level (level 0 = bilateral output)
 *   vmap_pyr[l]    : camera-space vertices per level (float3 AoS)               */
int nuif_frame_preprocess(const uint16_t* raw, float* depth_out, float* const* depth_pyr, float* const* vmap_pyr,
                          int levels, int width, int height, const NuiCLCameraParams* cam, const float* gaussian,
                          int radius, float sigma_depth2_inv_half, float depth_threshold, int scale_for_tsdf,
                          nuif_stream_t s);

/* ---- projective ICP (K/icp.cl:162-326, OCL/NuiKinfuOpenCLDepthTracker.cpp:240-504) ---- */

#define NUIF_ICP_TERMS 29
#define NUIF_ICP_MAX_LEVELS 4

/* Host-visible view of the device tracking state (nuif_icp_state_read). */
typedef struct nuif_icp_result {
    float R[9];          /* current pose estimate, camera -> world */
    float t[3];
    float error;         /* m_error of the last executed iteration */
    float count;         /* m_count of the last executed iteration */
    int32_t status;      /* 0 ok, 1 tracking failure (singular system), 2 iterations left unrun
                            (fewer launched than the frame's schedule needs) */
    int32_t iterations;  /* iterations that ran (not skipped by a break) */
    int32_t integrate;   /* integration gate decision of the last frame */
    int32_t frame;       /* frames finalised */
    double sums[NUIF_ICP_TERMS]; /* 29 sums of the last executed iteration */
    double Rd[9], td[3]; /* compensated order: the pose in FP64 */
    int32_t level_iterations[NUIF_ICP_MAX_LEVELS]; /* iterations that ran per level, last frame */
} nuif_icp_result;

/* Summation order of the 29 ICP sums */
#define NUIF_ICP_REDUCE_REFERENCE 0    /* the reference's float trees and LM host loop, bit for bit */
/* Kahan FP32 per lane, FP64 across lanes and tiles; Gauss-Newton on the undamped normal
 * equations (FP64 LDL^T), the increment applied as an SE(3) exponential to an FP64 pose,
 * a step below 1e-6 ends a level.  It replaces the reference's LM loop, whose accept /
 * reject decisions at the noise floor of the residual make float and FP64 runs diverge
 * by millimetres; this loop stays within hundredths of a millimetre of its FP64 run. */
#define NUIF_ICP_REDUCE_COMPENSATED 1
/* The compensated order with a photometric row per pixel beside the geometric one
 * (nuif_icp_photometric): joint geometric + photometric tracking.  The reference's joint
 * kernel (K/icp.cl:352-480) has no host loop; this is the project's own formulation. */
#define NUIF_ICP_REDUCE_JOINT 2

typedef struct nuif_icp_config {
    int levels;                            /* pyramid levels (<= 4) */
    int iterations[NUIF_ICP_MAX_LEVELS];   /* per level, level 0 = finest */
    float dist_threshold;                  /* NuiTrackerConfig::dist_threshold */
    float gate_threshold;                  /* integration metric threshold (0 = always) */
    int reduction;                         /* NUIF_ICP_REDUCE_* */
} nuif_icp_config;

/* Inputs of the joint order's photometric rows.  For a pixel whose geometric association
 * holds (current vertex p at the estimate, projected through the previous pose to the
 * full-resolution pixel (u, v)): I the current intensity at the pixel's level, I_ref and
 * g = (dI/du, dI/dv) the previous frame's intensity and its gradient bilinearly at (u, v)
 * (the previous frame's colours were taken at the previous pose); with
 * |I - I_ref| < max_difference and |g|^2 > min_gradient^2 the row is
 * sqrt(weight) [J x p, J, I - I_ref], J the gradient carried to world coordinates through
 * the previous camera, added to the normal equations beside the geometric row
 * [n x p, n, n . (q - p)]. */
typedef struct nuif_icp_photometric {
    const float* intensity_pyr[NUIF_ICP_MAX_LEVELS]; /* current intensities, W>>l x H>>l per level */
    const float* ref_model;                          /* W x H float4: the previous frame's
                                                        intensity, dI/du, dI/dv, 0 */
    float weight;                                    /* photometric weight (0.1) */
    float max_difference;
    float min_gradient;
} nuif_icp_photometric;
/* The joint order's inputs of a frame in one launch (no counterpart in the reference):
 * intensity_pyr[0] = nuif_bgra_to_intensity of the frame's colours, level l the
 * nuif_intensity_half_sample of level l - 1; ref_model = the intensities of ref_bgra (the
 * previous frame's colours) with their nuif_intensity_derivatives.  width and height
 * must be multiples of 2^(levels - 1). */
int nuif_icp_photometric_inputs(const uint8_t* bgra, const uint8_t* ref_bgra, int width, int height, int levels,
                                const nuif_icp_photometric* photometric, nuif_stream_t s);

/* State block: the tracking state plus the current iteration ("epoch") of the frame
 * (device memory, initialised by nuif_icp_state_init). */
size_t nuif_icp_state_bytes(void);
/* Workspace of nuif_icp_iteration for a width x height level-0 image. */
size_t nuif_icp_partials_bytes(int width, int height);
/* Reset tracking state: estimate = pose, last-integration pose = pose. */
int nuif_icp_state_init(void* state, const NuiCLRigidTransform* pose, nuif_stream_t s);
int nuif_icp_state_read(const void* state, nuif_icp_result* out, nuif_stream_t s); /* synchronises s */
/* Begin a frame: estimate <- pose, clear break/failure flags. */
int nuif_icp_begin(void* state, const NuiCLRigidTransform* pose, nuif_stream_t s);
/* One iteration at `level` (two launches: association + the tiles' partial sums, then one
 * block that reduces them in a fixed order and runs the update of the chosen order
 * (reference: the LM accept/reject, FP64 LLT solve and Euler increment of
 * OCL/NuiKinfuOpenCLDepthTracker.cpp:240-497)).  Skipped after a break at this level or a
 * failure.  vmap_level: current-frame vertices at the level; ref_vmap/ref_nmap: world maps
 * at full resolution; prev_pose: previous frame's transform; count_threshold: see
 * nuif_icp_count_thresholds.  At most 2^22 pixels per level.  Launches sharing one state
 * block or one partials workspace must not overlap. */
int nuif_icp_iteration(void* state, float* partials, const NuiCLCameraParams* cam, const float* vmap_level,
                       int level, const float* ref_vmap, const float* ref_nmap,
                       const NuiCLRigidTransform* prev_pose, float dist_threshold, uint32_t count_threshold,
                       int iter, int reduction, nuif_stream_t s);
/* A frame without tracking (the first frame, VO/NuiKinfuTrackingEngine.cpp:93-104): the
 * pose is kept, a history row is written and the frame integrates. */
int nuif_icp_record(void* state, NuiCLRigidTransform* pose, float* history, int capacity, nuif_stream_t s);
/* Finish the frame: on success pose <- estimate (with R_inv), and the integration gate
 * (VO/NuiKinfuTrackingEngine.cpp:113-127) is evaluated into the state.  If history is not
 * NULL, row (frame % capacity) of NUIF_POSE_ROW floats receives
 * [R (9, row-major), t (3), ok (1/0), error, count, integrated (1/0)]. */
#define NUIF_POSE_ROW 16
int nuif_icp_end(void* state, NuiCLRigidTransform* pose, float gate_threshold, float* history, int capacity,
                 nuif_stream_t s);
/* Shifting volume (OCL/NuiKinfuOpenCLShiftScene.cpp:353-364): after a shift of the local frame,
 * t3 (host float[3]) becomes the pose translation, the last integration translation and the
 * translation of the last history row; offset3 (host float[3], metres) is added to both
 * translations of the next integration gates (global = local + offset). */
int nuif_icp_set_translation(void* state, NuiCLRigidTransform* pose, float* history, int capacity, const float* t3,
                             const float* offset3, nuif_stream_t s);
/* All levels and iterations of one frame (OCL/NuiKinfuOpenCLDepthTracker.cpp:240-504
 * IterativeClosestPoint), no host synchronisation: begin, two launches per iteration of the
 * frame's schedule (an iteration after a break or a failure returns at once), end. */
int nuif_icp_estimate(void* state, float* partials, const NuiCLCameraParams* cam, const float* const* vmap_pyr,
                      const float* ref_vmap, const float* ref_nmap, NuiCLRigidTransform* pose,
                      const nuif_icp_config* cfg, float* history, int capacity, nuif_stream_t s);

/* nuif_icp_estimate without the partials workspace, for callers that cannot allocate
 * device memory (code under graph capture, tight serving loops, harnesses that forbid
 * hipMalloc); no counterpart in the reference.  The partial sums live in a static
 * device array of nuif_icp_static_scratch_bytes() (1 MiB) per GPU, held until the
 * process ends; an image whose nuif_icp_partials_bytes exceed it is refused with
 * NUIF_ERR_INVALID (1280 x 720 fits).  Nothing is allocated or freed.  Every call on a
 * GPU shares that GPU's array: calls that overlap on one GPU (different streams) race. */
/* nuif_icp_estimate in the joint order (cfg->reduction NUIF_ICP_REDUCE_JOINT; photometric
 * inputs from nuif_icp_photometric_inputs).  No counterpart in the reference: its
 * photometric terms (K/icp.cl:352-480) have no host loop. */
int nuif_icp_estimate_joint(void* state, float* partials, const NuiCLCameraParams* cam, const float* const* vmap_pyr,
                            const float* ref_vmap, const float* ref_nmap, NuiCLRigidTransform* pose,
                            const nuif_icp_config* cfg, const nuif_icp_photometric* photometric, float* history,
                            int capacity, nuif_stream_t s);

size_t nuif_icp_static_scratch_bytes(void);
int nuif_icp_estimate_static(void* state, const NuiCLCameraParams* cam, const float* const* vmap_pyr,
                             const float* ref_vmap, const float* ref_nmap, NuiCLRigidTransform* pose,
                             const nuif_icp_config* cfg, float* history, int capacity, nuif_stream_t s);

/* One session of nuif_icp_estimate_batch (the arguments of nuif_icp_estimate). */
typedef struct nuif_icp_session {
    void* state;                               /* nuif_icp_state_bytes() */
    void* partials;                            /* nuif_icp_partials_bytes(W, H) */
    NuiCLCameraParams cam;
    const float* vmap_pyr[NUIF_ICP_MAX_LEVELS];
    const float* ref_vmap;
    const float* ref_nmap;
    NuiCLRigidTransform* pose;                 /* device camera pose (previous frame in, new out) */
    nuif_icp_config cfg;
    float* history;
    int capacity;
    int32_t* outcome;                          /* NULL, or an int32 that the frame's end sets to 1
                                                  (tracked) or 0 (failed): device memory or pinned
                                                  host memory (nuif_host_alloc) */
    nuif_icp_photometric photometric;          /* cfg.reduction NUIF_ICP_REDUCE_JOINT only */
} nuif_icp_session;
/* Device bytes of the session table of nuif_icp_estimate_batch. */
size_t nuif_icp_batch_table_bytes(int n);
/* The frame's ICP of n <= 4096 sessions (any image sizes, any NUIF_ICP_REDUCE_* order) in lockstep steps:
 * step k is two launches for all sessions of an order, the 512-pixel tiles of every
 * session's k-th iteration (whatever its level: each session moves through its levels on
 * its own and is done after a break at its finest level or a failure) spread over the
 * grid, then one block per session for the reduction and the update.  The steps are as
 * many as the most iterations a session is configured for.  Results are bit for bit
 * those of nuif_icp_estimate (nuif_icp_estimate_joint for a joint-order session, whose
 * photometric inputs are made beforehand, e.g. by nuif_icp_photometric_inputs).  No host
 * synchronisation.  table:
 * nuif_icp_batch_table_bytes(n) of device memory, written on the stream (calls on one
 * stream may reuse it). */
int nuif_icp_estimate_batch(const nuif_icp_session* sessions, int n, void* table, nuif_stream_t s);
/* Device int32 flags inside a state block: integration gate of the last frame, and
 * tracking success of the last frame (raycast runs only if set). */
const int32_t* nuif_icp_gate_flag(const void* state);
const int32_t* nuif_icp_ok_flag(const void* state);
/* Count thresholds of the reference's stale-global-size rule. */
int nuif_icp_count_thresholds(int width, int height, const nuif_icp_config* cfg, uint32_t* thresholds);

/* ---- dense TSDF scene (OCL/NuiKinfuOpenCLScene.cpp, K/tsdfVolume.cl, K/raycast.cl) ---- */
/* (nuif_fusion_round, the batched frame of many dense sessions, follows the scene entry points) */

size_t nuif_tsdf_voxel_bytes(int storage);
/* K/tsdfVolume.cl:192-211 initializeVolume: (32767,0) / (0,0) and colour 0 */
int nuif_tsdf_reset(void* volume, uint8_t* color_volume, const TsdfParams* params, int storage, nuif_stream_t s);
/* Depth bounds of a frame for nuif_tsdf_integrate: the largest valid depth over square
 * pixel cells of 8, 16, 32, ... pixels (device floats, nuif_tsdf_depth_bounds_bytes). */
size_t nuif_tsdf_depth_bounds_bytes(int width, int height);
int nuif_tsdf_depth_bounds(const float* depth, int width, int height, float* bounds, const int32_t* gate,
                           nuif_stream_t s);
/* K/tsdfVolume.cl:31-190 integrateTsdfVolumeKernel.  depth: ray-length depth (scaled);
 * bgra/color_volume may be NULL; gate: optional device int32, integrate only if non-zero;
 * updated: optional device counter, incremented by the number of voxels updated;
 * depth_bounds: optional nuif_tsdf_depth_bounds of the same depth -- 32-voxel runs of a
 * column that lie more than the truncation behind every depth they can project to are
 * skipped (the reference's kernel skips each of their voxels; results are identical). */
int nuif_tsdf_integrate(const float* depth, const uint8_t* bgra, void* volume, uint8_t* color_volume,
                        const TsdfParams* params, const NuiCLCameraParams* cam, const NuiCLRigidTransform* pose,
                        int storage, int max_color_weight, const int32_t* gate, unsigned long long* updated,
                        const float* depth_bounds, uint8_t* surface_blocks, nuif_stream_t s);
/* K/raycast.cl:128-277 raycastKernel: world vmap/nmap (float3 AoS) and optional cmap */
int nuif_tsdf_raycast(const void* volume, const uint8_t* color_volume, const TsdfParams* params,
                      const NuiCLCameraParams* cam, const NuiCLRigidTransform* pose, float* vmap, float* nmap,
                      uint8_t* cmap, int storage, nuif_stream_t s);
/* Measurement, no counterpart in the reference: the number of distinct voxels the raycast
 * reads at this pose (with these surface blocks), i.e. U of its algorithmic bytes
 * B * U + 28 * W * H.  The same march as nuif_tsdf_raycast (no wrap), with every voxel
 * read marked in a bitset; vmap / nmap receive the raycast.  workspace:
 * nuif_tsdf_footprint_bytes(params) of device memory.  Synchronises s; returns U. */
size_t nuif_tsdf_footprint_bytes(const TsdfParams* params);
/* The launch schedule of the library: raycast samples in flight, integration walk steps in
 * flight, integration column chunks in flight (compiled in), and the integration z-segment
 * voxels of a single-session launch; nuif_schedule_segment gives the segment of a launch
 * shared by that many sessions (a batched round). */
int nuif_schedule(int32_t* out4);
int nuif_schedule_segment(int sessions);
long long nuif_tsdf_raycast_footprint(const void* volume, const TsdfParams* params, const NuiCLCameraParams* cam,
                                      const NuiCLRigidTransform* pose, int storage, const uint8_t* surface_blocks,
                                      float* vmap, float* nmap, void* workspace, nuif_stream_t s);
/* Same, skipped entirely when the device int32 *enable is zero (tracking failure). */
int nuif_tsdf_raycast_gated(const void* volume, const uint8_t* color_volume, const TsdfParams* params,
                            const NuiCLCameraParams* cam, const NuiCLRigidTransform* pose, float* vmap,
                            float* nmap, uint8_t* cmap, int storage, const int32_t* enable,
                            const uint8_t* surface_blocks, nuif_stream_t s);
/* ---- shifting volume (OCL/NuiKinfuOpenCLShiftScene.cpp) ----------------------------
 * wrap: int32[3] voxel wrap (any integers, reduced modulo the resolution): the voxel at
 * logical (x, y, z) is stored at ((x + w.x) % Nx, (y + w.y) % Ny, (z + w.z) % Nz)
 * (K/tsdfVolume.cl:71,104, K/raycast.cl:62,184); NULL or 0 = the plain volume. */
/* Device bytes of the walk-state workspace of nuif_tsdf_integrate_wrapped (8 B per column
 * and z segment after the first, plus 8 B per such segment). */
size_t nuif_tsdf_walk_states_bytes(const TsdfParams* params);
/* nuif_tsdf_integrate in a wrapped volume; the surface blocks are marked at stored positions.
 * walk_states: optional workspace (nuif_tsdf_walk_states_bytes) -- the integration then
 * computes every column's incremental walk state at the start of each z segment once, instead
 * of replaying it per segment (same values).  The workspace is overwritten by every call that
 * gets it: calls that may overlap (different streams) need workspaces of their own. */
int nuif_tsdf_integrate_wrapped(const float* depth, const uint8_t* bgra, void* volume, uint8_t* color_volume,
                                const TsdfParams* params, const NuiCLCameraParams* cam,
                                const NuiCLRigidTransform* pose, int storage, int max_color_weight,
                                const int32_t* wrap, const int32_t* gate, unsigned long long* updated,
                                const float* depth_bounds, uint8_t* surface_blocks, float* walk_states,
                                nuif_stream_t s);
/* nuif_tsdf_raycast_gated in a wrapped volume.  Every trilinear corner is wrapped on its own
 * (the reference wraps the first corner only); the surface-block skip is used only when the
 * wrap is 0. */
int nuif_tsdf_raycast_wrapped(const void* volume, const uint8_t* color_volume, const TsdfParams* params,
                              const NuiCLCameraParams* cam, const NuiCLRigidTransform* pose, float* vmap, float* nmap,
                              uint8_t* cmap, int storage, const int32_t* wrap, const int32_t* enable,
                              const uint8_t* surface_blocks, nuif_stream_t s);
/* K/tsdfVolume.cl:192-211 initializeVolume over range[3] voxels from offset `wrap`
 * (clearSlice, OCL/NuiKinfuOpenCLShiftScene.cpp:72-117) */
int nuif_tsdf_reset_slice(void* volume, uint8_t* color_volume, const TsdfParams* params, int storage,
                          const int32_t* wrap, const int32_t* range, nuif_stream_t s);
/* K/extract.cl:80-190 fetchShiftVolumeKernel (fetchSlice, OCL/NuiKinfuOpenCLShiftScene.cpp:119-245,
 * and the shifting scene's Volume2CLVertices, :366-): range[0] x range[1] columns from stored
 * column (x + wrap.x, y + wrap.y), walking voxel_z = wrap.z .. wrap.z + range[2] - 1 at stored
 * z (voxel_z + wrap.z) % Nz as the reference does, points placed by the volume's voxel offset
 * `offset` (int32[3]) as the reference computes it.  wrap in [0, N]; output as
 * nuif_points_dense, in (y, x, z) order; scratch: nuif_points_dense_scratch_bytes. */
int nuif_points_dense_slice(const void* volume, const uint8_t* color_volume, const TsdfParams* params, int storage,
                            const int32_t* wrap, const int32_t* offset, const int32_t* range, float* points,
                            float* colors, uint32_t capacity, uint32_t* total, void* scratch, nuif_stream_t s);

/* Surface blocks of a volume: one byte per 8 x 8 x 8 voxel block (y-major, z fastest),
 * then one byte per 32 x 32 x 32 superblock (the OR of its blocks), set by
 * nuif_tsdf_integrate (surface_blocks argument) when it stores a tsdf <= 0 and cleared
 * only by nuif_tsdf_surface_blocks_clear (with the volume's reset).  The raycast does not
 * read the voxels of blocks left at zero where its exit tests cannot need their values,
 * and crosses clear superblocks without locating its samples; its output is identical
 * with or without them (NULL). */
size_t nuif_tsdf_surface_blocks_bytes(const TsdfParams* params);
int nuif_tsdf_surface_blocks_clear(uint8_t* surface_blocks, const TsdfParams* params, nuif_stream_t s);
/* Effective truncation distance (OCL/NuiKinfuOpenCLScene.cpp:41: >= 2.1 * max cell). */
int nuif_tsdf_params(int nx, int ny, int nz, float dx, float dy, float dz, float trunc, TsdfParams* out);

/* ---- the batched frame of many dense sessions (the fusion server's round) ---------------
 * The reference runs one session's frame on one queue (VO/NuiKinfuTrackingEngine.cpp:68-142
 * over CLU/NuiOpenCLGlobal.cpp:175).  nuif_fusion_round runs one tracked frame of each of n
 * dense sessions with one launch per stage for all of them: preprocessing (passing +
 * bilateral filter, ray-length depth, vertex pyramid), the ICP of every session in
 * lockstep ICP steps (nuif_icp_estimate_batch), depth bounds, column walk states,
 * integration (gated by each session's integration flag) and raycast into the feedback maps
 * (gated by its tracking flag).  Every session's results equal those of its own
 * nuif_frame_preprocess / nuif_icp_estimate / nuif_tsdf_depth_bounds /
 * nuif_tsdf_integrate_wrapped / nuif_tsdf_raycast_gated sequence, bit for bit.  Volumes need
 * Nx, Ny multiples of 8 and Nz a multiple of 32, no wrap. */
typedef struct nuif_round_session {
    const uint16_t* depth;                     /* W*H raw depth, mm (device) */
    float* depth_out;                          /* W*H passing-filter depth scaled to ray length */
    float* depth_pyr[NUIF_ICP_MAX_LEVELS];     /* filtered depth per level */
    float* bounds;                             /* nuif_tsdf_depth_bounds_bytes(W, H) */
    const float* gaussian;                     /* 2*radius+1 host floats */
    int radius;
    float sigma_depth2_inv_half, depth_threshold;
    nuif_icp_session icp;                      /* vmap_pyr: written by the preprocessing; ref maps:
                                                  the feedback maps below (previous frame) */
    void* volume;
    uint8_t* color_volume;                     /* NULL: no colour integration */
    const uint8_t* bgra;                       /* W*H BGRA frame or NULL */
    TsdfParams params;
    int storage;                               /* NUIF_TSDF_* */
    int max_color_weight;
    unsigned long long* updated;               /* device counter or NULL */
    uint8_t* surface_blocks;                   /* nuif_tsdf_surface_blocks_bytes or NULL */
    float* walk_states;                        /* nuif_tsdf_walk_states_bytes or NULL */
    float* vmap;                               /* feedback maps: raycast output, W*H float3 */
    float* nmap;
    uint8_t* cmap;                             /* W*H uchar4 or NULL */
    uint8_t* prev_bgra;                        /* joint ICP order: the previous frame's colours
                                                  (W*H BGRA), the photometric rows' model;
                                                  replaced by this frame's colours */
} nuif_round_session;
/* Bytes of the round's job table (device memory) and of its host staging copy (pinned host
 * memory, nuif_host_alloc) for n sessions. */
size_t nuif_fusion_round_table_bytes(int n);
/* table / staging: nuif_fusion_round_table_bytes(n) each; the staging copy may be reused once
 * the stream has passed this call's upload (e.g. after waiting for the round). */
int nuif_fusion_round(const nuif_round_session* sessions, int n, void* table, void* staging, nuif_stream_t s);
/* The host part of nuif_fusion_round alone: writes the job table into staging and returns
 * a signature of the round's launches (grids, groups, kernel choices, the table and staging
 * addresses).  A round captured into a graph (nuif_graph_begin / nuif_fusion_round /
 * nuif_graph_end) replays a later round whose signature is the same, its table written by
 * this call first.  No device work.  No counterpart in the reference. */
int nuif_fusion_round_prepare(const nuif_round_session* sessions, int n, void* table, void* staging,
                              uint64_t* signature);
/* nuif_fusion_round with 6 events recorded on s around its stages (no counterpart in the
 * reference): [0] start, [1] after preprocessing and the pyramid, [2] after the depth bounds,
 * [3] after the ICP, [4] after the integration, [5] after the raycast. */
int nuif_fusion_round_timed(const nuif_round_session* sessions, int n, void* table, void* staging,
                            nuif_event_t* events6, nuif_stream_t s);

/* ---- prefix sum and compaction (OCL/NuiOpenCLPrefixSum.cpp:75-300, K/prefixSum.cl:163-232) ---- */

size_t nuif_scan_scratch_bytes(size_t n);
/* out[i] = sum of in[0..i-1]; total (device, optional) = sum of all; in == out allowed */
int nuif_exclusive_scan_u32(const uint32_t* in, uint32_t* out, size_t n, uint32_t* total, void* scratch,
                            nuif_stream_t s);
/* the same over (flags[i] > 0) */
int nuif_exclusive_scan_flags(const uint8_t* flags, uint32_t* out, size_t n, uint32_t* total, void* scratch,
                              nuif_stream_t s);
/* ids[scan[i]] = i for every i with flags[i] > 0 (order preserving; the reference's
 * compactifyValidPrefixSum_kernel is racy) */
int nuif_compact_flags(const uint8_t* flags, const uint32_t* scan, size_t n, int32_t* ids, nuif_stream_t s);

/* ---- dense marching cubes (K/marchingCube.cl:100-256, OCL/NuiKinfuOpenCLScene.cpp:617-786) ---- */

/* Reference cap on output vertices (K/gpu_def.h:3) */
#define NUIF_MC_MAX_VERTICES 5000000
size_t nuif_mc_scratch_bytes(const TsdfParams* params);
/* Triangles of the zero level set: vertices float3 (3 per triangle) and colours float4 RGBA
 * (nearest corner of the colour volume, or (0,0,0,1) without one) written in a fixed tile
 * order; *total (device uint32) receives the vertex count before capping at `capacity`. */
int nuif_mc_dense(const void* volume, const uint8_t* color_volume, const TsdfParams* params, int storage,
                  float* vertices, float* colors, uint32_t capacity, uint32_t* total, void* scratch,
                  nuif_stream_t s);

/* ---- dense point cloud (K/extract.cl:4-78, OCL/NuiKinfuOpenCLScene.cpp:436-610) ---- */

size_t nuif_points_dense_scratch_bytes(const TsdfParams* params);
/* fetchVolumeKernel (FetchVolume / Volume2CLVertices): per (x, y) column along z, a point at
 * the voxel centre where the tsdf is exactly 0, else at the linear zero crossing with the
 * voxel below where the sign changes (weights > 0 on both).  Points float3 (x and y centred,
 * y flipped); colours float4 RGBA from the colour volume (bytes 0,1,2 for a zero, 2,1,0 for a
 * crossing, as the reference reads them; (0,0,0,1) without one; colors may be NULL).  Written
 * in (y, x, z) order; *total (device uint32) receives the count before capping at `capacity`.
 * surface_blocks: optional flags maintained by nuif_tsdf_integrate for this volume -- steps
 * whose voxels lie in clear blocks are skipped unread (same result). */
int nuif_points_dense(const void* volume, const uint8_t* color_volume, const TsdfParams* params, int storage,
                      const uint8_t* surface_blocks, float* points, float* colors, uint32_t capacity,
                      uint32_t* total, void* scratch, nuif_stream_t s);

/* OCL/NuiKinfuOpenCLScene.cpp:788-984 Volume2Mesh on the device: marching cubes in the
 * reference's (y, x, z) cube order with its weight test, vertices keyed by the ordered voxel
 * pair of their cube edge and numbered in first-use order, positions centred on all axes
 * (getNodeCoo, :84-91).  points: float3 x max_vertices, indices: uint32 x max_vertices
 * (device); *n_points / *n_indices are host counts.  Synchronous (the count decides the
 * work); fails with NUIF_ERR_INVALID when the corners exceed max_vertices (*n_indices then
 * holds the needed count). */
size_t nuif_mesh_dense_scratch_bytes(const TsdfParams* params, uint32_t max_vertices);
int nuif_mesh_dense(const void* volume, const TsdfParams* params, int storage, uint32_t max_vertices, float* points,
                    uint32_t* indices, uint32_t* n_points, uint32_t* n_indices, void* scratch, nuif_stream_t s);

/* ---- voxel hashing (OCL/NuiKinfuOpenCLHashScene.cpp, K/KinfuHashScene.cl,
 * K/KinfuHashRaycast.cl, K/KinfuHashVisualization.cl) ----------------------------
 * One device allocation holds the whole scene: header, hash table (2^b buckets of
 * 4 entries, a full bucket overflowing into the next ones), per-entry flags, the
 * visible list and the brick pool (8^3 voxels per brick, voxel index x + 8y + 64z).
 * Replaces NuiKinfuVoxelBlockHash (OCL/NuiKinfuVoxelBlockHash.cpp:24-213). */
typedef struct nuif_hash_params {
    float voxel_size;   /* m_virtualVoxelSize, metres (VO/NuiHashingSDFConfig.cpp:3-15) */
    float trunc_scale;  /* m_truncScale: band half-width and sdf normalisation, metres */
    float truncation;   /* m_truncation: raycast step scale, metres */
    int weight_max;     /* m_integrationWeightMax, 1..255 */
    int buckets_log2;   /* 2^b buckets of 4 entries, b in 4..24 */
    int bricks;         /* brick pool capacity */
    int storage;        /* NUIF_TSDF_S16: the reference's short sdf*32767 + weight;
                           NUIF_TSDF_F16: FP16 sdf + uint16 weight */
    int color;          /* colour per voxel (uchar4) */
} nuif_hash_params;

size_t nuif_hash_scene_bytes(const nuif_hash_params* p);
/* Byte offsets inside the scene: header, entries (16 B: uint64 key, int32 brick, pad),
 * flags, visible list (int32 entry ids), bricks (uint32 per voxel), colours (0 if none),
 * then the entry count and the total size. */
int nuif_hash_layout(const nuif_hash_params* p, size_t* offsets8);
/* Empty table, empty pool (sdf 0, weight 0, colour 0): K/KinfuHashScene.cl:5-40 reset kernels. */
int nuif_hash_reset(void* scene, const nuif_hash_params* p, nuif_stream_t s);
/* OCL/NuiKinfuOpenCLHashScene.cpp:404-424 AllocateSceneFromDepth: every block that a pixel's
 * truncation band touches is inserted (K/KinfuHashScene.cl:52-194); the visible list becomes
 * the touched blocks plus the previously visible blocks with a corner inside the image
 * (:197-285), compacted by the device-wide scan.  depth: passing-filter depth (metres, NaN
 * invalid); gate: device flag, 0 skips the frame (NULL: always). */
int nuif_hash_allocate(void* scene, const nuif_hash_params* p, const float* depth, const NuiCLCameraParams* cam,
                       const NuiCLRigidTransform* pose, const int32_t* gate, nuif_stream_t s);
/* K/KinfuHashScene.cl:387-470 integrateIntoScene_kernel over the visible list; bgra: W*H*4
 * or NULL; updated: device counter of voxels written (NULL: none). */
int nuif_hash_integrate(void* scene, const nuif_hash_params* p, const float* depth, const uint8_t* bgra,
                        const NuiCLCameraParams* cam, const NuiCLRigidTransform* pose, const int32_t* gate,
                        unsigned long long* updated, nuif_stream_t s);
/* Ray ranges of the visible blocks (K/KinfuHashRaycast.cl:4-76, exact min / max) into
 * range2 (W*H float2), then castRay (:155-448): world vertex map, normal map (may be NULL)
 * and colour map (W*H uchar4, may be NULL).  enable: device flag, 0 skips (NULL: always). */
int nuif_hash_raycast(void* scene, const nuif_hash_params* p, const NuiCLCameraParams* cam,
                      const NuiCLRigidTransform* pose, float* range2, float* vmap, float* nmap, uint8_t* cmap,
                      const int32_t* enable, nuif_stream_t s);
/* Workspace of nuif_hash_forward_render for a W x H image. */
size_t nuif_hash_forward_render_bytes(int width, int height);
/* ForwardRender (OCL/NuiKinfuOpenCLHashScene.cpp:991-1012) after the ray ranges of
 * nuif_hash_raycast: the points of the previous world vertex map vmap (and colours cmap,
 * may be NULL) move to the pixels they project to at `pose` (forwardProject_kernel,
 * K/KinfuHashRaycast.cl:473-506; the nearest point wins a pixel); pixels that received
 * none and have a non-empty ray range are raycast (findMissingPoints_kernel :508-529,
 * genericRaycastMissingPoints_kernel :531-559); normals are recomputed from the vertex
 * map with a 0.3 m threshold (Vertex2Normal, :943-989).  vmap and cmap are updated in
 * place, nmap is overwritten.  reference_transform != 0 projects as the reference does
 * (camera->world applied to the world points, scaled by the voxel size); 0 uses
 * world->camera.  work: nuif_hash_forward_render_bytes(W, H) bytes of device memory. */
int nuif_hash_forward_render(void* scene, const nuif_hash_params* p, const NuiCLCameraParams* cam,
                             const NuiCLRigidTransform* pose, float* range2, float* vmap, float* nmap, uint8_t* cmap,
                             void* work, int reference_transform, const int32_t* enable, nuif_stream_t s);
size_t nuif_hash_mesh_scratch_bytes(const nuif_hash_params* p);
/* K/KinfuHashVisualization.cl:88-271 meshHashScene_kernel over the visible list: count,
 * scan, emit; vertices float3 in metres (3 per triangle), colours float4 (may be NULL);
 * *total (device uint32) receives the vertex count before capping at `capacity`. */
int nuif_hash_mesh(void* scene, const nuif_hash_params* p, float* vertices, float* colors, uint32_t capacity,
                   uint32_t* total, void* scratch, nuif_stream_t s);
/* ---- host swapping of hash bricks (OCL/NuiKinfuOpenCLHashGlobalCache.cpp:121-538,
 * K/KinfuHashSwapping.cl).  Every entry has a swap state (0: nothing to do, 1: visible and
 * waiting for its stored data, 2: synchronised); bricks come from a free-brick stack.  Per
 * integrated frame: nuif_hash_swap_prepare after nuif_hash_allocate (visible entries not in
 * state 2 go to 1; visible entries without a brick take one), then after nuif_hash_integrate
 * the swap-in list (state 1 with a brick, visible-list order) receives the host's stored
 * blocks through nuif_hash_swap_in (weight-averaged into the brick, state 2), and the
 * swap-out list (state 2 with a brick, not visible this frame) leaves through
 * nuif_hash_swap_out (brick copied out, zeroed and freed, ptr -1, state 0).  The reference
 * moves at most SDF_TRANSFER_BLOCK_NUM blocks per frame; here every listed entry moves. */
#define NUIF_HASH_SWAP_IN 0
#define NUIF_HASH_SWAP_OUT 1
/* Byte offsets of the swap states (uint8 per entry) and the free-brick stack (int32 per brick). */
int nuif_hash_swap_layout(const nuif_hash_params* p, size_t* offsets2);
/* K/KinfuHashScene.cl:383-384 (swap states of the visible list) and K/KinfuHashSwapping.cl:6-21
 * reAllocateSwappedOutVoxelBlocks_kernel; gate: device flag (NULL: always). */
int nuif_hash_swap_prepare(void* scene, const nuif_hash_params* p, const int32_t* gate, nuif_stream_t s);
size_t nuif_hash_swap_scratch_bytes(const nuif_hash_params* p);
/* which = NUIF_HASH_SWAP_IN / _OUT: entry ids (device int32, capacity = entry count) and their
 * number (*count, device uint32). */
int nuif_hash_swap_list(void* scene, const nuif_hash_params* p, int which, int32_t* ids, uint32_t* count,
                        void* scratch, nuif_stream_t s);
/* K/KinfuHashSwapping.cl:57-118 integrateOldIntoActiveData_kernel for n listed entries: words
 * (device, 512 per entry, the scene's voxel words; zeros where nothing was stored) and colours
 * (uchar4 per voxel or NULL) are the stored blocks. */
int nuif_hash_swap_in(void* scene, const nuif_hash_params* p, const int32_t* ids, uint32_t n, const uint32_t* words,
                      const uint8_t* colors, nuif_stream_t s);
/* K/KinfuHashSwapping.cl:135-180 moveActiveDataToTransferBuffer + cleanVoxelMemory for n listed
 * entries: their bricks to words / colours (device, 512 per entry; colours may be NULL). */
int nuif_hash_swap_out(void* scene, const nuif_hash_params* p, const int32_t* ids, uint32_t n, uint32_t* words,
                       uint8_t* colors, nuif_stream_t s);

size_t nuif_hash_points_scratch_bytes(const nuif_hash_params* p);
/* K/KinfuHashVisualization.cl:5-52 fetchHashScene_kernel (OCL/NuiKinfuOpenCLHashScene.cpp:1066-1190
 * Volume2CLVertices): every voxel of a visible block with weight > 0 and |short sdf| <
 * 32767 * trunc_scale (FP16 storage: |sdf| < trunc_scale) becomes a point at its centre
 * ((8 * block + local) + 0.5) * voxel_size; colours float4 RGBA (voxel bytes 0..2 / 255, or
 * (0,0,0,1) without colour; may be NULL).  Written in visible-list order, voxel index
 * ascending inside a block; *total (device uint32) receives the count before the cap. */
int nuif_hash_points(void* scene, const nuif_hash_params* p, float* points, float* colors, uint32_t capacity,
                     uint32_t* total, void* scratch, nuif_stream_t s);
/* Host copy of [bricks handed out, visible-list length, failed insertions, occupied
 * entries]; synchronises s. */
int nuif_hash_stats(const void* scene, uint32_t* out4, nuif_stream_t s);

/* ---- photometric (colour) tracker (OCL/NuiKinfuOpenCLColorTracker.cpp) ---- */

/* K/transformMaps.cl:405-446 half_sample_bgra_kernel: dst (dst_width x dst_height BGRA)
 * = integer mean of the (2r+1)^2 window around (2x, 2y) of src (2*dst_width x
 * 2*dst_height) clipped to the image, alpha 255. */
int nuif_color_half_sample(const uint8_t* src, uint8_t* dst, int dst_width, int dst_height, int radius,
                           nuif_stream_t s);
/* K/transformMaps.cl:513-563 bgra_gradient_kernel: Sobel / 8 per channel, alpha 255
 * inside and all zero on the border.  sign = 0: uint8 values (saturating, as the
 * reference's conversion on the GPU); sign = 1: int8 values (the corrected tracker). */
int nuif_color_gradient(const uint8_t* bgra, uint8_t* gx, uint8_t* gy, int width, int height, int sign,
                        nuif_stream_t s);
/* SubSampleColors + GradientBuffers (OCL/NuiKinfuOpenCLColorTracker.cpp:125-242): level 0
 * = a copy of bgra, levels 1.. by nuif_color_half_sample (radius 1), then the gradients
 * of every level.  colors / gx / gy: `levels` device pointers each (host arrays). */
int nuif_color_pyramid(const uint8_t* bgra, uint8_t* const* colors, uint8_t* const* gx, uint8_t* const* gy,
                       int width, int height, int levels, int sign, nuif_stream_t s);
/* K/colorIcp.cl:7-57 color_difference_kernel (GetColorsDifference,
 * OCL/NuiKinfuOpenCLColorTracker.cpp:244-340): out2 (device uint64[2]) = exact
 * {sum over valid points of |obs - known|^2, valid count}, the model points
 * transformed by `pose` (R p + t, as the reference passes the camera pose).  The
 * reference's result is out2[0] * (W*H / out2[1]). */
int nuif_color_difference(const float* vmap, const uint8_t* colors, const uint8_t* level_colors,
                          const NuiCLCameraParams* cam, int level, const NuiCLRigidTransform* pose,
                          unsigned long long* out2, nuif_stream_t s);

/* Host view of the colour tracker state (nuif_color_icp_read). */
typedef struct nuif_color_icp_result {
    int32_t ok;          /* 0: the normal matrix was singular or not positive definite */
    int32_t iterations;  /* iterations that ran */
    float error;         /* sqrt(sum |diff|) / count of the last iteration */
    float count;
    float R[9], t[3];    /* the transform applied to the model points */
    float sums[29];      /* the last iteration's sums in the kernel's slot order */
} nuif_color_icp_result;

size_t nuif_color_icp_state_bytes(void);
size_t nuif_color_icp_partials_bytes(int width, int height);
/* Starts a track from the camera pose.  corrected = 0: the reference's host decode
 * (OCL/NuiKinfuOpenCLColorTracker.cpp:341-578 as written); 1: the sums decoded in the
 * kernel's order, world -> camera transform, signed gradients (nuif_color_gradient
 * sign = 1), Gauss-Newton step. */
int nuif_color_icp_begin(void* state, const NuiCLRigidTransform* pose, int corrected, nuif_stream_t s);
/* One iteration at `level` over all W*H model points (world vertices vmap, colours
 * with alpha 255 where valid): color_icp_block_kernel + compute_color_sums + the host
 * update, in one launch.  Skipped after a break at this level or a failure. */
int nuif_color_icp_iteration(void* state, float* partials, const NuiCLCameraParams* cam, int level,
                             int rotation_only, float dist_threshold, const float* vmap, const uint8_t* colors,
                             const uint8_t* level_colors, const uint8_t* level_gx, const uint8_t* level_gy,
                             nuif_stream_t s);
/* UpdateCameraTransform: writes the estimate into pose (R, R_inv, t) unless the track failed. */
int nuif_color_icp_end(const void* state, NuiCLRigidTransform* pose, nuif_stream_t s);
/* ColorIterativeClosestPoint: begin, levels - 1 .. 0 x iterations[level], end; no host sync. */
int nuif_color_icp_estimate(void* state, float* partials, const NuiCLCameraParams* cam, int levels,
                            const int* iterations, const int* rotation_only, float dist_threshold, const float* vmap,
                            const uint8_t* colors, const uint8_t* const* level_colors,
                            const uint8_t* const* level_gx, const uint8_t* const* level_gy,
                            NuiCLRigidTransform* pose, int corrected, nuif_stream_t s);
/* Synchronises s and copies the state out. */
int nuif_color_icp_read(const void* state, nuif_color_icp_result* out, nuif_stream_t s);

/* ---- joint depth + intensity ICP terms (K/icp.cl:328-480; the reference has no host loop) ---- */

/* K/transformMaps.cl:372-384 bgra_to_intensity_kernel: 0.299 R + 0.587 G + 0.114 B in [0, 1],
 * NaN where alpha is 0. */
int nuif_bgra_to_intensity(const uint8_t* bgra, float* intensities, size_t n, nuif_stream_t s);
/* K/transformMaps.cl:448-511 intensity_derivatives_kernel: Sobel / 8 (float2 per pixel) with
 * the neighbours one row apart, NaN on the border and next to NaN intensities (the
 * reference steps by the image height and leaves those pixels unwritten). */
int nuif_intensity_derivatives(const float* intensities, float* derivatives, int width, int height,
                               nuif_stream_t s);
/* Intensities one pyramid level down (no counterpart in the reference): dst (dst_width x
 * dst_height) = the mean of the valid (non-NaN) ones of the 2 x 2 source pixels, summed in
 * row order, NaN when none is valid; src is 2 dst_width x 2 dst_height. */
int nuif_intensity_half_sample(const float* src, float* dst, int dst_width, int dst_height, nuif_stream_t s);
size_t nuif_intensity_icp_workspace_bytes(int width, int height);
/* K/icp.cl:352-480 intensity_icp_block_kernel + the 128-wide reduction trees + the
 * sequential sum of the second-level results, at `level` (all maps W >> level by H >>
 * level: current vertices / normals / intensities, previous vertices / normals /
 * intensities / derivatives).  pose: current estimate applied to the current vertices;
 * prev_pose: the previous camera (transformInverse); euler3 (host): the Euler angles of
 * the evalR_d* derivatives; thresholds4 (host): distance, angle (sine), intensity
 * difference, minimum gradient.  sums29: device float[29], the reference's slot order.
 * workspace: nuif_intensity_icp_workspace_bytes(W >> level, H >> level) bytes. */
int nuif_intensity_icp_sums(const NuiCLCameraParams* cam, int level, const float* vmap, const float* nmap,
                            const float* intensities, const NuiCLRigidTransform* pose, const float* vmap_prev,
                            const float* nmap_prev, const float* intensities_prev, const float* derivatives_prev,
                            const NuiCLRigidTransform* prev_pose, const float* euler3, const float* thresholds4,
                            void* workspace, float* sums29, nuif_stream_t s);

#ifdef __cplusplus
}
#endif

#endif /* NUIF_H */
