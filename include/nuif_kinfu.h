/*
 * nuif_kinfu.h - C ABI of the tracking engine and the multi-session fusion
 * server (libnuif_kinfu.so), for hosts that are not C++ (ctypes, cgo, JNI).
 *
 * The engine entry points drive NuiKinfuTrackingEngine::RunTracking with a
 * dense scene as the reference's tracking manager does
 * (src/SLAM/VisualOdometry/NuiKinfuTrackingManager.cpp:53-108 calling
 * src/SLAM/VisualOdometry/NuiKinfuTrackingEngine.cpp:68-142).  The reference
 * runs one session on one global in-order queue
 * (src/OpenCLUtilities/NuiOpenCLGlobal.cpp:175); the server entry points keep
 * many sessions resident on one GPU, each on its own stream, admitted against
 * a device-memory budget.
 *
 * Every function returns 0 (or a documented non-negative value) on success and
 * a negative status on failure; nuifk_last_error() gives the calling thread's
 * last message.  No C++ exception crosses this boundary.
 */
#ifndef NUIF_KINFU_H
#define NUIF_KINFU_H

#include <stddef.h>
#include <stdint.h>

#include "nuif.h"

#ifdef __cplusplus
extern "C" {
#endif

/* One session: camera, tracker (VO/NuiTrackerConfig.cpp:5-24) and dense volume
 * (VO/NuiKinfuVolumeConfig.cpp:5-15) settings. */
typedef struct nuifk_desc {
    int width, height;
    float fx, fy, cx, cy;
    float depth_min, depth_max;     /* metres */
    int filter_radius;
    float sigma_space2_inv_half, sigma_depth2_inv_half, depth_threshold, dist_threshold;
    int levels;
    int iterations[4];              /* level 0 = finest */
    float integration_threshold;    /* 0 = integrate every frame */
    int resolution[3];
    float dimensions[3];
    float tranc_dist;
    int color;                      /* colour volume on / off */
    int storage;                    /* NUIF_TSDF_* */
    int use_graph;                  /* replay tracked frames from a captured hipGraph */
    int icp_reduction;              /* NUIF_ICP_REDUCE_* */
    int scene;                      /* NUIFK_SCENE_DENSE or NUIFK_SCENE_HASH */
    /* voxel hashing (VO/NuiHashingSDFConfig.cpp:3-15); storage S16 or F16, color as above */
    float hash_voxel_size, hash_trunc_scale, hash_truncation;
    int hash_weight_max;
    int hash_bricks;                /* brick pool (m_numSDFBlocks) */
    int hash_forward_render;        /* NUIFK_RENDER_*: how the hash scene renders the ICP maps */
    /* shifting volume (VO/NuiKinfuVolumeConfig.cpp:13-14): shift threshold in voxels (> 0) and
     * translateBasis (metres) */
    int voxel_shift;
    float volume_basis[3];
    int hash_swapping;              /* m_bUseSwapping: host swapping of hash bricks (no graph replay) */
    int hash_buckets;               /* m_hashNumBuckets: the table has at least this many buckets of 4
                                     * entries, and enough for a load of at most 1/2 of hash_bricks
                                     * (0: the reference's default 512 x 512) */
} nuifk_desc;

#define NUIFK_SCENE_DENSE 0         /* OCL/NuiKinfuOpenCLScene: dense TSDF volume */
#define NUIFK_SCENE_HASH 1          /* OCL/NuiKinfuOpenCLHashScene: voxel hashing */
#define NUIFK_SCENE_SHIFT 2         /* OCL/NuiKinfuOpenCLShiftScene: dense volume that follows the camera
                                       (no graph replay: each integrated frame waits for its pose) */
#define NUIFK_RENDER_CASTRAY 0      /* castRay at every pixel (m_bUseForwardRender false) */
#define NUIFK_RENDER_FORWARD 1      /* ForwardRender, world->camera projection */
#define NUIFK_RENDER_FORWARD_REFERENCE 2 /* ForwardRender with the reference's camera->world projection */

const char* nuifk_last_error(void);

/* ---- one session: NuiKinfuTrackingEngine + its dense or hash NuiKinfuScene ---- */
int nuifk_create(const nuifk_desc* d, void** out);
int nuifk_destroy(void* h);
/* One frame (VO/NuiKinfuTrackingEngine.h:36-42 RunTracking).  depth: W*H uint16 mm,
 * a device pointer if on_device else a host pointer; bgra: W*H*4 bytes or NULL.
 * sync != 0 waits and returns 1 (tracked) / 0 (tracking failure); sync == 0 returns 1. */
int nuifk_run(void* h, const uint16_t* depth, int on_device, const uint8_t* bgra, int sync);
int nuifk_synchronize(void* h);                    /* 1: no failure so far, 0: a failure */
int nuifk_reset(void* h);
/* Pose history rows (NUIF_POSE_ROW floats) of the first n frames; returns the count.  The
 * translation is global (local + the shifting volume's offset, NuiCameraPos::getGlobalTranslation).
 * Every frame since the last reset has a row, a failed one too (ok = 0, the pose it kept); the
 * reference returns before m_poses.push_back on a failed frame (VO/NuiKinfuTrackingEngine.cpp:107-139). */
int nuifk_poses(void* h, float* out, int n);
/* Rows of the device pose ring (default 65536) are copied to the host before the ring wraps
 * over them, which synchronises the session at most once every `capacity` frames. */
int nuifk_set_history_capacity(void* h, int capacity);
int nuifk_state(void* h, nuif_icp_result* out);
/* Copies the scene's device allocation: the dense volume, or the whole hash scene
 * (layout of nuif_hash_layout); nuifk_volume_bytes gives its size. */
int nuifk_read_volume(void* h, void* out, size_t bytes);
int nuifk_read_maps(void* h, float* vmap, float* nmap);
/* The feedback frame's colours (W*H BGRA): the last raycast's, or the first frame's
 * (m_colorsCL of OCL/NuiKinfuOpenCLFeedbackFrame.cpp:48, filled by CopyColors :90,186-200
 * or by the raycast). */
int nuifk_read_feedback_colors(void* h, uint8_t* out);
/* The joint ICP order's photometric weight (default 0.1), largest intensity difference
 * (0.2) and minimum gradient (0.002) (nuif_icp_photometric); before the next frame.  No
 * counterpart in the reference. */
int nuifk_set_photometric(void* h, float weight, float max_difference, float min_gradient);
/* Shifting volume: voxel offset (int32[3]) and its size in metres (float[3]). */
int nuifk_shift_state(void* h, int32_t* offset, float* offset_size);
/* Shifting volume: points of the evicted slabs (NuiKinfuVertexCache); returns their number
 * and copies up to cap points (float3) and colours (float4, may be NULL). */
int nuifk_shift_cache(void* h, float* points, float* colors, size_t cap);
/* Hash scene with swapping: blocks held by the host store (NuiKinfuOpenCLHashGlobalCache). */
int nuifk_hash_stored_blocks(void* h, size_t* out);
/* Hash scene: insertions that found no brick (pool exhausted) or no free bucket slot so far,
 * which the reference drops silently; 0 for a dense scene.  Synchronises the session. */
int nuifk_scene_failures(void* h, unsigned long long* out);
/* Hash scene: [bricks handed out, visible-list length of the last frame, failed insertions,
 * occupied entries], as nuif_hash_stats; all zero for a dense scene.  Synchronises the session. */
int nuifk_hash_stats(void* h, uint32_t* out4);
/* DS/NuiKinfuScene.h:43 Volume2Mesh(NuiMeshShape*) (src/SLAM/Backend/NuiSLAMController.cpp:121-127):
 * the shared-vertex mesh of the scene, points (float3) then triangle indices.  Copies up to
 * cap_points points and cap_indices indices (either pointer may be NULL) and returns the counts. */
int nuifk_mesh_shape(void* h, float* points, size_t cap_points, int32_t* indices, size_t cap_indices,
                     size_t* n_points, size_t* n_indices);
/* DS/NuiKinfuFrame.h:22 UpdateColorBuffers(ColorSpacePoint*, UINT*, UINT, const NuiColorImage&): the
 * frame's colours from a colour image through the depth-to-colour map (host arrays: W*H (x, y)
 * floats, W*H uint32 lookup table or NULL, iw*ih BGRA image); nuifk_frame_colors reads the
 * frame's W*H BGRA colours back. */
int nuifk_map_color(void* h, const float* map_xy, const uint32_t* lut, const uint8_t* image, int iw, int ih);
int nuifk_frame_colors(void* h, uint8_t* out);
/* Dense scene: distinct voxels the raycast reads at the current camera pose
 * (nuif_tsdf_raycast_footprint), for the raycast's algorithmic bytes. */
int nuifk_raycast_footprint(void* h, unsigned long long* out);
/* VO/NuiTrackerConfig.cpp:26-115 key=value files; -1 if the file cannot be opened. */
int nuifk_load_tracker_config(const char* path, nuifk_desc* d);
int nuifk_save_tracker_config(const char* path, const nuifk_desc* d);
int nuifk_set_stage_timing(void* h, int on);
/* Sums (ms) of preprocess / ICP / integrate / raycast over the timed frames; returns their number. */
int nuifk_take_stage_times(void* h, double* ms4);
int nuifk_updated_voxels(void* h, unsigned long long* out);
void* nuifk_stream(void* h);
void* nuifk_volume_ptr(void* h);
size_t nuifk_volume_bytes(void* h);
float nuifk_trunc(void* h);

/* ---- multi-session fusion server ------------------------------------------
 * Sessions stay resident (volume, pyramids, pose, tracking state).  A round
 * submits one frame for each listed session (see nuifs_submit).  Latency of a
 * frame = device time from the round's start event to the end of that frame's
 * pipeline. */
int nuifs_create(size_t hbm_budget_bytes, void** out);
int nuifs_destroy(void* srv);
/* Device bytes a session would hold (volume + frame / pyramid / map buffers). */
size_t nuifs_session_bytes(const nuifk_desc* d);
/* Admits a session; returns its id (>= 0) or NUIF_ERR_OOM when the budget is exceeded. */
int nuifs_open(void* srv, const nuifk_desc* d);
int nuifs_close(void* srv, int id);
size_t nuifs_bytes_in_use(void* srv);
int nuifs_session_count(void* srv);
/* The engine handle of a session (for the nuifk_* read-back calls). */
void* nuifs_engine(void* srv, int id);
/* Enqueue one round: n sessions, device-resident depth frames (W*H uint16 each).  The
 * tracked frames of plain dense sessions run as one batched round (nuif_fusion_round: one
 * launch per stage for all of them, the ICP of all of them in lockstep steps) on
 * the server's stream; every other frame (a session's first frame, hash and shifting
 * scenes) runs on its session's own stream, overlapping the batch. */
int nuifs_submit(void* srv, int n, const int* ids, const uint16_t* const* depths);
/* nuifs_submit with each frame's device-resident BGRA colours (W*H*4 bytes, or NULL for a
 * frame without): integrated into sessions with a colour volume, and the photometric
 * rows' input of sessions in the joint ICP order (icp_reduction NUIF_ICP_REDUCE_JOINT,
 * which require them).  No counterpart in the reference. */
int nuifs_submit_color(void* srv, int n, const int* ids, const uint16_t* const* depths,
                       const uint8_t* const* bgras);
/* Batched rounds captured once as a graph and replayed while their launches stay the same
 * (default, non-zero) or enqueued launch by launch (0).  No counterpart in the reference. */
int nuifs_set_round_graph(void* srv, int on);
/* Batched rounds on (default, non-zero) or off: every frame on its session's stream. */
int nuifs_set_batching(void* srv, int on);
/* Frames of the last round that ran in the batch. */
int nuifs_last_batch(void* srv);
/* Device time of the batched rounds' stages, event-timed on the server stream: sums (ms) of
 * [preprocess + pyramid, depth bounds, ICP, integration, raycast] over the rounds timed since
 * the last call; returns their number. */
int nuifs_set_stage_timing(void* srv, int on);
int nuifs_take_stage_times(void* srv, double* ms5);
/* Wait for the last round; latency_ms[i] for the i-th submitted frame (may be NULL).  Then
 * every frame's tracking outcome is known (nuifs_last_outcomes) and the failure policy is
 * applied to the sessions whose frame failed. */
int nuifs_wait(void* srv, float* latency_ms);
/* What a tracking failure does to its session.  The reference's tracking manager either
 * resets the engine (m_bAutoReset) or stops (VO/NuiKinfuTrackingManager.cpp:94-106). */
#define NUIFS_ON_FAILURE_KEEP 0    /* default: the next frame tracks from the last good pose */
#define NUIFS_ON_FAILURE_RESET 1   /* engine reset and scene cleared: the next frame starts over */
#define NUIFS_ON_FAILURE_EVICT 2   /* the session is closed and its device memory released */
int nuifs_set_failure_policy(void* srv, int policy);
/* Outcomes of the frames of the last waited round, in submission order: 1 tracked, 0 failed
 * (a session's first frame always tracks).  Copies min(cap, n) and returns n. */
int nuifs_last_outcomes(void* srv, int32_t* out, int cap);

#ifdef __cplusplus
}
#endif

#endif /* NUIF_KINFU_H */
