// C ABI of the tracking engine (libnuif_kinfu.so): lets non-C++ hosts (the
// Python tests, the benchmark, a ctypes/cgo binding) drive
// NuiKinfuTrackingEngine::RunTracking with a dense scene, as the reference's
// tracking manager does (VO/NuiKinfuTrackingManager.cpp:53-108).  No
// exception crosses this boundary: failures return a negative status and the
// message is available from nuifk_last_error().
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>

#include "NuiKinfuHIP.h"

using namespace NuiKinfuEngine;
using namespace nuif_kinfu;

namespace {
thread_local std::string t_err;

struct Session {
    std::unique_ptr<NuiKinfuTrackingEngine> engine;
    std::unique_ptr<NuiKinfuHIPScene> scene;
    NuiCameraParams cam;
    UINT W = 0, H = 0;
};

template <class F> int guarded(F&& f) {
    try {
        return f();
    } catch (const std::exception& e) {
        t_err = e.what();
        return -2;
    }
}
}  // namespace

extern "C" {

typedef struct nuifk_desc {
    int width, height;
    float fx, fy, cx, cy;
    float depth_min, depth_max;     // metres
    // tracker (NuiTrackerConfig)
    int filter_radius;
    float sigma_space2_inv_half, sigma_depth2_inv_half, depth_threshold, dist_threshold;
    int levels;
    int iterations[4];              // level 0 = finest
    float integration_threshold;    // 0 = integrate every frame
    // dense volume (NuiKinfuVolumeConfig)
    int resolution[3];
    float dimensions[3];
    float tranc_dist;
    int color;                      // colour volume on/off
    int storage;                    // NUIF_TSDF_*
    int use_graph;
    int icp_reduction;              // NUIF_ICP_REDUCE_*
} nuifk_desc;

const char* nuifk_last_error(void) { return t_err.c_str(); }

int nuifk_create(const nuifk_desc* d, void** out) {
    return guarded([&] {
        if (!d || !out) throw std::runtime_error("nuifk_create: null argument");
        auto s = std::make_unique<Session>();
        NuiTrackerConfig tc;
        tc.filter_radius = (UINT)d->filter_radius;
        tc.sigma_space2_inv_half = d->sigma_space2_inv_half;
        tc.sigma_depth2_inv_half = d->sigma_depth2_inv_half;
        tc.depth_threshold = d->depth_threshold;
        tc.dist_threshold = d->dist_threshold;
        tc.iterations.clear();
        for (int l = 0; l < d->levels; ++l)
            tc.iterations.push_back({(UINT)d->iterations[l], NuiTrackerConfig::eTracker_Iteration_Both});
        NuiKinfuVolumeConfig vc;
        vc.resolution = Vector3i(d->resolution[0], d->resolution[1], d->resolution[2]);
        vc.dimensions = Vector3f(d->dimensions[0], d->dimensions[1], d->dimensions[2]);
        vc.tranc_dist = d->tranc_dist;
        vc.bHas_color_volume = d->color != 0;
        s->engine = std::make_unique<NuiKinfuTrackingEngine>();
        s->engine->setIntegrationMetricThreshold(d->integration_threshold);
        s->engine->initialize(tc, false, (UINT)d->width, (UINT)d->height);
        s->engine->setUseGraph(d->use_graph != 0);
        s->engine->tracker()->SetReduction(d->icp_reduction);
        s->scene = std::make_unique<NuiKinfuHIPScene>(vc, d->storage, s->engine->stream());
        s->cam.m_intrinsics = NuiCameraIntrinsics(d->fx, d->fy, d->cx, d->cy);
        s->cam.m_sensorDepthMin = d->depth_min;
        s->cam.m_sensorDepthMax = d->depth_max;
        s->W = (UINT)d->width;
        s->H = (UINT)d->height;
        *out = s.release();
        return 0;
    });
}

int nuifk_destroy(void* h) {
    return guarded([&] {
        delete static_cast<Session*>(h);
        return 0;
    });
}

// One frame.  depth: W*H uint16 (mm); on_device selects a device or host pointer.
// bgra: W*H*4 bytes or NULL.  sync != 0 waits and returns 1 (tracked) / 0 (failure);
// sync == 0 returns 1 immediately.
int nuifk_run(void* h, const uint16_t* depth, int on_device, const uint8_t* bgra, int sync) {
    return guarded([&] {
        Session* s = static_cast<Session*>(h);
        if (!s) throw std::runtime_error("nuifk_run: null session");
        bool ok;
        if (on_device)
            ok = s->engine->RunTrackingDevice(0, depth, bgra, s->scene.get(), s->cam, sync != 0);
        else
            ok = s->engine->RunTracking(0, const_cast<uint16_t*>(depth), (BGRQUAD*)const_cast<uint8_t*>(bgra),
                                        s->W * s->H, s->scene.get(), s->cam);
        return ok ? 1 : 0;
    });
}

int nuifk_synchronize(void* h) {
    return guarded([&] { return static_cast<Session*>(h)->engine->synchronize() ? 1 : 0; });
}

int nuifk_reset(void* h) {
    return guarded([&] {
        Session* s = static_cast<Session*>(h);
        s->engine->reset();
        s->scene->reset();
        return 0;
    });
}

// Pose history rows (NUIF_POSE_ROW floats per frame) of the first n frames.
int nuifk_poses(void* h, float* out, int n) {
    return guarded([&] {
        Session* s = static_cast<Session*>(h);
        s->engine->synchronize();
        const int m = std::min(n, std::min(s->engine->getFrameID(), s->engine->poseHistoryCapacity()));
        if (m > 0) {
            check(nuif_memcpy_d2h(out, s->engine->poseHistoryDevice(), (size_t)m * NUIF_POSE_ROW * sizeof(float),
                                  s->engine->stream()),
                  "pose history");
            check(nuif_stream_synchronize(s->engine->stream()), "pose history");
        }
        return m;
    });
}

int nuifk_state(void* h, nuif_icp_result* out) {
    return guarded([&] {
        *out = static_cast<Session*>(h)->engine->tracker()->ReadState();
        return 0;
    });
}

// Copies the raw TSDF volume (voxel bytes per nuif_tsdf_voxel_bytes) to host.
int nuifk_read_volume(void* h, void* out, size_t bytes) {
    return guarded([&] {
        Session* s = static_cast<Session*>(h);
        if (bytes != s->scene->VolumeBytes()) throw std::runtime_error("nuifk_read_volume: size mismatch");
        check(nuif_memcpy_d2h(out, s->scene->GetVolume(), bytes, s->engine->stream()), "volume read");
        check(nuif_stream_synchronize(s->engine->stream()), "volume read");
        return 0;
    });
}

// Reference maps used by the next ICP (world vertex / normal maps, W*H*3 floats each).
int nuifk_read_maps(void* h, float* vmap, float* nmap) {
    return guarded([&] {
        Session* s = static_cast<Session*>(h);
        const size_t n = (size_t)s->W * s->H * 12;
        if (vmap) check(nuif_memcpy_d2h(vmap, s->engine->feedbackFrame()->GetVertexBuffer(), n, s->engine->stream()), "vmap");
        if (nmap) check(nuif_memcpy_d2h(nmap, s->engine->feedbackFrame()->GetNormalBuffer(), n, s->engine->stream()), "nmap");
        check(nuif_stream_synchronize(s->engine->stream()), "maps");
        return 0;
    });
}

// NuiTrackerConfig::load / log (VO/NuiTrackerConfig.cpp:26-115) into / from a descriptor.
// Keys absent from the file keep the descriptor's values; returns -1 if the file cannot be opened.
int nuifk_load_tracker_config(const char* path, nuifk_desc* d) {
    return guarded([&] {
        if (!path || !d) throw std::runtime_error("nuifk_load_tracker_config: null argument");
        NuiTrackerConfig tc;
        tc.filter_radius = (UINT)d->filter_radius;
        tc.sigma_space2_inv_half = d->sigma_space2_inv_half;
        tc.sigma_depth2_inv_half = d->sigma_depth2_inv_half;
        tc.depth_threshold = d->depth_threshold;
        tc.dist_threshold = d->dist_threshold;
        if (!tc.load(path)) return -1;
        if (tc.iterations.empty() || tc.iterations.size() > 4)
            throw std::runtime_error("tracker config needs 1..4 icp_iteration entries");
        d->filter_radius = (int)tc.filter_radius;
        d->sigma_space2_inv_half = tc.sigma_space2_inv_half;
        d->sigma_depth2_inv_half = tc.sigma_depth2_inv_half;
        d->depth_threshold = tc.depth_threshold;
        d->dist_threshold = tc.dist_threshold;
        d->levels = (int)tc.iterations.size();
        for (int l = 0; l < d->levels; ++l) d->iterations[l] = (int)tc.iterations[l].m_num;
        return 0;
    });
}

int nuifk_save_tracker_config(const char* path, const nuifk_desc* d) {
    return guarded([&] {
        if (!path || !d) throw std::runtime_error("nuifk_save_tracker_config: null argument");
        NuiTrackerConfig tc;
        tc.filter_radius = (UINT)d->filter_radius;
        tc.sigma_space2_inv_half = d->sigma_space2_inv_half;
        tc.sigma_depth2_inv_half = d->sigma_depth2_inv_half;
        tc.depth_threshold = d->depth_threshold;
        tc.dist_threshold = d->dist_threshold;
        tc.iterations.clear();
        for (int l = 0; l < d->levels; ++l)
            tc.iterations.push_back({(UINT)d->iterations[l], NuiTrackerConfig::eTracker_Iteration_Both});
        std::remove(path);
        return tc.log(path) ? 0 : -1;
    });
}

int nuifk_set_stage_timing(void* h, int on) {
    return guarded([&] {
        static_cast<Session*>(h)->engine->setStageTiming(on != 0);
        return 0;
    });
}

// Sums (ms) of preprocess / ICP / integrate / raycast over the frames timed since
// the last call; returns the number of frames.
int nuifk_take_stage_times(void* h, double* ms4) {
    return guarded([&] { return static_cast<Session*>(h)->engine->takeStageTimes(ms4); });
}

// Total voxels updated by integration so far.
int nuifk_updated_voxels(void* h, unsigned long long* out) {
    return guarded([&] {
        *out = static_cast<Session*>(h)->scene->UpdatedVoxels();
        return 0;
    });
}

void* nuifk_stream(void* h) { return h ? (void*)static_cast<Session*>(h)->engine->stream() : nullptr; }
void* nuifk_volume_ptr(void* h) { return h ? static_cast<Session*>(h)->scene->GetVolume() : nullptr; }
size_t nuifk_volume_bytes(void* h) { return h ? static_cast<Session*>(h)->scene->VolumeBytes() : 0; }
float nuifk_trunc(void* h) { return h ? static_cast<Session*>(h)->scene->GetParams().tranc_dist : 0.f; }

}  // extern "C"
