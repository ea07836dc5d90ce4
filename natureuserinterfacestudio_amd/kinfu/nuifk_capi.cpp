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

#include "../../include/nuif_kinfu.h"
#include "NuiFrameIO.h"
#include "NuiKinfuSession.h"

using namespace NuiKinfuEngine;
using namespace nuif_kinfu;

namespace {
thread_local std::string t_err;

using Session = NuiKinfuSession;

template <class F> int guarded(F&& f) {
    try {
        return f();
    } catch (const IoError& e) {
        t_err = e.what();
        return NUIF_ERR_IO;
    } catch (const std::invalid_argument& e) {
        t_err = e.what();
        return NUIF_ERR_INVALID;
    } catch (const std::exception& e) {
        t_err = e.what();
        return -2;
    }
}
}  // namespace

extern "C" {

const char* nuifk_last_error(void) { return t_err.c_str(); }

int nuifk_create(const nuifk_desc* d, void** out) {
    return guarded([&] {
        if (!d || !out) throw std::runtime_error("nuifk_create: null argument");
        *out = Session::create(*d).release();
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
        s->afterFrame();
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
        s->offsets.clear();
        return 0;
    });
}

// Pose history rows (NUIF_POSE_ROW floats per frame) of the first n frames.
int nuifk_poses(void* h, float* out, int n) {
    return guarded([&] {
        Session* s = static_cast<Session*>(h);
        s->engine->synchronize();
        const std::vector<float>& rows = s->engine->poseRows();
        const int m = std::min(n, (int)(rows.size() / NUIF_POSE_ROW));
        if (m > 0) {
            std::memcpy(out, rows.data(), (size_t)m * NUIF_POSE_ROW * sizeof(float));
            // global translation = local + the shifting volume's offset after that frame
            if (s->offsets.size() >= (size_t)m * 3)
                for (int i = 0; i < m; ++i)
                    for (int k = 0; k < 3; ++k) out[(size_t)i * NUIF_POSE_ROW + 9 + k] += s->offsets[(size_t)i * 3 + k];
        }
        return m;
    });
}

int nuifk_set_history_capacity(void* h, int capacity) {
    return guarded([&] {
        static_cast<Session*>(h)->engine->setPoseHistoryCapacity(capacity);
        return 0;
    });
}

int nuifk_shift_state(void* h, int32_t* offset, float* offset_size) {
    return guarded([&] {
        auto* sh = dynamic_cast<NuiKinfuHIPShiftScene*>(static_cast<Session*>(h)->scene.get());
        if (!sh) throw std::invalid_argument("nuifk_shift_state: not a shifting volume");
        for (int k = 0; k < 3; ++k)
            if (offset) offset[k] = sh->VoxelOffset()[k];
        if (offset_size) sh->OffsetSize(offset_size);
        return 0;
    });
}

int nuifk_shift_cache(void* h, float* points, float* colors, size_t cap) {
    return guarded([&] {
        auto* sh = dynamic_cast<NuiKinfuHIPShiftScene*>(static_cast<Session*>(h)->scene.get());
        if (!sh) throw std::invalid_argument("nuifk_shift_cache: not a shifting volume");
        const size_t n = sh->CachedPositions().size() / 3, m = std::min(n, cap);
        if (points && m) std::memcpy(points, sh->CachedPositions().data(), m * 3 * sizeof(float));
        if (colors && m) std::memcpy(colors, sh->CachedColors().data(), m * 4 * sizeof(float));
        return (int)n;
    });
}

int nuifk_hash_stored_blocks(void* h, size_t* out) {
    return guarded([&] {
        auto* hs = dynamic_cast<NuiKinfuHIPHashScene*>(static_cast<Session*>(h)->scene.get());
        if (!hs || !out) throw std::invalid_argument("nuifk_hash_stored_blocks: not a hash scene");
        *out = hs->StoredBlocks();
        return 0;
    });
}

int nuifk_mesh_shape(void* h, float* points, size_t cap_points, int32_t* indices, size_t cap_indices,
                     size_t* n_points, size_t* n_indices) {
    return guarded([&] {
        if (!n_points || !n_indices) throw std::invalid_argument("nuifk_mesh_shape: null count");
        Session* s = static_cast<Session*>(h);
        s->engine->synchronize();
        NuiMeshShape mesh;
        NuiKinfuScene* scene = s->scene.get();
        if (!scene->Volume2Mesh(&mesh)) throw std::runtime_error("nuifk_mesh_shape: no mesh");
        *n_points = mesh.pointsNum();
        *n_indices = mesh.trianglesNum() * 3;
        if (points)
            for (size_t i = 0; i < std::min(cap_points, mesh.pointsNum()); ++i) {
                const SgVec3f& p = mesh.getPoint((unsigned)i);
                points[3 * i] = p[0];
                points[3 * i + 1] = p[1];
                points[3 * i + 2] = p[2];
            }
        if (indices)
            for (size_t i = 0; i < std::min(cap_indices, *n_indices); ++i) indices[i] = mesh.triangleIndex((unsigned)i);
        return 0;
    });
}

int nuifk_map_color(void* h, const float* map_xy, const uint32_t* lut, const uint8_t* image, int iw, int ih) {
    return guarded([&] {
        Session* s = static_cast<Session*>(h);
        if (!map_xy || !image || iw <= 0 || ih <= 0) throw std::invalid_argument("nuifk_map_color: bad argument");
        NuiColorImage img;
        BGRQUAD* dst = img.AllocateBuffer((UINT)iw, (UINT)ih);
        std::memcpy(dst, image, (size_t)iw * ih * 4);
        auto* frame = s->engine->frame();
        frame->UpdateColorBuffers(reinterpret_cast<ColorSpacePoint*>(const_cast<float*>(map_xy)),
                                  const_cast<UINT*>(reinterpret_cast<const UINT*>(lut)),
                                  frame->GetWidth() * frame->GetHeight(), img);
        return 0;
    });
}

int nuifk_frame_colors(void* h, uint8_t* out) {
    return guarded([&] {
        Session* s = static_cast<Session*>(h);
        auto* frame = s->engine->frame();
        const uint8_t* c = frame->GetColorBuffer();
        if (!out || !c) throw std::invalid_argument("nuifk_frame_colors: no colour buffer");
        check(nuif_memcpy_d2h(out, c, (size_t)frame->GetWidth() * frame->GetHeight() * 4, s->engine->stream()),
              "frame colours");
        check(nuif_stream_synchronize(s->engine->stream()), "frame colours");
        return 0;
    });
}

int nuifk_raycast_footprint(void* h, unsigned long long* out) {
    return guarded([&] {
        Session* s = static_cast<Session*>(h);
        auto* ds = dynamic_cast<NuiKinfuHIPScene*>(s->scene.get());
        if (!ds || !out) throw std::invalid_argument("nuifk_raycast_footprint: not a dense scene");
        s->engine->synchronize();
        *out = (unsigned long long)ds->RaycastFootprint(s->engine->cameraState());
        return 0;
    });
}

int nuifk_scene_failures(void* h, unsigned long long* out) {
    return guarded([&] {
        if (!out) throw std::invalid_argument("nuifk_scene_failures: null output");
        auto* hs = dynamic_cast<NuiKinfuHIPHashScene*>(static_cast<Session*>(h)->scene.get());
        *out = 0;
        if (hs) {
            uint32_t st[4];
            hs->Stats(st);
            *out = st[2];
        }
        return 0;
    });
}

int nuifk_hash_stats(void* h, uint32_t* out4) {
    return guarded([&] {
        if (!out4) throw std::invalid_argument("nuifk_hash_stats: null output");
        auto* hs = dynamic_cast<NuiKinfuHIPHashScene*>(static_cast<Session*>(h)->scene.get());
        for (int k = 0; k < 4; ++k) out4[k] = 0;
        if (hs) hs->Stats(out4);
        return 0;
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
        if (bytes != s->scene->DeviceBytes()) throw std::runtime_error("nuifk_read_volume: size mismatch");
        check(nuif_memcpy_d2h(out, s->scene->DevicePointer(), bytes, s->engine->stream()), "volume read");
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

int nuifk_set_photometric(void* h, float weight, float max_difference, float min_gradient) {
    return guarded([&] {
        Session* s = static_cast<Session*>(h);
        s->engine->tracker()->SetPhotometric(weight, max_difference, min_gradient);
        return 0;
    });
}

int nuifk_read_feedback_colors(void* h, uint8_t* out) {
    return guarded([&] {
        Session* s = static_cast<Session*>(h);
        if (!out) throw std::runtime_error("nuifk_read_feedback_colors: null output");
        check(nuif_memcpy_d2h(out, s->engine->feedbackFrame()->GetColorBuffer(), (size_t)s->W * s->H * 4,
                              s->engine->stream()),
              "feedback colours");
        check(nuif_stream_synchronize(s->engine->stream()), "feedback colours");
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
        NuiTrackerConfig tc = Session::trackerConfig(*d);
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
void* nuifk_volume_ptr(void* h) { return h ? static_cast<Session*>(h)->scene->DevicePointer() : nullptr; }
size_t nuifk_volume_bytes(void* h) { return h ? static_cast<Session*>(h)->scene->DeviceBytes() : 0; }
float nuifk_trunc(void* h) { return h ? static_cast<Session*>(h)->scene->Truncation() : 0.f; }

// ---- multi-session server -------------------------------------------------------

int nuifs_create(size_t hbm_budget_bytes, void** out) {
    return guarded([&] {
        if (!out) throw std::runtime_error("nuifs_create: null argument");
        *out = new NuiFusionServer(hbm_budget_bytes);
        return 0;
    });
}

int nuifs_destroy(void* srv) {
    return guarded([&] {
        delete static_cast<NuiFusionServer*>(srv);
        return 0;
    });
}

size_t nuifs_session_bytes(const nuifk_desc* d) { return d ? Session::deviceBytes(*d) : 0; }

int nuifs_open(void* srv, const nuifk_desc* d) {
    return guarded([&] {
        if (!srv || !d) throw std::runtime_error("nuifs_open: null argument");
        const int id = static_cast<NuiFusionServer*>(srv)->open(*d);
        if (id < 0) t_err = "nuifs_open: session does not fit the device-memory budget";
        return id;
    });
}

int nuifs_close(void* srv, int id) {
    return guarded([&] {
        static_cast<NuiFusionServer*>(srv)->close(id);
        return 0;
    });
}

size_t nuifs_bytes_in_use(void* srv) { return srv ? static_cast<NuiFusionServer*>(srv)->bytesInUse() : 0; }

int nuifs_session_count(void* srv) { return srv ? static_cast<NuiFusionServer*>(srv)->sessionCount() : 0; }

void* nuifs_engine(void* srv, int id) {
    try {
        return srv ? static_cast<NuiFusionServer*>(srv)->session(id) : nullptr;
    } catch (const std::exception& e) {
        t_err = e.what();
        return nullptr;
    }
}

int nuifs_submit(void* srv, int n, const int* ids, const uint16_t* const* depths) {
    return guarded([&] {
        if (!srv) throw std::runtime_error("nuifs_submit: null server");
        static_cast<NuiFusionServer*>(srv)->submit(n, ids, depths);
        return 0;
    });
}

int nuifs_submit_color(void* srv, int n, const int* ids, const uint16_t* const* depths,
                       const uint8_t* const* bgras) {
    return guarded([&] {
        if (!srv) throw std::runtime_error("nuifs_submit_color: null server");
        static_cast<NuiFusionServer*>(srv)->submit(n, ids, depths, bgras);
        return 0;
    });
}

int nuifs_set_round_graph(void* srv, int on) {
    return guarded([&] {
        if (!srv) throw std::runtime_error("nuifs_set_round_graph: null server");
        static_cast<NuiFusionServer*>(srv)->setRoundGraph(on != 0);
        return 0;
    });
}

int nuifs_set_batching(void* srv, int on) {
    return guarded([&] {
        if (!srv) throw std::runtime_error("nuifs_set_batching: null server");
        static_cast<NuiFusionServer*>(srv)->setBatching(on != 0);
        return 0;
    });
}

int nuifs_last_batch(void* srv) { return srv ? static_cast<NuiFusionServer*>(srv)->lastBatchSize() : 0; }

int nuifs_wait(void* srv, float* latency_ms) {
    return guarded([&] {
        if (!srv) throw std::runtime_error("nuifs_wait: null server");
        return static_cast<NuiFusionServer*>(srv)->wait(latency_ms);
    });
}

int nuifs_set_stage_timing(void* srv, int on) {
    return guarded([&] {
        if (!srv) throw std::runtime_error("nuifs_set_stage_timing: null server");
        static_cast<NuiFusionServer*>(srv)->setStageTiming(on != 0);
        return 0;
    });
}

int nuifs_take_stage_times(void* srv, double* ms5) {
    return guarded([&] {
        if (!srv || !ms5) throw std::runtime_error("nuifs_take_stage_times: null argument");
        return static_cast<NuiFusionServer*>(srv)->takeStageTimes(ms5);
    });
}

int nuifs_set_failure_policy(void* srv, int policy) {
    return guarded([&] {
        if (!srv) throw std::runtime_error("nuifs_set_failure_policy: null server");
        static_cast<NuiFusionServer*>(srv)->setFailurePolicy(policy);
        return 0;
    });
}

int nuifs_last_outcomes(void* srv, int32_t* out, int cap) {
    return guarded([&] {
        if (!srv) throw std::runtime_error("nuifs_last_outcomes: null server");
        const std::vector<int32_t>& o = static_cast<NuiFusionServer*>(srv)->lastOutcomes();
        const int n = std::min(cap, (int)o.size());
        if (n > 0 && !out) throw std::invalid_argument("nuifs_last_outcomes: null output");
        for (int i = 0; i < n; ++i) out[i] = o[i];
        return (int)o.size();
    });
}

// ---- frame ingest (include/nuif_io.h) ------------------------------------------

int nuif_io_write_frame_image(const char* path, int compressed, int64_t timestamp, uint32_t width,
                              uint32_t height, const void* buffer, uint32_t bytes) {
    return guarded([&] {
        if (!path || (bytes && !buffer)) throw std::invalid_argument("nuif_io_write_frame_image: null argument");
        NuiFileIOUtilities::writeFrameImage(path, compressed != 0, timestamp, width, height, buffer, bytes);
        return 0;
    });
}

int nuif_io_read_frame_header(const char* path, int64_t* timestamp, uint32_t* width, uint32_t* height) {
    return guarded([&] {
        if (!path) throw std::invalid_argument("nuif_io_read_frame_header: null path");
        if (!NuiFileIOUtilities::readFrameImageHeader(path, timestamp, width, height))
            throw IoError(std::string(path) + ": could not open file");
        return 0;
    });
}

int nuif_io_read_frame_buffer(const char* path, void* buffer, uint32_t bytes) {
    return guarded([&] {
        if (!path || (bytes && !buffer)) throw std::invalid_argument("nuif_io_read_frame_buffer: null argument");
        NuiFileIOUtilities::readFrameImageBuffer(path, buffer, bytes);
        return 0;
    });
}

int nuif_io_write_camera(const char* path, float fx, float fy, float cx, float cy) {
    return guarded([&] {
        if (!path) throw std::invalid_argument("nuif_io_write_camera: null path");
        NuiFileIOUtilities::writeCamera(path, fx, fy, cx, cy);
        return 0;
    });
}

int nuif_io_read_camera(const char* path, float* fx, float* fy, float* cx, float* cy) {
    return guarded([&] {
        if (!path || !fx || !fy) throw std::invalid_argument("nuif_io_read_camera: null argument");
        if (!NuiFileIOUtilities::readCamera(path, fx, fy, cx, cy))
            throw IoError(std::string(path) + ": could not open file");
        return 0;
    });
}

int nuif_frame_loader_create(const char* dir, int slots, int threads, int upload, void** out) {
    return guarded([&] {
        if (!dir || !out) throw std::invalid_argument("nuif_frame_loader_create: null argument");
        *out = new NuiKinfuFrameLoader(dir, slots, threads, upload != 0);
        return 0;
    });
}

int nuif_frame_loader_next(void* loader, nuif_loaded_frame* out) {
    return guarded([&] {
        if (!loader || !out) throw std::invalid_argument("nuif_frame_loader_next: null argument");
        return static_cast<NuiKinfuFrameLoader*>(loader)->next(out);
    });
}

int nuif_frame_loader_release(void* loader, int slot) {
    return guarded([&] {
        if (!loader) throw std::invalid_argument("nuif_frame_loader_release: null loader");
        static_cast<NuiKinfuFrameLoader*>(loader)->release(slot);
        return 0;
    });
}

int nuif_frame_loader_reset(void* loader) {
    return guarded([&] {
        if (!loader) throw std::invalid_argument("nuif_frame_loader_reset: null loader");
        static_cast<NuiKinfuFrameLoader*>(loader)->reset();
        return 0;
    });
}

int nuif_frame_loader_destroy(void* loader) {
    return guarded([&] {
        delete static_cast<NuiKinfuFrameLoader*>(loader);
        return 0;
    });
}

int nuif_frame_save(const char* dir, const nuif_loaded_frame* frame, int compressed) {
    return guarded([&] {
        if (!dir || !frame) throw std::invalid_argument("nuif_frame_save: null argument");
        saveFrame(dir, *frame, compressed != 0);
        return 0;
    });
}

}  // extern "C"
