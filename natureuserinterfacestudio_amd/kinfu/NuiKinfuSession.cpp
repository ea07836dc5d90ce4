#include "NuiKinfuSession.h"

#include <algorithm>
#include <stdexcept>

using namespace NuiKinfuEngine;

namespace nuif_kinfu {

NuiTrackerConfig NuiKinfuSession::trackerConfig(const nuifk_desc& d) {
    NuiTrackerConfig tc;
    tc.filter_radius = (UINT)d.filter_radius;
    tc.sigma_space2_inv_half = d.sigma_space2_inv_half;
    tc.sigma_depth2_inv_half = d.sigma_depth2_inv_half;
    tc.depth_threshold = d.depth_threshold;
    tc.dist_threshold = d.dist_threshold;
    tc.iterations.clear();
    for (int l = 0; l < d.levels; ++l)
        tc.iterations.push_back({(UINT)d.iterations[l], NuiTrackerConfig::eTracker_Iteration_Both});
    return tc;
}

std::unique_ptr<NuiKinfuSession> NuiKinfuSession::create(const nuifk_desc& d) {
    if (d.width <= 0 || d.height <= 0) throw std::runtime_error("session: empty image");
    if (d.levels < 1 || d.levels > NUIF_ICP_MAX_LEVELS) throw std::runtime_error("session: 1..4 pyramid levels");
    auto s = std::make_unique<NuiKinfuSession>();
    NuiKinfuVolumeConfig vc;
    vc.resolution = Vector3i(d.resolution[0], d.resolution[1], d.resolution[2]);
    vc.dimensions = Vector3f(d.dimensions[0], d.dimensions[1], d.dimensions[2]);
    vc.tranc_dist = d.tranc_dist;
    vc.bHas_color_volume = d.color != 0;
    s->engine = std::make_unique<NuiKinfuTrackingEngine>();
    s->engine->setIntegrationMetricThreshold(d.integration_threshold);
    s->engine->initialize(trackerConfig(d), false, (UINT)d.width, (UINT)d.height);
    s->engine->setUseGraph(d.use_graph != 0);
    if (d.icp_reduction == NUIF_ICP_REDUCE_JOINT && !d.color)
        throw std::runtime_error("session: the joint ICP needs the colour volume (color = 1)");
    s->engine->tracker()->SetReduction(d.icp_reduction);
    if (d.hash_forward_render < NUIFK_RENDER_CASTRAY || d.hash_forward_render > NUIFK_RENDER_FORWARD_REFERENCE)
        throw std::runtime_error("session: unknown hash_forward_render mode");
    if (d.scene == NUIFK_SCENE_HASH) {
        s->scene = std::make_unique<NuiKinfuHIPHashScene>(hashConfig(d), d.storage, d.color != 0, s->engine->stream());
        if (d.hash_swapping) s->engine->setUseGraph(false);  // the host store waits for each frame's lists
    } else if (d.scene == NUIFK_SCENE_DENSE) {
        s->scene = std::make_unique<NuiKinfuHIPScene>(vc, d.storage, s->engine->stream());
    } else if (d.scene == NUIFK_SCENE_SHIFT) {
        if (d.voxel_shift <= 0) throw std::runtime_error("session: the shifting volume needs voxel_shift > 0");
        vc.bIsDynamic = true;
        vc.voxel_shift = d.voxel_shift;
        vc.translateBasis = Vector3f(d.volume_basis[0], d.volume_basis[1], d.volume_basis[2]);
        auto sh = std::make_unique<NuiKinfuHIPShiftScene>(vc, d.storage, s->engine->stream());
        sh->SetTracker(s->engine->tracker());
        s->scene = std::move(sh);
        s->engine->setUseGraph(false);  // the shift decision waits for each integrated frame's pose
    } else {
        throw std::runtime_error("session: unknown scene kind");
    }
    s->cam.m_intrinsics = NuiCameraIntrinsics(d.fx, d.fy, d.cx, d.cy);
    s->cam.m_sensorDepthMin = d.depth_min;
    s->cam.m_sensorDepthMax = d.depth_max;
    s->W = (UINT)d.width;
    s->H = (UINT)d.height;
    return s;
}

void NuiKinfuSession::afterFrame() {
    auto* sh = dynamic_cast<NuiKinfuHIPShiftScene*>(scene.get());
    if (!sh) return;
    float o[3];
    sh->OffsetSize(o);
    offsets.insert(offsets.end(), o, o + 3);
}

NuiHashingSDFConfig NuiKinfuSession::hashConfig(const nuifk_desc& d) {
    NuiHashingSDFConfig c;
    c.m_virtualVoxelSize = d.hash_voxel_size;
    c.m_truncScale = d.hash_trunc_scale;
    c.m_truncation = d.hash_truncation;
    c.m_integrationWeightMax = (UINT)d.hash_weight_max;
    c.m_numSDFBlocks = (UINT)d.hash_bricks;
    c.m_bUseForwardRender = d.hash_forward_render != NUIFK_RENDER_CASTRAY;
    c.m_bForwardReferenceProjection = d.hash_forward_render == NUIFK_RENDER_FORWARD_REFERENCE;
    c.m_bUseSwapping = d.hash_swapping != 0;
    if (d.hash_buckets > 0) c.m_hashNumBuckets = (UINT)d.hash_buckets;
    return c;
}

size_t NuiKinfuSession::deviceBytes(const nuifk_desc& d) {
    const size_t n = (size_t)d.width * d.height;
    size_t bytes = 256;  // update counter
    if (d.scene == NUIFK_SCENE_HASH) {
        const nuif_hash_params hp = NuiKinfuHIPHashScene::ParamsOf(hashConfig(d), d.storage, d.color != 0);
        bytes += nuif_hash_scene_bytes(&hp) + n * 8;  // + ray-range image
        if (d.hash_forward_render != NUIFK_RENDER_CASTRAY) bytes += nuif_hash_forward_render_bytes(d.width, d.height);
        if (d.hash_swapping) bytes += nuif_hash_swap_scratch_bytes(&hp) + (1u << 20) * 4;
    } else {  // dense or shifting volume
        const size_t voxels = (size_t)d.resolution[0] * d.resolution[1] * d.resolution[2];
        if (d.scene == NUIFK_SCENE_SHIFT) bytes += ((size_t)d.resolution[0] * d.resolution[1] + 64) * 4 + 65536;
        TsdfParams tp{};
        if (nuif_tsdf_params(d.resolution[0], d.resolution[1], d.resolution[2], d.dimensions[0], d.dimensions[1],
                             d.dimensions[2], d.tranc_dist, &tp) == NUIF_OK)
            bytes += nuif_tsdf_walk_states_bytes(&tp);
        bytes += voxels * nuif_tsdf_voxel_bytes(d.storage) + (d.color ? voxels * 4 : 0);
        bytes += (size_t)((d.resolution[0] + 7) / 8) * ((d.resolution[1] + 7) / 8) * ((d.resolution[2] + 7) / 8) +
                 (size_t)((d.resolution[0] + 31) / 32) * ((d.resolution[1] + 31) / 32) * ((d.resolution[2] + 31) / 32);
    }
    bytes += n * (2 + 4 + 4);  // raw, passing-filter depth, colours
    for (int l = 0; l < d.levels; ++l) bytes += (size_t)(d.width >> l) * (d.height >> l) * (4 + 12);
    bytes += n * (12 + 12 + 4);  // feedback vertex / normal / colour maps
    bytes += nuif_icp_state_bytes() + nuif_icp_partials_bytes(d.width, d.height);
    if (d.icp_reduction == NUIF_ICP_REDUCE_JOINT) {  // intensity pyramid, feedback intensities and gradients
        for (int l = 0; l < d.levels; ++l) bytes += (size_t)(d.width >> l) * (d.height >> l) * 4;
        bytes += n * 16;  // the model: intensity and gradient per pixel
    }
    bytes += (size_t)(1 << 16) * NUIF_POSE_ROW * sizeof(float) + 256;  // pose history, camera pose
    return bytes;
}

NuiFusionServer::NuiFusionServer(size_t hbmBudget) : m_budget(hbmBudget) {
    check(nuif_stream_create(&m_ctl), "server stream");
    check(nuif_event_create(&m_start), "server event");
    check(nuif_event_create(&m_batchEnd), "server event");
    for (nuif_event_t& e : m_stageEv) check(nuif_event_create(&e), "server event");
}

int NuiFusionServer::takeStageTimes(double ms[5]) {
    for (int k = 0; k < 5; ++k) {
        ms[k] = m_stageMs[k];
        m_stageMs[k] = 0;
    }
    const int n = m_stageRounds;
    m_stageRounds = 0;
    return n;
}

NuiFusionServer::~NuiFusionServer() {
    if (m_ctl) nuif_stream_synchronize(m_ctl);
    for (auto& s : m_sessions)
        if (s) s->engine->synchronize();
    m_sessions.clear();
    for (nuif_event_t e : m_ends)
        if (e) nuif_event_destroy(e);
    if (m_start) nuif_event_destroy(m_start);
    if (m_batchEnd) nuif_event_destroy(m_batchEnd);
    for (nuif_event_t e : m_stageEv)
        if (e) nuif_event_destroy(e);
    if (m_roundGraph) nuif_graph_destroy(m_roundGraph);
    if (m_staging) nuif_host_free(m_staging);
    if (m_status) nuif_host_free(m_status);
    if (m_ctl) nuif_stream_destroy(m_ctl);
}

void NuiFusionServer::setFailurePolicy(int policy) {
    if (policy != NUIFS_ON_FAILURE_KEEP && policy != NUIFS_ON_FAILURE_RESET && policy != NUIFS_ON_FAILURE_EVICT)
        throw std::invalid_argument("server: unknown failure policy");
    m_policy = policy;
}

void NuiFusionServer::ensureTable(int n) {
    if (n <= m_tableCap) return;
    const int cap = std::max(n, 2 * m_tableCap);
    if (m_batchPending) check(nuif_event_synchronize(m_batchEnd), "batched round");
    m_batchPending = false;
    m_table = DeviceBuffer(nuif_fusion_round_table_bytes(cap));
    if (m_staging) nuif_host_free(m_staging);
    m_staging = nullptr;
    check(nuif_host_alloc(&m_staging, nuif_fusion_round_table_bytes(cap)), "round staging");
    m_tableCap = cap;
}

int NuiFusionServer::open(const nuifk_desc& d) {
    const size_t need = NuiKinfuSession::deviceBytes(d);
    if (m_inUse + need > m_budget) return NUIF_ERR_OOM;
    auto s = NuiKinfuSession::create(d);
    nuif_event_t e = nullptr;
    check(nuif_event_create(&e), "session event");
    int id = -1;
    for (size_t i = 0; i < m_sessions.size(); ++i)
        if (!m_sessions[i]) {
            id = (int)i;
            break;
        }
    if (id < 0) {
        id = (int)m_sessions.size();
        m_sessions.emplace_back();
        m_bytes.push_back(0);
        m_ends.push_back(nullptr);
    }
    m_sessions[id] = std::move(s);
    m_bytes[id] = need;
    m_ends[id] = e;
    m_inUse += need;
    return id;
}

void NuiFusionServer::close(int id) {
    NuiKinfuSession* s = session(id);
    s->engine->synchronize();
    m_sessions[id].reset();
    m_inUse -= m_bytes[id];
    m_bytes[id] = 0;
    nuif_event_destroy(m_ends[id]);
    m_ends[id] = nullptr;
}

NuiKinfuSession* NuiFusionServer::session(int id) const {
    if (id < 0 || id >= (int)m_sessions.size() || !m_sessions[id]) throw std::runtime_error("server: no such session");
    return m_sessions[id].get();
}

int NuiFusionServer::sessionCount() const {
    int n = 0;
    for (auto& s : m_sessions) n += s ? 1 : 0;
    return n;
}

void NuiFusionServer::submit(int n, const int* ids, const uint16_t* const* depths, const uint8_t* const* bgras) {
    if (n < 0 || (n > 0 && (!ids || !depths))) throw std::runtime_error("server submit: bad arguments");
    for (int i = 0; i < n; ++i) {
        session(ids[i]);
        if (!depths[i]) throw std::runtime_error("server submit: null frame");
        for (int j = 0; j < i; ++j)
            if (ids[j] == ids[i]) throw std::runtime_error("server submit: a session appears twice in a round");
    }
    m_round.assign(ids, ids + n);
    m_slot.assign((size_t)n, -1);
    m_outcomes.clear();
    check(nuif_event_record(m_start, m_ctl), "round start");
    // the staging copy and the outcome slots of the previous batched round are free
    if (m_batchPending) check(nuif_event_synchronize(m_batchEnd), "batched round");
    m_batchPending = false;
    if (n > m_statusCap) {
        if (m_status) nuif_host_free(m_status);
        m_status = nullptr;
        void* p = nullptr;
        check(nuif_host_alloc(&p, (size_t)std::max(n, 2 * m_statusCap) * sizeof(int32_t)), "outcome slots");
        m_status = static_cast<int32_t*>(p);
        m_statusCap = std::max(n, 2 * m_statusCap);
    }
    std::vector<nuif_round_session> batch;
    std::vector<int> batched;
    for (int i = 0; i < n; ++i) {
        NuiKinfuSession* s = m_sessions[ids[i]].get();
        nuif_round_session rs;
        const uint8_t* bgra = bgras ? bgras[i] : nullptr;
        if (m_batching && s->engine->PrepareRoundFrame(depths[i], bgra, s->scene.get(), s->cam, &rs)) {
            m_slot[i] = (int)batch.size();
            m_status[m_slot[i]] = -1;
            rs.icp.outcome = m_status + m_slot[i];
            batch.push_back(rs);
            batched.push_back(ids[i]);
            continue;
        }
        const nuif_stream_t st = s->engine->stream();
        check(nuif_stream_wait_event(st, m_start), "round wait");
        s->engine->RunTrackingDevice(0, depths[i], bgra, s->scene.get(), s->cam, false);
        s->afterFrame();
        check(nuif_event_record(m_ends[ids[i]], st), "frame end");
    }
    m_lastBatch = (int)batch.size();
    m_stagePending = false;
    if (batch.empty()) return;
    // the batch follows each session's previous frame; each session's stream follows the
    // batch (its later work and read-backs see the round's results)
    for (int id : batched) check(nuif_stream_wait_event(m_ctl, m_ends[id]), "batch wait");
    ensureTable((int)batch.size());
    launchRound(batch.data(), (int)batch.size());
    m_stagePending = m_stageTiming;
    check(nuif_event_record(m_batchEnd, m_ctl), "batch end");
    m_batchPending = true;
    for (int id : batched) {
        NuiKinfuSession* s = m_sessions[id].get();
        check(nuif_stream_wait_event(s->engine->stream(), m_batchEnd), "session wait");
        check(nuif_event_record(m_ends[id], m_ctl), "frame end");
        s->afterFrame();
    }
}

void NuiFusionServer::setRoundGraph(bool on) {
    m_roundGraphs = on;
    if (!on && m_roundGraph) {
        if (m_batchPending) check(nuif_event_synchronize(m_batchEnd), "batched round");
        nuif_graph_destroy(m_roundGraph);
        m_roundGraph = nullptr;
    }
}

// The batched round: with stage timing, enqueued with its events; else the host table is
// written and the captured round replayed, captured anew when its launches changed (other
// sessions' sizes or orders, a reallocated table)
void NuiFusionServer::launchRound(const nuif_round_session* batch, int n) {
    if (m_stageTiming || !m_roundGraphs) {
        check(nuif_fusion_round_timed(batch, n, m_table.get(), m_staging, m_stageTiming ? m_stageEv : nullptr, m_ctl),
              "nuif_fusion_round");
        return;
    }
    uint64_t sig = 0;
    check(nuif_fusion_round_prepare(batch, n, m_table.get(), m_staging, &sig), "nuif_fusion_round_prepare");
    if (!m_roundGraph || sig != m_roundSig) {
        if (m_roundGraph) nuif_graph_destroy(m_roundGraph);
        m_roundGraph = nullptr;
        check(nuif_graph_begin(m_ctl), "round capture");
        const int rc = nuif_fusion_round(batch, n, m_table.get(), m_staging, m_ctl);
        nuif_graph_t g = nullptr;
        const int rc_end = nuif_graph_end(m_ctl, &g);
        check(rc, "nuif_fusion_round");
        check(rc_end, "round capture");
        m_roundGraph = g;
        m_roundSig = sig;
    }
    check(nuif_graph_launch(m_roundGraph, m_ctl), "round replay");
}

int NuiFusionServer::wait(float* latencyMs) {
    for (size_t i = 0; i < m_round.size(); ++i) {
        const int id = m_round[i];
        check(nuif_event_synchronize(m_ends[id]), "frame wait");
        if (latencyMs) check(nuif_event_elapsed_ms(m_start, m_ends[id], &latencyMs[i]), "frame latency");
    }
    if (m_stagePending) {
        check(nuif_event_synchronize(m_stageEv[5]), "stage events");
        for (int k = 0; k < 5; ++k) {
            float ms = 0.f;
            check(nuif_event_elapsed_ms(m_stageEv[k], m_stageEv[k + 1], &ms), "stage time");
            m_stageMs[k] += ms;
        }
        m_stageRounds++;
        m_stagePending = false;
    }
    // outcomes: the batched frames' from their slots, the others' from their session's state
    // (the first frame of a session records its pose and always tracks)
    m_outcomes.assign(m_round.size(), 1);
    for (size_t i = 0; i < m_round.size(); ++i) {
        NuiKinfuSession* s = m_sessions[m_round[i]].get();
        if (m_slot[i] >= 0) {
            m_outcomes[i] = m_status[m_slot[i]] == 1 ? 1 : 0;
        } else if (s->engine->getFrameID() > 1) {
            m_outcomes[i] = s->engine->tracker()->ReadState().status == 0 ? 1 : 0;
        }
    }
    if (m_policy == NUIFS_ON_FAILURE_KEEP) return 0;
    for (size_t i = 0; i < m_round.size(); ++i) {
        if (m_outcomes[i]) continue;
        const int id = m_round[i];
        if (m_policy == NUIFS_ON_FAILURE_EVICT) {
            close(id);
        } else {
            NuiKinfuSession* s = m_sessions[id].get();
            s->engine->synchronize();
            s->engine->reset();
            s->scene->reset();
            s->offsets.clear();
        }
    }
    return 0;
}

}  // namespace nuif_kinfu
