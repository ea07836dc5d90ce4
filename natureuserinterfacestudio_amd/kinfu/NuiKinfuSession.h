// One tracking session (engine + dense scene) and the multi-session fusion
// server that keeps many of them resident on one GPU.
//
// The reference builds one engine and one scene per application
// (src/SLAM/Backend/NuiSLAMController.cpp:33-69) and runs every OpenCL kernel
// on one global in-order queue (CLU/NuiOpenCLGlobal.cpp:175).  Here a session
// owns its stream, volume, pyramids, feedback maps and tracking state, and the
// server admits sessions against a device-memory budget and enqueues one frame
// per session per round on the sessions' own streams, so that the GPU overlaps
// independent sessions.
#pragma once
#include <memory>
#include <vector>

#include "../../include/nuif_kinfu.h"
#include "NuiKinfuHIP.h"

namespace nuif_kinfu {

struct NuiKinfuSession {
    std::unique_ptr<NuiKinfuEngine::NuiKinfuTrackingEngine> engine;
    std::unique_ptr<NuiKinfuHIPSceneBase> scene;
    NuiCameraParams cam;
    UINT W = 0, H = 0;
    std::vector<float> offsets;  // shifting volume: offset size after each frame (3 per frame)
    // after each frame: keeps the shifting volume's offset for the pose history
    void afterFrame();

    static std::unique_ptr<NuiKinfuSession> create(const nuifk_desc& d);
    // device bytes the session allocates (volume, colour volume, frame buffers,
    // pyramids, feedback maps, ICP state and workspace, pose history)
    static size_t deviceBytes(const nuifk_desc& d);
    // converts the descriptor's tracker fields (VO/NuiTrackerConfig.cpp:5-24)
    static NuiTrackerConfig trackerConfig(const nuifk_desc& d);
    static NuiHashingSDFConfig hashConfig(const nuifk_desc& d);
};

class NuiFusionServer {
public:
    explicit NuiFusionServer(size_t hbmBudget);
    ~NuiFusionServer();
    NuiFusionServer(const NuiFusionServer&) = delete;
    NuiFusionServer& operator=(const NuiFusionServer&) = delete;

    // admits a session; returns its id, or NUIF_ERR_OOM when it does not fit the budget
    int open(const nuifk_desc& d);
    void close(int id);
    NuiKinfuSession* session(int id) const;
    size_t bytesInUse() const { return m_inUse; }
    int sessionCount() const;

    // one round: a frame (device-resident uint16 depth) for each listed session.  The
    // tracked frames of plain dense sessions run together as one batched round
    // (nuif_fusion_round) on the server's stream; every other frame (a session's first,
    // hash and shifting scenes) runs on its session's own stream, overlapping the batch.
    void submit(int n, const int* ids, const uint16_t* const* depths, const uint8_t* const* bgras = nullptr);
    // waits for the round; latencyMs[i] = round start -> end of the i-th frame.  Then
    // each frame's tracking outcome is known (lastOutcomes) and the failure policy is
    // applied to the sessions whose frame failed.
    int wait(float* latencyMs);
    // what a failed frame does to its session (VO/NuiKinfuTrackingManager.cpp:94-106):
    // NUIFS_ON_FAILURE_KEEP (track the next frame from the last good pose), _RESET (the
    // reference's m_bAutoReset, with a cleared scene: the next frame starts the session
    // over), _EVICT (the reference's quit: the session is closed, its memory released)
    void setFailurePolicy(int policy);
    // outcome of each frame of the last waited round: 1 tracked, 0 failed
    const std::vector<int32_t>& lastOutcomes() const { return m_outcomes; }
    // device time of the batched round's stages (preprocess + pyramid, depth bounds, ICP,
    // integration, raycast), event-timed on the server stream
    void setStageTiming(bool on) { m_stageTiming = on; }
    // per-stage sums (ms) over the batched rounds timed since the last call; returns their number
    int takeStageTimes(double ms[5]);
    // batched rounds on (default) or off (every frame on its session's stream)
    void setBatching(bool on) { m_batching = on; }
    // frames of the last round that ran in the batch
    int lastBatchSize() const { return m_lastBatch; }
    // batched rounds captured once as a graph and replayed while the round's launches stay
    // the same (default on; stage timing runs rounds un-captured)
    void setRoundGraph(bool on);

private:
    void ensureTable(int n);
    size_t m_budget;
    size_t m_inUse = 0;
    std::vector<std::unique_ptr<NuiKinfuSession>> m_sessions;  // nullptr: closed id
    std::vector<size_t> m_bytes;
    std::vector<nuif_event_t> m_ends;  // per id: end of its last submitted frame
    std::vector<int> m_round;          // ids of the last round, in submission order
    nuif_stream_t m_ctl = nullptr;     // carries the round start event and the batched round
    nuif_event_t m_start = nullptr;
    nuif_event_t m_batchEnd = nullptr; // end of the last batched round
    bool m_batching = true;
    bool m_batchPending = false;       // a batched round's upload may still read the staging copy
    int m_lastBatch = 0;
    DeviceBuffer m_table;              // job table of the batched round (device)
    void* m_staging = nullptr;         // its pinned host copy
    int m_tableCap = 0;
    int32_t* m_status = nullptr;       // pinned: the batched frames' outcomes, written by the ICP end
    int m_statusCap = 0;
    std::vector<int> m_slot;           // per frame of the round: its outcome slot (-1: not batched)
    std::vector<int32_t> m_outcomes;
    int m_policy = NUIFS_ON_FAILURE_KEEP;
    bool m_stageTiming = false;
    bool m_stagePending = false;       // the last batched round recorded the stage events
    nuif_event_t m_stageEv[6] = {};
    double m_stageMs[5] = {0, 0, 0, 0, 0};
    int m_stageRounds = 0;
    bool m_roundGraphs = true;
    nuif_graph_t m_roundGraph = nullptr;  // the captured batched round and its launch signature
    uint64_t m_roundSig = 0;
    void launchRound(const nuif_round_session* batch, int n);
};

}  // namespace nuif_kinfu
