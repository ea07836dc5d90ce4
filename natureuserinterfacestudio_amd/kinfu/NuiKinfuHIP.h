// GPU implementations of the fusion interfaces on top of the nuif C ABI
// (replacing OCL/NuiKinfuOpenCL{Frame,FeedbackFrame,CameraState,DepthTracker,
// Scene}.*).  Every object of one tracking session shares the session's HIP
// stream; nothing is global.  All device work is enqueued asynchronously; the
// engine decides where to synchronise.
#pragma once

#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/nuif.h"
#include "NuiKinfuConfig.h"
#include "NuiKinfuInterfaces.h"

namespace nuif_kinfu {

// RAII device buffer (replaces NuiGPUMemManager::CreateBufferCL / ReleaseMemObjectCL)
class DeviceBuffer {
public:
    DeviceBuffer() = default;
    explicit DeviceBuffer(size_t bytes);
    ~DeviceBuffer();
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    DeviceBuffer(DeviceBuffer&& o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
    DeviceBuffer& operator=(DeviceBuffer&& o) noexcept;
    void* get() const { return p_; }
    template <class T> T* as() const { return static_cast<T*>(p_); }
    size_t bytes() const { return n_; }

private:
    void* p_ = nullptr;
    size_t n_ = 0;
};

// Throws std::runtime_error with nuif_last_error() on a non-zero status.
void check(int rc, const char* what);

// OCL/NuiKinfuOpenCLCameraState.*: device camera params + rigid transform
class NuiKinfuHIPCameraState : public NuiKinfuCameraState {
public:
    explicit NuiKinfuHIPCameraState(nuif_stream_t s);
    ~NuiKinfuHIPCameraState() override;
    void UpdateCameraTransform(const Matrix3frm& rot, const Vector3f& tran) override;
    void UpdateCameraTransform(const Matrix3frm& rot, const Vector3f& tran, const Vector3f& offsetTran) override;
    void UpdateCameraParams(const NuiCameraParams& cameraParams, UINT nWidth, UINT nHeight) override;
    // host pose <- device pose (after asynchronous tracking)
    void SyncFromDevice();
    NuiCLRigidTransform* GetCameraTransformBuffer() const { return m_pose.as<NuiCLRigidTransform>(); }
    const NuiCLCameraParams& GetCameraParams() const { return m_cam; }

private:
    void WriteToBuffer(const Matrix3frm& rot, const Vector3f& tran);
    nuif_stream_t m_stream;
    DeviceBuffer m_pose;
    NuiCLRigidTransform* m_staging = nullptr;  // pinned
    NuiCLCameraParams m_cam{};
};

// OCL/NuiKinfuOpenCLFrame.*: raw depth upload, passing + bilateral filter,
// vertex pyramid (the pyramid of OCL/NuiKinfuOpenCLDepthTracker.cpp:140-236 is
// produced here in the same fused preprocessing pass).
class NuiKinfuHIPFrame : public NuiKinfuFrame {
public:
    NuiKinfuHIPFrame(const NuiTrackerConfig& config, UINT nWidth, UINT nHeight, nuif_stream_t s);
    ~NuiKinfuHIPFrame() override;
    void UpdateVertexBuffers(UINT16* pDepths, UINT nNum, NuiKinfuCameraState* pCameraState) override;
    void UpdateColorBuffers(BGRQUAD* pColors, UINT nNum) override;
    // OCL/NuiKinfuOpenCLFrame.cpp:405-448 on the device (nuif_map_depth_to_color)
    void UpdateColorBuffers(ColorSpacePoint* pDepthToColor, UINT* pDepthDistortionLT, UINT nNum,
                            const NuiColorImage& image) override;
    // device-resident inputs (no host copy)
    void UpdateVertexBuffersDevice(const UINT16* dDepths, NuiKinfuCameraState* pCameraState);
    void UpdateColorBuffersDevice(const uint8_t* dBgra);
    UINT GetWidth() const override { return m_W; }
    UINT GetHeight() const override { return m_H; }
    int GetLevels() const { return m_levels; }
    // the scene needing ray-length depth (dense TSDF) sets this before the first frame
    void SetScaleDepthForTsdf(bool v) { m_scaleForTsdf = v; }
    bool IsDepthScaled() const { return m_scaleForTsdf; }
    float* GetDepthBuffer() const { return m_depth.as<float>(); }
    // max-depth pyramid of the ray-length depth (nuif_tsdf_depth_bounds), when scaled
    const float* GetDepthBounds() const { return m_bounds.as<float>(); }
    float* GetFilteredDepthBuffer() const { return m_dpyr[0].as<float>(); }
    float* GetDepthPyramidLevel(int level) const { return m_dpyr[level].as<float>(); }
    const float* GetGaussian() const { return m_gaussian.data(); }
    const NuiTrackerConfig& GetConfig() const { return m_cfg; }
    // the depth-bounds buffer of the ray-length depth (allocated on first use)
    float* EnsureDepthBounds();
    float* GetVertexBuffer(int level = 0) const { return m_vpyr[level].as<float>(); }
    const float* const* GetVertexPyramid() const { return m_vptrs.data(); }
    uint8_t* GetColorBuffer() const { return m_hasColor ? m_colors.as<uint8_t>() : nullptr; }
    UINT16* GetRawBuffer() const { return m_raw.as<UINT16>(); }
    float GetDepthThreshold() const { return m_cfg.depth_threshold; }

private:
    void Preprocess(NuiKinfuCameraState* pCameraState);
    DeviceBuffer m_colorMap, m_colorLut, m_colorImage;  // staging of the depth-to-colour mapping
    NuiTrackerConfig m_cfg;
    UINT m_W, m_H;
    int m_levels;
    nuif_stream_t m_stream;
    bool m_scaleForTsdf = false;
    bool m_hasColor = false;
    DeviceBuffer m_raw, m_depth, m_colors, m_bounds;
    std::vector<DeviceBuffer> m_dpyr, m_vpyr;
    std::vector<float*> m_dptrs;
    std::vector<const float*> m_vptrs;
    std::vector<float> m_gaussian;
};

// OCL/NuiKinfuOpenCLFeedbackFrame.*: the reference (model) maps used by ICP
class NuiKinfuHIPFeedbackFrame : public NuiKinfuFeedbackFrame {
public:
    NuiKinfuHIPFeedbackFrame(UINT nWidth, UINT nHeight, nuif_stream_t s);
    void UpdateBuffers(NuiKinfuFrame* frame, NuiKinfuCameraState* pCameraState) override;
    UINT GetWidth() const override { return m_W; }
    UINT GetHeight() const override { return m_H; }
    float* GetVertexBuffer() const { return m_vmap.as<float>(); }
    float* GetNormalBuffer() const { return m_nmap.as<float>(); }
    uint8_t* GetColorBuffer() const { return m_colors.as<uint8_t>(); }

private:
    UINT m_W, m_H;
    nuif_stream_t m_stream;
    DeviceBuffer m_vmap, m_nmap, m_colors;
};

// OCL/NuiKinfuOpenCLDepthTracker.*: pyramid projective ICP, all on device
class NuiKinfuHIPDepthTracker : public NuiKinfuTracker {
public:
    NuiKinfuHIPDepthTracker(const NuiTrackerConfig& config, UINT nWidth, UINT nHeight, nuif_stream_t s);
    bool EstimatePose(NuiKinfuFrame* pFrame, NuiKinfuFeedbackFrame* pFeedbackFrame, NuiKinfuCameraState* pCameraState,
                      Affine3f* hint) override;
    void log(const std::string& fileName) const override { m_cfg.log(fileName); }
    float getError() const override;
    int getCount() const override;
    // tracking state (device) and its flags, consumed by the scene
    void* GetState() const { return m_state.get(); }
    const int32_t* GateFlag() const { return nuif_icp_gate_flag(m_state.get()); }
    const int32_t* OkFlag() const { return nuif_icp_ok_flag(m_state.get()); }
    void ResetState(NuiKinfuHIPCameraState* cam);
    // frame without ICP (first frame): records the pose and advances the frame count
    void RecordPose(NuiKinfuHIPCameraState* cam);
    void SetIntegrationThreshold(float v) { m_icfg.gate_threshold = v; }
    // NUIF_ICP_REDUCE_REFERENCE (default: the reference's float summation order),
    // NUIF_ICP_REDUCE_COMPENSATED (Kahan FP32 + FP64) or NUIF_ICP_REDUCE_JOINT (the
    // compensated order with photometric rows against the feedback colours: the frame
    // and the feedback frame must carry colours)
    void SetReduction(int mode);
    int Reduction() const { return m_icfg.reduction; }
    // the joint order's model: the previous frame's colours (kept from each frame)
    void KeepColors(NuiKinfuHIPFrame* f);
    uint8_t* PreviousColors() const { return m_prevColors.as<uint8_t>(); }
    // the joint order's photometric weight, intensity difference and minimum gradient
    void SetPhotometric(float weight, float maxDifference, float minGradient);
    void SetHistory(float* dHistory, int capacity) { m_history = dHistory; m_historyCap = capacity; }
    nuif_icp_result ReadState() const;
    // the arguments of this tracker's nuif_icp_estimate for a frame (a batched round's entry)
    void FillIcpSession(NuiKinfuHIPFrame* f, NuiKinfuHIPFeedbackFrame* fb, NuiKinfuHIPCameraState* cam,
                        nuif_icp_session* out) const;
    // a shifting volume moved the local frame: new local translation and frame offset
    void SetTranslation(NuiKinfuHIPCameraState* cam, const float t[3], const float offset[3]);

private:
    NuiTrackerConfig m_cfg;
    nuif_icp_config m_icfg{};
    UINT m_W, m_H;
    nuif_stream_t m_stream;
    DeviceBuffer m_state, m_partials;
    float* m_history = nullptr;
    int m_historyCap = 0;
    // the joint order: intensity pyramid of the frame, the feedback intensities and gradients
    DeviceBuffer m_intensity[NUIF_ICP_MAX_LEVELS], m_refModel, m_prevColors;
    nuif_icp_photometric m_photo{};
    void preparePhotometric(NuiKinfuHIPFrame* f, NuiKinfuHIPFeedbackFrame* fb);
};

// OCL/NuiKinfuOpenCLColorTracker.*: photometric tracker against the feedback
// frame's world vertices and colours.  The reference's factory never builds it
// (DS/NuiKinfuTrackingFactory.h:36-43); here it is a tracker of its own that a
// caller constructs directly.  corrected selects the decode of the sums
// (nuif_color_icp_begin): false = the reference's host loop as written.
class NuiKinfuHIPColorTracker : public NuiKinfuTracker {
public:
    NuiKinfuHIPColorTracker(const NuiTrackerConfig& config, UINT nWidth, UINT nHeight, bool corrected,
                            nuif_stream_t s);
    // OCL/NuiKinfuOpenCLColorTracker.cpp:580-605: SubSampleColors, GradientBuffers,
    // ColorIterativeClosestPoint; synchronises the stream (as the reference's reads do)
    bool EstimatePose(NuiKinfuFrame* pFrame, NuiKinfuFeedbackFrame* pFeedbackFrame, NuiKinfuCameraState* pCameraState,
                      Affine3f* hint) override;
    bool hasColorData() const override { return true; }
    void log(const std::string& fileName) const override { m_cfg.log(fileName); }
    float getError() const override { return m_error; }
    int getCount() const override { return (int)m_count; }
    const nuif_color_icp_result& LastResult() const { return m_last; }

private:
    NuiTrackerConfig m_cfg;
    UINT m_W, m_H;
    bool m_corrected;
    nuif_stream_t m_stream;
    std::vector<DeviceBuffer> m_colors, m_gx, m_gy;
    DeviceBuffer m_state, m_partials, m_pose;
    float m_error = 0.f, m_count = 0.f;
    nuif_color_icp_result m_last{};
};

// What the tracking engine needs from a GPU scene beyond NuiKinfuScene: the
// device predicates of the tracker (integration gate, tracking success) that
// guard its launches, and whether it integrates ray-length depth.
class NuiKinfuHIPSceneBase : public NuiKinfuScene {
public:
    void SetGates(const int32_t* integrate, const int32_t* raycast) {
        m_gateIntegrate = integrate;
        m_gateRaycast = raycast;
    }
    // the dense scene integrates depth scaled to ray length (K/tsdfVolume.cl:8-29);
    // the hash scene integrates the passing-filter depth as it is
    virtual bool NeedsRayLengthDepth() const = 0;
    // the scene's main device allocation (dense volume / hash scene) and its size
    virtual void* DevicePointer() const = 0;
    virtual size_t DeviceBytes() const = 0;
    // truncation distance (dense: clamped tranc_dist; hash: truncScale), metres
    virtual float Truncation() const = 0;
    // running count of voxels written by integration (synchronises)
    virtual unsigned long long UpdatedVoxels() const = 0;
    // the scene's integration waits for the frame's pose and decides on the host
    // (shifting volume, host swapping): such frames cannot be replayed from a graph
    virtual bool RequiresHostWork() const { return false; }

protected:
    const int32_t* m_gateIntegrate = nullptr;
    const int32_t* m_gateRaycast = nullptr;
};

// OCL/NuiKinfuOpenCLScene.*: dense TSDF volume
class NuiKinfuHIPScene : public NuiKinfuHIPSceneBase {
public:
    NuiKinfuHIPScene(const NuiKinfuVolumeConfig& config, int storage, nuif_stream_t s);
    bool log(const std::string& fileName) const override { return m_config.log(fileName); }
    bool hasColorData() const override { return m_config.bHas_color_volume; }
    float getVoxelLeafSize() const override;
    void reset() override;
    bool integrateVolume(NuiKinfuFrame* pFrame, NuiKinfuCameraState* pCameraState) override;
    void raycastRender(NuiKinfuFeedbackFrame* pFeedbackFrame, NuiKinfuCameraState* pCameraState) override;
    // OCL/NuiKinfuOpenCLScene.cpp:572-615 via nuif_points_dense: the point cloud
    // (positions and colours, no indices); false when the scene is unchanged or empty
    bool Volume2CLVertices(NuiKinfuMesh* pData) override;
    // K/marchingCube.cl via nuif_mc_dense: triangle soup (3 vertices per triangle)
    bool Volume2CLMesh(NuiKinfuMesh* pData) override;
    // OCL/NuiKinfuOpenCLScene.cpp:788-984 via nuif_mesh_dense: shared vertices + indices
    bool Volume2Mesh(NuiKinfuMesh* pMesh) override;
    bool NeedsRayLengthDepth() const override { return true; }
    void* DevicePointer() const override { return m_volume.get(); }
    size_t DeviceBytes() const override { return m_volume.bytes(); }
    float Truncation() const override { return m_params.tranc_dist; }
    const TsdfParams& GetParams() const { return m_params; }
    int GetStorage() const { return m_storage; }
    void* GetVolume() const { return m_volume.get(); }
    uint8_t* GetColorVolume() const { return m_color.as<uint8_t>(); }
    size_t VolumeBytes() const { return m_volume.bytes(); }
    unsigned long long UpdatedVoxels() const override;
    // distinct voxels the raycast reads at the camera's current pose (nuif_tsdf_raycast_footprint;
    // synchronises); not for a wrapped volume
    long long RaycastFootprint(NuiKinfuHIPCameraState* cam) const;
    using NuiKinfuScene::Volume2Mesh;
    // the scene's part of a batched round's entry (volume, surface blocks, walk states);
    // false when the volume cannot be batched (a shifting volume, or dimensions the tile
    // integration does not take)
    virtual bool FillRoundScene(nuif_round_session* out);

protected:
    NuiKinfuVolumeConfig m_config;
    int m_storage;
    nuif_stream_t m_stream;
    TsdfParams m_params{};
    DeviceBuffer m_volume, m_color, m_counter;
    DeviceBuffer m_surface;  // surface blocks of the volume (nuif_tsdf_surface_blocks_bytes)
    int32_t m_wrap[3] = {0, 0, 0};  // voxel wrap (non-zero only in a shifting volume)
    DeviceBuffer m_walk;            // per-column walk states of the integration (nuif_tsdf_walk_states_bytes)
};

// OCL/NuiKinfuOpenCLShiftScene.*: the dense volume as a cyclic buffer that follows the
// camera.  Before an integration the camera's local translation is compared with the
// volume basis; a move of voxel_shift voxels or more along an axis evicts the slab the
// volume leaves (its points go to the vertex cache), clears it, and advances the voxel
// offset; integration and raycast then address the volume through the wrap.  The
// decision is the reference's host-side one: an integrated frame waits for its tracked
// pose (no graph capture).
class NuiKinfuHIPShiftScene : public NuiKinfuHIPScene {
public:
    NuiKinfuHIPShiftScene(const NuiKinfuVolumeConfig& config, int storage, nuif_stream_t s);
    void reset() override;
    bool integrateVolume(NuiKinfuFrame* pFrame, NuiKinfuCameraState* pCameraState) override;
    bool RequiresHostWork() const override { return true; }
    // the whole volume placed at its offset (the reference's fetchShiftVolumeKernel with wrap 0)
    bool Volume2CLVertices(NuiKinfuMesh* pData) override;
    void SetTracker(NuiKinfuHIPDepthTracker* tracker) { m_tracker = tracker; }
    const int* VoxelOffset() const { return m_offset; }
    // getVoxelOffsetSize (metres)
    void OffsetSize(float out[3]) const;
    // NuiKinfuVertexCache: points of the evicted slabs
    const std::vector<float>& CachedPositions() const { return m_cachePos; }
    const std::vector<float>& CachedColors() const { return m_cacheCol; }

private:
    void fetchAndClear(int axis, int trans);
    void shiftVolume(float t[3]);
    void updateWrap();
    int m_offset[3] = {0, 0, 0};
    NuiKinfuHIPDepthTracker* m_tracker = nullptr;
    std::vector<float> m_cachePos, m_cacheCol;
    DeviceBuffer m_scratch;
};

// OCL/NuiKinfuOpenCLHashScene.*: voxel-hashing scene.  The render is castRay over
// ray ranges from the visible blocks (OCL/NuiKinfuOpenCLHashScene.cpp:659-736) or,
// with m_bUseForwardRender, the forward render (:991-1012).
class NuiKinfuHIPHashScene : public NuiKinfuHIPSceneBase {
public:
    NuiKinfuHIPHashScene(const NuiHashingSDFConfig& config, int storage, bool color, nuif_stream_t s);
    bool log(const std::string&) const override { return true; }
    bool hasColorData() const override { return m_params.color != 0; }
    bool needAcceleratedFeedback() const override { return true; }
    float getVoxelLeafSize() const override { return m_config.m_virtualVoxelSize; }
    void reset() override;
    bool integrateVolume(NuiKinfuFrame* pFrame, NuiKinfuCameraState* pCameraState) override;
    void raycastRender(NuiKinfuFeedbackFrame* pFeedbackFrame, NuiKinfuCameraState* pCameraState) override;
    // OCL/NuiKinfuOpenCLHashScene.cpp:1066-1190 via nuif_hash_points over the visible
    // list; false when the scene is unchanged or nothing is visible
    bool Volume2CLVertices(NuiKinfuMesh* pData) override;
    // K/KinfuHashVisualization.cl via nuif_hash_mesh over the visible list
    bool Volume2CLMesh(NuiKinfuMesh* pData) override;
    bool Volume2Mesh(NuiKinfuMesh* pMesh) override { return Volume2CLMesh(pMesh); }
    using NuiKinfuScene::Volume2Mesh;
    bool NeedsRayLengthDepth() const override { return false; }
    void* DevicePointer() const override { return m_scene.get(); }
    size_t DeviceBytes() const override { return m_scene.bytes(); }
    float Truncation() const override { return m_params.trunc_scale; }
    const nuif_hash_params& GetParams() const { return m_params; }
    void* GetScene() const { return m_scene.get(); }
    float* GetRangeImage() const { return m_range.as<float>(); }
    // bricks, visible-list length, failed insertions, occupied entries (synchronises)
    void Stats(uint32_t out[4]) const;
    unsigned long long UpdatedVoxels() const override;
    static nuif_hash_params ParamsOf(const NuiHashingSDFConfig& config, int storage, bool color);
    // host swapping (m_bUseSwapping): blocks held by the host store
    size_t StoredBlocks() const;
    bool RequiresHostWork() const override { return m_cache != nullptr; }

private:
    NuiHashingSDFConfig m_config;
    nuif_hash_params m_params{};
    nuif_stream_t m_stream;
    DeviceBuffer m_scene, m_range, m_counter, m_forwardWork;
    UINT m_rangeW = 0, m_rangeH = 0;
    std::unique_ptr<class NuiKinfuHIPHashGlobalCache> m_cache;
};

// OCL/NuiKinfuOpenCLHashGlobalCache.*: the host store of swapped-out bricks (keyed by
// hash entry) and the transfers of one frame: the swap-in list receives its stored
// blocks (zeros when none) and merges them into the active bricks, the swap-out list
// leaves the device into the store.  Synchronous: each list length is read back.
class NuiKinfuHIPHashGlobalCache {
public:
    NuiKinfuHIPHashGlobalCache(const nuif_hash_params& p, nuif_stream_t s);
    void reset() { m_store.clear(); }
    // integrateGlobalIntoLocal + saveToGlobalMemory (OCL/NuiKinfuOpenCLHashScene.cpp:576-580)
    void update(void* scene);
    size_t StoredBlocks() const { return m_store.size(); }

private:
    struct Stored {
        std::vector<uint32_t> words;
        std::vector<uint8_t> colors;
    };
    uint32_t list(void* scene, int which);
    void grow(uint32_t n);
    nuif_hash_params m_p;
    nuif_stream_t m_stream;
    DeviceBuffer m_scratch, m_ids, m_count, m_words, m_colors;
    uint32_t m_cap = 0;
    std::vector<int32_t> m_hostIds;
    std::vector<uint32_t> m_hostWords;
    std::vector<uint8_t> m_hostColors;
    std::unordered_map<int32_t, Stored> m_store;
};

}  // namespace nuif_kinfu

namespace NuiKinfuEngine {

// DS/NuiKinfuTrackingFactory.h:23-50
class NuiKinfuTrackingFactory {
public:
    static NuiKinfuTrackingFactory& Instance() {
        static NuiKinfuTrackingFactory s;
        return s;
    }
    static void BuildTrackingEngine(NuiKinfuTracker** pTracker, NuiKinfuFrame** pFrame,
                                    NuiKinfuFeedbackFrame** pFeedbackFrame, NuiKinfuCameraState** pCameraState,
                                    const NuiTrackerConfig& trackerConfig, bool bAcceleratedFeedback,
                                    unsigned int nWidth, unsigned int nHeight, nuif_stream_t stream);
};

// VO/NuiKinfuTrackingEngine.*: per-frame orchestration
class NuiKinfuTrackingEngine {
public:
    explicit NuiKinfuTrackingEngine(nuif_stream_t stream = nullptr);
    ~NuiKinfuTrackingEngine();
    bool isInit() const { return m_pTracker != nullptr; }
    void initialize(const NuiTrackerConfig& trackerConfig, bool bAcceleratedFeedback, UINT nWidth, UINT nHeight);
    void log(const std::string& fileName) const;
    void setTranslateBasis(const Vector3f& basis) { m_translateBasis = basis; }
    const Vector3f& getTranslateBasis() const { return m_translateBasis; }
    void reset();
    // VO/NuiKinfuTrackingEngine.cpp:68-142; synchronous: returns the tracking outcome
    bool RunTracking(INT64 timeStamp, UINT16* pDepths, BGRQUAD* pColors, UINT nPointNum, NuiKinfuScene* pVolume,
                     const NuiCameraParams& cameraParams);
    // Device-resident inputs; sync=false enqueues the frame and returns at once
    // (outcome via getStatus()); with useGraph the frame pipeline is replayed
    // from a captured hipGraph after its first execution.
    bool RunTrackingDevice(INT64 timeStamp, const UINT16* dDepths, const uint8_t* dBgra, NuiKinfuScene* pVolume,
                           const NuiCameraParams& cameraParams, bool sync);
    void setUseGraph(bool v) { m_useGraph = v; }
    bool synchronize();  // waits for queued frames; false if any tracking failure occurred
    const NuiCameraPos& getCameraPose(int time = -1) const;
    float getTrackerError() const;
    int getTrackerCount() const;
    int getFrameID() const { return m_frames; }
    void setIntegrationMetricThreshold(float threshold);
    nuif_stream_t stream() const { return m_stream; }
    nuif_kinfu::NuiKinfuHIPFeedbackFrame* feedbackFrame() const { return m_pFeedbackFrame; }
    nuif_kinfu::NuiKinfuHIPDepthTracker* tracker() const { return m_pTracker; }
    nuif_kinfu::NuiKinfuHIPCameraState* cameraState() const { return m_pCameraState; }
    nuif_kinfu::NuiKinfuHIPFrame* frame() const { return m_pFrame; }
    const float* poseHistoryDevice() const { return m_history.as<float>(); }
    int poseHistoryCapacity() const { return m_historyCap; }
    // Rows of the device pose ring (written by the frame's ICP end, row frame % capacity)
    // are copied to the host before the ring can wrap over them; capacity >= 1.
    void setPoseHistoryCapacity(int capacity);
    // Every frame's row (NUIF_POSE_ROW floats: R, t, ok, error, count, integrated) since
    // the last reset, in frame order; synchronises the stream.
    const std::vector<float>& poseRows() const;
    // A tracked frame of a plain dense scene as one session of a batched round
    // (nuif_fusion_round): fills *out from the session's buffers and counts the frame.
    // Returns false, doing nothing, when the frame must run alone: the first frame, a
    // scene other than a plain dense volume of batchable size, or stage timing on.
    // dBgra: the frame's colours (device) or NULL.
    bool PrepareRoundFrame(const UINT16* dDepths, const uint8_t* dBgra, NuiKinfuScene* pScene,
                           const NuiCameraParams& cameraParams, nuif_round_session* out);
    // Per-stage device timing of tracked frames (events around preprocessing, ICP,
    // integration and raycast on the session stream; frames then run un-captured).
    void setStageTiming(bool on);
    // Sums of the stage durations (ms) of the frames timed since the last call,
    // and their number; synchronises the stream.
    int takeStageTimes(double ms[4]);

private:
    void enqueueTrackedFrame(nuif_kinfu::NuiKinfuHIPSceneBase* scene);
    void drainHistory() const;
    void beforeFrame();
    void recordStage(int k);
    bool m_stageTiming = false;
    std::vector<nuif_event_t> m_events;  // 5 per timed frame
    size_t m_eventsUsed = 0;
    nuif_stream_t m_stream;
    bool m_ownStream = false;
    nuif_kinfu::NuiKinfuHIPDepthTracker* m_pTracker = nullptr;
    nuif_kinfu::NuiKinfuHIPFrame* m_pFrame = nullptr;
    nuif_kinfu::NuiKinfuHIPFeedbackFrame* m_pFeedbackFrame = nullptr;
    nuif_kinfu::NuiKinfuHIPCameraState* m_pCameraState = nullptr;
    Vector3f m_translateBasis;
    float m_integrationThreshold = 0.15f;  // VO/NuiKinfuTrackingEngine.cpp:22,32
    int m_frames = 0;
    bool m_useGraph = false;
    nuif_graph_t m_graph = nullptr;
    nuif_kinfu::NuiKinfuHIPSceneBase* m_graphScene = nullptr;
    nuif_kinfu::DeviceBuffer m_history;
    int m_historyCap = 1 << 16;
    mutable std::vector<float> m_rows;  // host copy of the rows of frames [0, m_drained)
    mutable int m_drained = 0;
    mutable bool m_anyFailed = false;
    mutable std::vector<NuiCameraPos> m_poses;
};

}  // namespace NuiKinfuEngine
