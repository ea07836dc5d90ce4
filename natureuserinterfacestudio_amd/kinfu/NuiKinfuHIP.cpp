// Host side of the fusion path: the reference's frame / feedback frame /
// camera state / depth tracker / dense scene / engine, implemented with the
// nuif C ABI.  Reference behaviour is cited per method.
#include "NuiKinfuHIP.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>

namespace nuif_kinfu {

void check(int rc, const char* what) {
    if (rc != NUIF_OK) throw std::runtime_error(std::string(what) + ": " + nuif_last_error());
}

DeviceBuffer::DeviceBuffer(size_t bytes) : n_(bytes) { check(nuif_malloc(&p_, bytes), "nuif_malloc"); }
DeviceBuffer::~DeviceBuffer() {
    if (p_) nuif_free(p_);
}
DeviceBuffer& DeviceBuffer::operator=(DeviceBuffer&& o) noexcept {
    if (this != &o) {
        if (p_) nuif_free(p_);
        p_ = o.p_;
        n_ = o.n_;
        o.p_ = nullptr;
        o.n_ = 0;
    }
    return *this;
}

// ------------------------------------------------------------ camera state

NuiKinfuHIPCameraState::NuiKinfuHIPCameraState(nuif_stream_t s) : m_stream(s), m_pose(sizeof(NuiCLRigidTransform)) {
    check(nuif_host_alloc((void**)&m_staging, sizeof(NuiCLRigidTransform)), "nuif_host_alloc");
    WriteToBuffer(Matrix3frm::Identity(), Vector3f::Zero());
}

NuiKinfuHIPCameraState::~NuiKinfuHIPCameraState() {
    if (m_staging) {
        nuif_stream_synchronize(m_stream);
        nuif_host_free(m_staging);
    }
}

// OCL/NuiKinfuOpenCLCameraState.cpp:64-90 WriteToBuffer: R, R.inverse(), t
void NuiKinfuHIPCameraState::WriteToBuffer(const Matrix3frm& rot, const Vector3f& tran) {
    check(nuif_stream_synchronize(m_stream), "camera staging");  // staging may still be in flight
    const Matrix3frm inv = rot.inverse();
    std::memcpy(m_staging->R, rot.data(), 9 * sizeof(float));
    std::memcpy(m_staging->R_inv, inv.data(), 9 * sizeof(float));
    std::memcpy(m_staging->t, tran.data(), 3 * sizeof(float));
    check(nuif_memcpy_h2d(m_pose.get(), m_staging, sizeof(NuiCLRigidTransform), m_stream), "pose upload");
}

void NuiKinfuHIPCameraState::UpdateCameraTransform(const Matrix3frm& rot, const Vector3f& tran) {
    WriteToBuffer(rot, tran);
    NuiKinfuCameraState::UpdateCameraTransform(rot, tran);
}

void NuiKinfuHIPCameraState::UpdateCameraTransform(const Matrix3frm& rot, const Vector3f& tran,
                                                   const Vector3f& offsetTran) {
    WriteToBuffer(rot, tran);
    NuiKinfuCameraState::UpdateCameraTransform(rot, tran, offsetTran);
}

// OCL/NuiKinfuOpenCLCameraState.cpp:33-62
void NuiKinfuHIPCameraState::UpdateCameraParams(const NuiCameraParams& p, UINT nWidth, UINT nHeight) {
    m_cam.fx = p.m_intrinsics.m_fx;
    m_cam.fx_inv = 1 / m_cam.fx;
    m_cam.fy = p.m_intrinsics.m_fy;
    m_cam.fy_inv = 1 / m_cam.fy;
    m_cam.cx = p.m_intrinsics.m_cx;
    m_cam.cy = p.m_intrinsics.m_cy;
    m_cam.depthImageWidth = nWidth;
    m_cam.depthImageHeight = nHeight;
    m_cam.sensorDepthWorldMin = p.m_sensorDepthMin;
    m_cam.sensorDepthWorldMax = p.m_sensorDepthMax;
    NuiKinfuCameraState::UpdateCameraParams(p, nWidth, nHeight);
}

void NuiKinfuHIPCameraState::SyncFromDevice() {
    NuiCLRigidTransform h;
    check(nuif_memcpy_d2h(&h, m_pose.get(), sizeof(h), m_stream), "pose read");
    check(nuif_stream_synchronize(m_stream), "pose read");
    Matrix3frm R;
    std::memcpy(R.data(), h.R, sizeof(h.R));
    NuiKinfuCameraState::UpdateCameraTransform(R, Vector3f(h.t[0], h.t[1], h.t[2]));
}

// ------------------------------------------------------------ frame

NuiKinfuHIPFrame::NuiKinfuHIPFrame(const NuiTrackerConfig& config, UINT nWidth, UINT nHeight, nuif_stream_t s)
    : m_cfg(config), m_W(nWidth), m_H(nHeight), m_stream(s) {
    m_levels = std::max<int>(1, (int)config.iterations.size());
    if (m_levels > NUIF_ICP_MAX_LEVELS) throw std::runtime_error("too many pyramid levels");
    const size_t n = (size_t)m_W * m_H;
    m_raw = DeviceBuffer(n * sizeof(UINT16));
    m_depth = DeviceBuffer(n * sizeof(float));
    m_colors = DeviceBuffer(n * 4);
    for (int l = 0; l < m_levels; ++l) {
        const size_t nl = (size_t)(m_W >> l) * (m_H >> l);
        m_dpyr.emplace_back(nl * sizeof(float));
        m_vpyr.emplace_back(nl * 3 * sizeof(float));
        m_dptrs.push_back(m_dpyr.back().as<float>());
        m_vptrs.push_back(m_vpyr.back().as<float>());
    }
    // OCL/NuiKinfuOpenCLFrame.cpp:165-189 GenerateGaussianBuffer (host table)
    const int r = (int)config.filter_radius;
    m_gaussian.resize(2 * r + 1);
    for (int i = 0; i < 2 * r + 1; ++i) {
        const int x = i - r;
        m_gaussian[i] = std::exp(-(float)(x * x) * config.sigma_space2_inv_half);
    }
}

NuiKinfuHIPFrame::~NuiKinfuHIPFrame() { nuif_stream_synchronize(m_stream); }

float* NuiKinfuHIPFrame::EnsureDepthBounds() {
    const size_t nb = nuif_tsdf_depth_bounds_bytes((int)m_W, (int)m_H);
    if (m_bounds.bytes() != nb) m_bounds = DeviceBuffer(nb);
    return m_bounds.as<float>();
}

// OCL/NuiKinfuOpenCLFrame.cpp:314-367 (+ the tracker's pyramid)
void NuiKinfuHIPFrame::UpdateVertexBuffers(UINT16* pDepths, UINT nNum, NuiKinfuCameraState* pCameraState) {
    if (!pDepths) return;
    if (nNum != m_W * m_H) throw std::runtime_error("UpdateVertexBuffers: frame size mismatch");
    check(nuif_memcpy_h2d(m_raw.get(), pDepths, (size_t)nNum * sizeof(UINT16), m_stream), "depth upload");
    UpdateVertexBuffersDevice(m_raw.as<UINT16>(), pCameraState);
}

void NuiKinfuHIPFrame::UpdateVertexBuffersDevice(const UINT16* dDepths, NuiKinfuCameraState* pCameraState) {
    auto* cam = dynamic_cast<NuiKinfuHIPCameraState*>(pCameraState);
    if (!dDepths || !cam) return;
    const NuiCLCameraParams& c = cam->GetCameraParams();
    check(nuif_frame_preprocess(dDepths, m_depth.as<float>(), m_dptrs.data(), m_vpyr.empty() ? nullptr : (float* const*)m_vptrs.data(),
                                m_levels, (int)m_W, (int)m_H, &c, m_gaussian.data(), (int)m_cfg.filter_radius,
                                m_cfg.sigma_depth2_inv_half, m_cfg.depth_threshold, m_scaleForTsdf ? 1 : 0, m_stream),
          "nuif_frame_preprocess");
    if (m_scaleForTsdf) {
        // depth bounds of the ray-length depth, used by the dense integration to skip
        // runs of voxels behind every observed surface
        EnsureDepthBounds();
        check(nuif_tsdf_depth_bounds(m_depth.as<float>(), (int)m_W, (int)m_H, m_bounds.as<float>(), nullptr, m_stream),
              "nuif_tsdf_depth_bounds");
    }
}

void NuiKinfuHIPFrame::UpdateColorBuffers(BGRQUAD* pColors, UINT nNum) {
    if (!pColors) return;
    if (nNum != m_W * m_H) throw std::runtime_error("UpdateColorBuffers: frame size mismatch");
    check(nuif_memcpy_h2d(m_colors.get(), pColors, (size_t)nNum * 4, m_stream), "color upload");
    m_hasColor = true;
}

void NuiKinfuHIPFrame::UpdateColorBuffers(ColorSpacePoint* pDepthToColor, UINT* pDepthDistortionLT, UINT nNum,
                                          const NuiColorImage& image) {
    if (!pDepthToColor || !image.GetBuffer()) return;
    if (nNum != m_W * m_H) throw std::runtime_error("UpdateColorBuffers: frame size mismatch");
    const size_t img = (size_t)image.GetWidth() * image.GetHeight() * 4;
    if (m_colorMap.bytes() < (size_t)nNum * 8) m_colorMap = DeviceBuffer((size_t)nNum * 8);
    if (pDepthDistortionLT && m_colorLut.bytes() < (size_t)nNum * 4) m_colorLut = DeviceBuffer((size_t)nNum * 4);
    if (m_colorImage.bytes() < img) m_colorImage = DeviceBuffer(img);
    check(nuif_memcpy_h2d(m_colorMap.get(), pDepthToColor, (size_t)nNum * 8, m_stream), "colour map upload");
    if (pDepthDistortionLT)
        check(nuif_memcpy_h2d(m_colorLut.get(), pDepthDistortionLT, (size_t)nNum * 4, m_stream), "lut upload");
    check(nuif_memcpy_h2d(m_colorImage.get(), image.GetBuffer(), img, m_stream), "colour image upload");
    check(nuif_map_depth_to_color(m_colorMap.as<float>(), pDepthDistortionLT ? m_colorLut.as<uint32_t>() : nullptr,
                                  nNum, m_colorImage.as<uint8_t>(), (int)image.GetWidth(), (int)image.GetHeight(),
                                  m_colors.as<uint8_t>(), m_stream),
          "nuif_map_depth_to_color");
    m_hasColor = true;
}

void NuiKinfuHIPFrame::UpdateColorBuffersDevice(const uint8_t* dBgra) {
    if (!dBgra) return;
    if (dBgra != m_colors.as<uint8_t>())
        check(nuif_memcpy_d2d(m_colors.get(), dBgra, (size_t)m_W * m_H * 4, m_stream), "color copy");
    m_hasColor = true;
}

// ------------------------------------------------------------ feedback frame

NuiKinfuHIPFeedbackFrame::NuiKinfuHIPFeedbackFrame(UINT nWidth, UINT nHeight, nuif_stream_t s)
    : m_W(nWidth), m_H(nHeight), m_stream(s) {
    const size_t n = (size_t)m_W * m_H;
    m_vmap = DeviceBuffer(n * 12);
    m_nmap = DeviceBuffer(n * 12);
    m_colors = DeviceBuffer(n * 4);
}

// OCL/NuiKinfuOpenCLFeedbackFrame.cpp:65-84: normals of the frame's vertices,
// then camera -> world for vertices and normals
void NuiKinfuHIPFeedbackFrame::UpdateBuffers(NuiKinfuFrame* frame, NuiKinfuCameraState* pCameraState) {
    auto* f = dynamic_cast<NuiKinfuHIPFrame*>(frame);
    auto* cam = dynamic_cast<NuiKinfuHIPCameraState*>(pCameraState);
    if (!f || !cam || f->GetWidth() != m_W || f->GetHeight() != m_H) return;
    check(nuif_estimate_normals_simple(f->GetVertexBuffer(), m_nmap.as<float>(), (int)m_W, (int)m_H,
                                       f->GetDepthThreshold(), m_stream),
          "nuif_estimate_normals_simple");
    check(nuif_transform_maps(f->GetVertexBuffer(), m_nmap.as<float>(), m_vmap.as<float>(), m_nmap.as<float>(),
                              (int)m_W, (int)m_H, cam->GetCameraTransformBuffer(), m_stream),
          "nuif_transform_maps");
    if (f->GetColorBuffer())
        check(nuif_memcpy_d2d(m_colors.get(), f->GetColorBuffer(), (size_t)m_W * m_H * 4, m_stream), "color copy");
}

// ------------------------------------------------------------ depth tracker

NuiKinfuHIPDepthTracker::NuiKinfuHIPDepthTracker(const NuiTrackerConfig& config, UINT nWidth, UINT nHeight,
                                                 nuif_stream_t s)
    : m_cfg(config), m_W(nWidth), m_H(nHeight), m_stream(s) {
    m_icfg.levels = (int)config.iterations.size();
    if (m_icfg.levels < 1 || m_icfg.levels > NUIF_ICP_MAX_LEVELS) throw std::runtime_error("bad ICP level count");
    for (int l = 0; l < m_icfg.levels; ++l) m_icfg.iterations[l] = (int)config.iterations[l].m_num;
    m_icfg.dist_threshold = config.dist_threshold;
    m_icfg.gate_threshold = 0.15f;
    m_state = DeviceBuffer(nuif_icp_state_bytes());
    m_partials = DeviceBuffer(nuif_icp_partials_bytes((int)nWidth, (int)nHeight));
}

void NuiKinfuHIPDepthTracker::ResetState(NuiKinfuHIPCameraState* cam) {
    check(nuif_icp_state_init(m_state.get(), cam->GetCameraTransformBuffer(), m_stream), "nuif_icp_state_init");
}

void NuiKinfuHIPDepthTracker::SetTranslation(NuiKinfuHIPCameraState* cam, const float t[3], const float offset[3]) {
    check(nuif_icp_set_translation(m_state.get(), cam->GetCameraTransformBuffer(), m_history, m_historyCap, t, offset,
                                   m_stream),
          "nuif_icp_set_translation");
}

void NuiKinfuHIPDepthTracker::RecordPose(NuiKinfuHIPCameraState* cam) {
    // a frame without tracking: the pose is unchanged and the frame always integrates
    check(nuif_icp_record(m_state.get(), cam->GetCameraTransformBuffer(), m_history, m_historyCap, m_stream),
          "nuif_icp_record");
}

// OCL/NuiKinfuOpenCLDepthTracker.cpp:103-138: pyramid (already built by the
// frame's preprocessing) + IterativeClosestPoint, enqueued without host sync.
bool NuiKinfuHIPDepthTracker::EstimatePose(NuiKinfuFrame* pFrame, NuiKinfuFeedbackFrame* pFeedbackFrame,
                                           NuiKinfuCameraState* pCameraState, Affine3f* hint) {
    auto* f = dynamic_cast<NuiKinfuHIPFrame*>(pFrame);
    auto* fb = dynamic_cast<NuiKinfuHIPFeedbackFrame*>(pFeedbackFrame);
    auto* cam = dynamic_cast<NuiKinfuHIPCameraState*>(pCameraState);
    if (!f || !fb || !cam) return false;
    if (hint) {  // start from the hint instead of the previous pose
        cam->UpdateCameraTransform(hint->rotation(), hint->translation());
    }
    const NuiCLCameraParams& c = cam->GetCameraParams();
    if (m_icfg.reduction == NUIF_ICP_REDUCE_JOINT) {
        preparePhotometric(f, fb);
        check(nuif_icp_estimate_joint(m_state.get(), m_partials.as<float>(), &c, f->GetVertexPyramid(),
                                      fb->GetVertexBuffer(), fb->GetNormalBuffer(), cam->GetCameraTransformBuffer(),
                                      &m_icfg, &m_photo, m_history, m_historyCap, m_stream),
              "nuif_icp_estimate_joint");
        return true;
    }
    check(nuif_icp_estimate(m_state.get(), m_partials.as<float>(), &c, f->GetVertexPyramid(), fb->GetVertexBuffer(),
                            fb->GetNormalBuffer(), cam->GetCameraTransformBuffer(), &m_icfg, m_history, m_historyCap,
                            m_stream),
          "nuif_icp_estimate");
    return true;
}

void NuiKinfuHIPDepthTracker::SetReduction(int mode) {
    if (mode != NUIF_ICP_REDUCE_REFERENCE && mode != NUIF_ICP_REDUCE_COMPENSATED && mode != NUIF_ICP_REDUCE_JOINT)
        throw std::runtime_error("unknown ICP reduction");
    m_icfg.reduction = mode;
    if (mode != NUIF_ICP_REDUCE_JOINT || m_refModel.get()) return;
    const UINT m = 1u << (m_icfg.levels - 1);
    if (m_W % m || m_H % m) throw std::runtime_error("the joint ICP needs an image size divisible by 2^(levels - 1)");
    for (int l = 0; l < m_icfg.levels; ++l) {
        m_intensity[l] = DeviceBuffer((size_t)(m_W >> l) * (m_H >> l) * sizeof(float));
        m_photo.intensity_pyr[l] = m_intensity[l].as<float>();
    }
    m_refModel = DeviceBuffer((size_t)m_W * m_H * 4 * sizeof(float));
    m_photo.ref_model = m_refModel.as<float>();
    m_prevColors = DeviceBuffer((size_t)m_W * m_H * 4);
    if (m_photo.weight == 0.f) SetPhotometric(0.1f, 0.2f, 0.002f);
}

void NuiKinfuHIPDepthTracker::SetPhotometric(float weight, float maxDifference, float minGradient) {
    if (!(weight >= 0.f) || !(maxDifference > 0.f) || !(minGradient >= 0.f))
        throw std::runtime_error("bad photometric parameters");
    m_photo.weight = weight;
    m_photo.max_difference = maxDifference;
    m_photo.min_gradient = minGradient;
}

// The joint order's inputs: intensities of the frame's colours at every level, and of
// the previous frame's colours with their gradients; then this frame's colours are kept
// for the next frame.
void NuiKinfuHIPDepthTracker::preparePhotometric(NuiKinfuHIPFrame* f, NuiKinfuHIPFeedbackFrame*) {
    if (!f->GetColorBuffer())
        throw std::runtime_error("the joint ICP needs the frame's colours (and a scene with a colour volume)");
    check(nuif_icp_photometric_inputs(f->GetColorBuffer(), m_prevColors.as<uint8_t>(), (int)m_W, (int)m_H,
                                      m_icfg.levels, &m_photo, m_stream),
          "nuif_icp_photometric_inputs");
    KeepColors(f);
}

void NuiKinfuHIPDepthTracker::KeepColors(NuiKinfuHIPFrame* f) {
    if (!m_prevColors.get()) return;
    if (!f->GetColorBuffer())
        throw std::runtime_error("the joint ICP needs the frame's colours (and a scene with a colour volume)");
    check(nuif_memcpy_d2d(m_prevColors.get(), f->GetColorBuffer(), (size_t)m_W * m_H * 4, m_stream), "colour copy");
}

void NuiKinfuHIPDepthTracker::FillIcpSession(NuiKinfuHIPFrame* f, NuiKinfuHIPFeedbackFrame* fb,
                                             NuiKinfuHIPCameraState* cam, nuif_icp_session* out) const {
    *out = nuif_icp_session{};
    out->state = m_state.get();
    out->partials = m_partials.get();
    out->cam = cam->GetCameraParams();
    for (int l = 0; l < NUIF_ICP_MAX_LEVELS; ++l) out->vmap_pyr[l] = l < m_icfg.levels ? f->GetVertexBuffer(l) : nullptr;
    out->ref_vmap = fb->GetVertexBuffer();
    out->ref_nmap = fb->GetNormalBuffer();
    out->pose = cam->GetCameraTransformBuffer();
    out->cfg = m_icfg;
    out->history = m_history;
    out->capacity = m_historyCap;
    if (m_icfg.reduction == NUIF_ICP_REDUCE_JOINT) out->photometric = m_photo;  // inputs made by the round
}

nuif_icp_result NuiKinfuHIPDepthTracker::ReadState() const {
    nuif_icp_result r;
    check(nuif_icp_state_read(m_state.get(), &r, m_stream), "nuif_icp_state_read");
    return r;
}

float NuiKinfuHIPDepthTracker::getError() const { return ReadState().error; }
int NuiKinfuHIPDepthTracker::getCount() const { return (int)ReadState().count; }

// ------------------------------------------------------------ dense scene

NuiKinfuHIPScene::NuiKinfuHIPScene(const NuiKinfuVolumeConfig& config, int storage, nuif_stream_t s)
    : m_config(config), m_storage(storage), m_stream(s) {
    check(nuif_tsdf_params(config.resolution[0], config.resolution[1], config.resolution[2], config.dimensions[0],
                           config.dimensions[1], config.dimensions[2], config.tranc_dist, &m_params),
          "nuif_tsdf_params");
    const size_t n = (size_t)config.resolution[0] * config.resolution[1] * config.resolution[2];
    const size_t vb = nuif_tsdf_voxel_bytes(storage);
    if (!vb) throw std::runtime_error("unknown TSDF storage");
    m_volume = DeviceBuffer(n * vb);
    if (config.bHas_color_volume) m_color = DeviceBuffer(n * 4);
    m_counter = DeviceBuffer(sizeof(unsigned long long));
    m_surface = DeviceBuffer(nuif_tsdf_surface_blocks_bytes(&m_params));
    m_walk = DeviceBuffer(nuif_tsdf_walk_states_bytes(&m_params));
    reset();
}

unsigned long long NuiKinfuHIPScene::UpdatedVoxels() const {
    unsigned long long h = 0;
    check(nuif_memcpy_d2h(&h, m_counter.get(), sizeof(h), m_stream), "counter read");
    check(nuif_stream_synchronize(m_stream), "counter read");
    return h;
}

float NuiKinfuHIPScene::getVoxelLeafSize() const {
    return std::max(m_params.cell_size[0], std::max(m_params.cell_size[1], m_params.cell_size[2]));
}

// OCL/NuiKinfuOpenCLScene.cpp:93-137
void NuiKinfuHIPScene::reset() {
    check(nuif_tsdf_reset(m_volume.get(), m_color.as<uint8_t>(), &m_params, m_storage, m_stream), "nuif_tsdf_reset");
    check(nuif_tsdf_surface_blocks_clear(m_surface.as<uint8_t>(), &m_params, m_stream),
          "nuif_tsdf_surface_blocks_clear");
    setDirty();
}

// OCL/NuiKinfuOpenCLScene.cpp:199-336: scaleDepths + integrateTsdfVolumeKernel
bool NuiKinfuHIPScene::integrateVolume(NuiKinfuFrame* pFrame, NuiKinfuCameraState* pCameraState) {
    auto* f = dynamic_cast<NuiKinfuHIPFrame*>(pFrame);
    auto* cam = dynamic_cast<NuiKinfuHIPCameraState*>(pCameraState);
    if (!f || !cam) return false;
    const NuiCLCameraParams& c = cam->GetCameraParams();
    if (!f->IsDepthScaled())
        check(nuif_scale_depths(f->GetDepthBuffer(), (int)f->GetWidth(), (int)f->GetHeight(), &c, m_stream),
              "nuif_scale_depths");
    const uint8_t* colors = m_config.bHas_color_volume ? f->GetColorBuffer() : nullptr;
    // the frame's depth bounds exist when it was preprocessed for this scene
    const float* bounds = f->IsDepthScaled() ? f->GetDepthBounds() : nullptr;
    check(nuif_tsdf_integrate_wrapped(f->GetDepthBuffer(), colors, m_volume.get(),
                                      colors ? m_color.as<uint8_t>() : nullptr, &m_params, &c,
                                      cam->GetCameraTransformBuffer(), m_storage, m_config.max_color_weight, m_wrap,
                                      m_gateIntegrate, m_counter.as<unsigned long long>(), bounds,
                                      m_surface.as<uint8_t>(), m_walk.as<float>(), m_stream),
          "nuif_tsdf_integrate");
    setDirty();
    return true;
}

bool NuiKinfuHIPScene::FillRoundScene(nuif_round_session* out) {
    const int Nx = (int)m_params.resolution[0], Ny = (int)m_params.resolution[1], Nz = (int)m_params.resolution[2];
    if (Nx % 8 || Ny % 8 || Nz % 32 || m_wrap[0] || m_wrap[1] || m_wrap[2]) return false;
    out->volume = m_volume.get();
    out->color_volume = m_color.as<uint8_t>();
    out->bgra = nullptr;
    out->params = m_params;
    out->storage = m_storage;
    out->max_color_weight = m_config.max_color_weight;
    out->updated = m_counter.as<unsigned long long>();
    out->surface_blocks = m_surface.as<uint8_t>();
    out->walk_states = m_walk.as<float>();
    setDirty();
    return true;
}

// OCL/NuiKinfuOpenCLScene.cpp:339-434
void NuiKinfuHIPScene::raycastRender(NuiKinfuFeedbackFrame* pFeedbackFrame, NuiKinfuCameraState* pCameraState) {
    auto* fb = dynamic_cast<NuiKinfuHIPFeedbackFrame*>(pFeedbackFrame);
    auto* cam = dynamic_cast<NuiKinfuHIPCameraState*>(pCameraState);
    if (!fb || !cam) return;
    const NuiCLCameraParams& c = cam->GetCameraParams();
    check(nuif_tsdf_raycast_wrapped(m_volume.get(), m_color.as<uint8_t>(), &m_params, &c,
                                    cam->GetCameraTransformBuffer(), fb->GetVertexBuffer(), fb->GetNormalBuffer(),
                                    m_color.get() ? fb->GetColorBuffer() : nullptr, m_storage, m_wrap, m_gateRaycast,
                                    m_surface.as<uint8_t>(), m_stream),
          "nuif_tsdf_raycast");
}

long long NuiKinfuHIPScene::RaycastFootprint(NuiKinfuHIPCameraState* cam) const {
    if (m_wrap[0] || m_wrap[1] || m_wrap[2]) throw std::runtime_error("raycast footprint: wrapped volume");
    const NuiCLCameraParams& c = cam->GetCameraParams();
    const size_t n = (size_t)c.depthImageWidth * c.depthImageHeight;
    DeviceBuffer ws(nuif_tsdf_footprint_bytes(&m_params)), v(n * 12), nm(n * 12);
    const long long u = nuif_tsdf_raycast_footprint(m_volume.get(), &m_params, &c, cam->GetCameraTransformBuffer(),
                                                    m_storage, m_surface.as<uint8_t>(), v.as<float>(), nm.as<float>(),
                                                    ws.get(), m_stream);
    if (u < 0) check((int)u, "nuif_tsdf_raycast_footprint");
    return u;
}

// ------------------------------------------------------------ shifting volume

NuiKinfuHIPShiftScene::NuiKinfuHIPShiftScene(const NuiKinfuVolumeConfig& config, int storage, nuif_stream_t s)
    : NuiKinfuHIPScene(config, storage, s) {
    m_scratch = DeviceBuffer(nuif_points_dense_scratch_bytes(&m_params));
}

// OCL/NuiKinfuOpenCLShiftScene.cpp:35-41
void NuiKinfuHIPShiftScene::reset() {
    for (int k = 0; k < 3; ++k) m_offset[k] = 0;
    updateWrap();
    m_cachePos.clear();
    m_cacheCol.clear();
    NuiKinfuHIPScene::reset();
}

// :44-56 getVoxelWrap
void NuiKinfuHIPShiftScene::updateWrap() {
    for (int k = 0; k < 3; ++k) {
        const int n = m_config.resolution[k];
        m_wrap[k] = m_offset[k] < 0 ? n - ((-m_offset[k]) % n) : m_offset[k];
    }
}

void NuiKinfuHIPShiftScene::OffsetSize(float out[3]) const {
    for (int k = 0; k < 3; ++k) out[k] = (float)m_offset[k] * m_params.cell_size[k];
}

// :247-287 fetchAndClear{X,Y,Z}: fetchSlice (:119-245) then clearSlice (:72-117)
void NuiKinfuHIPShiftScene::fetchAndClear(int axis, int trans) {
    const int n = m_config.resolution[axis];
    int o = m_offset[axis];
    if (trans < 0) o += trans;
    o = o >= 0 ? o % n : n - (-o) % n;
    int32_t wrap[3] = {0, 0, 0};
    int32_t range[3] = {m_config.resolution[0], m_config.resolution[1], m_config.resolution[2]};
    wrap[axis] = o;
    range[axis] = std::abs(trans);
    DeviceBuffer total(sizeof(uint32_t));
    const uint8_t* cv = m_color.as<uint8_t>();
    check(nuif_points_dense_slice(m_volume.get(), cv, &m_params, m_storage, wrap, m_offset, range, nullptr, nullptr, 0,
                                  total.as<uint32_t>(), m_scratch.get(), m_stream),
          "nuif_points_dense_slice");
    uint32_t cnt = 0;
    check(nuif_memcpy_d2h(&cnt, total.get(), sizeof(cnt), m_stream), "slab count");
    check(nuif_stream_synchronize(m_stream), "slab count");
    if (cnt) {
        DeviceBuffer pts((size_t)cnt * 3 * sizeof(float)), cols((size_t)cnt * 4 * sizeof(float));
        check(nuif_points_dense_slice(m_volume.get(), cv, &m_params, m_storage, wrap, m_offset, range,
                                      pts.as<float>(), cols.as<float>(), cnt, total.as<uint32_t>(), m_scratch.get(),
                                      m_stream),
              "nuif_points_dense_slice");
        const size_t at = m_cachePos.size() / 3;
        m_cachePos.resize((at + cnt) * 3);
        m_cacheCol.resize((at + cnt) * 4);
        check(nuif_memcpy_d2h(m_cachePos.data() + at * 3, pts.get(), pts.bytes(), m_stream), "slab points");
        check(nuif_memcpy_d2h(m_cacheCol.data() + at * 4, cols.get(), cols.bytes(), m_stream), "slab colours");
        check(nuif_stream_synchronize(m_stream), "slab points");
    }
    check(nuif_tsdf_reset_slice(m_volume.get(), m_color.as<uint8_t>(), &m_params, m_storage, wrap, range, m_stream),
          "nuif_tsdf_reset_slice");
    m_offset[axis] += trans;
}

// :289-351 shiftVolume on the local translation t (float, as the reference computes it)
void NuiKinfuHIPShiftScene::shiftVolume(float t[3]) {
    float sh[3];
    int vt[3];
    for (int k = 0; k < 3; ++k) {
        sh[k] = t[k] - m_config.translateBasis[k];
        const int f = (int)std::floor(sh[k] / m_params.cell_size[k]);
        vt[k] = sh[k] < 0 ? std::max(-m_config.voxel_shift, f) : std::min(m_config.voxel_shift, f);
    }
    for (int k = 0; k < 3; ++k)
        if (std::abs(vt[k]) >= m_config.voxel_shift) {
            fetchAndClear(k, vt[k]);
            sh[k] -= (float)vt[k] * m_params.cell_size[k];
        }
    for (int k = 0; k < 3; ++k) t[k] = sh[k] + m_config.translateBasis[k];
}

// :353-364 integrateVolume: shift on the tracked pose, then the dense integration
bool NuiKinfuHIPShiftScene::integrateVolume(NuiKinfuFrame* pFrame, NuiKinfuCameraState* pCameraState) {
    auto* cam = dynamic_cast<NuiKinfuHIPCameraState*>(pCameraState);
    if (!cam) return false;
    if (m_gateIntegrate) {  // the frame integrates only when the device gate says so
        int32_t gate = 0;
        check(nuif_memcpy_d2h(&gate, m_gateIntegrate, sizeof(gate), m_stream), "gate read");
        check(nuif_stream_synchronize(m_stream), "gate read");
        if (!gate) return true;
    }
    cam->SyncFromDevice();
    const NuiCameraPos& pos = cam->GetCameraPos();
    float t[3] = {pos.getLocalTranslation()[0], pos.getLocalTranslation()[1], pos.getLocalTranslation()[2]};
    const int before[3] = {m_offset[0], m_offset[1], m_offset[2]};
    const float t0[3] = {t[0], t[1], t[2]};
    shiftVolume(t);
    updateWrap();
    float off[3];
    OffsetSize(off);
    const bool moved = std::memcmp(before, m_offset, sizeof(before)) != 0 || std::memcmp(t0, t, sizeof(t0)) != 0;
    if (moved && m_tracker) m_tracker->SetTranslation(cam, t, off);
    cam->NuiKinfuCameraState::UpdateCameraTransform(pos.getRotation(), Vector3f(t[0], t[1], t[2]),
                                                    Vector3f(off[0], off[1], off[2]));
    return NuiKinfuHIPScene::integrateVolume(pFrame, pCameraState);
}

// :366-450 Volume2CLVertices: fetchShiftVolumeKernel over the whole volume, wrap 0,
// placed by the voxel offset
bool NuiKinfuHIPShiftScene::Volume2CLVertices(NuiKinfuMesh* pData) {
    if (!pData || !m_dirty) return false;
    clearDirty();
    const int32_t wrap[3] = {0, 0, 0};
    const int32_t range[3] = {m_config.resolution[0], m_config.resolution[1], m_config.resolution[2]};
    DeviceBuffer total(sizeof(uint32_t));
    uint32_t n = 0;
    check(nuif_points_dense_slice(m_volume.get(), m_color.as<uint8_t>(), &m_params, m_storage, wrap, m_offset, range,
                                  nullptr, nullptr, 0, total.as<uint32_t>(), m_scratch.get(), m_stream),
          "nuif_points_dense_slice");
    check(nuif_memcpy_d2h(&n, total.get(), sizeof(n), m_stream), "point count");
    check(nuif_stream_synchronize(m_stream), "point count");
    pData->positions.assign((size_t)n * 3, 0.f);
    pData->colors.assign((size_t)n * 4, 0.f);
    pData->indices.clear();
    if (n == 0) return false;
    DeviceBuffer pts((size_t)n * 3 * sizeof(float)), cols((size_t)n * 4 * sizeof(float));
    check(nuif_points_dense_slice(m_volume.get(), m_color.as<uint8_t>(), &m_params, m_storage, wrap, m_offset, range,
                                  pts.as<float>(), cols.as<float>(), n, total.as<uint32_t>(), m_scratch.get(),
                                  m_stream),
          "nuif_points_dense_slice");
    check(nuif_memcpy_d2h(pData->positions.data(), pts.get(), pts.bytes(), m_stream), "points");
    check(nuif_memcpy_d2h(pData->colors.data(), cols.get(), cols.bytes(), m_stream), "points");
    check(nuif_stream_synchronize(m_stream), "points");
    return true;
}

}  // namespace nuif_kinfu

namespace NuiKinfuEngine {

using namespace nuif_kinfu;

void NuiKinfuTrackingFactory::BuildTrackingEngine(NuiKinfuTracker** pTracker, NuiKinfuFrame** pFrame,
                                                  NuiKinfuFeedbackFrame** pFeedbackFrame,
                                                  NuiKinfuCameraState** pCameraState,
                                                  const NuiTrackerConfig& trackerConfig, bool, unsigned int nWidth,
                                                  unsigned int nHeight, nuif_stream_t stream) {
    delete *pFrame;
    delete *pTracker;
    delete *pCameraState;
    delete *pFeedbackFrame;
    *pFrame = new NuiKinfuHIPFrame(trackerConfig, nWidth, nHeight, stream);
    *pFeedbackFrame = new NuiKinfuHIPFeedbackFrame(nWidth, nHeight, stream);
    *pTracker = new NuiKinfuHIPDepthTracker(trackerConfig, nWidth, nHeight, stream);
    *pCameraState = new NuiKinfuHIPCameraState(stream);
}

NuiKinfuTrackingEngine::NuiKinfuTrackingEngine(nuif_stream_t stream) : m_stream(stream) {
    if (!m_stream) {
        check(nuif_stream_create(&m_stream), "nuif_stream_create");
        m_ownStream = true;
    }
}

NuiKinfuTrackingEngine::~NuiKinfuTrackingEngine() {
    nuif_stream_synchronize(m_stream);
    if (m_graph) nuif_graph_destroy(m_graph);
    for (nuif_event_t e : m_events) nuif_event_destroy(e);
    delete m_pFrame;
    delete m_pFeedbackFrame;
    delete m_pTracker;
    delete m_pCameraState;
    m_history = DeviceBuffer();
    if (m_ownStream) nuif_stream_destroy(m_stream);
}

void NuiKinfuTrackingEngine::initialize(const NuiTrackerConfig& trackerConfig, bool bAcceleratedFeedback,
                                        UINT nWidth, UINT nHeight) {
    NuiKinfuTracker* t = m_pTracker;
    NuiKinfuFrame* f = m_pFrame;
    NuiKinfuFeedbackFrame* fb = m_pFeedbackFrame;
    NuiKinfuCameraState* c = m_pCameraState;
    NuiKinfuTrackingFactory::BuildTrackingEngine(&t, &f, &fb, &c, trackerConfig, bAcceleratedFeedback, nWidth, nHeight,
                                                 m_stream);
    m_pTracker = static_cast<NuiKinfuHIPDepthTracker*>(t);
    m_pFrame = static_cast<NuiKinfuHIPFrame*>(f);
    m_pFeedbackFrame = static_cast<NuiKinfuHIPFeedbackFrame*>(fb);
    m_pCameraState = static_cast<NuiKinfuHIPCameraState*>(c);
    if (!m_history.get()) m_history = DeviceBuffer((size_t)m_historyCap * NUIF_POSE_ROW * sizeof(float));
    m_pTracker->SetHistory(m_history.as<float>(), m_historyCap);
    m_pTracker->SetIntegrationThreshold(m_integrationThreshold);
    reset();
}

void NuiKinfuTrackingEngine::log(const std::string& fileName) const {
    if (m_pTracker) m_pTracker->log(fileName);
}

// VO/NuiKinfuTrackingEngine.cpp:46-58
void NuiKinfuTrackingEngine::reset() {
    if (m_frames > m_drained) check(nuif_stream_synchronize(m_stream), "reset");
    m_frames = 0;
    m_rows.clear();
    m_drained = 0;
    m_anyFailed = false;
    m_poses.clear();
    if (m_graph) {
        nuif_graph_destroy(m_graph);
        m_graph = nullptr;
        m_graphScene = nullptr;
    }
    if (!m_pCameraState) return;
    m_pCameraState->UpdateCameraTransform(Matrix3frm::Identity(), m_translateBasis);
    m_pTracker->ResetState(m_pCameraState);
}

void NuiKinfuTrackingEngine::setIntegrationMetricThreshold(float threshold) {
    m_integrationThreshold = threshold;
    if (m_pTracker) m_pTracker->SetIntegrationThreshold(threshold);
    if (m_graph) {  // the threshold is baked into the captured frame
        nuif_graph_destroy(m_graph);
        m_graph = nullptr;
        m_graphScene = nullptr;
    }
}

bool NuiKinfuTrackingEngine::RunTracking(INT64 timeStamp, UINT16* pDepths, BGRQUAD* pColors, UINT nPointNum,
                                         NuiKinfuScene* pScene, const NuiCameraParams& cameraParams) {
    if (!m_pTracker || !m_pFrame || !m_pCameraState) return true;
    if (!pDepths || 0 == nPointNum) return true;  // VO/NuiKinfuTrackingEngine.cpp:79-80
    if (nPointNum != m_pFrame->GetWidth() * m_pFrame->GetHeight())
        throw std::runtime_error("RunTracking: frame size mismatch");
    check(nuif_memcpy_h2d(m_pFrame->GetRawBuffer(), pDepths, (size_t)nPointNum * sizeof(UINT16), m_stream),
          "depth upload");
    const uint8_t* colors = nullptr;
    if (pColors && pScene && pScene->hasColorData()) {
        m_pFrame->UpdateColorBuffers(pColors, nPointNum);
        colors = m_pFrame->GetColorBuffer();
    }
    return RunTrackingDevice(timeStamp, m_pFrame->GetRawBuffer(), colors, pScene, cameraParams, true);
}

void NuiKinfuTrackingEngine::recordStage(int k) {
    if (!m_stageTiming) return;
    const size_t idx = m_eventsUsed + (size_t)k;
    while (m_events.size() <= idx) {
        nuif_event_t e = nullptr;
        check(nuif_event_create(&e), "nuif_event_create");
        m_events.push_back(e);
    }
    check(nuif_event_record(m_events[idx], m_stream), "nuif_event_record");
    if (k == 4) m_eventsUsed += 5;
}

void NuiKinfuTrackingEngine::setStageTiming(bool on) {
    m_stageTiming = on;
    m_eventsUsed = 0;
}

int NuiKinfuTrackingEngine::takeStageTimes(double ms[4]) {
    check(nuif_stream_synchronize(m_stream), "stage timing");
    for (int k = 0; k < 4; ++k) ms[k] = 0.0;
    const int frames = (int)(m_eventsUsed / 5);
    for (int f = 0; f < frames; ++f)
        for (int k = 0; k < 4; ++k) {
            float t = 0.f;
            check(nuif_event_elapsed_ms(m_events[5 * f + k], m_events[5 * f + k + 1], &t), "elapsed");
            ms[k] += t;
        }
    m_eventsUsed = 0;
    return frames;
}

void NuiKinfuTrackingEngine::enqueueTrackedFrame(NuiKinfuHIPSceneBase* scene) {
    recordStage(0);
    m_pFrame->UpdateVertexBuffersDevice(m_pFrame->GetRawBuffer(), m_pCameraState);
    recordStage(1);
    m_pTracker->EstimatePose(m_pFrame, m_pFeedbackFrame, m_pCameraState, nullptr);
    recordStage(2);
    if (scene) {
        // VO/NuiKinfuTrackingEngine.cpp:109-131: the gate decision and the tracking
        // outcome are device flags, so integrate/raycast are enqueued unconditionally
        scene->SetGates(m_pTracker->GateFlag(), m_pTracker->OkFlag());
        scene->integrateVolume(m_pFrame, m_pCameraState);
        recordStage(3);
        scene->raycastRender(m_pFeedbackFrame, m_pCameraState);
        scene->SetGates(nullptr, nullptr);
    } else {
        recordStage(3);
        m_pFeedbackFrame->UpdateBuffers(m_pFrame, m_pCameraState);  // ICP-only mode
    }
    recordStage(4);
}

bool NuiKinfuTrackingEngine::RunTrackingDevice(INT64, const UINT16* dDepths, const uint8_t* dBgra,
                                               NuiKinfuScene* pScene, const NuiCameraParams& cameraParams, bool sync) {
    if (!m_pTracker || !m_pFrame || !m_pCameraState) return true;
    if (!dDepths) return true;
    auto* scene = dynamic_cast<NuiKinfuHIPSceneBase*>(pScene);
    if (pScene && !scene) throw std::runtime_error("RunTracking: scene is not a GPU scene");
    m_pCameraState->UpdateCameraParams(cameraParams, m_pFrame->GetWidth(), m_pFrame->GetHeight());
    m_pFrame->SetScaleDepthForTsdf(scene != nullptr && scene->NeedsRayLengthDepth());
    const size_t n = (size_t)m_pFrame->GetWidth() * m_pFrame->GetHeight();
    // inputs go to the frame's own buffers so that a captured graph can replay
    if (dDepths != m_pFrame->GetRawBuffer())
        check(nuif_memcpy_d2d(m_pFrame->GetRawBuffer(), dDepths, n * sizeof(UINT16), m_stream), "depth copy");
    if (dBgra && scene && scene->hasColorData()) m_pFrame->UpdateColorBuffersDevice(dBgra);

    beforeFrame();
    if (m_frames == 0) {
        // VO/NuiKinfuTrackingEngine.cpp:93-104: first frame: feedback maps from the
        // frame itself, integrate, no raycast
        m_pFrame->UpdateVertexBuffersDevice(m_pFrame->GetRawBuffer(), m_pCameraState);
        m_pFeedbackFrame->UpdateBuffers(m_pFrame, m_pCameraState);
        if (scene) {
            scene->SetGates(nullptr, nullptr);
            scene->integrateVolume(m_pFrame, m_pCameraState);
        }
        if (m_pTracker->Reduction() == NUIF_ICP_REDUCE_JOINT) m_pTracker->KeepColors(m_pFrame);
        m_pTracker->RecordPose(m_pCameraState);
    } else if (m_useGraph && !m_stageTiming && !(scene && scene->RequiresHostWork())) {
        if (!m_graph || m_graphScene != scene) {
            if (m_graph) nuif_graph_destroy(m_graph);
            m_graph = nullptr;
            check(nuif_graph_begin(m_stream), "nuif_graph_begin");
            enqueueTrackedFrame(scene);
            check(nuif_graph_end(m_stream, &m_graph), "nuif_graph_end");
            m_graphScene = scene;
        }
        check(nuif_graph_launch(m_graph, m_stream), "nuif_graph_launch");
    } else {
        enqueueTrackedFrame(scene);
    }
    m_frames++;
    if (!sync) return true;
    check(nuif_stream_synchronize(m_stream), "frame");
    const nuif_icp_result r = m_pTracker->ReadState();
    m_pCameraState->SyncFromDevice();
    return r.status == 0;
}

bool NuiKinfuTrackingEngine::PrepareRoundFrame(const UINT16* dDepths, const uint8_t* dBgra, NuiKinfuScene* pScene,
                                               const NuiCameraParams& cameraParams, nuif_round_session* out) {
    if (!m_pTracker || !m_pFrame || !m_pCameraState || !dDepths || m_frames == 0 || m_stageTiming) return false;
    auto* scene = dynamic_cast<NuiKinfuHIPScene*>(pScene);
    if (!scene || dynamic_cast<NuiKinfuHIPShiftScene*>(pScene)) return false;
    *out = nuif_round_session{};
    if (!scene->FillRoundScene(out)) return false;
    beforeFrame();
    m_pCameraState->UpdateCameraParams(cameraParams, m_pFrame->GetWidth(), m_pFrame->GetHeight());
    m_pFrame->SetScaleDepthForTsdf(true);
    out->depth = dDepths;
    out->depth_out = m_pFrame->GetDepthBuffer();
    for (int l = 0; l < NUIF_ICP_MAX_LEVELS; ++l)
        out->depth_pyr[l] = l < m_pFrame->GetLevels() ? m_pFrame->GetDepthPyramidLevel(l) : nullptr;
    out->bounds = m_pFrame->EnsureDepthBounds();
    out->gaussian = m_pFrame->GetGaussian();
    out->radius = (int)m_pFrame->GetConfig().filter_radius;
    out->sigma_depth2_inv_half = m_pFrame->GetConfig().sigma_depth2_inv_half;
    out->depth_threshold = m_pFrame->GetConfig().depth_threshold;
    m_pTracker->FillIcpSession(m_pFrame, m_pFeedbackFrame, m_pCameraState, &out->icp);
    out->vmap = m_pFeedbackFrame->GetVertexBuffer();
    out->nmap = m_pFeedbackFrame->GetNormalBuffer();
    out->cmap = out->color_volume ? m_pFeedbackFrame->GetColorBuffer() : nullptr;
    out->bgra = out->color_volume ? dBgra : nullptr;
    if (m_pTracker->Reduction() == NUIF_ICP_REDUCE_JOINT) {
        if (!out->bgra) throw std::runtime_error("the joint ICP needs the frame's colours");
        out->prev_bgra = m_pTracker->PreviousColors();
    }
    m_frames++;
    return true;
}

bool NuiKinfuTrackingEngine::synchronize() {
    check(nuif_stream_synchronize(m_stream), "synchronize");
    if (m_pCameraState) m_pCameraState->SyncFromDevice();
    drainHistory();
    return !m_anyFailed;
}

// The next frame writes ring row m_frames % capacity: rows not yet copied to the host
// are drained first when the ring is full.
void NuiKinfuTrackingEngine::beforeFrame() {
    if (m_frames - m_drained >= m_historyCap) drainHistory();
}

void NuiKinfuTrackingEngine::drainHistory() const {
    if (m_drained >= m_frames || !m_history.get()) return;
    check(nuif_stream_synchronize(m_stream), "history");
    const size_t row = NUIF_POSE_ROW;
    const size_t old = m_rows.size();
    m_rows.resize(old + (size_t)(m_frames - m_drained) * row);
    for (int f = m_drained; f < m_frames;) {
        const int r = f % m_historyCap;
        const int k = std::min(m_frames - f, m_historyCap - r);  // contiguous rows up to the ring's end
        check(nuif_memcpy_d2h(&m_rows[old + (size_t)(f - m_drained) * row], m_history.as<float>() + (size_t)r * row,
                              (size_t)k * row * sizeof(float), m_stream),
              "history");
        f += k;
    }
    check(nuif_stream_synchronize(m_stream), "history");
    for (size_t i = old; i < m_rows.size(); i += row) m_anyFailed = m_anyFailed || m_rows[i + 12] == 0.f;
    m_drained = m_frames;
}

void NuiKinfuTrackingEngine::setPoseHistoryCapacity(int capacity) {
    if (capacity < 1) throw std::invalid_argument("pose history capacity must be at least 1");
    if (m_frames > m_drained) {
        check(nuif_stream_synchronize(m_stream), "history");
        drainHistory();
    }
    m_historyCap = capacity;
    if (m_history.get()) {
        m_history = DeviceBuffer((size_t)m_historyCap * NUIF_POSE_ROW * sizeof(float));
        if (m_pTracker) m_pTracker->SetHistory(m_history.as<float>(), m_historyCap);
        if (m_graph) {  // the captured frame holds the old ring's address
            nuif_graph_destroy(m_graph);
            m_graph = nullptr;
            m_graphScene = nullptr;
        }
    }
}

const std::vector<float>& NuiKinfuTrackingEngine::poseRows() const {
    drainHistory();
    return m_rows;
}

const NuiCameraPos& NuiKinfuTrackingEngine::getCameraPose(int time) const {
    drainHistory();
    const size_t row = NUIF_POSE_ROW;
    for (size_t i = m_poses.size(); i < m_rows.size() / row; ++i) {
        NuiCameraPos p;
        Matrix3frm R;
        std::memcpy(R.data(), &m_rows[i * row], 9 * sizeof(float));
        p.setRotation(R);
        const float* t = &m_rows[i * row + 9];
        p.setLocalTranslation(Vector3f(t[0], t[1], t[2]));
        m_poses.push_back(p);
    }
    if (time == -1 || time >= (int)m_poses.size()) {
        if (m_poses.empty()) return m_pCameraState->GetCameraPos();
        return m_poses.back();
    }
    return m_poses[time];
}

float NuiKinfuTrackingEngine::getTrackerError() const { return m_pTracker ? m_pTracker->getError() : 0.0f; }
int NuiKinfuTrackingEngine::getTrackerCount() const { return m_pTracker ? m_pTracker->getCount() : 0; }

}  // namespace NuiKinfuEngine
