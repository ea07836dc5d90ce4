// Photometric tracker (OCL/NuiKinfuOpenCLColorTracker.cpp) on the nuif_color_* entry
// points: the frame's colour pyramid and gradients, then the level loop on the device.
#include <cstring>
#include <stdexcept>

#include "NuiKinfuHIP.h"

namespace nuif_kinfu {

// OCL/NuiKinfuOpenCLColorTracker.cpp:18-92 (constructor, AcquireBuffers)
NuiKinfuHIPColorTracker::NuiKinfuHIPColorTracker(const NuiTrackerConfig& config, UINT nWidth, UINT nHeight,
                                                 bool corrected, nuif_stream_t s)
    : m_cfg(config), m_W(nWidth), m_H(nHeight), m_corrected(corrected), m_stream(s) {
    const size_t levels = m_cfg.iterations.size();
    if (levels < 1 || levels > NUIF_ICP_MAX_LEVELS) throw std::runtime_error("colour tracker: bad level count");
    for (size_t l = 0; l < levels; ++l) {
        const size_t bytes = (size_t)(m_W >> l) * (m_H >> l) * 4;
        m_colors.emplace_back(bytes);
        m_gx.emplace_back(bytes);
        m_gy.emplace_back(bytes);
    }
    m_state = DeviceBuffer(nuif_color_icp_state_bytes());
    m_partials = DeviceBuffer(nuif_color_icp_partials_bytes((int)m_W, (int)m_H));
    m_pose = DeviceBuffer(sizeof(NuiCLRigidTransform));
}

bool NuiKinfuHIPColorTracker::EstimatePose(NuiKinfuFrame* pFrame, NuiKinfuFeedbackFrame* pFeedbackFrame,
                                           NuiKinfuCameraState* pCameraState, Affine3f* hint) {
    auto* f = dynamic_cast<NuiKinfuHIPFrame*>(pFrame);
    auto* fb = dynamic_cast<NuiKinfuHIPFeedbackFrame*>(pFeedbackFrame);
    auto* cam = dynamic_cast<NuiKinfuHIPCameraState*>(pCameraState);
    if (!f || !fb || !cam || !f->GetColorBuffer()) return false;
    const int levels = (int)m_cfg.iterations.size();
    std::vector<uint8_t*> c(levels), gx(levels), gy(levels);
    for (int l = 0; l < levels; ++l) {
        c[l] = m_colors[l].as<uint8_t>();
        gx[l] = m_gx[l].as<uint8_t>();
        gy[l] = m_gy[l].as<uint8_t>();
    }
    // SubSampleColors + GradientBuffers (:125-242)
    check(nuif_color_pyramid(f->GetColorBuffer(), c.data(), gx.data(), gy.data(), (int)m_W, (int)m_H, levels,
                             m_corrected ? 1 : 0, m_stream),
          "nuif_color_pyramid");
    // the start pose: the hint or the camera's (:374-387)
    if (hint) {
        NuiCLRigidTransform h{};
        const Matrix3frm inv = hint->rotation().inverse();
        std::memcpy(h.R, hint->rotation().data(), sizeof(h.R));
        std::memcpy(h.R_inv, inv.data(), sizeof(h.R_inv));
        std::memcpy(h.t, hint->translation().data(), sizeof(h.t));
        check(nuif_memcpy_h2d(m_pose.get(), &h, sizeof(h), m_stream), "pose upload");
        check(nuif_stream_synchronize(m_stream), "pose upload");
    } else {
        check(nuif_memcpy_d2d(m_pose.get(), cam->GetCameraTransformBuffer(), sizeof(NuiCLRigidTransform), m_stream),
              "pose copy");
    }
    std::vector<int> its(levels), ro(levels);
    for (int l = 0; l < levels; ++l) {
        its[l] = (int)m_cfg.iterations[l].m_num;
        ro[l] = m_cfg.iterations[l].m_type == NuiTrackerConfig::eTracker_Iteration_Rotation ? 1 : 0;
    }
    std::vector<const uint8_t*> cc(c.begin(), c.end()), cgx(gx.begin(), gx.end()), cgy(gy.begin(), gy.end());
    check(nuif_color_icp_estimate(m_state.get(), m_partials.as<float>(), &cam->GetCameraParams(), levels, its.data(),
                                  ro.data(), m_cfg.dist_threshold, fb->GetVertexBuffer(), fb->GetColorBuffer(),
                                  cc.data(), cgx.data(), cgy.data(), m_pose.as<NuiCLRigidTransform>(),
                                  m_corrected ? 1 : 0, m_stream),
          "nuif_color_icp_estimate");
    check(nuif_color_icp_read(m_state.get(), &m_last, m_stream), "nuif_color_icp_read");
    m_error = m_last.error;
    m_count = m_last.count;
    if (!m_last.ok) return false;
    // UpdateCameraTransform(Rcurr, tcurr) (:576)
    check(nuif_memcpy_d2d(cam->GetCameraTransformBuffer(), m_pose.get(), sizeof(NuiCLRigidTransform), m_stream),
          "pose copy");
    cam->SyncFromDevice();
    return true;
}

}  // namespace nuif_kinfu
