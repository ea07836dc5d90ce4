// Voxel-hashing scene (OCL/NuiKinfuOpenCLHashScene.cpp) on the nuif_hash_* entry
// points, and the meshing of both GPU scenes.
#include <algorithm>
#include <stdexcept>

#include "NuiKinfuHIP.h"

namespace nuif_kinfu {

nuif_hash_params NuiKinfuHIPHashScene::ParamsOf(const NuiHashingSDFConfig& c, int storage, bool color) {
    nuif_hash_params p{};
    p.voxel_size = c.m_virtualVoxelSize;
    p.trunc_scale = c.m_truncScale;
    p.truncation = c.m_truncation;
    p.weight_max = (int)std::min<UINT>(c.m_integrationWeightMax, 255u);
    p.buckets_log2 = c.bucketsLog2();
    p.bricks = (int)c.m_numSDFBlocks;
    p.storage = storage;
    p.color = color ? 1 : 0;
    return p;
}

NuiKinfuHIPHashScene::NuiKinfuHIPHashScene(const NuiHashingSDFConfig& config, int storage, bool color,
                                           nuif_stream_t s)
    : m_config(config), m_stream(s) {
    m_params = ParamsOf(config, storage, color);
    const size_t bytes = nuif_hash_scene_bytes(&m_params);
    if (!bytes) throw std::runtime_error(std::string("hash scene: ") + nuif_last_error());
    m_scene = DeviceBuffer(bytes);
    m_counter = DeviceBuffer(sizeof(unsigned long long));
    if (config.m_bUseSwapping) m_cache = std::make_unique<NuiKinfuHIPHashGlobalCache>(m_params, m_stream);
    reset();
}

// OCL/NuiKinfuOpenCLHashScene.cpp:134-144 (NuiKinfuVoxelBlockHash::reset, the global cache's reset)
void NuiKinfuHIPHashScene::reset() {
    check(nuif_hash_reset(m_scene.get(), &m_params, m_stream), "nuif_hash_reset");
    if (m_cache) m_cache->reset();
    setDirty();
}

size_t NuiKinfuHIPHashScene::StoredBlocks() const { return m_cache ? m_cache->StoredBlocks() : 0; }

// ------------------------------------------------------------ global cache

NuiKinfuHIPHashGlobalCache::NuiKinfuHIPHashGlobalCache(const nuif_hash_params& p, nuif_stream_t s)
    : m_p(p), m_stream(s) {
    size_t off[8];
    check(nuif_hash_layout(&m_p, off), "nuif_hash_layout");
    m_scratch = DeviceBuffer(nuif_hash_swap_scratch_bytes(&m_p));
    m_ids = DeviceBuffer(off[6] * sizeof(int32_t));
    m_count = DeviceBuffer(sizeof(uint32_t));
}

void NuiKinfuHIPHashGlobalCache::grow(uint32_t n) {
    if (n <= m_cap) return;
    m_cap = std::max(n, 2 * m_cap);
    m_words = DeviceBuffer((size_t)m_cap * 512 * sizeof(uint32_t));
    if (m_p.color) m_colors = DeviceBuffer((size_t)m_cap * 512 * 4);
}

uint32_t NuiKinfuHIPHashGlobalCache::list(void* scene, int which) {
    check(nuif_hash_swap_list(scene, &m_p, which, m_ids.as<int32_t>(), m_count.as<uint32_t>(), m_scratch.get(),
                              m_stream),
          "nuif_hash_swap_list");
    uint32_t n = 0;
    check(nuif_memcpy_d2h(&n, m_count.get(), sizeof(n), m_stream), "swap list");
    check(nuif_stream_synchronize(m_stream), "swap list");
    m_hostIds.resize(n);
    if (n) {
        check(nuif_memcpy_d2h(m_hostIds.data(), m_ids.get(), (size_t)n * sizeof(int32_t), m_stream), "swap list");
        check(nuif_stream_synchronize(m_stream), "swap list");
    }
    return n;
}

void NuiKinfuHIPHashGlobalCache::update(void* scene) {
    // OCL/NuiKinfuOpenCLHashGlobalCache.cpp:212-319 LoadFromGlobalMemory + integrateGlobalIntoLocal
    uint32_t n = list(scene, NUIF_HASH_SWAP_IN);
    if (n) {
        grow(n);
        m_hostWords.assign((size_t)n * 512, 0u);
        m_hostColors.assign(m_p.color ? (size_t)n * 512 * 4 : 0, 0);
        for (uint32_t i = 0; i < n; ++i) {
            auto it = m_store.find(m_hostIds[i]);
            if (it == m_store.end()) continue;
            std::copy(it->second.words.begin(), it->second.words.end(), m_hostWords.begin() + (size_t)i * 512);
            if (m_p.color)
                std::copy(it->second.colors.begin(), it->second.colors.end(),
                          m_hostColors.begin() + (size_t)i * 512 * 4);
        }
        check(nuif_memcpy_h2d(m_words.get(), m_hostWords.data(), m_hostWords.size() * sizeof(uint32_t), m_stream),
              "swap-in upload");
        if (m_p.color)
            check(nuif_memcpy_h2d(m_colors.get(), m_hostColors.data(), m_hostColors.size(), m_stream),
                  "swap-in upload");
        check(nuif_hash_swap_in(scene, &m_p, m_ids.as<int32_t>(), n, m_words.as<uint32_t>(),
                                m_p.color ? m_colors.as<uint8_t>() : nullptr, m_stream),
              "nuif_hash_swap_in");
    }
    // :321-538 BuildSwapOutList + MoveActiveDataToTransferBuffer + CleanVoxelMemory + saveToGlobalMemory
    const uint32_t m = list(scene, NUIF_HASH_SWAP_OUT);
    if (!m) return;
    grow(m);
    check(nuif_hash_swap_out(scene, &m_p, m_ids.as<int32_t>(), m, m_words.as<uint32_t>(),
                             m_p.color ? m_colors.as<uint8_t>() : nullptr, m_stream),
          "nuif_hash_swap_out");
    m_hostWords.resize((size_t)m * 512);
    m_hostColors.resize(m_p.color ? (size_t)m * 512 * 4 : 0);
    check(nuif_memcpy_d2h(m_hostWords.data(), m_words.get(), m_hostWords.size() * sizeof(uint32_t), m_stream),
          "swap-out read");
    if (m_p.color)
        check(nuif_memcpy_d2h(m_hostColors.data(), m_colors.get(), m_hostColors.size(), m_stream), "swap-out read");
    check(nuif_stream_synchronize(m_stream), "swap-out read");
    for (uint32_t i = 0; i < m; ++i) {
        Stored& st = m_store[m_hostIds[i]];
        st.words.assign(m_hostWords.begin() + (size_t)i * 512, m_hostWords.begin() + (size_t)(i + 1) * 512);
        if (m_p.color)
            st.colors.assign(m_hostColors.begin() + (size_t)i * 512 * 4,
                             m_hostColors.begin() + (size_t)(i + 1) * 512 * 4);
    }
}

// OCL/NuiKinfuOpenCLHashScene.cpp:533-583 integrateVolume: AllocateSceneFromDepth +
// IntegrateIntoScene, guarded by the tracker's integration gate
bool NuiKinfuHIPHashScene::integrateVolume(NuiKinfuFrame* pFrame, NuiKinfuCameraState* pCameraState) {
    auto* f = dynamic_cast<NuiKinfuHIPFrame*>(pFrame);
    auto* cam = dynamic_cast<NuiKinfuHIPCameraState*>(pCameraState);
    if (!f || !cam) return false;
    if (f->IsDepthScaled()) throw std::runtime_error("hash scene: the frame depth is scaled to ray length");
    const NuiCLCameraParams& c = cam->GetCameraParams();
    if (m_cache && m_gateIntegrate) {  // the host side of swapping runs only for integrated frames
        int32_t gate = 0;
        check(nuif_memcpy_d2h(&gate, m_gateIntegrate, sizeof(gate), m_stream), "gate read");
        check(nuif_stream_synchronize(m_stream), "gate read");
        if (!gate) return true;
    }
    check(nuif_hash_allocate(m_scene.get(), &m_params, f->GetDepthBuffer(), &c, cam->GetCameraTransformBuffer(),
                             m_gateIntegrate, m_stream),
          "nuif_hash_allocate");
    if (m_cache)
        check(nuif_hash_swap_prepare(m_scene.get(), &m_params, m_gateIntegrate, m_stream), "nuif_hash_swap_prepare");
    const uint8_t* colors = m_params.color ? f->GetColorBuffer() : nullptr;
    check(nuif_hash_integrate(m_scene.get(), &m_params, f->GetDepthBuffer(), colors, &c,
                              cam->GetCameraTransformBuffer(), m_gateIntegrate,
                              m_counter.as<unsigned long long>(), m_stream),
          "nuif_hash_integrate");
    if (m_cache) m_cache->update(m_scene.get());  // OCL/NuiKinfuOpenCLHashScene.cpp:576-580
    setDirty();
    return true;
}

// OCL/NuiKinfuOpenCLHashScene.cpp:1014-1064 raycastRender: ray ranges of the
// visible blocks, then the forward render (m_bUseForwardRender) or castRay at every
// pixel into the feedback frame's maps
void NuiKinfuHIPHashScene::raycastRender(NuiKinfuFeedbackFrame* pFeedbackFrame, NuiKinfuCameraState* pCameraState) {
    auto* fb = dynamic_cast<NuiKinfuHIPFeedbackFrame*>(pFeedbackFrame);
    auto* cam = dynamic_cast<NuiKinfuHIPCameraState*>(pCameraState);
    if (!fb || !cam) return;
    if (!m_range.get() || m_rangeW != fb->GetWidth() || m_rangeH != fb->GetHeight()) {
        m_rangeW = fb->GetWidth();
        m_rangeH = fb->GetHeight();
        m_range = DeviceBuffer((size_t)m_rangeW * m_rangeH * 2 * sizeof(float));
        if (m_config.m_bUseForwardRender)
            m_forwardWork = DeviceBuffer(nuif_hash_forward_render_bytes((int)m_rangeW, (int)m_rangeH));
    }
    const NuiCLCameraParams& c = cam->GetCameraParams();
    if (m_config.m_bUseForwardRender) {
        check(nuif_hash_forward_render(m_scene.get(), &m_params, &c, cam->GetCameraTransformBuffer(),
                                       m_range.as<float>(), fb->GetVertexBuffer(), fb->GetNormalBuffer(),
                                       m_params.color ? fb->GetColorBuffer() : nullptr, m_forwardWork.get(),
                                       m_config.m_bForwardReferenceProjection ? 1 : 0, m_gateRaycast, m_stream),
              "nuif_hash_forward_render");
        return;
    }
    check(nuif_hash_raycast(m_scene.get(), &m_params, &c, cam->GetCameraTransformBuffer(), m_range.as<float>(),
                            fb->GetVertexBuffer(), fb->GetNormalBuffer(),
                            m_params.color ? fb->GetColorBuffer() : nullptr, m_gateRaycast, m_stream),
          "nuif_hash_raycast");
}

void NuiKinfuHIPHashScene::Stats(uint32_t out[4]) const {
    check(nuif_hash_stats(m_scene.get(), out, m_stream), "nuif_hash_stats");
}

unsigned long long NuiKinfuHIPHashScene::UpdatedVoxels() const {
    unsigned long long h = 0;
    check(nuif_memcpy_d2h(&h, m_counter.get(), sizeof(h), m_stream), "counter read");
    check(nuif_stream_synchronize(m_stream), "counter read");
    return h;
}

namespace {
// count pass, then the emit pass into device buffers of the exact size, then the
// host copy (the reference maps the result into GL buffers instead); used for the
// triangle soups and the point clouds alike
template <class CountEmit>
bool mesh_to_host(NuiKinfuMesh* out, nuif_stream_t s, CountEmit&& run) {
    if (!out) return false;
    DeviceBuffer total(sizeof(uint32_t));
    run(nullptr, nullptr, 0u, total.as<uint32_t>());
    uint32_t n = 0;
    check(nuif_memcpy_d2h(&n, total.get(), sizeof(n), s), "mesh count");
    check(nuif_stream_synchronize(s), "mesh count");
    out->positions.assign((size_t)n * 3, 0.f);
    out->colors.assign((size_t)n * 4, 0.f);
    out->indices.clear();
    if (n == 0) return true;
    DeviceBuffer v((size_t)n * 3 * sizeof(float)), c((size_t)n * 4 * sizeof(float));
    run(v.as<float>(), c.as<float>(), n, total.as<uint32_t>());
    check(nuif_memcpy_d2h(out->positions.data(), v.get(), v.bytes(), s), "mesh read");
    check(nuif_memcpy_d2h(out->colors.data(), c.get(), c.bytes(), s), "mesh read");
    check(nuif_stream_synchronize(s), "mesh read");
    return true;
}
}  // namespace

bool NuiKinfuHIPHashScene::Volume2CLVertices(NuiKinfuMesh* pData) {
    if (!pData || !m_dirty) return false;
    clearDirty();
    DeviceBuffer scratch(nuif_hash_points_scratch_bytes(&m_params));
    const bool ok = mesh_to_host(pData, m_stream, [&](float* v, float* c, uint32_t cap, uint32_t* total) {
        check(nuif_hash_points(m_scene.get(), &m_params, v, c, cap, total, scratch.get(), m_stream),
              "nuif_hash_points");
    });
    return ok && !pData->positions.empty();
}

bool NuiKinfuHIPHashScene::Volume2CLMesh(NuiKinfuMesh* pData) {
    DeviceBuffer scratch(nuif_hash_mesh_scratch_bytes(&m_params));
    return mesh_to_host(pData, m_stream, [&](float* v, float* c, uint32_t cap, uint32_t* total) {
        check(nuif_hash_mesh(m_scene.get(), &m_params, v, c, cap, total, scratch.get(), m_stream), "nuif_hash_mesh");
    });
}

bool NuiKinfuHIPScene::Volume2CLVertices(NuiKinfuMesh* pData) {
    if (!pData || !m_dirty) return false;
    clearDirty();
    DeviceBuffer scratch(nuif_points_dense_scratch_bytes(&m_params));
    const bool ok = mesh_to_host(pData, m_stream, [&](float* v, float* c, uint32_t cap, uint32_t* total) {
        check(nuif_points_dense(m_volume.get(), m_color.as<uint8_t>(), &m_params, m_storage, m_surface.as<uint8_t>(),
                                v, c, cap, total, scratch.get(), m_stream),
              "nuif_points_dense");
    });
    return ok && !pData->positions.empty();
}

// OCL/NuiKinfuOpenCLScene.cpp:617-786 Volume2CLMesh
bool NuiKinfuHIPScene::Volume2CLMesh(NuiKinfuMesh* pData) {
    DeviceBuffer scratch(nuif_mc_scratch_bytes(&m_params));
    return mesh_to_host(pData, m_stream, [&](float* v, float* c, uint32_t cap, uint32_t* total) {
        check(nuif_mc_dense(m_volume.get(), m_color.as<uint8_t>(), &m_params, m_storage, v, c, cap, total,
                            scratch.get(), m_stream),
              "nuif_mc_dense");
    });
}

bool NuiKinfuHIPScene::Volume2Mesh(NuiKinfuMesh* pMesh) {
    if (!pMesh) return false;
    // count first (a zero-capacity call reports the corner count), then build
    DeviceBuffer total(sizeof(uint32_t));
    DeviceBuffer scratch(nuif_mc_scratch_bytes(&m_params));
    check(nuif_mc_dense(m_volume.get(), nullptr, &m_params, m_storage, nullptr, nullptr, 0, total.as<uint32_t>(),
                        scratch.get(), m_stream),
          "nuif_mc_dense");
    uint32_t n = 0;
    check(nuif_memcpy_d2h(&n, total.get(), sizeof(n), m_stream), "mesh count");
    check(nuif_stream_synchronize(m_stream), "mesh count");
    pMesh->positions.clear();
    pMesh->colors.clear();
    pMesh->indices.clear();
    if (n == 0) return true;
    DeviceBuffer pts((size_t)n * 3 * sizeof(float)), idx((size_t)n * sizeof(uint32_t));
    DeviceBuffer work(nuif_mesh_dense_scratch_bytes(&m_params, n));
    uint32_t np = 0, ni = 0;
    check(nuif_mesh_dense(m_volume.get(), &m_params, m_storage, n, pts.as<float>(), idx.as<uint32_t>(), &np, &ni,
                          work.get(), m_stream),
          "nuif_mesh_dense");
    pMesh->positions.resize((size_t)np * 3);
    pMesh->indices.resize(ni);
    check(nuif_memcpy_d2h(pMesh->positions.data(), pts.get(), (size_t)np * 3 * sizeof(float), m_stream), "mesh read");
    check(nuif_memcpy_d2h(pMesh->indices.data(), idx.get(), (size_t)ni * sizeof(uint32_t), m_stream), "mesh read");
    check(nuif_stream_synchronize(m_stream), "mesh read");
    return true;
}

}  // namespace nuif_kinfu
