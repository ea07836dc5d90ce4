// key=value configuration files, same keys and defaults as
// VO/NuiTrackerConfig.cpp:5-115 and VO/NuiKinfuVolumeConfig.cpp:5-125.
#include "NuiKinfuConfig.h"

#include <cmath>
#include <cstdlib>
#include <fstream>

NuiTrackerConfig::NuiTrackerConfig() : bHasColor(false) {
    filter_radius = 3;
    sigma_space2_inv_half = 0.08f;
    sigma_depth2_inv_half = 50.0f;
    depth_threshold = 0.3f;
    dist_threshold = 0.1f;
    normal_threshold = std::sin(20.f * 3.14159254f / 180.f);
    color_dist_threshold = 0.1f;
    color_gradiant_min = 0.2f;
    track_threshold = 1e-3f;
    TrackerIterationParams p;
    p.m_num = 10;
    p.m_type = eTracker_Iteration_Both;
    iterations.assign(3, p);
}

bool NuiTrackerConfig::load(const std::string& fileName) {
    std::ifstream in(fileName.c_str());
    if (!in.is_open()) return false;
    iterations.clear();
    std::string line;
    while (std::getline(in, line)) {
        const size_t pos = line.find('=');
        if (pos == std::string::npos) continue;
        const std::string key = line.substr(0, pos);
        const char* val = line.c_str() + pos + 1;
        if (key == "icp_filter_radius") filter_radius = (UINT)std::atoi(val);
        else if (key == "icp_sigma_space2_inv_half") sigma_space2_inv_half = (float)std::atof(val);
        else if (key == "icp_sigma_depth2_inv_half") sigma_depth2_inv_half = (float)std::atof(val);
        else if (key == "icp_depth_threshold") depth_threshold = (float)std::atof(val);
        else if (key == "icp_dist_threshold") dist_threshold = (float)std::atof(val);
        else if (key == "icp_normal_threshold") normal_threshold = (float)std::atof(val);
        else if (key == "icp_track_threshold") track_threshold = (float)std::atof(val);
        else if (key == "icp_iteration") {
            const UINT n = (UINT)std::atoi(val);
            if (n > 0) iterations.push_back(TrackerIterationParams{n, eTracker_Iteration_Both});
        }
    }
    return true;
}

bool NuiTrackerConfig::log(const std::string& fileName) const {
    std::ofstream out(fileName.c_str(), std::ios::out | std::ios::app);
    if (!out.is_open()) return false;
    out << "icp_filter_radius=" << filter_radius << "\n";
    out << "icp_sigma_space2_inv_half=" << sigma_space2_inv_half << "\n";
    out << "icp_sigma_depth2_inv_half=" << sigma_depth2_inv_half << "\n";
    out << "icp_depth_threshold=" << depth_threshold << "\n";
    out << "icp_dist_threshold=" << dist_threshold << "\n";
    out << "icp_normal_threshold=" << normal_threshold << "\n";
    out << "icp_track_threshold=" << track_threshold << "\n";
    for (const auto& it : iterations) out << "icp_iteration=" << it.m_num << "\n";
    return true;
}

NuiKinfuVolumeConfig::NuiKinfuVolumeConfig() {
    dimensions = Vector3f::Constant(3.0f);
    resolution = Vector3i::Constant(256);
    tranc_dist = 0.03f;
    bHas_color_volume = true;
    max_color_weight = 128;
    bIsDynamic = false;
    voxel_shift = 15;
    translateBasis = Vector3f::Zero();
}

bool NuiKinfuVolumeConfig::load(const std::string& fileName) {
    std::ifstream in(fileName.c_str());
    if (!in.is_open()) return false;
    std::string line;
    while (std::getline(in, line)) {
        const size_t pos = line.find('=');
        if (pos == std::string::npos) continue;
        const std::string key = line.substr(0, pos);
        const char* val = line.c_str() + pos + 1;
        if (key == "VolumeSizeX") dimensions[0] = (float)std::atof(val);
        else if (key == "VolumeSizeY") dimensions[1] = (float)std::atof(val);
        else if (key == "VolumeSizeZ") dimensions[2] = (float)std::atof(val);
        else if (key == "VolumeResolutionX") resolution[0] = std::atoi(val);
        else if (key == "VolumeResolutionY") resolution[1] = std::atoi(val);
        else if (key == "VolumeResolutionZ") resolution[2] = std::atoi(val);
        else if (key == "VolumeTrancDist") tranc_dist = (float)std::atof(val);
        else if (key == "VolumeHadColor") bHas_color_volume = std::atoi(val) != 0;
        else if (key == "MaxColorWeight") max_color_weight = std::atoi(val);
        else if (key == "IsDynamicVolume") bIsDynamic = std::atoi(val) != 0;
        else if (key == "DynamicVoxelShift") voxel_shift = std::atoi(val);
        else if (key == "VolumeBasisX") translateBasis[0] = (float)std::atof(val);
        else if (key == "VolumeBasisY") translateBasis[1] = (float)std::atof(val);
        else if (key == "VolumeBasisZ") translateBasis[2] = (float)std::atof(val);
    }
    return true;
}

bool NuiKinfuVolumeConfig::log(const std::string& fileName) const {
    std::ofstream out(fileName.c_str(), std::ios::out | std::ios::app);
    if (!out.is_open()) return false;
    out << "VolumeSizeX=" << dimensions[0] << "\nVolumeSizeY=" << dimensions[1] << "\nVolumeSizeZ=" << dimensions[2]
        << "\n";
    out << "VolumeResolutionX=" << resolution[0] << "\nVolumeResolutionY=" << resolution[1]
        << "\nVolumeResolutionZ=" << resolution[2] << "\n";
    out << "VolumeTrancDist=" << tranc_dist << "\nVolumeHadColor=" << (bHas_color_volume ? 1 : 0)
        << "\nMaxColorWeight=" << max_color_weight << "\nIsDynamicVolume=" << (bIsDynamic ? 1 : 0)
        << "\nDynamicVoxelShift=" << voxel_shift << "\nVolumeBasisX=" << translateBasis[0]
        << "\nVolumeBasisY=" << translateBasis[1] << "\nVolumeBasisZ=" << translateBasis[2] << "\n";
    return true;
}

NuiHashingSDFConfig::NuiHashingSDFConfig()
    : m_hashNumBuckets(512 * 512),
      m_numSDFBlocks(400000),
      m_hashMaxCollisionLinkedListSize(7),
      m_virtualVoxelSize(0.005f),
      m_truncScale(0.02f),
      m_truncation(0.06f),
      m_integrationWeightSample(3),
      m_integrationWeightMax(255),
      m_bUseSwapping(false),
      m_bUseForwardRender(false),
      m_bForwardReferenceProjection(false) {}

int NuiHashingSDFConfig::bucketsLog2() const {
    int b = 4;
    while (b < 24 && (((size_t)4 << b) < (size_t)2 * m_numSDFBlocks || ((size_t)1 << b) < m_hashNumBuckets)) ++b;
    return b;
}

