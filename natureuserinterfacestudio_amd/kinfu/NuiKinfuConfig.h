// Tracker and volume configuration (src/SLAM/VisualOdometry/NuiTrackerConfig.*,
// NuiKinfuVolumeConfig.*, NuiHashingSDFConfig.*): same fields, defaults and
// key=value text format.
#pragma once

#include <string>
#include <vector>

#include "NuiKinfuTypes.h"

struct NuiTrackerConfig {
    typedef enum {
        eTracker_Iteration_Rotation = 1,
        eTracker_Iteration_Translation = 2,
        eTracker_Iteration_Both = 3,
        eTracker_Iteration_None = 4
    } TrackerIterationType;
    struct TrackerIterationParams {
        UINT m_num;
        TrackerIterationType m_type;
    };
    typedef std::vector<TrackerIterationParams> ITERATION_CLASS;

    bool bHasColor;
    UINT filter_radius;
    float sigma_space2_inv_half;
    float sigma_depth2_inv_half;
    float depth_threshold;
    float dist_threshold;
    float normal_threshold;
    float color_dist_threshold;
    float color_gradiant_min;
    float track_threshold;
    ITERATION_CLASS iterations;  // iterations[0] = finest level

    NuiTrackerConfig();  // VO/NuiTrackerConfig.cpp:5-24
    bool load(const std::string& fileName);
    bool log(const std::string& fileName) const;
};

struct NuiKinfuVolumeConfig {
    Vector3f dimensions;  // metres
    Vector3i resolution;
    float tranc_dist;     // metres
    bool bHas_color_volume;
    int max_color_weight;
    bool bIsDynamic;
    int voxel_shift;
    Vector3f translateBasis;

    NuiKinfuVolumeConfig();  // VO/NuiKinfuVolumeConfig.cpp:5-15
    bool load(const std::string& fileName);
    bool log(const std::string& fileName) const;
};

// VO/NuiHashingSDFConfig.cpp:3-15.  The table here has 2^b buckets of 4 entries, the
// smallest power of two with at least m_hashNumBuckets buckets and a load of at most
// 1/2 for m_numSDFBlocks bricks; a full bucket probes the next ones instead of the
// reference's linked excess list, so m_hashMaxCollisionLinkedListSize is kept for
// the configuration interface only.
struct NuiHashingSDFConfig {
    UINT m_hashNumBuckets;
    UINT m_numSDFBlocks;
    UINT m_hashMaxCollisionLinkedListSize;
    float m_virtualVoxelSize;  // metres
    float m_truncScale;        // metres: band half-width and sdf normalisation
    float m_truncation;        // metres: raycast step scale
    UINT m_integrationWeightSample;
    UINT m_integrationWeightMax;
    bool m_bUseSwapping;
    bool m_bUseForwardRender;
    // Not in the reference: true projects the forward render as forwardProject_kernel
    // does (camera->world transform applied to world points); false uses world->camera.
    bool m_bForwardReferenceProjection;

    NuiHashingSDFConfig();
    int bucketsLog2() const;
};

