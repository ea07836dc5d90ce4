// Launch schedule of the dense integration and raycast.  The values are the ones
// natureuserinterfacestudio_amd/tune_schedules.py measured best on MI355X at
// 640x480 / 512^3 (server round and single session); a build may override the
// compile-time ones (-DNUIF_RAY_BATCH=... etc.) and nuif_schedule() reports them.
#pragma once

#ifndef NUIF_RAY_BATCH
#define NUIF_RAY_BATCH 8     // raycast samples whose loads are in flight together
#endif
#ifndef NUIF_WALK_BATCH
#define NUIF_WALK_BATCH 8    // integration walk steps whose depth reads are in flight together
#endif
#ifndef NUIF_RMW_BATCH
#define NUIF_RMW_BATCH 2     // integration column chunks whose volume loads are in flight together
#endif
#ifndef NUIF_CULL_GROUP
#define NUIF_CULL_GROUP 4    // integration chunks per coarse cull test
#endif
#include <cstdlib>

namespace nuif {
constexpr int kRayBatch = NUIF_RAY_BATCH;
constexpr int kWalkBatch = NUIF_WALK_BATCH;
constexpr int kRmwBatch = NUIF_RMW_BATCH;
constexpr int kCullGroup = NUIF_CULL_GROUP;

// Columns are split into z segments walked by separate waves: tiles that look through
// long runs of free space would otherwise set the kernel's duration alone.  The length
// is chosen per launch by the number of sessions sharing it: one session needs the
// parallelism of short segments, a batched round has waves enough and gains from fewer
// walk states.  Rules: from `min_sessions` sessions on, `voxels` per segment.  The
// environment variable NUIF_SEG_VOXELS forces a length (the tuner's sweep).
struct SegmentRule {
    int min_sessions, voxels;
};
// measured at 640x480 / 512^3: 1 session 64, 8 and 32 sessions 512 (a whole column per
// wave: 7-10 % faster than 256 there); 2 to 7 sessions keep 128
constexpr SegmentRule kSegmentRules[] = {{1, 64}, {2, 128}, {8, 512}};
constexpr int kMinSegVoxels = 64;  // walk-state buffers are sized for this length

// XCD-aware block numbering (xcd_block: consecutive blocks of a session share one XCD's
// L2) per stage: bit 0 the ICP tile runs, bit 1 the batched integration, bit 2 the
// batched raycast.  The environment variable NUIF_XCD_REMAP (0..7) forces a mask (the
// tuner's sweep).  Measured on the 32-session round at 640x480 / 512^3: the integration
// gains 11 % (each XCD reads 4 sessions' depth images instead of all 32), the raycast
// loses 5 % although its L2 misses fall by 43 %, the ICP does not change.
enum : unsigned { kXcdIcp = 1u, kXcdIntegrate = 2u, kXcdRaycast = 4u };
constexpr unsigned kXcdRemapDefault = kXcdIntegrate;

inline bool xcd_remap(unsigned stage) {
    static const unsigned mask = [] {
        const char* e = std::getenv("NUIF_XCD_REMAP");
        const int v = e ? std::atoi(e) : -1;
        return (v >= 0 && v <= 7) ? (unsigned)v : kXcdRemapDefault;
    }();
    return (mask & stage) != 0;
}

inline int segment_voxels(int sessions) {
    static const int forced = [] {
        const char* e = std::getenv("NUIF_SEG_VOXELS");
        const int v = e ? std::atoi(e) : 0;
        return (v >= kMinSegVoxels && v % 32 == 0) ? v : 0;
    }();
    if (forced) return forced;
    int v = kSegmentRules[0].voxels;
    for (const SegmentRule& r : kSegmentRules)
        if (sessions >= r.min_sessions) v = r.voxels;
    return v;
}
}  // namespace nuif
