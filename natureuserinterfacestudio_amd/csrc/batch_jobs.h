// Per-session jobs of the batched fusion round (nuif_fusion_round): one launch per
// stage covers every session of the round.  A batch kernel's flat block index is
// mapped to its session by a binary search over the jobs' first block (block0),
// then to the session's own block coordinates; the kernel bodies are the ones the
// single-session launches run, so a session's results do not depend on batching.
#pragma once

#include "dense_volume.h"
#include "nuif_internal.h"
#include "schedule.h"

namespace nuif {

// The integration's walk-state workspace: per-column states at the starts of z
// segments 1, 2, ... of the shortest segments (kMinSegVoxels), then the z table: the z
// terms (vgz, zs) at every z, float2[Nz], the same for every column -- the integration
// takes a walked chunk's z terms from it and estimates the columns' numerators to test
// chunks it then skips, instead of stepping every column through them.
__host__ __device__ __forceinline__ size_t ztable_offset(int Nx, int Ny, int Nz) {
    const size_t nseg = (size_t)((Nz + kMinSegVoxels - 1) / kMinSegVoxels);
    return (nseg - 1) * ((size_t)Nx * Ny + 1);
}
// blocks of the column-states launch: the per-column states only with several segments
inline unsigned column_state_blocks(bool walk, int Nx, int Ny, unsigned gz) {
    return walk ? (gz > 1 ? div_up((size_t)Nx * Ny, 256) : 1u) : 0u;
}

// the job whose blocks contain flat block b (jobs sorted by block0, jobs[0].block0 == 0)
template <class Job>
__device__ __forceinline__ int job_of(const Job* __restrict__ jobs, int n, unsigned b) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].block0 <= b)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

struct Gauss {
    float g[15];  // 2*radius+1 <= 15
};

// passing filter + bilateral + level-0 vertices + ray-length depth (preprocess_l0)
struct PreJob {
    const uint16_t* raw;
    float* depth_out;
    float* depth_f;
    float* vmap;
    int W, H;
    NuiCLCameraParams cam;
    Gauss g;
    int r;
    float sigma_depth, thr;
    int scale;
    unsigned gx;      // tiles along x
    unsigned block0;
};

// one pyramid level (half_sample + vertices)
struct HalfJob {
    const float* src;
    float* dst;
    float* vmap;
    int Wd, Hd, r;
    float thr;
    NuiCLCameraParams cam;
    int div;
    unsigned gx;
    unsigned block0;
};

// depth bounds of a frame (max valid depth over 8-pixel cells, then the levels)
struct BoundsJob {
    const float* depth;
    int W, H;
    float* out;
    const int32_t* gate;
    unsigned block0;
};

// dense integration of one session (column walk states + tile kernel)
struct IntJob {
    const float* depth;
    const uchar4* bgra;
    void* vol;
    uchar4* cvol;
    TsdfParams p;
    NuiCLCameraParams cam;
    const NuiCLRigidTransform* pose;
    int maxcw;
    const int32_t* gate;
    unsigned long long* updated;
    const float* bounds;
    SurfaceBlocks surf;
    float2* walk;     // walk states (NULL: one z segment)
    int seg;          // z-segment voxels (segment_voxels)
    unsigned gx, gy, gz;
    unsigned block0;  // tile kernel blocks
    unsigned cblock0; // column_states blocks
};

// dense raycast of one session
struct RayJob {
    const void* vol;
    const uchar4* cvol;
    TsdfParams p;
    NuiCLCameraParams cam;
    const NuiCLRigidTransform* pose;
    float* vmap;
    float* nmap;
    uchar4* cmap;
    const int32_t* enable;
    SurfaceBlocks surf;
    unsigned gx;
    unsigned block0;
};

__device__ __forceinline__ SurfaceBlocks in_global(SurfaceBlocks s) {
    s.b = in_global(s.b);
    return s;
}

// host-side launchers of the batch kernels (jobs: device arrays inside the round table,
// hjobs: the same jobs on the host, for the grid sizes)
int launch_preprocess_batch(const PreJob* jobs, const PreJob* hjobs, int n, hipStream_t st);
int launch_half_batch(const HalfJob* jobs, const HalfJob* hjobs, int n, hipStream_t st);
int launch_bounds_batch(const BoundsJob* jobs, const BoundsJob* hjobs, int n, hipStream_t st);
int launch_integrate_batch(const IntJob* jobs, const IntJob* hjobs, int n, int storage, bool color, hipStream_t st);
int launch_raycast_batch(const RayJob* jobs, const RayJob* hjobs, int n, int storage, hipStream_t st);
// the ICP of n sessions: make_icp_jobs writes the jobs (icp_job_bytes() each) to host
// memory, reference order first (*nref of them); launch_icp_jobs runs them from the
// device copy dtable
size_t icp_job_bytes();
int make_icp_jobs(const nuif_icp_session* sessions, int n, void* host_out, int* counts);
int launch_icp_jobs(const void* dtable, const void* hjobs, int n, const int* counts, hipStream_t st);
// the joint ICP order's inputs (color.hip): job k of a host table for a frame (its blocks
// numbered from *blocks, which advances past them; ref_bgra, the previous frame's
// colours, receives this frame's after the inputs are made), and the launches for n jobs
size_t photo_job_bytes();
int make_photo_job_at(void* hjobs, int k, const uint8_t* bgra, uint8_t* ref_bgra, const nuif_icp_photometric& ph,
                      int W, int H, int levels, unsigned* blocks);
int launch_photo_jobs(const void* djobs, int n, unsigned blocks, hipStream_t st);
int make_gauss(const float* gaussian, int radius, Gauss* g);

}  // namespace nuif
