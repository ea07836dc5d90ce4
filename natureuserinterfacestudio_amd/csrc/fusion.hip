// The batched frame of many dense sessions (nuif_fusion_round): the fusion
// server's round.  The reference runs one session's frame on one in-order queue
// (VO/NuiKinfuTrackingEngine.cpp:68-142 over CLU/NuiOpenCLGlobal.cpp:175); here
// one launch per stage covers every session of the round, in the order of the
// single-session pipeline:
//   preprocess (level 0), half samples (levels 1..), depth bounds,
//   ICP (begin, the lockstep steps of each reduction order, end),
//   column walk states + integration tiles (per storage / colour group),
//   raycast (per storage group).
// All per-session parameters of all stages go to the device in one copy of a job
// table from pinned host memory.

#include "batch_jobs.h"
#include "schedule.h"
#include "nuif_internal.h"

#include <algorithm>
#include <vector>

namespace nuif {

namespace {

constexpr int kLevels = NUIF_ICP_MAX_LEVELS;
constexpr int kTileW = 32, kTileH = 8;  // preprocess_l0 tiles (preprocess.hip)

size_t align256(size_t b) { return (b + 255) / 256 * 256; }

// byte offsets of the table's parts for n sessions
struct Layout {
    size_t pre, half[kLevels], icp, photo, bounds, integ, ray, total;
    explicit Layout(int n) {
        size_t o = 0;
        pre = o;
        o = align256(o + (size_t)n * sizeof(PreJob));
        for (int l = 0; l < kLevels; ++l) {
            half[l] = o;
            o = align256(o + (size_t)n * sizeof(HalfJob));
        }
        icp = o;
        o = align256(o + (size_t)n * icp_job_bytes());
        photo = o;
        o = align256(o + (size_t)n * photo_job_bytes());
        bounds = o;
        o = align256(o + (size_t)n * sizeof(BoundsJob));
        integ = o;
        o = align256(o + (size_t)n * sizeof(IntJob));
        ray = o;
        o = align256(o + (size_t)n * sizeof(RayJob));
        total = o;
    }
};

int check_session(const nuif_round_session& r, int i) {
    NUIF_CHECK_ARG(r.depth && r.depth_out && r.bounds && r.gaussian && r.volume && r.vmap && r.nmap,
                   "nuif_fusion_round: session %d has a null buffer", i);
    const int W = (int)r.icp.cam.depthImageWidth, H = (int)r.icp.cam.depthImageHeight;
    NUIF_CHECK_ARG(W > 0 && H > 0, "nuif_fusion_round: session %d has an empty image", i);
    const int L = r.icp.cfg.levels;
    NUIF_CHECK_ARG(L >= 1 && L <= kLevels, "nuif_fusion_round: session %d has %d levels", i, L);
    for (int l = 0; l < L; ++l)
        NUIF_CHECK_ARG(r.depth_pyr[l] && r.icp.vmap_pyr[l], "nuif_fusion_round: session %d level %d is NULL", i, l);
    NUIF_CHECK_ARG((W >> (L - 1)) > 0 && (H >> (L - 1)) > 0, "nuif_fusion_round: session %d: too many levels", i);
    NUIF_CHECK_ARG(r.icp.ref_vmap == r.vmap && r.icp.ref_nmap == r.nmap,
                   "nuif_fusion_round: session %d: the ICP reference maps must be its feedback maps", i);
    const int Nx = (int)r.params.resolution[0], Ny = (int)r.params.resolution[1], Nz = (int)r.params.resolution[2];
    NUIF_CHECK_ARG(Nx > 0 && Ny > 0 && Nz > 0 && Nx % 8 == 0 && Ny % 8 == 0 && Nz % 32 == 0,
                   "nuif_fusion_round: session %d: volume %d x %d x %d (needs multiples of 8, 8, 32)", i, Nx, Ny, Nz);
    NUIF_CHECK_ARG(r.storage == NUIF_TSDF_S16 || r.storage == NUIF_TSDF_F16 || r.storage == NUIF_TSDF_F32,
                   "nuif_fusion_round: session %d: unknown storage %d", i, r.storage);
    NUIF_CHECK_ARG(Nz <= kMinSegVoxels || r.walk_states, "nuif_fusion_round: session %d needs walk states", i);
    if (r.icp.cfg.reduction == NUIF_ICP_REDUCE_JOINT) {
        const nuif_icp_photometric& ph = r.icp.photometric;
        NUIF_CHECK_ARG(r.bgra && r.prev_bgra && ph.ref_model && (const void*)r.bgra != r.prev_bgra,
                       "nuif_fusion_round: session %d: the joint ICP needs the frame's colours, the previous "
                       "frame's colours and the intensity buffers", i);
        for (int l = 0; l < L; ++l)
            NUIF_CHECK_ARG(ph.intensity_pyr[l], "nuif_fusion_round: session %d: intensity level %d is NULL", i, l);
    }
    return NUIF_OK;
}

}  // namespace

}  // namespace nuif

using namespace nuif;

extern "C" {

size_t nuif_fusion_round_table_bytes(int n) { return n > 0 ? Layout(n).total : 0; }

int nuif_fusion_round(const nuif_round_session* S, int n, void* table, void* staging, nuif_stream_t s) {
    return nuif_fusion_round_timed(S, n, table, staging, nullptr, s);
}

}  // extern "C"

namespace {

// FNV-1a over the values that decide a round's launches (grids, groups, kernel choices and
// the table and staging addresses baked into them)
struct Sig {
    uint64_t h = 1469598103934665603ull;
    void add(uint64_t v) {
        for (int k = 0; k < 8; ++k) {
            h ^= (v >> (8 * k)) & 0xff;
            h *= 1099511628211ull;
        }
    }
};

uint64_t round_signature(const nuif_round_session* S, int n, const void* table, const void* staging) {
    Sig g;
    g.add((uint64_t)n);
    g.add((uint64_t)(uintptr_t)table);
    g.add((uint64_t)(uintptr_t)staging);
    for (int i = 0; i < n; ++i) {
        const nuif_round_session& r = S[i];
        g.add(r.icp.cam.depthImageWidth);
        g.add(r.icp.cam.depthImageHeight);
        g.add((uint64_t)r.icp.cfg.levels);
        for (int l = 0; l < NUIF_ICP_MAX_LEVELS; ++l) g.add((uint64_t)r.icp.cfg.iterations[l]);
        g.add((uint64_t)r.icp.cfg.reduction);
        g.add((uint64_t)r.storage);
        g.add((uint64_t)(r.color_volume && r.bgra));
        g.add((uint64_t)(r.color_volume != nullptr));
        for (int k = 0; k < 3; ++k) g.add((uint64_t)r.params.resolution[k]);
        g.add((uint64_t)(r.walk_states != nullptr));
        g.add((uint64_t)(r.surface_blocks != nullptr));
        g.add((uint64_t)r.radius);
    }
    return g.h;
}

int fusion_round_impl(const nuif_round_session* S, int n, void* table, void* staging, nuif_event_t* ev,
                      nuif_stream_t s, bool launch);

}  // namespace

extern "C" {

int nuif_fusion_round_prepare(const nuif_round_session* S, int n, void* table, void* staging, uint64_t* signature) {
    NUIF_CHECK_ARG(signature, "nuif_fusion_round_prepare: NULL signature");
    const int rc = fusion_round_impl(S, n, table, staging, nullptr, nullptr, false);
    if (rc) return rc;
    *signature = round_signature(S, n, table, staging);
    return NUIF_OK;
}

int nuif_fusion_round_timed(const nuif_round_session* S, int n, void* table, void* staging, nuif_event_t* ev,
                            nuif_stream_t s) {
    return fusion_round_impl(S, n, table, staging, ev, s, true);
}

}  // extern "C"

namespace {

int fusion_round_impl(const nuif_round_session* S, int n, void* table, void* staging, nuif_event_t* ev,
                      nuif_stream_t s, bool launch) {
    NUIF_REQUIRE_INIT();
    NUIF_CHECK_ARG(n >= 0 && (n == 0 || (S && table && staging)), "nuif_fusion_round: bad arguments");
    if (n == 0) return NUIF_OK;
    for (int i = 0; i < n; ++i) {
        const int rc = check_session(S[i], i);
        if (rc) return rc;
    }
    const Layout lay(n);
    char* h = static_cast<char*>(staging);
    const char* d = static_cast<const char*>(table);
    hipStream_t st = stream_of(s);

    // preprocessing: level 0 and each further level
    PreJob* pre = reinterpret_cast<PreJob*>(h + lay.pre);
    unsigned b = 0;
    for (int i = 0; i < n; ++i) {
        const nuif_round_session& r = S[i];
        PreJob& J = pre[i];
        J = PreJob{};
        J.raw = r.depth;
        J.depth_out = r.depth_out;
        J.depth_f = r.depth_pyr[0];
        J.vmap = const_cast<float*>(r.icp.vmap_pyr[0]);
        J.W = (int)r.icp.cam.depthImageWidth;
        J.H = (int)r.icp.cam.depthImageHeight;
        J.cam = r.icp.cam;
        const int rc = make_gauss(r.gaussian, r.radius, &J.g);
        if (rc) return rc;
        J.r = r.radius;
        J.sigma_depth = r.sigma_depth2_inv_half;
        J.thr = r.depth_threshold;
        J.scale = 1;
        J.gx = div_up(J.W, kTileW);
        J.block0 = b;
        b += J.gx * div_up(J.H, kTileH);
    }
    int nhalf[kLevels] = {0};
    for (int l = 1; l < kLevels; ++l) {
        HalfJob* hj = reinterpret_cast<HalfJob*>(h + lay.half[l]);
        b = 0;
        for (int i = 0; i < n; ++i) {
            const nuif_round_session& r = S[i];
            if (l >= r.icp.cfg.levels) continue;
            HalfJob& J = hj[nhalf[l]++];
            J = HalfJob{};
            J.src = r.depth_pyr[l - 1];
            J.dst = r.depth_pyr[l];
            J.vmap = const_cast<float*>(r.icp.vmap_pyr[l]);
            J.Wd = (int)r.icp.cam.depthImageWidth >> l;
            J.Hd = (int)r.icp.cam.depthImageHeight >> l;
            J.r = 1;  // OCL/NuiKinfuOpenCLDepthTracker.cpp:158 subSampleRadius
            J.thr = r.depth_threshold;
            J.cam = r.icp.cam;
            J.div = 1 << l;
            J.gx = div_up(J.Wd, 16);
            J.block0 = b;
            b += J.gx * div_up(J.Hd, 16);
        }
    }
    // depth bounds of the ray-length depth (computed before tracking, as the frame does)
    BoundsJob* bj = reinterpret_cast<BoundsJob*>(h + lay.bounds);
    b = 0;
    for (int i = 0; i < n; ++i) {
        const nuif_round_session& r = S[i];
        BoundsJob& J = bj[i];
        J = BoundsJob{r.depth_out, (int)r.icp.cam.depthImageWidth, (int)r.icp.cam.depthImageHeight, r.bounds,
                      nullptr, b};
        b += div_up((size_t)((J.W - 1) / 8 + 1) * ((J.H - 1) / 8 + 1), 256);
    }
    // ICP
    std::vector<nuif_icp_session> icp((size_t)n);
    for (int i = 0; i < n; ++i) icp[i] = S[i].icp;
    int counts[2] = {0, 0};
    int rc = make_icp_jobs(icp.data(), n, h + lay.icp, counts);
    if (rc) return rc;
    // the joint order's photometric inputs
    int nphoto = 0;
    unsigned pblocks = 0;
    for (int i = 0; i < n; ++i) {
        const nuif_round_session& r = S[i];
        if (r.icp.cfg.reduction != NUIF_ICP_REDUCE_JOINT) continue;
        rc = make_photo_job_at(h + lay.photo, nphoto++, r.bgra, r.prev_bgra, r.icp.photometric,
                               (int)r.icp.cam.depthImageWidth, (int)r.icp.cam.depthImageHeight, r.icp.cfg.levels,
                               &pblocks);
        if (rc) return rc;
    }
    // integration and raycast, grouped by storage (and colour for the integration)
    const int seg = segment_voxels(n);  // integration z segment for a launch of n sessions
    std::vector<int> order((size_t)n);
    for (int i = 0; i < n; ++i) order[i] = i;
    auto color_of = [&](int i) { return S[i].color_volume && S[i].bgra; };
    std::stable_sort(order.begin(), order.end(), [&](int a, int c) {
        return S[a].storage != S[c].storage ? S[a].storage < S[c].storage : color_of(a) < color_of(c);
    });
    IntJob* ij = reinterpret_cast<IntJob*>(h + lay.integ);
    RayJob* rj = reinterpret_cast<RayJob*>(h + lay.ray);
    struct Group {
        int first, count, storage;
        bool color;
    };
    std::vector<Group> igroups, rgroups;
    unsigned ib = 0, icb = 0, rb = 0;
    for (int k = 0; k < n; ++k) {
        const nuif_round_session& r = S[order[k]];
        const bool col = color_of(order[k]);
        if (igroups.empty() || igroups.back().storage != r.storage || igroups.back().color != col) {
            igroups.push_back(Group{k, 0, r.storage, col});
            ib = icb = 0;
        }
        if (rgroups.empty() || rgroups.back().storage != r.storage) {
            rgroups.push_back(Group{k, 0, r.storage, false});
            rb = 0;
        }
        igroups.back().count++;
        rgroups.back().count++;
        const int Nx = (int)r.params.resolution[0], Ny = (int)r.params.resolution[1], Nz = (int)r.params.resolution[2];
        IntJob& J = ij[k];
        J = IntJob{};
        J.depth = r.depth_out;
        J.bgra = col ? reinterpret_cast<const uchar4*>(r.bgra) : nullptr;
        J.vol = r.volume;
        J.cvol = col ? reinterpret_cast<uchar4*>(r.color_volume) : nullptr;
        J.p = r.params;
        J.cam = r.icp.cam;
        J.pose = r.icp.pose;
        J.maxcw = r.max_color_weight;
        J.gate = nuif_icp_gate_flag(r.icp.state);
        J.updated = r.updated;
        J.bounds = r.bounds;
        J.surf = surface_blocks_of(r.surface_blocks, Nx, Ny, Nz);
        J.gx = (unsigned)Nx / 8;
        J.gy = (unsigned)Ny / 8;
        J.seg = seg;
        J.gz = div_up(Nz, seg);
        J.walk = reinterpret_cast<float2*>(r.walk_states);
        J.block0 = ib;
        J.cblock0 = icb;
        ib += J.gx * J.gy * J.gz;
        icb += column_state_blocks(J.walk != nullptr, Nx, Ny, J.gz);
        RayJob& R = rj[k];
        R = RayJob{};
        R.vol = r.volume;
        R.cvol = r.color_volume ? reinterpret_cast<const uchar4*>(r.color_volume) : nullptr;
        R.p = r.params;
        R.cam = r.icp.cam;
        R.pose = r.icp.pose;
        R.vmap = r.vmap;
        R.nmap = r.nmap;
        R.cmap = r.color_volume ? reinterpret_cast<uchar4*>(r.cmap) : nullptr;
        R.enable = nuif_icp_ok_flag(r.icp.state);
        R.surf = J.surf;
        R.gx = div_up(r.icp.cam.depthImageWidth, 8);
        R.block0 = rb;
        rb += R.gx * div_up(r.icp.cam.depthImageHeight, 8);
    }

    if (!launch) return NUIF_OK;  // the host table only (a captured round replays the rest)
    NUIF_HIP(hipMemcpyAsync(table, staging, lay.total, hipMemcpyHostToDevice, st));
    auto mark = [&](int k) -> int {
        if (ev) NUIF_HIP(hipEventRecord(reinterpret_cast<hipEvent_t>(ev[k]), st));
        return NUIF_OK;
    };
    if ((rc = mark(0))) return rc;
    rc = launch_preprocess_batch(reinterpret_cast<const PreJob*>(d + lay.pre), pre, n, st);
    if (rc) return rc;
    for (int l = 1; l < kLevels && nhalf[l] > 0; ++l) {
        rc = launch_half_batch(reinterpret_cast<const HalfJob*>(d + lay.half[l]),
                               reinterpret_cast<const HalfJob*>(h + lay.half[l]), nhalf[l], st);
        if (rc) return rc;
    }
    if ((rc = mark(1))) return rc;
    rc = launch_bounds_batch(reinterpret_cast<const BoundsJob*>(d + lay.bounds), bj, n, st);
    if (rc) return rc;
    if ((rc = mark(2))) return rc;
    rc = launch_photo_jobs(d + lay.photo, nphoto, pblocks, st);
    if (rc) return rc;
    rc = launch_icp_jobs(d + lay.icp, h + lay.icp, n, counts, st);
    if (rc) return rc;
    if ((rc = mark(3))) return rc;
    for (const Group& g : igroups) {
        rc = launch_integrate_batch(reinterpret_cast<const IntJob*>(d + lay.integ) + g.first, ij + g.first, g.count,
                                    g.storage, g.color, st);
        if (rc) return rc;
    }
    if ((rc = mark(4))) return rc;
    for (const Group& g : rgroups) {
        rc = launch_raycast_batch(reinterpret_cast<const RayJob*>(d + lay.ray) + g.first, rj + g.first, g.count,
                                  g.storage, st);
        if (rc) return rc;
    }
    return mark(5);
}

}  // namespace

