// TEST INFRASTRUCTURE ONLY (see oracle_common.h).  CPU restatement of the
// reference's dense KinectFusion path: frame preprocessing, projective ICP
// with its Levenberg-Marquardt host loop, dense TSDF integration and raycast,
// and the per-frame orchestration of RunTracking.  Each function names the
// reference file:line it follows.  Exported with C linkage in two flavours:
// *_f (float, reference-compatible) and *_d (FP64 oracle).

#include "kinfu_core.h"

#include <chrono>
#include <type_traits>


// ============================================================ C interface
using namespace oracle;

namespace {
template <class T> Cam<T> make_cam(const float* c10, int W, int H) {
    // c10: fx, fy, cx, cy, dmin, dmax (+ unused)
    Cam<T> c;
    c.fx = (T)c10[0];
    c.fy = (T)c10[1];
    c.fx_inv = T(1) / c.fx;
    c.fy_inv = T(1) / c.fy;
    c.cx = (T)c10[2];
    c.cy = (T)c10[3];
    c.W = W;
    c.H = H;
    c.dmin = (T)c10[4];
    c.dmax = (T)c10[5];
    return c;
}
template <class T> Tsdf<T> make_tsdf(int N, float dim, float trunc) {
    Tsdf<T> p;
    for (int k = 0; k < 3; ++k) {
        p.res[k] = N;
        p.dim[k] = (T)dim;
        p.cell[k] = (T)dim / (T)N;
    }
    // OCL/NuiKinfuOpenCLScene.cpp:41: trunc >= 2.1 * max cell
    T c = std::max(p.cell[0], std::max(p.cell[1], p.cell[2]));
    p.trunc = std::max((T)trunc, (T)2.1f * c);
    return p;
}
}  // namespace

#define OR_EXPORT extern "C" __attribute__((visibility("default")))

// ---- per-stage entry points, float (compat) flavour
OR_EXPORT void or_passing_filter_f(const uint16_t* raw, float* out, long long n, float dmin, float dmax) {
    passing_filter<float>(raw, out, (size_t)n, dmin, dmax);
}
OR_EXPORT void or_passing_filter_d(const uint16_t* raw, double* out, long long n, double dmin, double dmax) {
    passing_filter<double>(raw, out, (size_t)n, dmin, dmax);
}
OR_EXPORT void or_gaussian_table_f(int radius, float sigma, float* g) { gaussian_table<float>(radius, sigma, g); }
OR_EXPORT void or_bilateral_f(const float* in, float* out, int W, int H, const float* g, int radius, float sd,
                              float thr) {
    bilateral<float>(in, out, W, H, g, radius, sd, thr);
}
OR_EXPORT void or_bilateral_d(const double* in, double* out, int W, int H, const double* g, int radius, double sd,
                              double thr) {
    bilateral<double>(in, out, W, H, g, radius, sd, thr);
}
OR_EXPORT void or_half_sample_f(const float* s, float* d, int Wd, int Hd, int radius, float thr) {
    half_sample<float>(s, d, Wd, Hd, radius, thr);
}
OR_EXPORT void or_half_sample_d(const double* s, double* d, int Wd, int Hd, int radius, double thr) {
    half_sample<double>(s, d, Wd, Hd, radius, thr);
}
OR_EXPORT void or_depth2vertex_f(const float* depth, float* vmap, int W, int H, const float* cam6, int Wfull,
                                 int Hfull, int div) {
    depth2vertex<float>(depth, vmap, W, H, make_cam<float>(cam6, Wfull, Hfull), div);
}
OR_EXPORT void or_depth2vertex_d(const double* depth, double* vmap, int W, int H, const float* cam6, int Wfull,
                                 int Hfull, int div) {
    depth2vertex<double>(depth, vmap, W, H, make_cam<double>(cam6, Wfull, Hfull), div);
}
OR_EXPORT void or_normals_simple_f(const float* v, float* n, int W, int H, float thr) {
    normals_simple<float>(v, n, W, H, thr);
}
OR_EXPORT void or_normals_simple_d(const double* v, double* n, int W, int H, double thr) {
    normals_simple<double>(v, n, W, H, thr);
}
OR_EXPORT void or_transform_maps_f(const float* v, const float* n, float* vd, float* nd, long long npx,
                                   const float* R9, const float* t3) {
    transform_maps<float>(v, n, vd, nd, (size_t)npx, make_rigid<float>(R9, t3));
}
OR_EXPORT void or_scale_depths_f(float* depth, int W, int H, const float* cam6) {
    scale_depths<float>(depth, W, H, make_cam<float>(cam6, W, H));
}
OR_EXPORT void or_scale_depths_d(double* depth, int W, int H, const float* cam6) {
    scale_depths<double>(depth, W, H, make_cam<double>(cam6, W, H));
}
// 29 ICP sums (reference reduction order) at level `level` for pose (R9,t3)
OR_EXPORT void or_icp_sums_f(const float* cam6, int W, int H, const float* vcur, int level, const float* R9,
                             const float* t3, const float* vref, const float* nref, const float* prevR9,
                             const float* prevt3, float distThr, double* sums29) {
    Cam<float> cam = make_cam<float>(cam6, W, H);
    std::vector<float> blocks;
    int nb = icp_reduce<float>(cam, vcur, W >> level, H >> level, R9, t3, vref, nref,
                               make_rigid<float>(prevR9, prevt3), distThr, blocks);
    for (int k = 0; k < kTerms; ++k) {
        float acc = 0;
        for (int n = 0; n < nb; ++n) acc += blocks[(size_t)n * kTerms + k];
        sums29[k] = acc;
    }
}
OR_EXPORT void or_icp_sums_d(const float* cam6, int W, int H, const double* vcur, int level, const double* R9,
                             const double* t3, const double* vref, const double* nref, const double* prevR9,
                             const double* prevt3, double distThr, double* sums29) {
    Cam<double> cam = make_cam<double>(cam6, W, H);
    std::vector<double> blocks;
    int nb = icp_reduce<double>(cam, vcur, W >> level, H >> level, R9, t3, vref, nref,
                                make_rigid<double>(prevR9, prevt3), distThr, blocks);
    for (int k = 0; k < kTerms; ++k) {
        double acc = 0;
        for (int n = 0; n < nb; ++n) acc += blocks[(size_t)n * kTerms + k];
        sums29[k] = acc;
    }
}
// Eigen-equivalent helpers exposed for their own closed-form tests
OR_EXPORT int or_cholesky6(const double* A, const double* b, double* x, double* det) {
    return cholesky6(A, b, x, det) ? 1 : 0;
}
OR_EXPORT double or_det6(const double* A) { return det6(A); }
OR_EXPORT void or_axis_rot_d(double angle, int axis, double* R9) { axis_rot<double>(angle, axis, R9); }
OR_EXPORT void or_inverse3_d(const double* m, double* out) { inverse3<double>(m, out); }
OR_EXPORT double or_rodrigues_norm_d(const double* R9) { return rodrigues_norm<double>(R9); }

// ================================================================ engine

namespace {

struct EngineCfg {
    int W, H;
    float cam[6];          // fx, fy, cx, cy, dmin, dmax
    int filter_radius;
    float sigma_space, sigma_depth, depth_threshold, dist_threshold;
    int levels;
    int iters[8];
    int N;                 // dense volume resolution (cubic)
    float dim;             // metres
    float trunc;           // metres (before the 2.1-cell clamp)
    float gate;            // integration metric threshold (0: integrate every frame)
    int max_color_weight;
    int voxel_shift;       // > 0: shifting volume (VO/NuiKinfuVolumeConfig.cpp:13 voxel_shift)
    float basis[3];        // translateBasis
    int icp_mode = 0;      // 0: the reference's ICP loop (icp), 1: the compensated mode (icp_compensated),
                           // 2: the joint order (icp_compensated with photometric rows; float engines)
    float photo[3] = {0.1f, 0.2f, 0.002f};  // joint order: weight, intensity difference, minimum gradient
};

// OCL/NuiKinfuOpenCLShiftScene.cpp:119-245 fetchSlice reads the slab with
// K/extract.cl:80-190; the restatement lives in points_oracle.cpp.
// K/transformMaps.cl:372-384 intensities in double (the FP64 engine's joint order)
void intensities_d(const uint8_t* bgra, double* out, size_t n) {
    for (size_t i = 0; i < n; ++i) {
        const uint8_t* q = bgra + 4 * i;
        out[i] = q[3] > 0 ? 0.299 * (q[2] / 255.0) + 0.587 * (q[1] / 255.0) + 0.114 * (q[0] / 255.0)
                          : qnan<double>();
    }
}
// K/transformMaps.cl:448-511 derivatives in double: Sobel / 8 at the centre, NaN on the
// border and next to a NaN intensity (as the float restatement in color_oracle.cpp)
void derivatives_d(const double* I, double* out, int W, int H) {
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            const size_t q = (size_t)y * W + x;
            out[2 * q] = out[2 * q + 1] = qnan<double>();
            if (std::isnan(I[q]) || !(x > 0 && x < W - 1 && y > 0 && y < H - 1)) continue;
            auto at = [&](int dx, int dy) { return I[(size_t)(y + dy) * W + x + dx]; };
            const double i00 = at(-1, -1), i10 = at(0, -1), i20 = at(1, -1), i01 = at(-1, 0), i21 = at(1, 0),
                         i02 = at(-1, 1), i12 = at(0, 1), i22 = at(1, 1);
            if (std::isnan(i00) || std::isnan(i10) || std::isnan(i20) || std::isnan(i01) || std::isnan(i21) ||
                std::isnan(i02) || std::isnan(i12) || std::isnan(i22))
                continue;
            out[2 * q] = ((i20 - i00) + 2.0 * (i21 - i01) + (i22 - i02)) / 8.0;
            out[2 * q + 1] = ((i02 - i00) + 2.0 * (i12 - i10) + (i22 - i20)) / 8.0;
        }
}

// K/transformMaps.cl:372-384 and :448-511, restated in color_oracle.cpp.
extern "C" void or_bgra_to_intensity(const uint8_t* bgra, float* out, long long n);
extern "C" void or_intensity_derivatives(const float* I, float* out, int W, int H);

extern "C" long long or_points_dense_slice(const void* vol, const uint8_t* cvol, int storage, int Nx, int Ny, int Nz,
                                           const float* cell, const int* wrap, const int* offset, const int* range,
                                           float* pts, float* cols, long long cap);

// Per-frame orchestration of src/SLAM/VisualOdometry/NuiKinfuTrackingEngine.cpp:68-142
// with the OpenCL frame / feedback frame / depth tracker / dense scene.
template <class T, class Vol> struct Engine {
    EngineCfg c;
    Cam<T> cam;
    Tsdf<T> prm;
    Vol vol;
    std::vector<uint8_t> cvol;
    std::vector<T> g, depth, filt, vref, nref;
    std::vector<std::vector<T>> dpyr, vpyr;
    T R[9], t[3];
    T Rlast[9], tlast[3];
    double Rd[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, td[3] = {0, 0, 0};  // icp_mode 1: the pose in FP64
    int frames = 0;
    IcpResult last{1, 0, 0, 0};
    long long nupd_last = 0;
    bool integrated_last = false;
    // shifting volume: m_voxel_offset and the evicted points (NuiKinfuVertexCache)
    int off[3] = {0, 0, 0};
    std::vector<float> cache_pts, cache_cols;
    // the feedback colours (first frame's, then the raycast's); joint order: the previous
    // frame's colours (the photometric model), the frame's intensity pyramid, the model's
    // intensities and their gradients
    std::vector<uint8_t> cfb, prevc;
    std::vector<std::vector<float>> ipyr;
    std::vector<float> iref, gref;
    std::vector<std::vector<double>> ipyrd;  // the FP64 engine's
    std::vector<double> irefd, grefd;
    // host seconds per stage summed over the frames run: preprocess, ICP (pyramid
    // included), integration, raycast
    double stage_s[4] = {0, 0, 0, 0};
    using Clock = std::chrono::steady_clock;
    static double since(Clock::time_point t0) {
        return std::chrono::duration<double>(Clock::now() - t0).count();
    }

    explicit Engine(const EngineCfg& cfg) : c(cfg) {
        cam = make_cam<T>(c.cam, c.W, c.H);
        prm = make_tsdf<T>(c.N, c.dim, c.trunc);
        vol.init((size_t)c.N * c.N * c.N);
        g.resize(2 * c.filter_radius + 1);
        {
            // host table in float as in OCL/NuiKinfuOpenCLFrame.cpp:181-186
            std::vector<float> gf(g.size());
            gaussian_table<float>(c.filter_radius, c.sigma_space, gf.data());
            for (size_t i = 0; i < g.size(); ++i) g[i] = (T)gf[i];
        }
        const size_t n = (size_t)c.W * c.H;
        depth.resize(n);
        filt.resize(n);
        vref.resize(3 * n);
        nref.resize(3 * n);
        dpyr.resize(c.levels);
        vpyr.resize(c.levels);
        for (int l = 0; l < c.levels; ++l) {
            size_t nl = (size_t)(c.W >> l) * (c.H >> l);
            dpyr[l].resize(nl);
            vpyr[l].resize(3 * nl);
        }
        const T I[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
        std::memcpy(R, I, sizeof(I));
        std::memcpy(Rlast, I, sizeof(I));
        t[0] = t[1] = t[2] = 0;
        tlast[0] = tlast[1] = tlast[2] = 0;
    }

    // OCL/NuiKinfuOpenCLShiftScene.cpp:44-56 getVoxelWrap
    Wrap3 wrap() const {
        Wrap3 w;
        for (int k = 0; k < 3; ++k) w.w[k] = off[k] < 0 ? c.N - ((-off[k]) % c.N) : off[k];
        return w;
    }
    // getVoxelOffsetSize: voxel offset in metres
    T offset_size(int k) const { return (T)off[k] * prm.cell[k]; }

    // OCL/NuiKinfuOpenCLShiftScene.cpp:247-287 fetchAndClear{X,Y,Z}
    void fetch_and_clear(int axis, int trans) {
        int o = off[axis];
        if (trans < 0) o += trans;
        o = o >= 0 ? o % c.N : c.N - (-o) % c.N;
        int w[3] = {0, 0, 0}, r[3] = {c.N, c.N, c.N};
        w[axis] = o;
        r[axis] = std::abs(trans);
        // fetchSlice: the slab's points into the cache (float volume of the compat engine only)
        if (std::is_same<Vol, VolS16>::value) {
            const float cell[3] = {(float)prm.cell[0], (float)prm.cell[1], (float)prm.cell[2]};
            const void* vp = static_cast<const void*>(volume_words());
            const long long n = or_points_dense_slice(vp, cvol.empty() ? nullptr : cvol.data(), 0, c.N, c.N, c.N, cell,
                                                      w, off, r, nullptr, nullptr, 0);
            const size_t at = cache_pts.size() / 3;
            cache_pts.resize(3 * (at + (size_t)n));
            cache_cols.resize(4 * (at + (size_t)n));
            or_points_dense_slice(vp, cvol.empty() ? nullptr : cvol.data(), 0, c.N, c.N, c.N, cell, w, off, r,
                                  cache_pts.data() + 3 * at, cache_cols.data() + 4 * at, n);
        }
        // clearSlice: initializeVolume over the slab (K/tsdfVolume.cl:192-211)
        Wrap3 cw;
        for (int k = 0; k < 3; ++k) cw.w[k] = w[k];
        for (int y = 0; y < r[1]; ++y)
            for (int x = 0; x < r[0]; ++x)
                for (int z = 0; z < r[2]; ++z) {
                    const size_t i = stored_index(x, y, z, c.N, c.N, c.N, cw);
                    vol.reset_voxel(i);
                    if (!cvol.empty()) std::memset(&cvol[4 * i], 0, 4);
                }
        off[axis] += trans;
    }
    const void* volume_words() const { return vol.words(); }

    // OCL/NuiKinfuOpenCLShiftScene.cpp:289-351 shiftVolume, applied to the local translation t
    void shift_volume(T* tr) {
        T sh[3];
        int vt[3];
        for (int k = 0; k < 3; ++k) {
            sh[k] = tr[k] - (T)c.basis[k];
            const int f = (int)std::floor(sh[k] / prm.cell[k]);
            vt[k] = sh[k] < 0 ? std::max(-c.voxel_shift, f) : std::min(c.voxel_shift, f);
        }
        for (int k = 0; k < 3; ++k)
            if (std::abs(vt[k]) >= c.voxel_shift) {
                fetch_and_clear(k, vt[k]);
                sh[k] -= (T)vt[k] * prm.cell[k];
            }
        for (int k = 0; k < 3; ++k) tr[k] = sh[k] + (T)c.basis[k];
    }

    // NuiKinfuOpenCLShiftScene::integrateVolume (:353-364) around the dense integration
    void integrate_frame(const uint8_t* bgra) {
        if (c.voxel_shift > 0) shift_volume(t);
        scale_depths<T>(depth.data(), c.W, c.H, cam);
        nupd_last = integrate<T, Vol>(depth.data(), bgra, vol, bgra ? cvol.data() : nullptr, prm, cam,
                                      make_rigid<T>(R, t), c.max_color_weight, wrap());
        integrated_last = true;
    }

    // The joint order's ICP: intensities of the frame's colours per level and of the
    // previous frame's colours with their gradients, then icp_compensated with photometric
    // rows.
    IcpResult joint_icp(const IcpConfig& ic, const T* const* vp, const Rigid<T>& pose, const uint8_t* bgra, double Rn[9],
                        double tn[3]) {
        if constexpr (!std::is_same<T, float>::value) {
            // FP64: the same inputs and rows evaluated in double
            if (!bgra || prevc.empty()) return IcpResult{0, 0, 0, 0};
            const size_t n = (size_t)c.W * c.H;
            ipyrd.resize(c.levels);
            ipyrd[0].resize(n);
            intensities_d(bgra, ipyrd[0].data(), n);
            for (int l = 1; l < c.levels; ++l) {
                ipyrd[l].resize((size_t)(c.W >> l) * (c.H >> l));
                intensity_half_sample<double>(ipyrd[l - 1].data(), ipyrd[l].data(), c.W >> l, c.H >> l);
            }
            irefd.resize(n);
            grefd.resize(2 * n);
            intensities_d(prevc.data(), irefd.data(), n);
            prevc.assign(bgra, bgra + 4 * n);  // the next frame's model, whatever this ICP's outcome
            derivatives_d(irefd.data(), grefd.data(), c.W, c.H);
            const double mg = (double)c.photo[2];
            PhotoIn<double> ph{nullptr, irefd.data(), grefd.data(), std::sqrt((double)c.photo[0]), (double)c.photo[1],
                               mg * mg};
            std::vector<const double*> ip(c.levels);
            for (int l = 0; l < c.levels; ++l) ip[l] = ipyrd[l].data();
            return icp_compensated<double>(cam, ic, vp, vref.data(), nref.data(), pose, Rn, tn, &ph, ip.data());
        } else {
            if (!bgra || prevc.empty()) return IcpResult{0, 0, 0, 0};  // the joint order needs colours
            const size_t n = (size_t)c.W * c.H;
            ipyr.resize(c.levels);
            ipyr[0].resize(n);
            or_bgra_to_intensity(bgra, ipyr[0].data(), (long long)n);
            for (int l = 1; l < c.levels; ++l) {
                ipyr[l].resize((size_t)(c.W >> l) * (c.H >> l));
                intensity_half_sample<float>(ipyr[l - 1].data(), ipyr[l].data(), c.W >> l, c.H >> l);
            }
            iref.resize(n);
            gref.resize(2 * n);
            or_bgra_to_intensity(prevc.data(), iref.data(), (long long)n);
            prevc.assign(bgra, bgra + 4 * n);  // the next frame's model, whatever this ICP's outcome
            or_intensity_derivatives(iref.data(), gref.data(), c.W, c.H);
            PhotoIn<float> ph{nullptr, iref.data(), gref.data(), std::sqrt(c.photo[0]), c.photo[1],
                              c.photo[2] * c.photo[2]};
            std::vector<const float*> ip(c.levels);
            for (int l = 0; l < c.levels; ++l) ip[l] = ipyr[l].data();
            return icp_compensated<float>(cam, ic, vp, vref.data(), nref.data(), pose, Rn, tn, &ph, ip.data());
        }
    }

    // returns 1 ok, 0 tracking failure
    int run(const uint16_t* raw, const uint8_t* bgra) {
        const int W = c.W, H = c.H;
        const size_t n = (size_t)W * H;
        auto t0 = Clock::now();
        // NuiKinfuOpenCLFrame::UpdateVertexBuffers (OCL/NuiKinfuOpenCLFrame.cpp:314-367)
        passing_filter<T>(raw, depth.data(), n, cam.dmin, cam.dmax);
        bilateral<T>(depth.data(), dpyr[0].data(), W, H, g.data(), c.filter_radius, (T)c.sigma_depth,
                     (T)c.depth_threshold);
        depth2vertex<T>(dpyr[0].data(), vpyr[0].data(), W, H, cam, 1);
        stage_s[0] += since(t0);
        const Rigid<T> pose = make_rigid<T>(R, t);
        integrated_last = false;
        if (frames == 0) {
            // NuiKinfuOpenCLFeedbackFrame::UpdateBuffers: normals + camera->world
            std::vector<T> nm(3 * n);
            normals_simple<T>(vpyr[0].data(), nm.data(), W, H, (T)c.depth_threshold);
            transform_maps<T>(vpyr[0].data(), nm.data(), vref.data(), nref.data(), n, pose);
            if (bgra) cfb.assign(bgra, bgra + 4 * n);
            if (bgra && c.icp_mode == 2) prevc.assign(bgra, bgra + 4 * n);
            integrate_frame(bgra);
        } else {
            t0 = Clock::now();
            // NuiKinfuOpenCLDepthTracker::EstimatePose: pyramid, vertices, ICP
            for (int l = 1; l < c.levels; ++l) {
                half_sample<T>(dpyr[l - 1].data(), dpyr[l].data(), W >> l, H >> l, 1, (T)c.depth_threshold);
                depth2vertex<T>(dpyr[l].data(), vpyr[l].data(), W >> l, H >> l, cam, 1 << l);
            }
            IcpConfig ic;
            ic.levels = c.levels;
            for (int l = 0; l < 8; ++l) ic.iters[l] = c.iters[l];
            ic.dist_threshold = c.dist_threshold;
            std::vector<const T*> vp(c.levels);
            for (int l = 0; l < c.levels; ++l) vp[l] = vpyr[l].data();
            if (c.icp_mode == 1 || c.icp_mode == 2) {
                double Rn[9], tn[3];
                std::memcpy(Rn, Rd, sizeof(Rn));
                std::memcpy(tn, td, sizeof(tn));
                if (c.icp_mode == 2)
                    last = joint_icp(ic, vp.data(), pose, bgra, Rn, tn);
                else
                    last = icp_compensated<T>(cam, ic, vp.data(), vref.data(), nref.data(), pose, Rn, tn);
                if (!last.ok) return 0;
                std::memcpy(Rd, Rn, sizeof(Rn));
                std::memcpy(td, tn, sizeof(tn));
                for (int k = 0; k < 9; ++k) R[k] = (T)Rd[k];
                for (int k = 0; k < 3; ++k) t[k] = (T)td[k];
            } else {
                T Rn[9], tn[3];
                std::memcpy(Rn, R, sizeof(Rn));
                std::memcpy(tn, t, sizeof(tn));
                last = icp<T>(cam, ic, vp.data(), vref.data(), nref.data(), pose, Rn, tn);
                if (!last.ok) return 0;
                std::memcpy(R, Rn, sizeof(Rn));
                std::memcpy(t, tn, sizeof(tn));
            }
            stage_s[1] += since(t0);
            t0 = Clock::now();
            // integration gate (VO/NuiKinfuTrackingEngine.cpp:113-127) on global translations
            // (local + the shifting volume's offset; NuiCameraPos::getGlobalTranslation)
            T Rinv[9], Rrel[9];
            inverse3(R, Rinv);
            matmul3(Rinv, Rlast, Rrel);
            T rnorm = rodrigues_norm<T>(Rrel);
            T d[3];
            for (int k = 0; k < 3; ++k) d[k] = (t[k] + offset_size(k)) - (tlast[k] + offset_size(k));
            T tnorm = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
            if ((rnorm + tnorm) / T(2) >= (T)c.gate) {
                integrate_frame(bgra);  // may shift the volume and move t (local)
                if (c.voxel_shift > 0)
                    for (int k = 0; k < 3; ++k) td[k] = (double)t[k];  // the shift moved the local frame
                std::memcpy(Rlast, R, sizeof(Rlast));
                std::memcpy(tlast, t, sizeof(tlast));
            }
            stage_s[2] += since(t0);
            t0 = Clock::now();
            const Rigid<T> np = make_rigid<T>(R, t);
            if (bgra && !cvol.empty() && cfb.size() != 4 * n) cfb.assign(4 * n, 0);
            raycast<T, Vol>(vol, bgra ? cvol.data() : nullptr, prm, cam, np, vref.data(), nref.data(),
                            bgra && !cvol.empty() ? cfb.data() : nullptr, wrap());
            stage_s[3] += since(t0);
        }
        frames++;
        return 1;
    }
};

template <class T, class Vol> struct EngineBox {
    Engine<T, Vol> e;
    explicit EngineBox(const EngineCfg& c) : e(c) {}
};

}  // namespace

// Engine handle API.  prec: 0 = float compat with the reference's short2 volume,
// 1 = FP64 oracle, 2 = float with FP16 tsdf + uint16 weight, 3 = float with FP32
// tsdf + FP32 weight (the GPU's storages).
struct OrEngine {
    int prec;
    void* p;
};

template <class F> auto with_engine(OrEngine* h, F&& f) {
    switch (h->prec) {
        case 0: return f(((EngineBox<float, VolS16>*)h->p)->e);
        case 2: return f(((EngineBox<float, VolF16>*)h->p)->e);
        case 3: return f(((EngineBox<float, VolF32>*)h->p)->e);
        default: return f(((EngineBox<double, VolF64>*)h->p)->e);
    }
}

OR_EXPORT void* or_engine_create(int prec, int W, int H, const float* cam6, int filter_radius, float sigma_space,
                                 float sigma_depth, float depth_threshold, float dist_threshold, int levels,
                                 const int* iters, int N, float dim, float trunc, float gate, int color) {
    EngineCfg c;
    c.W = W;
    c.H = H;
    std::memcpy(c.cam, cam6, sizeof(c.cam));
    c.filter_radius = filter_radius;
    c.sigma_space = sigma_space;
    c.sigma_depth = sigma_depth;
    c.depth_threshold = depth_threshold;
    c.dist_threshold = dist_threshold;
    c.levels = levels;
    for (int l = 0; l < 8; ++l) c.iters[l] = l < levels ? iters[l] : 0;
    c.N = N;
    c.dim = dim;
    c.trunc = trunc;
    c.gate = gate;
    c.max_color_weight = 128;
    c.voxel_shift = 0;
    c.basis[0] = c.basis[1] = c.basis[2] = 0.f;
    OrEngine* h = new OrEngine;
    h->prec = prec;
    switch (prec) {
        case 0: h->p = new EngineBox<float, VolS16>(c); break;
        case 2: h->p = new EngineBox<float, VolF16>(c); break;
        case 3: h->p = new EngineBox<float, VolF32>(c); break;
        default: h->prec = 1; h->p = new EngineBox<double, VolF64>(c); break;
    }
    with_engine(h, [&](auto& e) {
        if (color) e.cvol.assign((size_t)N * N * N * 4, 0);
        return 0;
    });
    return h;
}

OR_EXPORT void or_engine_destroy(void* hp) {
    OrEngine* h = (OrEngine*)hp;
    if (!h) return;
    switch (h->prec) {
        case 0: delete (EngineBox<float, VolS16>*)h->p; break;
        case 2: delete (EngineBox<float, VolF16>*)h->p; break;
        case 3: delete (EngineBox<float, VolF32>*)h->p; break;
        default: delete (EngineBox<double, VolF64>*)h->p; break;
    }
    delete h;
}

// Runs one frame. pose12 out: R (row-major) + t as double (the FP64 pose in the
// compensated mode). Returns 1 ok / 0 failure.
OR_EXPORT int or_engine_run(void* hp, const uint16_t* raw, const uint8_t* bgra, double* pose12, double* icp_info3) {
    return with_engine((OrEngine*)hp, [&](auto& e) {
        const int ok = e.run(raw, bgra);
        if (pose12) {
            for (int k = 0; k < 9; ++k) pose12[k] = e.c.icp_mode != 0 ? e.Rd[k] : (double)e.R[k];
            for (int k = 0; k < 3; ++k) pose12[9 + k] = e.c.icp_mode != 0 ? e.td[k] : (double)e.t[k];
        }
        if (icp_info3) {
            icp_info3[0] = e.last.error;
            icp_info3[1] = e.last.count;
            icp_info3[2] = (double)e.nupd_last;
        }
        return ok;
    });
}

// ICP loop: 0 the reference's (icp), 1 the compensated mode (icp_compensated), 2 the joint
// order (float engines, colour frames); before the first frame.
OR_EXPORT void or_engine_set_icp_mode(void* hp, int mode) {
    with_engine((OrEngine*)hp, [&](auto& e) {
        e.c.icp_mode = mode;
        return 0;
    });
}

// The joint order's photometric weight, intensity difference and minimum gradient.
OR_EXPORT void or_engine_set_photometric(void* hp, float weight, float max_difference, float min_gradient) {
    with_engine((OrEngine*)hp, [&](auto& e) {
        e.c.photo[0] = weight;
        e.c.photo[1] = max_difference;
        e.c.photo[2] = min_gradient;
        return 0;
    });
}

// Shifting volume (OCL/NuiKinfuOpenCLShiftScene.cpp): set before the first frame.
OR_EXPORT void or_engine_set_shift(void* hp, int voxel_shift, const float* basis) {
    with_engine((OrEngine*)hp, [&](auto& e) {
        e.c.voxel_shift = voxel_shift;
        for (int k = 0; k < 3; ++k) e.c.basis[k] = basis ? basis[k] : 0.f;
        return 0;
    });
}
// voxel offset (int[3]) and its size in metres (double[3], the engine's arithmetic type)
OR_EXPORT void or_engine_offset(void* hp, int* off3, double* size3) {
    with_engine((OrEngine*)hp, [&](auto& e) {
        for (int k = 0; k < 3; ++k) {
            if (off3) off3[k] = e.off[k];
            if (size3) size3[k] = (double)e.offset_size(k);
        }
        return 0;
    });
}
// the evicted points (compat engine): count, and up to cap points / colours
OR_EXPORT long long or_engine_cache(void* hp, float* pts, float* cols, long long cap) {
    OrEngine* h = (OrEngine*)hp;
    if (h->prec != 0) return 0;
    auto& e = ((EngineBox<float, VolS16>*)h->p)->e;
    const long long n = (long long)(e.cache_pts.size() / 3);
    const long long m = std::min(n, cap);
    if (pts && m > 0) std::memcpy(pts, e.cache_pts.data(), (size_t)m * 3 * sizeof(float));
    if (cols && m > 0) std::memcpy(cols, e.cache_cols.data(), (size_t)m * 4 * sizeof(float));
    return n;
}

// Volume readback: tsdf (float, normalised) and weight per voxel in the
// reference layout ((y*N + x)*N + z).
OR_EXPORT void or_engine_volume(void* hp, float* tsdf, int* weight) {
    with_engine((OrEngine*)hp, [&](auto& e) {
        size_t n = (size_t)e.c.N * e.c.N * e.c.N;
        for (size_t i = 0; i < n; ++i) {
            if (tsdf) tsdf[i] = e.vol.template tsdf<float>(i);
            if (weight) weight[i] = e.vol.weight(i);
        }
        return 0;
    });
}

// Raw volume words of a float engine, in the GPU's storage layout (S16: 2 x int16,
// F16: one uint32, F32: 2 x float per voxel); returns the byte count (-1 for FP64).
OR_EXPORT long long or_engine_volume_words(void* hp, void* out) {
    OrEngine* h = (OrEngine*)hp;
    if (h->prec == 1) return -1;
    const size_t n = with_engine(h, [&](auto& e) { return (size_t)e.c.N * e.c.N * e.c.N; });
    const size_t bytes = n * (h->prec == 3 ? 8 : 4);
    if (out) with_engine(h, [&](auto& e) {
        std::memcpy(out, e.vol.words(), bytes);
        return 0;
    });
    return (long long)bytes;
}

// Raw short2 volume of the compat engine (2 x int16 per voxel).
// 1 when the last frame integrated (VO/NuiKinfuTrackingEngine.cpp:113-127 gate passed or first frame)
OR_EXPORT int or_engine_integrated(void* hp) {
    return with_engine((OrEngine*)hp, [&](auto& e) { return e.integrated_last ? 1 : 0; });
}

OR_EXPORT int or_engine_volume_raw(void* hp, int16_t* out) {
    OrEngine* h = (OrEngine*)hp;
    if (h->prec != 0) return -1;
    auto& e = ((EngineBox<float, VolS16>*)h->p)->e;
    std::memcpy(out, e.vol.v.data(), e.vol.v.size() * sizeof(int16_t));
    return 0;
}

// Host seconds per stage summed over the frames run since the last call (preprocess, ICP,
// integration, raycast), then cleared.
OR_EXPORT void or_engine_stage_seconds(void* hp, double* out4) {
    with_engine((OrEngine*)hp, [&](auto& e) {
        for (int k = 0; k < 4; ++k) {
            out4[k] = e.stage_s[k];
            e.stage_s[k] = 0;
        }
        return 0;
    });
}

// The joint order's helpers on their own, in double (for the closed-form checks).
OR_EXPORT void or_intensity_half_sample(const float* src, float* dst, int Wd, int Hd) {
    intensity_half_sample<float>(src, dst, Wd, Hd);
}
// at8: pv (3), pp (3), px, py; Rinv9: the previous camera's inverse rotation; cam6: fx, fy,
// cx, cy (+ unused); iref W*H, gref 2*W*H; returns 1 and c7 when the pixel gives a row.
OR_EXPORT int or_photometric_row_d(double ic, const double* at8, const double* Rinv9, const float* cam6, int W, int H,
                                   const double* iref, const double* gref, double psw, double pdiff, double pgrad2,
                                   double* c7) {
    Cam<double> cam = make_cam<double>(cam6, W, H);
    Rigid<double> prev{};
    for (int k = 0; k < 9; ++k) prev.R_inv[k] = Rinv9[k];
    AssocAt<double> at{mk(at8[0], at8[1], at8[2]), mk(at8[3], at8[4], at8[5]), at8[6], at8[7]};
    PhotoIn<double> ph{nullptr, iref, gref, psw, pdiff, pgrad2};
    return photometric_row<double>(ic, at, prev, cam, ph, c7) ? 1 : 0;
}

// The feedback colours (W*H BGRA) after the last frame; 0 when there are none.
OR_EXPORT int or_engine_feedback_colors(void* hp, uint8_t* out) {
    return with_engine((OrEngine*)hp, [&](auto& e) {
        if (e.cfb.empty()) return 0;
        std::memcpy(out, e.cfb.data(), e.cfb.size());
        return 1;
    });
}

// The joint order's intensity pyramid level and the feedback intensities / gradients of
// the last tracked frame (float engines); returns the element count copied.
OR_EXPORT long long or_engine_photometric(void* hp, int which, float* out) {
    return with_engine((OrEngine*)hp, [&](auto& e) -> long long {
        const std::vector<float>* v = which < 0    ? &e.iref
                                      : which == 4 ? &e.gref
                                      : which < (int)e.ipyr.size() ? &e.ipyr[which] : nullptr;
        if (!v) return 0;
        if (out) std::memcpy(out, v->data(), v->size() * sizeof(float));
        return (long long)v->size();
    });
}

// Reference maps (world vmap/nmap, float) after the last frame.
OR_EXPORT void or_engine_refmaps(void* hp, float* vmap, float* nmap) {
    with_engine((OrEngine*)hp, [&](auto& e) {
        for (size_t i = 0; i < e.vref.size(); ++i) {
            if (vmap) vmap[i] = (float)e.vref[i];
            if (nmap) nmap[i] = (float)e.nref[i];
        }
        return 0;
    });
}

// The same maps in the engine's arithmetic (FP64 engine: doubles), for error reports.
OR_EXPORT void or_engine_refmaps_d(void* hp, double* vmap, double* nmap) {
    with_engine((OrEngine*)hp, [&](auto& e) {
        for (size_t i = 0; i < e.vref.size(); ++i) {
            if (vmap) vmap[i] = (double)e.vref[i];
            if (nmap) nmap[i] = (double)e.nref[i];
        }
        return 0;
    });
}

// ---- an FP64 dense volume driven with given poses: the precision reference for the
// TSDF and raycast of a GPU session (same frames, same poses, every operation in double)
struct OrVolD {
    Tsdf<double> prm;
    VolF64 vol;
};
OR_EXPORT void* or_vold_create(int N, float dim, float trunc) {
    OrVolD* v = new OrVolD;
    v->prm = make_tsdf<double>(N, dim, trunc);
    v->vol.init((size_t)N * N * N);
    return v;
}
OR_EXPORT void or_vold_destroy(void* h) { delete (OrVolD*)h; }
// one frame: raw depth -> passing filter -> ray-length scaling -> integration at (R, t), in double
OR_EXPORT long long or_vold_integrate(void* h, const uint16_t* raw, int W, int H, const float* cam6, const float* R9,
                                      const float* t3) {
    OrVolD* v = (OrVolD*)h;
    const Cam<double> cam = make_cam<double>(cam6, W, H);
    std::vector<double> depth((size_t)W * H);
    passing_filter<double>(raw, depth.data(), depth.size(), cam.dmin, cam.dmax);
    scale_depths<double>(depth.data(), W, H, cam);
    double R[9], t[3];
    for (int k = 0; k < 9; ++k) R[k] = R9[k];
    for (int k = 0; k < 3; ++k) t[k] = t3[k];
    return integrate<double, VolF64>(depth.data(), nullptr, v->vol, nullptr, v->prm, cam, make_rigid<double>(R, t), 128);
}
OR_EXPORT void or_vold_volume(void* h, double* tsdf, int* weight) {
    OrVolD* v = (OrVolD*)h;
    const size_t n = v->vol.f.size();
    for (size_t i = 0; i < n; ++i) {
        if (tsdf) tsdf[i] = v->vol.f[i];
        if (weight) weight[i] = v->vol.w[i];
    }
}
OR_EXPORT void or_vold_raycast(void* h, int W, int H, const float* cam6, const float* R9, const float* t3, double* vmap,
                               double* nmap) {
    OrVolD* v = (OrVolD*)h;
    double R[9], t[3];
    for (int k = 0; k < 9; ++k) R[k] = R9[k];
    for (int k = 0; k < 3; ++k) t[k] = t3[k];
    raycast<double, VolF64>(v->vol, nullptr, v->prm, make_cam<double>(cam6, W, H), make_rigid<double>(R, t), vmap, nmap,
                            nullptr);
}

// FP64 algebra of the compensated ICP mode (shared with the GPU, se3_math.h)
OR_EXPORT int or_se3_exp(const double* w3, const double* u3, double* R9, double* t3) {
    return nuif_se3::se3_exp(w3, u3, R9, t3) ? 1 : 0;
}
OR_EXPORT void or_se3_coefficients(double th2, double* abc) { nuif_se3::se3_coefficients(th2, abc, abc + 1, abc + 2); }
OR_EXPORT int or_ldlt6_solve(const double* A36, const double* b6, double* x6, double* det) {
    double F[36];
    std::memcpy(F, A36, sizeof(F));
    if (!nuif_se3::ldlt6(F, det)) return 0;
    nuif_se3::ldlt6_solve(F, b6, x6);
    return 1;
}
OR_EXPORT int or_gn_step(const double* sums29, double step2, double* R9, double* t3) {
    return nuif_se3::gn_step(sums29, step2, R9, t3);
}
// One iteration's sums of the compensated order (float instance: the GPU's order)
OR_EXPORT void or_icp_sums_comp_f(const float* cam6, int W, int H, const float* vcur, int level, const float* R9,
                                  const float* t3, const float* vref, const float* nref, const float* prevR9,
                                  const float* prevt3, float distThr, double* sums29) {
    Cam<float> cam = make_cam<float>(cam6, W, H);
    icp_sums_compensated(cam, vcur, (W >> level) * (H >> level), R9, t3, vref, nref, make_rigid<float>(prevR9, prevt3),
                         distThr, sums29);
}

// IEEE half conversions of the F16 storage (round to nearest even; exact back)
OR_EXPORT uint16_t or_f16_from_float(float f) { return f16_from_float(f); }
OR_EXPORT float or_f16_to_float(uint16_t h) { return f16_to_float(h); }

// ---- stateless dense stages on caller-provided short2 volumes (float compat)
OR_EXPORT long long or_integrate_s16_f(const float* depth_scaled, const uint8_t* bgra, int16_t* vol2,
                                       uint8_t* color_vol, int N, float dim, float trunc, const float* cam6, int W,
                                       int H, const float* R9, const float* t3, int max_color_weight) {
    VolS16 v;
    v.v.assign(vol2, vol2 + 2 * (size_t)N * N * N);
    long long nupd = integrate<float, VolS16>(depth_scaled, bgra, v, color_vol, make_tsdf<float>(N, dim, trunc),
                                              make_cam<float>(cam6, W, H), make_rigid<float>(R9, t3),
                                              max_color_weight);
    std::memcpy(vol2, v.v.data(), v.v.size() * sizeof(int16_t));
    return nupd;
}
OR_EXPORT void or_raycast_s16_f(const int16_t* vol2, const uint8_t* color_vol, int N, float dim, float trunc,
                                const float* cam6, int W, int H, const float* R9, const float* t3, float* vmap,
                                float* nmap, uint8_t* cmap) {
    VolS16 v;
    v.v.assign(vol2, vol2 + 2 * (size_t)N * N * N);
    raycast<float, VolS16>(v, color_vol, make_tsdf<float>(N, dim, trunc), make_cam<float>(cam6, W, H),
                           make_rigid<float>(R9, t3), vmap, nmap, cmap);
}
// the same in a wrapped volume (wrap: int[3], >= 0)
OR_EXPORT long long or_integrate_s16_wrap_f(const float* depth_scaled, const uint8_t* bgra, int16_t* vol2,
                                            uint8_t* color_vol, int N, float dim, float trunc, const float* cam6,
                                            int W, int H, const float* R9, const float* t3, int max_color_weight,
                                            const int* wrap) {
    VolS16 v;
    v.v.assign(vol2, vol2 + 2 * (size_t)N * N * N);
    Wrap3 wr;
    for (int k = 0; k < 3; ++k) wr.w[k] = wrap[k];
    long long nupd = integrate<float, VolS16>(depth_scaled, bgra, v, color_vol, make_tsdf<float>(N, dim, trunc),
                                              make_cam<float>(cam6, W, H), make_rigid<float>(R9, t3),
                                              max_color_weight, wr);
    std::memcpy(vol2, v.v.data(), v.v.size() * sizeof(int16_t));
    return nupd;
}
OR_EXPORT void or_raycast_s16_wrap_f(const int16_t* vol2, const uint8_t* color_vol, int N, float dim, float trunc,
                                     const float* cam6, int W, int H, const float* R9, const float* t3, float* vmap,
                                     float* nmap, uint8_t* cmap, const int* wrap) {
    VolS16 v;
    v.v.assign(vol2, vol2 + 2 * (size_t)N * N * N);
    Wrap3 wr;
    for (int k = 0; k < 3; ++k) wr.w[k] = wrap[k];
    raycast<float, VolS16>(v, color_vol, make_tsdf<float>(N, dim, trunc), make_cam<float>(cam6, W, H),
                           make_rigid<float>(R9, t3), vmap, nmap, cmap, wr);
}
OR_EXPORT float or_effective_trunc(int N, float dim, float trunc) { return make_tsdf<float>(N, dim, trunc).trunc; }
// K/raycast.cl:36-120 interpolateTrilineary at world point p3 of a short2 volume
OR_EXPORT float or_trilinear_s16_f(const int16_t* vol2, int N, float dim, float trunc, const float* p3) {
    VolS16 v;
    v.v.assign(vol2, vol2 + 2 * (size_t)N * N * N);
    V3<float> p;
    p.x = p3[0];
    p.y = p3[1];
    p.z = p3[2];
    return trilinear<float, VolS16>(p, v, make_tsdf<float>(N, dim, trunc));
}

// Reproducible float exp / sin / cos (exposed for their accuracy tests)
OR_EXPORT float or_expf_det(float x) { return nuif_det::expf_det(x); }
OR_EXPORT void or_sincosf_det(float a, float* s, float* c) { nuif_det::sincosf_det(a, s, c); }
