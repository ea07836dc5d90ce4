// Deterministic synthetic RGB-D workload generator.
//
// Renders uint16 depth (mm) and BGRA colour frames of analytic scenes (ground
// plane / room walls, spheres, axis-aligned boxes) by exact ray/primitive
// intersection, which is what sphere tracing of the same SDF union converges
// to.  Every random draw is a counter-based hash of (seed, stream, frame,
// index), so a frame is reproducible on its own and independent of the thread
// order.  This is workload input, not part of the fusion pipeline.
//
// Scene conventions follow the reference tracker: world coordinates are the
// frame-0 camera frame (the tracker starts at pose (I, 0),
// src/SLAM/VisualOdometry/NuiKinfuTrackingEngine.cpp:46-58), y points down,
// the camera looks along +z.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>

extern "C" {

typedef struct nsynth_params {
    int width, height;
    float fx, fy, cx, cy;
    int scene_kind;        // 0: ground plane + 6 spheres + 4 boxes (orbit camera)
                           // 1: 8x8x3 m room + 40 objects (random-walk camera)
    float orbit_radius;    // kind 0: camera orbit radius (m); orbit centre (0,0,r)
    float deg_per_frame;   // kind 0: orbit rate
    float cube_size;       // kind 0: objects placed uniformly in a cube of this side
    float jitter_pos;      // kind 0: N(0, jitter_pos) m per axis (not on frame 0)
    float jitter_deg;      // kind 0: N(0, jitter_deg) deg per axis (not on frame 0)
    float hole_frac;       // fraction of zero-depth holes
    int noise;             // 1: Nguyen-style axial noise + mm quantisation
    unsigned long long seed;
} nsynth_params;

}  // extern "C"

namespace {

inline uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

inline uint64_t key(uint64_t seed, uint64_t stream, uint64_t frame, uint64_t idx) {
    uint64_t h = splitmix64(seed ^ (stream * 0xD1B54A32D192ED03ull));
    h = splitmix64(h ^ (frame * 0x8CB92BA72F3D8DD7ull));
    return splitmix64(h ^ idx);
}

inline double u01(uint64_t h) { return (double)(h >> 11) * (1.0 / 9007199254740992.0); }

inline double gauss(uint64_t h) {  // Box-Muller on two derived uniforms
    double u1 = u01(splitmix64(h ^ 0x1234567ull));
    double u2 = u01(splitmix64(h ^ 0x89ABCDEFull));
    if (u1 < 1e-300) u1 = 1e-300;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

struct V3 { double x, y, z; };
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline V3 normalize(V3 a) { double n = std::sqrt(dot(a, a)); return a * (1.0 / n); }

enum PrimType { PLANE = 0, SPHERE = 1, BOX = 2, INNER_BOX = 3 };

struct Prim {
    int type;
    V3 a;        // plane normal / sphere centre / box min
    V3 b;        // box max
    double r;    // sphere radius / plane offset (n.p + r = 0)
    V3 albedo;   // linear RGB in [0,1]
};

// Ray/primitive intersection along o + t*d (t > eps), returns +inf on miss.
double intersect(const Prim& p, V3 o, V3 d) {
    const double inf = INFINITY;
    if (p.type == PLANE) {
        double den = dot(p.a, d);
        if (std::fabs(den) < 1e-12) return inf;
        double t = -(dot(p.a, o) + p.r) / den;
        return t > 1e-6 ? t : inf;
    }
    if (p.type == SPHERE) {
        V3 oc = o - p.a;
        double bq = dot(oc, d), aq = dot(d, d), cq = dot(oc, oc) - p.r * p.r;
        double disc = bq * bq - aq * cq;
        if (disc < 0) return inf;
        double s = std::sqrt(disc);
        double t0 = (-bq - s) / aq, t1 = (-bq + s) / aq;
        if (t0 > 1e-6) return t0;
        if (t1 > 1e-6) return t1;
        return inf;
    }
    // slab test; INNER_BOX returns the exit distance (camera inside a room)
    double tmin = -inf, tmax = inf;
    const double oo[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z};
    const double lo[3] = {p.a.x, p.a.y, p.a.z}, hi[3] = {p.b.x, p.b.y, p.b.z};
    for (int k = 0; k < 3; ++k) {
        if (std::fabs(dd[k]) < 1e-15) {
            if (oo[k] < lo[k] || oo[k] > hi[k]) return inf;
            continue;
        }
        double t0 = (lo[k] - oo[k]) / dd[k], t1 = (hi[k] - oo[k]) / dd[k];
        if (t0 > t1) std::swap(t0, t1);
        tmin = std::max(tmin, t0);
        tmax = std::min(tmax, t1);
    }
    if (tmax < tmin) return inf;
    if (p.type == INNER_BOX) return tmax > 1e-6 ? tmax : inf;
    if (tmin > 1e-6) return tmin;
    return inf;  // origin inside an object box: treat as miss
}

struct Scene {
    std::vector<Prim> prims;
};

V3 rand_albedo(uint64_t seed, int i) {
    return {0.2 + 0.7 * u01(key(seed, 7, 0, 3 * i)), 0.2 + 0.7 * u01(key(seed, 7, 0, 3 * i + 1)),
            0.2 + 0.7 * u01(key(seed, 7, 0, 3 * i + 2))};
}

// Orbit scene: ground plane, 6 spheres r in U(0.1,0.3), 4 boxes side U(0.2,0.5),
// positions uniform in a cube of side cube_size centred on the orbit centre.
// Objects that would touch the camera path (orbit circle in the y = 0 plane)
// are re-drawn so the camera never sits inside an object.
Scene build_orbit_scene(const nsynth_params& P) {
    Scene s;
    const double R = P.orbit_radius, half = 0.5 * P.cube_size;
    const V3 C = {0.0, 0.0, R};
    // y points down: the floor is 0.4 m below the orbit plane, so it is seen from
    // about 0.8 m on and the reconstruction volume holds a large floor patch
    const double floor_y = 0.4;
    Prim ground{PLANE, {0, -1, 0}, {0, 0, 0}, floor_y, {0.6, 0.6, 0.55}};
    s.prims.push_back(ground);
    auto clear_of_orbit = [&](V3 c, double rad) {
        double rho = std::sqrt((c.x - C.x) * (c.x - C.x) + (c.z - C.z) * (c.z - C.z));
        double dy = c.y;
        double dist = std::sqrt((rho - R) * (rho - R) + dy * dy);
        return dist > rad + 0.35;
    };
    int draw = 0;
    for (int i = 0; i < 10; ++i) {
        for (;;) {
            ++draw;
            V3 c = {C.x + (2 * u01(key(P.seed, 1, draw, 0)) - 1) * half,
                    C.y + (2 * u01(key(P.seed, 1, draw, 1)) - 1) * half,
                    C.z + (2 * u01(key(P.seed, 1, draw, 2)) - 1) * half};
            if (i < 6) {
                double r = 0.1 + 0.2 * u01(key(P.seed, 1, draw, 3));
                if (!clear_of_orbit(c, r) || c.y - 0.5 * r > floor_y) continue;  // not sunk in the floor
                s.prims.push_back({SPHERE, c, {0, 0, 0}, r, rand_albedo(P.seed, i)});
            } else {
                double side = 0.2 + 0.3 * u01(key(P.seed, 1, draw, 3));
                double h = 0.5 * side;
                if (!clear_of_orbit(c, h * 1.7320508) || c.y - 0.5 * h > floor_y) continue;
                s.prims.push_back({BOX, {c.x - h, c.y - h, c.z - h}, {c.x + h, c.y + h, c.z + h}, 0,
                                   rand_albedo(P.seed, i)});
            }
            break;
        }
    }
    return s;
}

// Room scene: 8 x 8 x 3 m room (walls, floor, ceiling) with 40 boxes/spheres.
// The frame-0 camera is at the origin, 1.5 m above the floor, looking +z.
Scene build_room_scene(const nsynth_params& P) {
    Scene s;
    s.prims.push_back({INNER_BOX, {-4.0, -1.5, -4.0}, {4.0, 1.5, 4.0}, 0, {0.7, 0.7, 0.65}});
    int draw = 0;
    for (int i = 0; i < 40; ++i) {
        for (;;) {
            ++draw;
            V3 c = {-3.6 + 7.2 * u01(key(P.seed, 2, draw, 0)), -1.3 + 2.6 * u01(key(P.seed, 2, draw, 1)),
                    -3.6 + 7.2 * u01(key(P.seed, 2, draw, 2))};
            double sz = 0.15 + 0.35 * u01(key(P.seed, 2, draw, 3));
            // keep a 1.2 m clear column around the camera's walk region
            if (std::sqrt(c.x * c.x + c.z * c.z) < sz + 1.2) continue;
            if (i % 2 == 0)
                s.prims.push_back({SPHERE, c, {0, 0, 0}, sz, rand_albedo(P.seed, 100 + i)});
            else
                s.prims.push_back({BOX, {c.x - sz, c.y - sz, c.z - sz}, {c.x + sz, c.y + sz, c.z + sz}, 0,
                                   rand_albedo(P.seed, 100 + i)});
            break;
        }
    }
    return s;
}

void rot_from_vec(const double w[3], double Rm[9]) {  // Rodrigues, row-major
    double th = std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    if (th < 1e-15) {
        const double I[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
        std::memcpy(Rm, I, sizeof(I));
        return;
    }
    double k[3] = {w[0] / th, w[1] / th, w[2] / th};
    double c = std::cos(th), s = std::sin(th), v = 1 - c;
    Rm[0] = c + k[0] * k[0] * v;        Rm[1] = k[0] * k[1] * v - k[2] * s; Rm[2] = k[0] * k[2] * v + k[1] * s;
    Rm[3] = k[1] * k[0] * v + k[2] * s; Rm[4] = c + k[1] * k[1] * v;        Rm[5] = k[1] * k[2] * v - k[0] * s;
    Rm[6] = k[2] * k[0] * v - k[1] * s; Rm[7] = k[2] * k[1] * v + k[0] * s; Rm[8] = c + k[2] * k[2] * v;
}

void matmul3(const double A[9], const double B[9], double C[9]) {
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) C[3 * i + j] = A[3 * i] * B[j] + A[3 * i + 1] * B[3 + j] + A[3 * i + 2] * B[6 + j];
}

// Camera-to-world pose of frame f: world = R * cam + t.
void camera_pose(const nsynth_params& P, int f, double Rm[9], double t[3]) {
    if (P.scene_kind == 0) {
        const double th = P.deg_per_frame * f * M_PI / 180.0, R = P.orbit_radius;
        const V3 C = {0, 0, R};
        V3 pos = {C.x - R * std::sin(th), 0.0, C.z - R * std::cos(th)};
        V3 z = {std::sin(th), 0.0, std::cos(th)};
        V3 y = {0, 1, 0};
        V3 x = cross(y, z);
        double Ro[9] = {x.x, y.x, z.x, x.y, y.y, z.y, x.z, y.z, z.z};
        if (f > 0) {
            double w[3];
            for (int k = 0; k < 3; ++k) w[k] = gauss(key(P.seed, 3, f, k)) * P.jitter_deg * M_PI / 180.0;
            double Rj[9];
            rot_from_vec(w, Rj);
            matmul3(Ro, Rj, Rm);
            pos.x += gauss(key(P.seed, 3, f, 3)) * P.jitter_pos;
            pos.y += gauss(key(P.seed, 3, f, 4)) * P.jitter_pos;
            pos.z += gauss(key(P.seed, 3, f, 5)) * P.jitter_pos;
        } else {
            std::memcpy(Rm, Ro, sizeof(Ro));
        }
        t[0] = pos.x; t[1] = pos.y; t[2] = pos.z;
        return;
    }
    // Random walk: integrate seeded per-frame steps (0.5-1.5 cm, 0.5 deg) from frame 0.
    double Rc[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, pc[3] = {0, 0, 0};
    for (int k = 1; k <= f; ++k) {
        double step = 0.005 + 0.010 * u01(key(P.seed, 4, k, 0));
        double ang = 2.0 * M_PI * u01(key(P.seed, 4, k, 1));
        double dir[3] = {std::cos(ang), 0.2 * (2 * u01(key(P.seed, 4, k, 2)) - 1), std::sin(ang)};
        double n = std::sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
        for (int a = 0; a < 3; ++a) pc[a] += step * dir[a] / n;
        // stay within a 1 m radius column and 0.3 m vertically
        double rho = std::sqrt(pc[0] * pc[0] + pc[2] * pc[2]);
        if (rho > 1.0) { pc[0] *= 1.0 / rho; pc[2] *= 1.0 / rho; }
        pc[1] = std::max(-0.3, std::min(0.3, pc[1]));
        double axis[3] = {0.15 * gauss(key(P.seed, 4, k, 3)), 1.0, 0.15 * gauss(key(P.seed, 4, k, 4))};
        double an = std::sqrt(axis[0] * axis[0] + axis[1] * axis[1] + axis[2] * axis[2]);
        double w[3];
        for (int a = 0; a < 3; ++a) w[a] = axis[a] / an * (0.5 * M_PI / 180.0);
        double Rs[9], Rn[9];
        rot_from_vec(w, Rs);
        matmul3(Rc, Rs, Rn);
        std::memcpy(Rc, Rn, sizeof(Rn));
    }
    std::memcpy(Rm, Rc, sizeof(Rc));
    std::memcpy(t, pc, sizeof(pc));
}

}  // namespace

extern "C" {

// Render frame `frame` of the workload.  depth: W*H uint16 mm (0 = no
// measurement); bgra: W*H*4 bytes or NULL; pose12: 12 floats, camera-to-world
// rotation (row-major) then translation, or NULL.  Returns 0 on success.
int nsynth_render(const nsynth_params* P, int frame, uint16_t* depth, uint8_t* bgra, float* pose12) {
    if (!P || P->width <= 0 || P->height <= 0) return -1;
    Scene scene = (P->scene_kind == 0) ? build_orbit_scene(*P) : build_room_scene(*P);
    double Rm[9], t[3];
    camera_pose(*P, frame, Rm, t);
    if (pose12) {
        for (int k = 0; k < 9; ++k) pose12[k] = (float)Rm[k];
        for (int k = 0; k < 3; ++k) pose12[9 + k] = (float)t[k];
    }
    const int W = P->width, H = P->height;
    const V3 o = {t[0], t[1], t[2]};
#pragma omp parallel for schedule(static)
    for (int v = 0; v < H; ++v) {
        for (int u = 0; u < W; ++u) {
            const int idx = v * W + u;
            V3 dc = {(u - P->cx) / P->fx, (v - P->cy) / P->fy, 1.0};
            V3 dw = {Rm[0] * dc.x + Rm[1] * dc.y + Rm[2] * dc.z, Rm[3] * dc.x + Rm[4] * dc.y + Rm[5] * dc.z,
                     Rm[6] * dc.x + Rm[7] * dc.y + Rm[8] * dc.z};
            double best = INFINITY;
            int hit = -1;
            for (size_t k = 0; k < scene.prims.size(); ++k) {
                double tk = intersect(scene.prims[k], o, dw);
                if (tk < best) { best = tk; hit = (int)k; }
            }
            // dw has unit camera-z component, so the ray parameter is the depth z
            double z = best;
            uint16_t dmm = 0;
            if (hit >= 0 && std::isfinite(z)) {
                double zm = z;
                if (P->noise) {
                    double sigma = 0.0012 + 0.0019 * (z - 0.4) * (z - 0.4);
                    zm = z + sigma * gauss(key(P->seed, 5, frame, idx));
                }
                double mm = std::floor(zm * 1000.0 + 0.5);
                if (mm > 0 && mm < 65535) dmm = (uint16_t)mm;
                if (u01(key(P->seed, 6, frame, idx)) < P->hole_frac) dmm = 0;
            }
            depth[idx] = dmm;
            if (bgra) {
                uint8_t* px = bgra + 4 * (size_t)idx;
                if (hit < 0 || !std::isfinite(z)) {
                    px[0] = px[1] = px[2] = 0; px[3] = 0;
                    continue;
                }
                V3 p = o + dw * z;
                const V3& al = scene.prims[hit].albedo;
                const double f = 2.0 * M_PI * 2.0;  // 2 cycles per metre
                double tex = 1.0 + 0.2 * std::sin(f * p.x) * std::sin(f * p.y) * std::sin(f * p.z);
                double rgb[3] = {al.x * tex, al.y * tex, al.z * tex};
                for (int c = 0; c < 3; ++c) {
                    double val = rgb[c] + (3.0 / 255.0) * gauss(key(P->seed, 8 + c, frame, idx));
                    val = std::max(0.0, std::min(1.0, val));
                    rgb[c] = std::floor(val * 255.0 + 0.5);
                }
                px[0] = (uint8_t)rgb[2]; px[1] = (uint8_t)rgb[1]; px[2] = (uint8_t)rgb[0]; px[3] = 255;
            }
        }
    }
    return 0;
}

// Ground-truth camera-to-world pose of a frame (12 floats).
int nsynth_pose(const nsynth_params* P, int frame, double* pose12) {
    if (!P) return -1;
    double Rm[9], t[3];
    camera_pose(*P, frame, Rm, t);
    for (int k = 0; k < 9; ++k) pose12[k] = Rm[k];
    for (int k = 0; k < 3; ++k) pose12[9 + k] = t[k];
    return 0;
}

}  // extern "C"
