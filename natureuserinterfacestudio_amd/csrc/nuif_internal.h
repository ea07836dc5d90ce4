// Internal helpers shared by the HIP translation units of libnuif.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>

#include "nuif.h"

namespace nuif {

// ---- error reporting (thread-local last error, status codes) ----------------
void set_error(const char* fmt, ...);
int hip_fail(hipError_t e, const char* what);
hipStream_t stream_of(nuif_stream_t s);
bool initialised();

#define NUIF_REQUIRE_INIT()                                      \
    do {                                                         \
        if (!nuif::initialised()) {                              \
            nuif::set_error("nuif_init() has not been called");  \
            return NUIF_ERR_NODEV;                               \
        }                                                        \
    } while (0)

#define NUIF_CHECK_ARG(cond, ...)        \
    do {                                 \
        if (!(cond)) {                   \
            nuif::set_error(__VA_ARGS__); \
            return NUIF_ERR_INVALID;     \
        }                                \
    } while (0)

#define NUIF_HIP(call)                                            \
    do {                                                          \
        hipError_t e_ = (call);                                   \
        if (e_ != hipSuccess) return nuif::hip_fail(e_, #call);   \
    } while (0)

#define NUIF_LAUNCHED(name)                                       \
    do {                                                          \
        hipError_t e_ = hipGetLastError();                        \
        if (e_ != hipSuccess) return nuif::hip_fail(e_, name);    \
    } while (0)

inline unsigned div_up(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

// The dispatcher hands the workgroups of a 1-D launch to the 8 XCDs in turn (workgroup b
// to XCD b % 8), and every XCD has an L2 of its own.  For a launch of G workgroups, G a
// multiple of 8, xcd_block(b, G) gives XCD x the logical blocks [x G/8, (x + 1) G/8): the
// consecutive blocks of one session, or of one band of an image, then share one L2
// instead of all eight fetching the same lines.  The map is a bijection of [0, G).
constexpr unsigned kXcds = 8;
__device__ __forceinline__ unsigned xcd_block(unsigned b, unsigned G) { return (b % kXcds) * (G / kXcds) + b / kXcds; }
inline unsigned xcd_grid(unsigned blocks) { return div_up(blocks, kXcds) * kXcds; }

// device address of the allocation-free ICP entry point's static scratch on the
// current device (resolved once per device)
int static_scratch(void** out);

// device-wide exclusive scan of n uint32 values (flags=false) or of (uint8 > 0)
// (flags=true); scratch of nuif_scan_scratch_bytes(n); in == out is allowed
int launch_scan(const void* in, bool flags, uint32_t* out, size_t n, uint32_t* total, void* scratch,
                hipStream_t st);
// Stable LSD radix sort of n (key, value) pairs over the low key_bits bits of the keys
// (scan.hip).  The result is in (keys, vals), or in (keys_alt, vals_alt) when
// *result_in_alt; scratch: sort_scratch_bytes(n) of device memory.
size_t sort_scratch_bytes(uint32_t n);
int launch_sort_pairs_u64(unsigned long long* keys, unsigned long long* keys_alt, uint32_t* vals, uint32_t* vals_alt,
                          uint32_t n, int key_bits, void* scratch, hipStream_t st, bool* result_in_alt);

// ---- device math, mirroring the reference's float evaluation order ----------
// Everything is compiled with -ffp-contract=off: no implicit FMA, so each
// expression rounds exactly as written (K/*.cl evaluation order).

struct f3 {
    float x, y, z;
};
__device__ __forceinline__ f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
__device__ __forceinline__ f3 add3(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 sub3(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ f3 mul3(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ f3 cross3(f3 a, f3 b) {
    return f3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ bool nan3(f3 v) { return isnan(v.x) || isnan(v.y) || isnan(v.z); }
// p, known to point into device (global) memory: accesses through the result
// compile to global_* instructions.  A pointer read from LDS or from a job table
// is otherwise generic, and its accesses compile to flat_* instructions (which
// also wait on the LDS counter).  The integer round trip keeps the compiler from
// folding the cast away.
template <class T>
__device__ __forceinline__ T* in_global(T* p) {
    return (T*)(__attribute__((address_space(1))) T*)(uintptr_t)p;
}
__device__ __forceinline__ f3 ld3(const float* p, size_t i) { return f3{p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }
__device__ __forceinline__ void st3(float* p, size_t i, f3 v) {
    p[3 * i] = v.x;
    p[3 * i + 1] = v.y;
    p[3 * i + 2] = v.z;
}
// normalize()/fast_normalize(): a zero vector is returned unchanged
__device__ __forceinline__ f3 normalize3(f3 v) {
    if (v.x == 0.f && v.y == 0.f && v.z == 0.f) return v;
    float d = dot3(v, v);
    float inv = 1.0f / sqrtf(d);
    return mul3(v, inv);
}
// OpenCL float -> int conversions: round toward zero; out of range / NaN -> INT_MIN
__device__ __forceinline__ int to_int(float v) {
    if (!(v > -2147483648.0f && v < 2147483647.0f)) return INT32_MIN;
    return (int)v;
}
__device__ __forceinline__ f3 rot3(const float* R, f3 v) {
    return f3{dot3(mk3(R[0], R[1], R[2]), v), dot3(mk3(R[3], R[4], R[5]), v), dot3(mk3(R[6], R[7], R[8]), v)};
}

constexpr float kNaN = __builtin_nanf("");

// Device-scope release / acquire fences for the last-block pattern: every block
// releases its results before counting itself in, only the block that goes on
// acquires.  (A full fence in every block would also invalidate its L2 while the
// other blocks are still reading through it.)
__device__ __forceinline__ void fence_release_agent() { __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent"); }
__device__ __forceinline__ void fence_acquire_agent() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent"); }

// K/normalEstimation.cl:770-822 estimate_normals_simple_kernel at pixel (x, y): the
// normalised cross product of the right-left and down-up differences; a neighbour
// outside the image, NaN or more than thr away in z is replaced by the centre; NaN
// centre -> NaN normal.
__device__ __forceinline__ f3 normal_simple_at(const float* __restrict__ vmap, int x, int y, int W, int H,
                                               float thr) {
    const size_t i = (size_t)y * W + x;
    const f3 c = ld3(vmap, i);
    if (nan3(c)) return mk3(kNaN, kNaN, kNaN);
    auto pick = [&](bool ok, size_t j) {
        if (!ok) return c;
        const f3 n = ld3(vmap, j);
        if (nan3(n) || fabsf(c.z - n.z) > thr) return c;
        return n;
    };
    const f3 l = pick(x > 0, i - 1), rr = pick(x < W - 1, i + 1);
    const f3 u = pick(y > 0, i - W), d = pick(y < H - 1, i + W);
    return normalize3(cross3(sub3(rr, l), sub3(d, u)));
}

}  // namespace nuif
