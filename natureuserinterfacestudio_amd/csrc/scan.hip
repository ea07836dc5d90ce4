// Device-wide exclusive prefix sum and flag compaction.
//
// The reference scans uchar flags with a warp-synchronous LDS scan written for
// 32-wide warps (K/prefixSum.cl:4-24, 163-232) and compacts with racy writes
//, then reads the count back to the host.  Here: a three-phase
// reduce / scan-of-partials / down-sweep over 1024-element tiles with 64-wide
// wavefront scans, deterministic results, and the total left in device memory
// so that later launches can size themselves without a host round trip.
// HBM-bound: 1 read of the input in the reduce phase, 1 read + 1 write in the
// down-sweep (the partials are 1/1024 of the data).

#include "block_scan.h"
#include "nuif_internal.h"

namespace nuif {

constexpr int kScanThreads = 256;
constexpr int kScanPer = 4;
constexpr int kScanTile = kScanThreads * kScanPer;

template <bool FLAGS>
__device__ __forceinline__ uint32_t scan_in(const void* in, size_t i, size_t n) {
    if (i >= n) return 0u;
    if (FLAGS) return static_cast<const uint8_t*>(in)[i] > 0 ? 1u : 0u;
    return static_cast<const uint32_t*>(in)[i];
}

template <bool FLAGS>
__global__ void __launch_bounds__(kScanThreads) scan_reduce_kernel(const void* __restrict__ in, size_t n,
                                                                   uint32_t* __restrict__ partials) {
    __shared__ uint32_t lds[kScanThreads / 64 + 1];
    const size_t base = (size_t)blockIdx.x * kScanTile + (size_t)threadIdx.x * kScanPer;
    uint32_t v = 0;
#pragma unroll
    for (int k = 0; k < kScanPer; ++k) v += scan_in<FLAGS>(in, base + k, n);
    uint32_t total;
    (void)block_exclusive_scan<kScanThreads>(v, lds, &total);
    if (threadIdx.x == 0) partials[blockIdx.x] = total;
}

// one workgroup scans all tile sums, tile by tile with a running carry
__global__ void __launch_bounds__(1024) scan_partials_kernel(uint32_t* __restrict__ partials, size_t nparts,
                                                             uint32_t* __restrict__ total_out) {
    __shared__ uint32_t lds[1024 / 64 + 1];
    uint32_t carry = 0;
    for (size_t c = 0; c < nparts; c += 1024) {
        const size_t i = c + threadIdx.x;
        const uint32_t v = i < nparts ? partials[i] : 0u;
        uint32_t tot;
        const uint32_t ex = block_exclusive_scan<1024>(v, lds, &tot);
        if (i < nparts) partials[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0 && total_out) *total_out = carry;
}

template <bool FLAGS>
__global__ void __launch_bounds__(kScanThreads)
    scan_down_kernel(const void* __restrict__ in, size_t n, const uint32_t* __restrict__ partials,
                     uint32_t* __restrict__ out) {
    __shared__ uint32_t lds[kScanThreads / 64 + 1];
    const size_t base = (size_t)blockIdx.x * kScanTile + (size_t)threadIdx.x * kScanPer;
    uint32_t v[kScanPer], s = 0;
#pragma unroll
    for (int k = 0; k < kScanPer; ++k) {
        v[k] = scan_in<FLAGS>(in, base + k, n);
        s += v[k];
    }
    uint32_t total;
    uint32_t run = block_exclusive_scan<kScanThreads>(s, lds, &total) + partials[blockIdx.x];
#pragma unroll
    for (int k = 0; k < kScanPer; ++k) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
}

// ids[scan[i]] = i for every flagged i (order-preserving compaction)
__global__ void compact_flags_kernel(const uint8_t* __restrict__ flags, const uint32_t* __restrict__ scan, size_t n,
                                     int32_t* __restrict__ ids) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && flags[i] > 0) ids[scan[i]] = (int32_t)i;
}

int launch_scan(const void* in, bool flags, uint32_t* out, size_t n, uint32_t* total, void* scratch,
                hipStream_t st) {
    const size_t nparts = (n + kScanTile - 1) / kScanTile;
    uint32_t* partials = static_cast<uint32_t*>(scratch);
    if (flags)
        scan_reduce_kernel<true><<<(unsigned)nparts, kScanThreads, 0, st>>>(in, n, partials);
    else
        scan_reduce_kernel<false><<<(unsigned)nparts, kScanThreads, 0, st>>>(in, n, partials);
    NUIF_LAUNCHED("scan_reduce_kernel");
    scan_partials_kernel<<<1, 1024, 0, st>>>(partials, nparts, total);
    NUIF_LAUNCHED("scan_partials_kernel");
    if (flags)
        scan_down_kernel<true><<<(unsigned)nparts, kScanThreads, 0, st>>>(in, n, partials, out);
    else
        scan_down_kernel<false><<<(unsigned)nparts, kScanThreads, 0, st>>>(in, n, partials, out);
    NUIF_LAUNCHED("scan_down_kernel");
    return NUIF_OK;
}

}  // namespace nuif

using namespace nuif;

extern "C" {

size_t nuif_scan_scratch_bytes(size_t n) { return ((n + kScanTile - 1) / kScanTile + 1) * sizeof(uint32_t); }

int nuif_exclusive_scan_u32(const uint32_t* in, uint32_t* out, size_t n, uint32_t* total, void* scratch,
                            nuif_stream_t s) {
    NUIF_REQUIRE_INIT();
    NUIF_CHECK_ARG(in && out && scratch, "nuif_exclusive_scan_u32: null argument");
    NUIF_CHECK_ARG(n > 0 && n < (1ull << 32), "nuif_exclusive_scan_u32: bad length %zu", n);
    return launch_scan(in, false, out, n, total, scratch, stream_of(s));
}

int nuif_exclusive_scan_flags(const uint8_t* flags, uint32_t* out, size_t n, uint32_t* total, void* scratch,
                              nuif_stream_t s) {
    NUIF_REQUIRE_INIT();
    NUIF_CHECK_ARG(flags && out && scratch, "nuif_exclusive_scan_flags: null argument");
    NUIF_CHECK_ARG(n > 0 && n < (1ull << 32), "nuif_exclusive_scan_flags: bad length %zu", n);
    return launch_scan(flags, true, out, n, total, scratch, stream_of(s));
}

int nuif_compact_flags(const uint8_t* flags, const uint32_t* scan, size_t n, int32_t* ids, nuif_stream_t s) {
    NUIF_REQUIRE_INIT();
    NUIF_CHECK_ARG(flags && scan && ids && n > 0, "nuif_compact_flags: bad argument");
    compact_flags_kernel<<<div_up(n, 256), 256, 0, stream_of(s)>>>(flags, scan, n, ids);
    NUIF_LAUNCHED("compact_flags_kernel");
    return NUIF_OK;
}

}  // extern "C"
