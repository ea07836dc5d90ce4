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

// ---- stable radix sort of (64-bit key, 32-bit value) pairs --------------------------
// Least significant digit first, 8-bit digits, one element per thread: per pass, every
// block counts its digits (digit-major table, so that one exclusive scan of the table
// gives each block's first output slot per digit), then scatters its elements; an
// element's rank among the block's elements of its digit is its rank among its wave's
// peers (lanes with the same digit, found with 8 ballots) plus the counts of the
// earlier waves.  Every pass keeps the order of equal digits: the sort is stable.
constexpr int kSortThreads = 256;  // == the number of digits

__global__ void __launch_bounds__(kSortThreads)
    sort_hist_kernel(const unsigned long long* __restrict__ keys, uint32_t n, int shift, uint32_t nblocks,
                     uint32_t* __restrict__ hist) {
    __shared__ uint32_t h[kSortThreads];
    h[threadIdx.x] = 0u;
    __syncthreads();
    const uint32_t i = blockIdx.x * kSortThreads + threadIdx.x;
    if (i < n) atomicAdd(&h[(uint32_t)(keys[i] >> shift) & 255u], 1u);
    __syncthreads();
    hist[(size_t)threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

__global__ void __launch_bounds__(kSortThreads)
    sort_scatter_kernel(const unsigned long long* __restrict__ keys, const uint32_t* __restrict__ vals,
                        unsigned long long* __restrict__ okeys, uint32_t* __restrict__ ovals, uint32_t n, int shift,
                        uint32_t nblocks, const uint32_t* __restrict__ offsets) {
    constexpr int kWaves = kSortThreads / 64;
    __shared__ uint32_t cnt[kWaves][256];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    for (int k = tid; k < kWaves * 256; k += kSortThreads) (&cnt[0][0])[k] = 0u;
    __syncthreads();
    const uint32_t i = blockIdx.x * kSortThreads + tid;
    const bool valid = i < n;
    const unsigned long long key = valid ? keys[i] : 0ull;
    const uint32_t d = (uint32_t)(key >> shift) & 255u;
    unsigned long long peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        const unsigned long long m = __ballot((d >> b) & 1u);
        peers &= ((d >> b) & 1u) ? m : ~m;
    }
    const unsigned long long below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    const uint32_t rank = (uint32_t)__popcll(peers & below);
    if (valid && (peers >> lane) == 1ull) cnt[wv][d] = (uint32_t)__popcll(peers);  // the wave's last peer
    __syncthreads();
    if (!valid) return;
    uint32_t before = 0;
    for (int w = 0; w < wv; ++w) before += cnt[w][d];
    const uint32_t pos = offsets[(size_t)d * nblocks + blockIdx.x] + before + rank;
    okeys[pos] = key;
    ovals[pos] = vals[i];
}

size_t sort_scratch_bytes(uint32_t n) {
    const size_t table = (size_t)kSortThreads * ((n + kSortThreads - 1) / kSortThreads);
    return (table * sizeof(uint32_t) + 255) / 256 * 256 + nuif_scan_scratch_bytes(table);
}

int launch_sort_pairs_u64(unsigned long long* keys, unsigned long long* keys_alt, uint32_t* vals, uint32_t* vals_alt,
                          uint32_t n, int key_bits, void* scratch, hipStream_t st, bool* result_in_alt) {
    const uint32_t nblocks = (n + kSortThreads - 1) / kSortThreads;
    const size_t table = (size_t)kSortThreads * nblocks;
    uint32_t* hist = static_cast<uint32_t*>(scratch);
    void* scan_scratch = static_cast<char*>(scratch) + (table * sizeof(uint32_t) + 255) / 256 * 256;
    bool alt = false;
    for (int shift = 0; shift < key_bits && n > 0; shift += 8) {
        const unsigned long long* ki = alt ? keys_alt : keys;
        const uint32_t* vi = alt ? vals_alt : vals;
        unsigned long long* ko = alt ? keys : keys_alt;
        uint32_t* vo = alt ? vals : vals_alt;
        sort_hist_kernel<<<nblocks, kSortThreads, 0, st>>>(ki, n, shift, nblocks, hist);
        NUIF_LAUNCHED("sort_hist_kernel");
        const int rc = launch_scan(hist, false, hist, table, nullptr, scan_scratch, st);
        if (rc) return rc;
        sort_scatter_kernel<<<nblocks, kSortThreads, 0, st>>>(ki, vi, ko, vo, n, shift, nblocks, hist);
        NUIF_LAUNCHED("sort_scatter_kernel");
        alt = !alt;
    }
    *result_in_alt = alt;
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
