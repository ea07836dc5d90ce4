// Wavefront / workgroup prefix sums on 64-wide wavefronts (deterministic: the
// same inputs always give the same outputs, no atomics).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace nuif {

__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t n = __shfl_up(v, d, 64);
        if (lane >= d) v += n;
    }
    return v;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

// Exclusive scan across a 1-D workgroup of NT threads (NT a multiple of 64).
// `lds` needs NT/64 + 1 words.  Returns the thread's exclusive prefix; *total
// receives the workgroup sum (in every thread).
template <int NT>
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* lds, uint32_t* total) {
    constexpr int NW = NT / 64;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const uint32_t incl = wave_inclusive_scan(v);
    if (lane == 63) lds[wid] = incl;
    __syncthreads();
    if (tid == 0) {
        uint32_t run = 0;
        for (int w = 0; w < NW; ++w) {
            const uint32_t s = lds[w];
            lds[w] = run;
            run += s;
        }
        lds[NW] = run;
    }
    __syncthreads();
    const uint32_t excl = lds[wid] + incl - v;
    *total = lds[NW];
    __syncthreads();
    return excl;
}

}  // namespace nuif
