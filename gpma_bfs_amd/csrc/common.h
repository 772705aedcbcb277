// Shared types, constants and wave64 helpers for the gfx950 GPMA kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gpma {

using gkey = uint64_t;
using gval = uint8_t;

// Encodings of the reference (gpma.cuh:15-19).
constexpr gkey KEY_NONE = ~0ull;
constexpr gkey KEY_MAX = KEY_NONE - 1;
constexpr uint32_t SIZE_NONE = 0xFFFFFFFFu;
constexpr uint32_t COL_NONE = 0xFFFFFFFFu;

constexpr int WAVE = 64;

__device__ __forceinline__ bool is_live(gkey k, gval v) { return k != KEY_NONE && v != 0; }
__device__ __forceinline__ bool is_guard(gkey k) { return (uint32_t)k == COL_NONE; }

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }
__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }
__device__ __forceinline__ uint64_t lanemask_lt() {
    uint32_t l = __lane_id();
    return l == 0 ? 0ull : (~0ull >> (64 - l));
}

// Slot of merged item i when a window of `width` slots (segment length S)
// receives m items: gpma.cuh:636-654, (f*i/m)*S + (f*i mod m)/f, f = W/S,
// in 64-bit arithmetic. Equivalently segment s holds items
// [ceil(s*m/f), ceil((s+1)*m/f)) packed at its front.
__device__ __host__ __forceinline__ uint32_t redistribute_slot(uint64_t f, uint32_t S, uint64_t m, uint64_t i) {
    uint64_t fi = f * i;
    return (uint32_t)(fi / m) * S + (uint32_t)((fi % m) / f);
}
// First item index of segment s.
__device__ __host__ __forceinline__ uint64_t seg_first_item(uint64_t f, uint64_t m, uint64_t s) {
    return (s * m + f - 1) / f;
}

// Block-wide exclusive scan of one uint32 per thread (blockDim.x <= 1024,
// multiple of 64). `scratch` must hold blockDim.x/64 + 1 words. Returns the
// exclusive prefix; *total receives the block sum. Contains barriers.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t x, uint32_t *scratch, uint32_t *total) {
    const uint32_t lane = lane_id();
    const uint32_t wid = threadIdx.x / WAVE;
    const uint32_t nw = blockDim.x / WAVE;
    // wave inclusive scan by shuffles
    uint32_t v = x;
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) {
        uint32_t y = __shfl_up(v, d, WAVE);
        if (lane >= (uint32_t)d) v += y;
    }
    if (lane == WAVE - 1) scratch[wid] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (uint32_t w = 0; w < nw; w++) {
            uint32_t t = scratch[w];
            scratch[w] = run;
            run += t;
        }
        scratch[nw] = run;
    }
    __syncthreads();
    uint32_t res = scratch[wid] + v - x;
    *total = scratch[nw];
    __syncthreads();
    return res;
}

}  // namespace gpma
