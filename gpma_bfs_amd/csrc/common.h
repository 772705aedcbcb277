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
// Orders one wave's LDS accesses before and after this point (the lanes of a
// wave then see each other's LDS writes).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
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

// Stable rank of this lane's digit d (BITS wide; ok = false for no item)
// among the lanes of its wave that hold the same digit, plus the count of
// that digit seen by the wave before this call: the leader lane of each
// digit reads and advances cnt[d] (a per-wave LDS counter row).
template <int BITS, typename C>
__device__ __forceinline__ uint32_t wave_rank(uint32_t d, bool ok, C *cnt) {
    uint64_t match = ballot(ok);
#pragma unroll
    for (int b = 0; b < BITS; b++) {
        const uint64_t bb = ballot(ok && ((d >> b) & 1));
        match &= ((d >> b) & 1) ? bb : ~bb;
    }
    const uint32_t leader = ok ? (uint32_t)__ffsll((unsigned long long)match) - 1 : 0;
    uint32_t base = 0;
    if (ok && lane_id() == leader) {
        base = cnt[d];
        cnt[d] = (C)(base + (uint32_t)__popcll(match));
    }
    base = __shfl(base, leader, WAVE);
    return base + (uint32_t)__popcll(match & lanemask_lt());
}

// Decoupled look-back of a chained scan, run by all lanes of one wave.
// status[t] = epoch << 32 | flag << 30 | value with flag 1 = the tile's own
// aggregate, 2 = inclusive prefix; words of another epoch read as not yet
// published. Publishes this tile's aggregate, adds predecessors' values 64
// per round trip until an inclusive prefix is reached, publishes its own
// inclusive prefix and returns the exclusive one (same in every lane). A
// spin limit sets *err and returns what was summed instead of hanging.
constexpr uint64_t LB_AGG = 1ull << 30, LB_INC = 2ull << 30, LB_VAL = (1ull << 30) - 1;
__device__ __forceinline__ uint32_t wave_lookback(unsigned long long *status, uint32_t tile, uint32_t epoch,
                                                  uint32_t agg, uint32_t *err) {
    const uint32_t lane = lane_id();
    const uint64_t tag = (uint64_t)epoch << 32;
    if (lane == 0)
        __hip_atomic_store(status + tile, tag | (tile == 0 ? LB_INC : LB_AGG) | agg, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    if (tile == 0) return 0;
    uint32_t excl = 0;
    int64_t t = (int64_t)tile - 1;  // next predecessor not yet added
    uint32_t spins = 0;
    while (true) {
        const int64_t q = t - (int64_t)lane;
        unsigned long long w = 0;
        if (q >= 0) w = __hip_atomic_load(status + q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const bool mine = q >= 0 && (w >> 32) == epoch;
        const uint64_t fl = w & (3ull << 30);
        const bool ready = q < 0 || (mine && fl != 0);
        const bool inc = q < 0 || (mine && fl == LB_INC);
        const uint64_t notready = ballot(!ready);
        const uint64_t incm = ballot(inc);
        // lanes before the first not-ready lane are usable; stop at the first inclusive one
        const uint64_t usable = notready ? ((notready & (~notready + 1)) - 1) : ~0ull;
        const uint64_t stop = incm & usable;
        const uint64_t take = stop ? (((stop & (~stop + 1)) << 1) - 1) : usable;
        uint32_t c = ((take >> lane) & 1ull) && q >= 0 ? (uint32_t)(w & LB_VAL) : 0u;
#pragma unroll
        for (uint32_t off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off, WAVE);
        excl += c;
        if (stop) break;
        const uint32_t used = (uint32_t)__popcll(take);
        t -= used;
        if (used == 0) {
            if (++spins > (1u << 22)) {
                if (lane == 0) atomicOr(err, 1u);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
    if (lane == 0)
        __hip_atomic_store(status + tile, tag | LB_INC | (excl + agg), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return excl;
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
