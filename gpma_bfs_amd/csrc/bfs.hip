// Frontier BFS over the gapped packed array.
//
// Reference semantics (gpma_bfs.cuh:149-395,446-509): level[start] = 1;
// round l expands every vertex of level l over the slots
// [row_offset[v], row_offset[v+1]), skipping guards and gaps (low word
// 0xFFFFFFFF) and tombstones (value 0), and gives level l+1 to every
// destination not yet labelled. The reference materialises an edge queue
// (gather) and culls it with hashes, a non-atomic bitmap and a CAS on the
// level (contract). Here a destination is claimed directly by a
// compare-and-swap of its own level word from 0 to l+1, after a plain read
// of that word filters the already labelled ones, so each vertex is labelled
// by exactly one lane and enters exactly one frontier. (A visited bitmap
// shares one atomic word among 32 vertices: on skewed destinations those
// words serialise at the memory-side atomic unit, 0.74 ms instead of 0.55 ms
// for the BFS from node 0 of the power-law 1.63M-node graph.)
//
// One kernel per level, whose mode is chosen on the device from the
// frontier's slot count (accumulated while the frontier was claimed):
//   push: vertices with short ranges: a workgroup scans the slot counts of 64
//         frontier vertices in LDS and walks the concatenated ranges 512
//         slots at a time, two independent slots per lane (owner by binary
//         search); vertices with long ranges (power-law hubs) sit in a second
//         queue whose ranges are cut into 2048-slot pieces spread over many
//         workgroups, so no single workgroup walks a hub alone;
//   pull: when the frontier's ranges cover a large share of the array, every
//         slot of the array is streamed once, coalesced, and a slot counts if
//         its source (the key's high word) has level l -- no adjacency
//         lookups, one sequential pass.
// The non-chosen kernel returns at once. The host launches levels in chunks
// without waiting; frontier sizes live in cnt[] on the device.
#include "common.h"
#include "kernels.h"

namespace gpma {

namespace {

constexpr uint32_t BFS_THREADS = 256;
constexpr uint32_t PULL_SLOTS_PER_THREAD = 16;

// Vertices whose slot range exceeds BIG_RANGE go to a second queue whose
// ranges are cut into BIG_CHUNK-slot pieces shared by many workgroups.
constexpr uint32_t BIG_RANGE = 2048;
constexpr uint32_t BIG_CHUNK = 2048;
constexpr uint32_t BIG_LIST_MAX = 2048;  // big vertices prefix-scanned per workgroup round

__device__ __forceinline__ uint32_t range_len(const uint32_t *row_offset, uint32_t v) {
    const uint32_t b = row_offset[v], e = row_offset[v + 1];
    return e > b ? e - b : 0;
}

// Per-level queue counters are packed u64 words: low 32 bits = vertices,
// high 32 bits = their total slot count (the next level's work), so one
// atomic reserves queue space and accounts work together.
__device__ __forceinline__ uint32_t pk_count(unsigned long long x) { return (uint32_t)x; }
__device__ __forceinline__ uint32_t pk_work(unsigned long long x) { return (uint32_t)(x >> 32); }

__global__ void k_bfs_init(uint32_t *__restrict__ levels, uint32_t N, uint32_t start,
                           uint32_t *__restrict__ frontier, uint32_t *__restrict__ big,
                           unsigned long long *__restrict__ cnt, unsigned long long *__restrict__ cnt_big,
                           const uint32_t *__restrict__ row_offset) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (uint64_t i = t; i < N; i += stride) levels[i] = (i == start) ? 1u : 0u;
    if (t == 0) {
        const unsigned long long len = range_len(row_offset, start);
        cnt[0] = 0;
        cnt_big[0] = 0;
        if (len > BIG_RANGE) {
            big[0] = start;
            cnt[1] = 0;
            cnt_big[1] = (len << 32) | 1ull;
        } else {
            frontier[0] = start;
            cnt[1] = (len << 32) | 1ull;
            cnt_big[1] = 0;
        }
    }
}

// Append claimed long-range vertices to the big queue (wave-aggregated),
// adding their slot counts to the counter's work half.
__device__ __forceinline__ void emit_big(bool claimed, uint32_t dst, uint32_t len, uint32_t *next,
                                         unsigned long long *next_cnt) {
    const uint64_t m = ballot(claimed);
    if (!m) return;
    uint64_t w = claimed ? len : 0;
    for (int d = WAVE / 2; d > 0; d >>= 1) w += __shfl_xor(w, d, WAVE);
    const uint32_t leader = (uint32_t)__ffsll((unsigned long long)m) - 1;
    uint32_t base = 0;
    if (lane_id() == leader) base = pk_count(atomicAdd(next_cnt, (w << 32) | (unsigned long long)__popcll(m)));
    base = __shfl(base, leader, WAVE);
    if (claimed) next[base + (uint32_t)__popcll(m & lanemask_lt())] = dst;
}

struct NextLevel {
    uint32_t *small, *big;
    unsigned long long *small_cnt, *big_cnt;
};

// Claims of short-range vertices are staged in an LDS queue and copied to
// the global queue with one counter reservation per flush (a global atomic
// per wave on one counter serialises at the widest levels); the rare
// long-range claims go straight to their global queue.
constexpr uint32_t LQ_CAP = 2048;

struct LdsQueue {
    uint32_t *buf;   // [LQ_CAP] in LDS
    uint32_t *n;     // LDS count
    uint32_t *work;  // LDS slot-count sum of the queued vertices
    uint32_t *base;  // LDS scratch for the reserved global offset
};

__device__ __forceinline__ void emit2(bool claimed, uint32_t dst, uint32_t len, const NextLevel &nx,
                                      const LdsQueue &lq) {
    const bool small = claimed && len <= BIG_RANGE;
    const uint64_t m = ballot(small);
    if (m) {
        const uint32_t leader = (uint32_t)__ffsll((unsigned long long)m) - 1;
        uint32_t base = 0;
        if (lane_id() == leader) base = atomicAdd(lq.n, (uint32_t)__popcll(m));
        base = __shfl(base, leader, WAVE);
        if (small) lq.buf[base + (uint32_t)__popcll(m & lanemask_lt())] = dst;
        // one LDS add per wave: same-word LDS atomics from many lanes serialise
        uint32_t wsum = small ? len : 0;
        for (int d = WAVE / 2; d > 0; d >>= 1) wsum += __shfl_xor(wsum, d, WAVE);
        if (lane_id() == leader) atomicAdd(lq.work, wsum);
    }
    emit_big(claimed && len > BIG_RANGE, dst, len, nx.big, nx.big_cnt);
}

// Flush the LDS queue when fewer than `room` entries are free (or always when
// room > LQ_CAP). Must be reached by the whole workgroup.
__device__ __forceinline__ void lq_flush(const LdsQueue &lq, const NextLevel &nx, uint32_t room) {
    __syncthreads();
    const uint32_t n = *lq.n;
    if (n == 0 || (room <= LQ_CAP && n + room <= LQ_CAP)) return;
    if (threadIdx.x == 0)
        *lq.base = pk_count(atomicAdd(nx.small_cnt, ((unsigned long long)*lq.work << 32) | n));
    __syncthreads();
    const uint32_t b = *lq.base;
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) nx.small[b + i] = lq.buf[i];
    __syncthreads();
    if (threadIdx.x == 0) {
        *lq.n = 0;
        *lq.work = 0;
    }
    __syncthreads();
}

// Expand U slots per lane in three stages, each issuing all of its memory
// operations before waiting on any: (1) the level words of candidate
// destinations (a plain read: it may miss a label set by another XCD during
// this kernel, never report a stale one, and the CAS decides), (2) the
// claiming compare-and-swap 0 -> newlev together with the destinations'
// row_offset pair (slot count for the next level's estimate and queue
// choice), (3) queue appends.
template <uint32_t U>
__device__ __forceinline__ void visit_batch(const gkey (&k)[U], const gval (&v)[U], const bool (&in)[U],
                                            uint32_t *levels, uint32_t newlev, const uint32_t *row_offset,
                                            const NextLevel &nx, const LdsQueue &lq) {
    uint32_t dst[U], word[U], old[U], b[U], e[U];
    bool cand[U];
#pragma unroll
    for (uint32_t u = 0; u < U; u++) {
        dst[u] = (uint32_t)k[u];
        cand[u] = in[u] && dst[u] != COL_NONE && v[u] != 0;
        word[u] = cand[u] ? *(volatile uint32_t *)(levels + dst[u]) : 1u;
    }
#pragma unroll
    for (uint32_t u = 0; u < U; u++) {
        cand[u] = cand[u] && word[u] == 0;
        old[u] = 1u;
        b[u] = e[u] = 0;
        if (cand[u]) {
            old[u] = atomicCAS(levels + dst[u], 0u, newlev);
            b[u] = row_offset[dst[u]];
            e[u] = row_offset[dst[u] + 1];
        }
    }
#pragma unroll
    for (uint32_t u = 0; u < U; u++) {
        const bool claimed = cand[u] && old[u] == 0;
        const uint32_t len = e[u] > b[u] ? e[u] - b[u] : 0;
        emit2(claimed, dst[u], len, nx, lq);
    }
}

// Top-down step over both queues in one launch.
//   Workgroups [0, small_grid): PUSH_TILE vertices of the small queue per
//   tile; their slot counts are scanned in LDS and the concatenated ranges
//   walked PUSH_UNROLL x 256 slots per round (owner by binary search).
//   Workgroups [small_grid, gridDim): BIG_CHUNK-slot pieces of the big
//   queue's ranges, so one hub is spread over many workgroups.
constexpr uint32_t PUSH_TILE = 64;      // most frontier vertices per tile (the tile_cap knob lowers it)
constexpr uint32_t PUSH_MIN_TILE = 16;  // fewest (short frontiers)
constexpr uint32_t PUSH_UNROLL = 4;

__device__ __forceinline__ void push_body(const gkey *__restrict__ keys, const gval *__restrict__ values,
                                          const uint32_t *__restrict__ row_offset,
                                          const uint32_t *__restrict__ front, const uint32_t *__restrict__ front_big,
                                          const NextLevel &nx, const unsigned long long *__restrict__ cnt,
                                          const unsigned long long *__restrict__ cnt_big, uint32_t *__restrict__ levels,
                                          uint32_t level, uint32_t small_grid, uint32_t tile_cap,
                                          uint32_t *s_beg, uint32_t *s_ofs, uint32_t *scratch, const LdsQueue &lq) {
    const uint32_t tid = threadIdx.x;
    const uint32_t newlev = level + 1;
    if (blockIdx.x < small_grid) {
        const uint32_t nf = pk_count(cnt[level]);
        // vertices per tile: a short frontier is spread over every workgroup
        const uint32_t tn =
            (uint32_t)min<uint64_t>(tile_cap, max<uint64_t>(PUSH_MIN_TILE, (nf + small_grid - 1) / small_grid));
        for (uint64_t tile = (uint64_t)blockIdx.x * tn; tile < nf; tile += (uint64_t)small_grid * tn) {
            const uint64_t fi = tile + tid;
            uint32_t b = 0, len = 0;
            if (tid < tn && fi < nf) {
                const uint32_t v = front[fi];
                b = row_offset[v];
                len = range_len(row_offset, v);
            }
            uint32_t total;
            const uint32_t ofs = block_exclusive_scan(len, scratch, &total);
            if (tid < tn) {
                s_beg[tid] = b;
                s_ofs[tid] = ofs;
            }
            __syncthreads();
            for (uint32_t j0 = 0; j0 < total; j0 += BFS_THREADS * PUSH_UNROLL) {
                gkey k[PUSH_UNROLL];
                gval v[PUSH_UNROLL];
                bool in[PUSH_UNROLL];
#pragma unroll
                for (uint32_t u = 0; u < PUSH_UNROLL; u++) {
                    const uint32_t j = j0 + u * BFS_THREADS + tid;
                    in[u] = j < total;
                    k[u] = KEY_NONE;
                    v[u] = 0;
                    if (in[u]) {
                        uint32_t lo = 0, hi = tn;
                        while (hi - lo > 1) {
                            const uint32_t mid = (lo + hi) >> 1;
                            if (s_ofs[mid] <= j) lo = mid;
                            else hi = mid;
                        }
                        const uint64_t slot = (uint64_t)s_beg[lo] + (j - s_ofs[lo]);
                        k[u] = keys[slot];
                        v[u] = values[slot];
                    }
                }
                visit_batch<PUSH_UNROLL>(k, v, in, levels, newlev, row_offset, nx, lq);
                lq_flush(lq, nx, BFS_THREADS * PUSH_UNROLL);
            }
            __syncthreads();
        }
    } else {
        const uint32_t nbig = pk_count(cnt_big[level]);
        const uint32_t bgrid = gridDim.x - small_grid;
        const uint32_t bid = blockIdx.x - small_grid;
        for (uint32_t l0 = 0; l0 < nbig; l0 += BIG_LIST_MAX) {
            const uint32_t lcount = min(BIG_LIST_MAX, nbig - l0);
            // chunk counts of this round's big vertices, scanned in LDS
            uint32_t run = 0;
            for (uint32_t i0 = 0; i0 < lcount; i0 += BFS_THREADS) {
                const uint32_t i = i0 + tid;
                uint32_t nch = 0, b = 0;
                if (i < lcount) {
                    const uint32_t v = front_big[l0 + i];
                    b = row_offset[v];
                    nch = (range_len(row_offset, v) + BIG_CHUNK - 1) / BIG_CHUNK;
                }
                uint32_t tot;
                const uint32_t pre = block_exclusive_scan(nch, scratch, &tot);
                if (i < lcount) {
                    s_beg[i] = b;
                    s_ofs[i] = run + pre;
                }
                run += tot;
            }
            if (tid == 0) s_ofs[lcount] = run;
            __syncthreads();
            for (uint32_t c = bid; c < run; c += bgrid) {
                uint32_t lo = 0, hi = lcount;
                while (hi - lo > 1) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (s_ofs[mid] <= c) lo = mid;
                    else hi = mid;
                }
                const uint32_t v = front_big[l0 + lo];
                const uint64_t beg = (uint64_t)s_beg[lo] + (uint64_t)(c - s_ofs[lo]) * BIG_CHUNK;
                const uint64_t end = min((uint64_t)row_offset[v + 1], beg + BIG_CHUNK);
                for (uint64_t p0 = beg; p0 < end; p0 += BFS_THREADS * PUSH_UNROLL) {
                    gkey k[PUSH_UNROLL];
                    gval vv[PUSH_UNROLL];
                    bool in[PUSH_UNROLL];
#pragma unroll
                    for (uint32_t u = 0; u < PUSH_UNROLL; u++) {
                        const uint64_t p = p0 + u * BFS_THREADS + tid;
                        in[u] = p < end;
                        k[u] = in[u] ? keys[p] : KEY_NONE;
                        vv[u] = in[u] ? values[p] : 0;
                    }
                    visit_batch<PUSH_UNROLL>(k, vv, in, levels, newlev, row_offset, nx, lq);
                    lq_flush(lq, nx, BFS_THREADS * PUSH_UNROLL);
                }
            }
            __syncthreads();
        }
    }
    lq_flush(lq, nx, LQ_CAP + 1);
}

__device__ __forceinline__ void pull_body(const gkey *__restrict__ keys, const gval *__restrict__ values, uint64_t C,
                                          uint32_t N, const uint32_t *__restrict__ row_offset, const NextLevel &nx,
                                          uint32_t *__restrict__ levels,
                                          uint32_t level, const LdsQueue &lq) {
    const uint32_t newlev = level + 1;
    const uint64_t per_block = (uint64_t)BFS_THREADS * PULL_SLOTS_PER_THREAD;
    for (uint64_t base = (uint64_t)blockIdx.x * per_block; base < C; base += (uint64_t)gridDim.x * per_block) {
        for (uint32_t r = 0; r < PULL_SLOTS_PER_THREAD; r += PUSH_UNROLL) {
            gkey k[PUSH_UNROLL];
            gval v[PUSH_UNROLL];
            bool in[PUSH_UNROLL];
#pragma unroll
            for (uint32_t u = 0; u < PUSH_UNROLL; u++) {
                const uint64_t p = base + (uint64_t)(r + u) * BFS_THREADS + threadIdx.x;
                k[u] = p < C ? keys[p] : KEY_NONE;
                v[u] = p < C ? values[p] : 0;
            }
#pragma unroll
            for (uint32_t u = 0; u < PUSH_UNROLL; u++) {
                const uint32_t src = (uint32_t)(k[u] >> 32);
                in[u] = src < N && levels[src] == level;
            }
            visit_batch<PUSH_UNROLL>(k, v, in, levels, newlev, row_offset, nx, lq);
            lq_flush(lq, nx, BFS_THREADS * PUSH_UNROLL);
        }
    }
    lq_flush(lq, nx, LQ_CAP + 1);
}

// One BFS level: streaming pull over all C slots when the frontier's slot
// count exceeds the threshold, else the push over both queues.
__global__ __launch_bounds__(BFS_THREADS) void k_bfs_step(const gkey *__restrict__ keys,
                                                          const gval *__restrict__ values, uint64_t C, uint32_t N,
                                                          const uint32_t *__restrict__ row_offset,
                                                          const uint32_t *__restrict__ front,
                                                          const uint32_t *__restrict__ front_big, NextLevel nx,
                                                          const unsigned long long *__restrict__ cnt,
                                                          const unsigned long long *__restrict__ cnt_big,
                                                          uint32_t *__restrict__ levels, uint32_t level,
                                                          unsigned long long pull_threshold,
                                                          uint32_t small_grid, uint32_t tile_cap) {
    __shared__ uint32_t s_beg[BIG_LIST_MAX];
    __shared__ uint32_t s_ofs[BIG_LIST_MAX + 1];
    __shared__ uint32_t scratch[BFS_THREADS / WAVE + 1];
    __shared__ uint32_t lq_buf[LQ_CAP];
    __shared__ uint32_t lq_n, lq_work, lq_base;
    const unsigned long long a = cnt[level], b = cnt_big[level];
    if (pk_count(a) == 0 && pk_count(b) == 0) return;
    const LdsQueue lq{lq_buf, &lq_n, &lq_work, &lq_base};
    if (threadIdx.x == 0) {
        lq_n = 0;
        lq_work = 0;
    }
    __syncthreads();
    const unsigned long long work = (unsigned long long)pk_work(a) + pk_work(b);
    if (C && work > pull_threshold)
        pull_body(keys, values, C, N, row_offset, nx, levels, level, lq);
    else
        push_body(keys, values, row_offset, front, front_big, nx, cnt, cnt_big, levels, level, small_grid, tile_cap,
                  s_beg, s_ofs, scratch, lq);
}

__global__ void k_bfs_clear(unsigned long long *__restrict__ cnt, unsigned long long *__restrict__ cnt_big,
                            uint32_t a, uint32_t b) {
    for (uint32_t i = a + blockIdx.x * blockDim.x + threadIdx.x; i < b; i += gridDim.x * blockDim.x) {
        cnt[i] = 0;
        cnt_big[i] = 0;
    }
}

}  // namespace

void launch_bfs_init(const BfsState &b, uint32_t N, uint32_t start, const uint32_t *row_offset, hipStream_t st) {
    uint32_t g = (uint32_t)std::min<uint64_t>(2048, ((uint64_t)N + 255) / 256);
    if (g == 0) g = 1;
    hipLaunchKernelGGL(k_bfs_init, dim3(g), dim3(256), 0, st, b.levels, N, start, b.q[1], b.qbig[1], b.cnt, b.cnt_big,
                       row_offset);
}

void launch_bfs_clear(const BfsState &b, uint32_t a, uint32_t e, hipStream_t st) {
    if (e <= a) return;
    hipLaunchKernelGGL(k_bfs_clear, dim3(std::max(1u, std::min(64u, (e - a + 255) / 256))), dim3(256), 0, st, b.cnt,
                       b.cnt_big, a, e);
}

void launch_bfs_level(const gkey *keys, const gval *values, uint64_t C, uint32_t N, const uint32_t *row_offset,
                      const BfsState &b, uint32_t level, unsigned long long pull_threshold, hipStream_t st) {
    const int cur = level & 1, nxt = (level + 1) & 1;
    NextLevel nx{b.q[nxt], b.qbig[nxt], b.cnt + level + 1, b.cnt_big + level + 1};
    const uint32_t small_grid = 1536, big_grid = 512;
    hipLaunchKernelGGL(k_bfs_step, dim3(small_grid + big_grid), dim3(BFS_THREADS), 0, st, keys, values, C, N,
                       row_offset, b.q[cur], b.qbig[cur], nx, b.cnt, b.cnt_big, b.levels, level, pull_threshold,
                       small_grid, std::min(PUSH_TILE, std::max(PUSH_MIN_TILE, b.tile_cap)));
}

}  // namespace gpma
