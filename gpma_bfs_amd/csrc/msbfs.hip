// Multi-source BFS: up to 64 sources at once over the gapped packed array,
// one bit per source in a 64-bit mask per vertex (the "MS-BFS" bit-parallel
// scheme of Then et al., VLDB 2015), so one traversal of an edge serves every
// source that reaches its tail at the same level.
//
// Per level l (masks: seen, fr = sources for which v is at level l):
//   expand:   for every frontier vertex v and live slot (v -> w),
//             add = fr[v] & ~seen[w]; atomicOr(nx[w], add); the first lane
//             to make nx[w] non-zero queues w as touched. Push walks the
//             frontier list's slot ranges (long ranges cut into 2048-slot
//             pieces); pull streams every slot of the array when the
//             frontier's slots exceed a share of it.
//   finalize: for every touched w: new = nx[w] & ~seen[w]; nx[w] = 0;
//             seen[w] |= new; fr_next[w] = new; level(s, w) = l + 1 for each
//             bit s of new; w joins the next frontier if new != 0. The old
//             frontier's fr entries are cleared in the same pass.
// Dense levels (frontier slots above N / 4) queue no touched vertices: the
// expand ORs without reading the old value back and the finalize scans nx
// over every vertex in order. On the power-law 1.63M-node graph with 64
// random sources this halves the widest levels' expand time (the queued
// form's ORs wait for their old values) for a somewhat slower finalize.
// Levels are recorded as events: each time a vertex w gets newly seen
// sources, the lane that owns w appends one 16-byte event (the fresh source
// mask and the level) to w's own list of MS_EVENTS slots, one vector store
// for the whole wave. Slots are slot-major (ev[k][v]), so the output pass
// reads each slot of 64 consecutive vertices as one coalesced load. The output pass rebuilds every source's level from a
// vertex's events (the sources in one event share its level) and transposes
// them into the caller's source-major u32 arrays, with per-source reachable
// counts. A vertex reached at more than MS_EVENTS distinct levels records
// the rest in a vertex-major row (a vertex's 64 levels as u16 in one 128-byte
// line, written by a loop over the wave's overflowing lanes). A traversal
// that reaches level 65,535 first widens those rows to u32 (k_ms_widen;
// 256-byte rows), so no depth wraps.
//
// The level of source s at w is 1 + hop distance, as in the single-source
// BFS (gpma_bfs.cuh:446-509 semantics), so every source's array equals the
// single-source result bit for bit.
#include "common.h"
#include "kernels.h"

namespace gpma {

namespace {

constexpr uint32_t T = 256;
constexpr uint32_t MS_TILE = 64;      // most frontier vertices per tile (the tile_cap knob lowers it)
constexpr uint32_t MS_MIN_TILE = 16;  // fewest (short frontiers)
constexpr uint32_t MS_UNROLL = 4;
constexpr uint32_t MS_BIG_RANGE = 2048;
constexpr uint32_t MS_BIG_CHUNK = 2048;
constexpr uint32_t MS_BIG_LIST = 2048;
constexpr uint32_t MS_PULL_PER_THREAD = 16;
constexpr uint32_t LQ = 2048;
constexpr uint32_t MS_EVENTS = 8;     // event slots per vertex (16 bytes each)
constexpr uint8_t EV_OVF = 0x80;      // event count flag: levels also in the overflow row

__device__ __forceinline__ uint32_t rlen(const uint32_t *ro, uint32_t v) {
    const uint32_t b = ro[v], e = ro[v + 1];
    return e > b ? e - b : 0;
}

// Masks start empty. The [N][64] level staging is not cleared: the output
// reads an entry only where the source's bit is set in seen[v], and every
// such entry was written when that bit was set.
__global__ void k_ms_init(uint64_t *__restrict__ seen, uint64_t *__restrict__ fr0, uint64_t *__restrict__ fr1,
                          uint64_t *__restrict__ nx, uint8_t *__restrict__ evn, uint32_t N) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (uint64_t i = t; i < N; i += stride) {
        seen[i] = 0;
        fr0[i] = 0;
        fr1[i] = 0;
        nx[i] = 0;
    }
    uint32_t *e4 = reinterpret_cast<uint32_t *>(evn);  // N rounded up to 4 bytes
    for (uint64_t i = t; i < ((uint64_t)N + 3) / 4; i += stride) e4[i] = 0;
}
// One lane per source: mark its bit, stage level 1, and queue the first
// occurrence of each start vertex (short or long range list).
__global__ void k_ms_sources(const uint32_t *__restrict__ starts, uint32_t nsrc, uint64_t *__restrict__ seen,
                             uint64_t *__restrict__ fr, uint4 *__restrict__ ev, uint8_t *__restrict__ evn,
                             uint32_t *__restrict__ q,
                             uint32_t *__restrict__ qbig, unsigned long long *__restrict__ cnt,
                             unsigned long long *__restrict__ cnt_big, uint32_t *__restrict__ tcnt,
                             const uint32_t *__restrict__ ro, uint32_t *__restrict__ gate, MsStepGate sg,
                             uint32_t ca, uint32_t cb) {
    if (sg.np0) {
        // a captured step: its host-work flag (the update left work for the
        // host), which also stands this traversal down
        const uint32_t g = (*sg.np0 >= sg.bulk || *sg.np_top != 0 || *sg.err != 0) ? 1u : 0u;
        if (threadIdx.x == 0) *gate = g;
        if (g) nsrc = 0;
    } else if (gate && *gate) {
        nsrc = 0;
    }
    for (uint32_t i = ca + threadIdx.x; i < cb; i += WAVE) {  // the counts of levels ca..cb-1
        cnt[i] = 0;
        cnt_big[i] = 0;
        tcnt[i] = 0;
    }
    const uint32_t lane = threadIdx.x;  // one wave
    const bool valid = lane < nsrc;
    const uint32_t s = valid ? starts[lane] : 0;
    // m: the lanes (sources) with this lane's start, from one ballot per bit
    uint64_t m = ballot(valid);
#pragma unroll
    for (uint32_t b = 0; b < 32; b++) {
        const bool bit = (s >> b) & 1u;
        const uint64_t bb = ballot(valid && bit);
        m &= bit ? bb : ~bb;
    }
    // the first lane of each distinct start writes its masks and its level-1
    // event (slot 0 of vertex s): all of its sources
    const bool first = valid && (m & lanemask_lt()) == 0;
    if (first) {
        seen[s] = m;
        fr[s] = m;
        ev[s] = make_uint4((uint32_t)m, (uint32_t)(m >> 32), 1u, 0u);
        evn[s] = 1;
    }
    const uint64_t len = first ? rlen(ro, s) : 0;
    const bool big = first && len > MS_BIG_RANGE, small = first && !big;
    const uint64_t ms = ballot(small), mb = ballot(big);
    if (small) q[__popcll(ms & lanemask_lt())] = s;
    if (big) qbig[__popcll(mb & lanemask_lt())] = s;
    uint64_t ws = small ? len : 0, wb = big ? len : 0;
    for (uint32_t d = WAVE / 2; d > 0; d >>= 1) {
        ws += __shfl_xor(ws, d, WAVE);
        wb += __shfl_xor(wb, d, WAVE);
    }
    if (lane == 0) {
        cnt[0] = cnt_big[0] = 0;
        tcnt[0] = tcnt[1] = 0;
        cnt[1] = (ws << 32) | (unsigned long long)__popcll(ms);
        cnt_big[1] = (wb << 32) | (unsigned long long)__popcll(mb);
    }
}

// A level whose frontier spans more than N / MS_DENSE_DIV slots is dense:
// its expand step queues no touched vertices and its finalize step scans
// every vertex's accumulated mask in order. Both steps decide from the same
// frontier count.
constexpr uint32_t MS_DENSE_DIV = 4;
__device__ __forceinline__ bool ms_dense(unsigned long long work, uint32_t N) {
    return work * MS_DENSE_DIV > N;
}

struct MsQueue {  // LDS staging of touched vertices
    uint32_t *buf, *n, *base;
};

__device__ __forceinline__ void touch(bool first, uint32_t w, const MsQueue &lq) {
    const uint64_t m = __ballot(first);
    if (!m) return;
    const uint32_t leader = (uint32_t)__ffsll((unsigned long long)m) - 1;
    uint32_t b = 0;
    if (lane_id() == leader) b = atomicAdd(lq.n, (uint32_t)__popcll(m));
    b = __shfl(b, leader, WAVE);
    if (first) lq.buf[b + (uint32_t)__popcll(m & lanemask_lt())] = w;
}

__device__ __forceinline__ void lq_flush(const MsQueue &lq, uint32_t *touched, uint32_t *tcnt, uint32_t room) {
    __syncthreads();
    const uint32_t n = *lq.n;
    if (n == 0 || (room <= LQ && n + room <= LQ)) return;
    if (threadIdx.x == 0) *lq.base = atomicAdd(tcnt, n);
    __syncthreads();
    const uint32_t b = *lq.base;
    for (uint32_t i = threadIdx.x; i < n; i += T) touched[b + i] = lq.buf[i];
    __syncthreads();
    if (threadIdx.x == 0) *lq.n = 0;
    __syncthreads();
}

// Three-stage batch of slots (source masks already known): seen words, then
// the atomic ORs, then touched-queue appends.
// Dense levels queue nothing: their ORs need no old value.
template <uint32_t U>
__device__ __forceinline__ void ms_visit(const gkey (&k)[U], const gval (&v)[U], const uint64_t (&m)[U],
                                         const uint64_t *seen, uint64_t *nx, const MsQueue &lq, bool dense) {
    uint32_t w[U];
    uint64_t add[U], old[U];
    bool cand[U];
#pragma unroll
    for (uint32_t u = 0; u < U; u++) {
        w[u] = (uint32_t)k[u];
        cand[u] = m[u] != 0 && w[u] != COL_NONE && v[u] != 0;
        add[u] = cand[u] ? m[u] & ~*(volatile const uint64_t *)(seen + w[u]) : 0;
    }
    if (dense) {
#pragma unroll
        for (uint32_t u = 0; u < U; u++)
            if (add[u]) (void)__hip_atomic_fetch_or(nx + w[u], add[u], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return;
    }
#pragma unroll
    for (uint32_t u = 0; u < U; u++) old[u] = add[u] ? atomicOr((unsigned long long *)(nx + w[u]), add[u]) : ~0ull;
#pragma unroll
    for (uint32_t u = 0; u < U; u++) touch(add[u] != 0 && old[u] == 0, w[u], lq);
}

__global__ __launch_bounds__(T) void k_ms_expand(const gkey *__restrict__ keys, const gval *__restrict__ values,
                                                 uint64_t C, uint32_t N, const uint32_t *__restrict__ ro,
                                                 const uint32_t *__restrict__ q, const uint32_t *__restrict__ qbig,
                                                 const unsigned long long *__restrict__ cnt,
                                                 const unsigned long long *__restrict__ cnt_big,
                                                 const uint64_t *__restrict__ fr, const uint64_t *__restrict__ seen,
                                                 uint64_t *__restrict__ nx, uint32_t *__restrict__ touched,
                                                 uint32_t *__restrict__ tcnt, uint32_t level,
                                                 unsigned long long pull_threshold, uint32_t small_grid,
                                                 uint32_t tile_cap) {
    __shared__ uint32_t s_beg[MS_BIG_LIST];
    __shared__ uint32_t s_ofs[MS_BIG_LIST + 1];
    __shared__ uint32_t scratch[T / WAVE + 1];
    __shared__ uint32_t lq_buf[LQ];
    __shared__ uint32_t lq_n, lq_base;
    const unsigned long long a = cnt[level], b = cnt_big[level];
    const uint32_t nf = (uint32_t)a, nbig = (uint32_t)b;
    if (nf == 0 && nbig == 0) return;
    const MsQueue lq{lq_buf, &lq_n, &lq_base};
    if (threadIdx.x == 0) lq_n = 0;
    __syncthreads();
    const uint32_t tid = threadIdx.x;
    const unsigned long long work = (a >> 32) + (b >> 32);
    const bool dense = ms_dense(work, N);
    if (C && work > pull_threshold) {
        // streaming pull over every slot
        const uint64_t per = (uint64_t)T * MS_PULL_PER_THREAD;
        for (uint64_t base = (uint64_t)blockIdx.x * per; base < C; base += (uint64_t)gridDim.x * per) {
            for (uint32_t r = 0; r < MS_PULL_PER_THREAD; r += MS_UNROLL) {
                gkey k[MS_UNROLL];
                gval v[MS_UNROLL];
                uint64_t m[MS_UNROLL];
#pragma unroll
                for (uint32_t u = 0; u < MS_UNROLL; u++) {
                    const uint64_t p = base + (uint64_t)(r + u) * T + tid;
                    k[u] = p < C ? keys[p] : KEY_NONE;
                    v[u] = p < C ? values[p] : 0;
                }
#pragma unroll
                for (uint32_t u = 0; u < MS_UNROLL; u++) {
                    const uint32_t src = (uint32_t)(k[u] >> 32);
                    m[u] = src < N ? fr[src] : 0;
                }
                ms_visit<MS_UNROLL>(k, v, m, seen, nx, lq, dense);
                if (!dense) lq_flush(lq, touched, tcnt, T * MS_UNROLL);
            }
        }
    } else if (blockIdx.x < small_grid) {
        // vertices per tile: a short frontier is spread over every workgroup
        const uint32_t tn = (uint32_t)min<uint64_t>(tile_cap, max<uint64_t>(MS_MIN_TILE, (nf + small_grid - 1) / small_grid));
        for (uint64_t tile = (uint64_t)blockIdx.x * tn; tile < nf; tile += (uint64_t)small_grid * tn) {
            const uint64_t fi = tile + tid;
            uint32_t bb = 0, len = 0;
            if (tid < tn && fi < nf) {
                const uint32_t vv = q[fi];
                bb = ro[vv];
                len = rlen(ro, vv);
                s_beg[MS_TILE + tid] = vv;  // vertex ids after the offsets
            }
            uint32_t total;
            const uint32_t ofs = block_exclusive_scan(len, scratch, &total);
            if (tid < tn) {
                s_beg[tid] = bb;
                s_ofs[tid] = ofs;
            }
            __syncthreads();
            for (uint32_t j0 = 0; j0 < total; j0 += T * MS_UNROLL) {
                gkey k[MS_UNROLL];
                gval v[MS_UNROLL];
                uint64_t m[MS_UNROLL];
#pragma unroll
                for (uint32_t u = 0; u < MS_UNROLL; u++) {
                    const uint32_t j = j0 + u * T + tid;
                    k[u] = KEY_NONE;
                    v[u] = 0;
                    m[u] = 0;
                    if (j < total) {
                        uint32_t lo = 0, hi = tn;
                        while (hi - lo > 1) {
                            const uint32_t mid = (lo + hi) >> 1;
                            if (s_ofs[mid] <= j) lo = mid;
                            else hi = mid;
                        }
                        const uint64_t slot = (uint64_t)s_beg[lo] + (j - s_ofs[lo]);
                        k[u] = keys[slot];
                        v[u] = values[slot];
                        m[u] = fr[s_beg[MS_TILE + lo]];
                    }
                }
                ms_visit<MS_UNROLL>(k, v, m, seen, nx, lq, dense);
                if (!dense) lq_flush(lq, touched, tcnt, T * MS_UNROLL);
            }
            __syncthreads();
        }
    } else {
        const uint32_t bgrid = gridDim.x - small_grid;
        const uint32_t bid = blockIdx.x - small_grid;
        for (uint32_t l0 = 0; l0 < nbig; l0 += MS_BIG_LIST) {
            const uint32_t lcount = min(MS_BIG_LIST, nbig - l0);
            uint32_t run = 0;
            for (uint32_t i0 = 0; i0 < lcount; i0 += T) {
                const uint32_t i = i0 + tid;
                uint32_t nch = 0, bb = 0;
                if (i < lcount) {
                    const uint32_t vv = qbig[l0 + i];
                    bb = ro[vv];
                    nch = (rlen(ro, vv) + MS_BIG_CHUNK - 1) / MS_BIG_CHUNK;
                }
                uint32_t tot;
                const uint32_t pre = block_exclusive_scan(nch, scratch, &tot);
                if (i < lcount) {
                    s_beg[i] = bb;
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
                const uint32_t vv = qbig[l0 + lo];
                const uint64_t mask = fr[vv];
                const uint64_t beg = (uint64_t)s_beg[lo] + (uint64_t)(c - s_ofs[lo]) * MS_BIG_CHUNK;
                const uint64_t end = min((uint64_t)ro[vv + 1], beg + MS_BIG_CHUNK);
                for (uint64_t p0 = beg; p0 < end; p0 += T * MS_UNROLL) {
                    gkey k[MS_UNROLL];
                    gval v[MS_UNROLL];
                    uint64_t m[MS_UNROLL];
#pragma unroll
                    for (uint32_t u = 0; u < MS_UNROLL; u++) {
                        const uint64_t p = p0 + u * T + tid;
                        const bool in = p < end;
                        k[u] = in ? keys[p] : KEY_NONE;
                        v[u] = in ? values[p] : 0;
                        m[u] = in ? mask : 0;
                    }
                    ms_visit<MS_UNROLL>(k, v, m, seen, nx, lq, dense);
                    if (!dense) lq_flush(lq, touched, tcnt, T * MS_UNROLL);
                }
            }
            __syncthreads();
        }
    }
    lq_flush(lq, touched, tcnt, LQ + 1);
}

// Finalize level `level`: clear the old frontier masks, turn touched
// vertices' accumulated masks into newly seen sources, record their levels,
// and queue the next frontier (short / long ranges, packed count + slots).
__global__ __launch_bounds__(T) void k_ms_finalize(const uint32_t *__restrict__ q, const uint32_t *__restrict__ qbig,
                                                   const unsigned long long *__restrict__ cnt,
                                                   const unsigned long long *__restrict__ cnt_big,
                                                   uint64_t *__restrict__ fr_cur, uint64_t *__restrict__ fr_next,
                                                   uint64_t *__restrict__ seen, uint64_t *__restrict__ nx,
                                                   uint16_t *__restrict__ lv16, uint32_t *__restrict__ lv32,
                                                   uint4 *__restrict__ ev, uint8_t *__restrict__ evn,
                                                   const uint32_t *__restrict__ touched,
                                                   const uint32_t *__restrict__ tcnt, uint32_t *__restrict__ q2,
                                                   uint32_t *__restrict__ qbig2, unsigned long long *__restrict__ cnt2,
                                                   unsigned long long *__restrict__ cnt_big2,
                                                   const uint32_t *__restrict__ ro, uint32_t N, uint32_t level) {
    __shared__ uint32_t lq_buf[LQ];
    __shared__ uint32_t lq_n, lq_work, lq_base;
    const unsigned long long ca = cnt[level], cb = cnt_big[level];
    const uint32_t nf = (uint32_t)ca, nbig = (uint32_t)cb;
    if (nf == 0 && nbig == 0) return;
    // dense: every vertex in order (nx is zero where nothing was added)
    const bool dense = ms_dense((ca >> 32) + (cb >> 32), N);
    const uint32_t nt = dense ? N : *tcnt;
    const uint64_t stride = (uint64_t)gridDim.x * T;
    const uint64_t t0 = (uint64_t)blockIdx.x * T + threadIdx.x;
    for (uint64_t i = t0; i < nf; i += stride) fr_cur[q[i]] = 0;
    for (uint64_t i = t0; i < nbig; i += stride) fr_cur[qbig[i]] = 0;
    if (threadIdx.x == 0) {
        lq_n = 0;
        lq_work = 0;
    }
    __syncthreads();
    // u16 staging while every level fits; lv32 (non-null once widened) past that
    const uint32_t newlev = level + 1;
    // a workgroup whose whole share fits the LDS queue never flushes early
    const bool roomy = (((uint64_t)nt + stride - 1) / stride) * T <= LQ;
    for (uint64_t i0 = (uint64_t)blockIdx.x * T; i0 < nt; i0 += stride) {
        const uint64_t i = i0 + threadIdx.x;
        bool small = false, big = false, ovf = false;
        uint32_t w = 0, len = 0;
        uint64_t fresh = 0;
        if (i < nt) {
            w = dense ? (uint32_t)i : touched[i];
            const uint64_t acc = nx[w];
            if (acc) nx[w] = 0;
            const uint64_t sn = acc ? seen[w] : ~0ull;
            const uint8_t en = acc ? evn[w] : 0;
            fresh = acc & ~sn;
            if (fresh) {
                seen[w] = sn | fresh;
                fr_next[w] = fresh;
                len = rlen(ro, w);
                small = len <= MS_BIG_RANGE;
                big = !small;
                // the level of the fresh sources: one event, or the overflow row
                const uint32_t k = en & ~EV_OVF;
                if (k < MS_EVENTS) {
                    ev[(uint64_t)k * N + w] = make_uint4((uint32_t)fresh, (uint32_t)(fresh >> 32), newlev, 0u);
                    evn[w] = (uint8_t)(en + 1);
                } else {
                    ovf = true;
                    if (!(en & EV_OVF)) evn[w] = (uint8_t)(en | EV_OVF);
                }
            }
        }
        {
            const uint32_t lane = lane_id();
            for (uint64_t todo = __ballot(ovf); todo; todo &= todo - 1) {
                const int j = __ffsll((unsigned long long)todo) - 1;
                const uint32_t wj = (uint32_t)__builtin_amdgcn_readlane((int)w, j);
                const uint32_t flo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)fresh, j);
                const uint32_t fhi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(fresh >> 32), j);
                if ((((uint64_t)fhi << 32 | flo) >> lane) & 1ull) {
                    if (lv32) lv32[(uint64_t)wj * 64 + lane] = newlev;
                    else lv16[(uint64_t)wj * 64 + lane] = (uint16_t)newlev;
                }
            }
        }
        // short ranges through LDS, long ranges directly
        const uint64_t ms = __ballot(small);
        if (ms) {
            const uint32_t leader = (uint32_t)__ffsll((unsigned long long)ms) - 1;
            uint32_t b = 0;
            if (lane_id() == leader) b = atomicAdd(&lq_n, (uint32_t)__popcll(ms));
            b = __shfl(b, leader, WAVE);
            if (small) lq_buf[b + (uint32_t)__popcll(ms & lanemask_lt())] = w;
            uint32_t wsum = small ? len : 0;
            for (int d = WAVE / 2; d > 0; d >>= 1) wsum += __shfl_xor(wsum, d, WAVE);
            if (lane_id() == leader) atomicAdd(&lq_work, wsum);
        }
        const uint64_t mb = __ballot(big);
        if (mb) {
            uint64_t wsum = big ? len : 0;
            for (int d = WAVE / 2; d > 0; d >>= 1) wsum += __shfl_xor(wsum, d, WAVE);
            const uint32_t leader = (uint32_t)__ffsll((unsigned long long)mb) - 1;
            uint32_t b = 0;
            if (lane_id() == leader)
                b = (uint32_t)atomicAdd(cnt_big2, (wsum << 32) | (unsigned long long)__popcll(mb));
            b = __shfl(b, leader, WAVE);
            if (big) qbig2[b + (uint32_t)__popcll(mb & lanemask_lt())] = w;
        }
        if (roomy) continue;
        __syncthreads();
        if (lq_n + T > LQ) {
            if (threadIdx.x == 0) lq_base = (uint32_t)atomicAdd(cnt2, ((unsigned long long)lq_work << 32) | lq_n);
            __syncthreads();
            for (uint32_t j = threadIdx.x; j < lq_n; j += T) q2[lq_base + j] = lq_buf[j];
            __syncthreads();
            if (threadIdx.x == 0) lq_n = lq_work = 0;
            __syncthreads();
        }
    }
    __syncthreads();
    if (lq_n) {
        if (threadIdx.x == 0) lq_base = (uint32_t)atomicAdd(cnt2, ((unsigned long long)lq_work << 32) | lq_n);
        __syncthreads();
        for (uint32_t j = threadIdx.x; j < lq_n; j += T) q2[lq_base + j] = lq_buf[j];
    }
}

// The 8 bits of x at bit positions 0, 4, 8, ..., 28.
__device__ __forceinline__ uint32_t spread4(uint32_t x) {
    x = (x | (x << 12)) & 0x000F000Fu;
    x = (x | (x << 6)) & 0x03030303u;
    return (x | (x << 3)) & 0x11111111u;
}

// Row of source S in a lane's level table: bit k of the index is bit S of
// plane k (the planes of the event index), bit 3 is bit S of nc (no event
// holds S).
__device__ __forceinline__ uint32_t ms_event_row(uint32_t S, uint64_t p0, uint64_t p1, uint64_t p2, uint64_t nc) {
    const uint32_t h = S & 32, b = S & 31;
    const uint32_t w0 = (uint32_t)(p0 >> h), w1 = (uint32_t)(p1 >> h), w2 = (uint32_t)(p2 >> h),
                   w3 = (uint32_t)(nc >> h);
    return ((w0 >> b) & 1u) | (((w1 >> b) & 1u) << 1) | (((w2 >> b) & 1u) << 2) | (((w3 >> b) & 1u) << 3);
}

// Source-major u32 levels from the events (and the overflow rows), with
// per-source reachable counts. One wave per 64-vertex group, no barrier:
// each lane loads its vertex's events (only if some source reached it) and
// keeps their levels in its own column of an LDS table (rows 0..7; rows
// 8..15 hold 0). Source s's level at the lane's vertex is then one table
// read at row idx(s) = e(s) | 8 * (no event holds s), e(s) = the index of the
// event whose mask holds bit s (a source sits in at most one event), taken
// from three bit planes of the event indices: no branch per source. For
// every source the wave writes the group's 64 levels as one coalesced store,
// and lane s counts source s with a ballot. A vertex reached at more than
// MS_EVENTS levels then overwrites its sources past the events from its
// overflow row. partial[workgroup][s]: per-workgroup counts. Workgroup 0 also zeroes
// reach[0 .. nsrc) for k_ms_reach. WIDE: the overflow rows hold u32 levels,
// else u16.
template <bool WIDE>
__global__ __launch_bounds__(T) void k_ms_output(const uint4 *__restrict__ ev, const uint8_t *__restrict__ evn,
                                                 const void *__restrict__ lv, const uint64_t *__restrict__ seen,
                                                 uint32_t N, uint32_t nsrc, uint32_t *__restrict__ out,
                                                 uint32_t *__restrict__ partial, uint32_t *__restrict__ reach) {
    __shared__ uint32_t tab[2 * MS_EVENTS][T];
    __shared__ uint32_t wcnt[T / WAVE][64];  // per wave: one group's count per source
    const uint32_t lane = lane_id();
    if (blockIdx.x == 0 && threadIdx.x < nsrc) reach[threadIdx.x] = 0;  // k_ms_reach adds into it
#pragma unroll
    for (uint32_t k = MS_EVENTS; k < 2 * MS_EVENTS; k++) tab[k][threadIdx.x] = 0;
    const uint32_t gw = (blockIdx.x * T + threadIdx.x) / WAVE;
    const uint32_t nw = gridDim.x * (T / WAVE);
    uint32_t acc = 0;
    for (uint64_t v0 = (uint64_t)gw * WAVE; v0 < N; v0 += (uint64_t)nw * WAVE) {
        const uint64_t v = v0 + lane;
        const bool in = v < N;
        const uint64_t mask = in ? seen[v] : 0;
        const uint8_t en = mask ? evn[v] : 0;
        const uint32_t ne = en & ~EV_OVF;
        uint64_t p0 = 0, p1 = 0, p2 = 0, cov = 0;
#pragma unroll
        for (uint32_t k = 0; k < MS_EVENTS; k++) {
            const uint4 e = k < ne ? ev[(uint64_t)k * N + v] : make_uint4(0, 0, 0, 0);
            const uint64_t m = (uint64_t)e.y << 32 | e.x;
            tab[k][threadIdx.x] = e.z;
            cov |= m;
            if (k & 1) p0 |= m;
            if (k & 2) p1 |= m;
            if (k & 4) p2 |= m;
        }
        static_assert(MS_EVENTS == 8, "three bit planes");
        const uint64_t nc = ~cov;
        if (nsrc == 64 && v0 + WAVE <= N && N <= (1u << 30)) {
            // a full group and every source: no per-lane or per-source
            // condition. nib[j]: the table rows of sources 8j..8j+7, 4 bits each
            uint32_t nib[8];
#pragma unroll
            for (uint32_t j = 0; j < 8; j++) {
                const uint32_t h = (j >> 2) * 32, b = (j & 3) * 8;
                nib[j] = spread4(((uint32_t)(p0 >> h) >> b) & 0xFFu) | spread4(((uint32_t)(p1 >> h) >> b) & 0xFFu) << 1 |
                         spread4(((uint32_t)(p2 >> h) >> b) & 0xFFu) << 2 | spread4(((uint32_t)(nc >> h) >> b) & 0xFFu) << 3;
            }
            const uint32_t voff = (uint32_t)v * 4u;  // byte offset in a source row
            uint32_t *wc = wcnt[threadIdx.x / WAVE];
#pragma unroll
            for (uint32_t s = 0; s < 64; s++) {
                const uint32_t idx = (nib[s >> 3] >> ((s & 7) * 4)) & 15u;
                char *row = reinterpret_cast<char *>(out + (uint64_t)s * N);
                *reinterpret_cast<uint32_t *>(row + voff) = tab[idx][threadIdx.x];
                wc[s] = (uint32_t)__popcll(ballot((uint32_t)(mask >> s) & 1u));  // same word from every lane
            }
            wave_sync();
            acc += wc[lane];
            wave_sync();
        } else {
            const uint32_t vi = in ? (uint32_t)v : 0u;
            for (uint32_t s = 0; s < nsrc; s++) {
                const uint32_t idx = ms_event_row(s, p0, p1, p2, nc);
                const uint32_t level = tab[idx][threadIdx.x];
                if (in) out[(uint64_t)s * N + vi] = level;
                const uint32_t c = (uint32_t)__popcll(ballot((uint32_t)(mask >> s) & 1u));
                acc += lane == s ? c : 0u;
            }
        }
        if (en & EV_OVF) {  // sources past the event slots: the overflow row
            for (uint64_t bits = mask & ~cov; bits; bits &= bits - 1) {
                const uint32_t s = (uint32_t)__ffsll((unsigned long long)bits) - 1;
                if (s < nsrc)
                    out[(uint64_t)s * N + v] = WIDE ? reinterpret_cast<const uint32_t *>(lv)[v * 64 + s]
                                                    : reinterpret_cast<const uint16_t *>(lv)[v * 64 + s];
            }
        }
    }
    __syncthreads();  // every wave is past its loop: wcnt is free
    wcnt[threadIdx.x / WAVE][lane] = acc;
    __syncthreads();
    if (threadIdx.x < WAVE) {
        uint32_t t = 0;
#pragma unroll
        for (uint32_t w = 0; w < T / WAVE; w++) t += wcnt[w][threadIdx.x];
        partial[(uint64_t)blockIdx.x * WAVE + threadIdx.x] = t;
    }
}

// reach[s] += the sum of column s over a share of the partials' rows: each
// workgroup reads whole 64-word rows (coalesced), reduces in LDS and adds
// one word per source (reach was zeroed by the output pass).
constexpr uint32_t RT = 1024, RB = 64;  // threads and workgroups of the reduction
__global__ __launch_bounds__(RT) void k_ms_reach(const uint32_t *__restrict__ partial, uint32_t nrows, uint32_t nsrc,
                                                 uint32_t *__restrict__ reach) {
    __shared__ uint32_t red[RT / 64][64];
    const uint32_t s = threadIdx.x & 63, r0 = threadIdx.x >> 6;
    const uint32_t per = (nrows + gridDim.x - 1) / gridDim.x;
    const uint32_t a = blockIdx.x * per, b = min(nrows, a + per);
    uint32_t sum = 0;
    for (uint32_t r = a + r0; r < b; r += RT / 64) sum += partial[(uint64_t)r * 64 + s];
    red[r0][s] = sum;
    __syncthreads();
    if (threadIdx.x < nsrc) {
        uint32_t t = 0;
        for (uint32_t j = 0; j < RT / 64; j++) t += red[j][threadIdx.x];
        if (t) atomicAdd(reach + threadIdx.x, t);
    }
}

// lv32[v][s] = lv16[v][s] wherever source s has reached v: the staging of a
// traversal about to pass level 65,535 moves to u32 rows.
__global__ void k_ms_widen(const uint16_t *__restrict__ lv16, const uint64_t *__restrict__ seen, uint32_t N,
                           uint32_t *__restrict__ lv32) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (uint64_t)N * 64; i += stride) {
        const uint64_t v = i >> 6;
        if ((seen[v] >> (i & 63)) & 1ull) lv32[i] = lv16[i];
    }
}

__global__ void k_ms_clear(unsigned long long *__restrict__ cnt, unsigned long long *__restrict__ cnt_big,
                           uint32_t *__restrict__ tcnt, uint32_t a, uint32_t b) {
    for (uint32_t i = a + blockIdx.x * blockDim.x + threadIdx.x; i < b; i += gridDim.x * blockDim.x) {
        cnt[i] = 0;
        cnt_big[i] = 0;
        tcnt[i] = 0;
    }
}

}  // namespace

void ms_init(const MsState &m, uint32_t N, const uint32_t *starts, uint32_t nsrc, const uint32_t *ro, hipStream_t st,
             const uint32_t *gate) {
    hipLaunchKernelGGL(k_ms_init, dim3(2048), dim3(256), 0, st, m.seen, m.fr[0], m.fr[1], m.nx, m.evn, N);
    hipLaunchKernelGGL(k_ms_sources, dim3(1), dim3(64), 0, st, starts, nsrc, m.seen, m.fr[1], m.ev, m.evn, m.q[1],
                       m.qbig[1], m.cnt, m.cnt_big, m.tcnt, ro, const_cast<uint32_t *>(gate), MsStepGate{}, 0u, 0u);
}

void ms_init_step(const MsState &m, uint32_t N, const uint32_t *starts, uint32_t nsrc, const uint32_t *ro,
                  const MsStepGate &sg, uint32_t *gate, uint32_t clear_a, uint32_t clear_b, hipStream_t st) {
    hipLaunchKernelGGL(k_ms_init, dim3(2048), dim3(256), 0, st, m.seen, m.fr[0], m.fr[1], m.nx, m.evn, N);
    hipLaunchKernelGGL(k_ms_sources, dim3(1), dim3(64), 0, st, starts, nsrc, m.seen, m.fr[1], m.ev, m.evn, m.q[1],
                       m.qbig[1], m.cnt, m.cnt_big, m.tcnt, ro, gate, sg, clear_a, clear_b);
}

void ms_clear(const MsState &m, uint32_t a, uint32_t b, hipStream_t st) {
    if (b <= a) return;
    hipLaunchKernelGGL(k_ms_clear, dim3(1), dim3(256), 0, st, m.cnt, m.cnt_big, m.tcnt, a, b);
}

void ms_level(const gkey *keys, const gval *values, uint64_t C, uint32_t N, const uint32_t *ro, const MsState &m,
              uint32_t level, unsigned long long pull_threshold, hipStream_t st) {
    const int cur = level & 1, nxt = (level + 1) & 1;
    const uint32_t small_grid = 1536, big_grid = 512;
    hipLaunchKernelGGL(k_ms_expand, dim3(small_grid + big_grid), dim3(T), 0, st, keys, values, C, N, ro, m.q[cur],
                       m.qbig[cur], m.cnt, m.cnt_big, m.fr[cur], m.seen, m.nx, m.touched, m.tcnt + level, level,
                       pull_threshold, small_grid, std::min(MS_TILE, std::max(MS_MIN_TILE, m.tile_cap)));
    hipLaunchKernelGGL(k_ms_finalize, dim3(1024), dim3(T), 0, st, m.q[cur], m.qbig[cur], m.cnt, m.cnt_big, m.fr[cur],
                       m.fr[nxt], m.seen, m.nx, m.lv16, m.lv32, m.ev, m.evn, m.touched, m.tcnt + level, m.q[nxt],
                       m.qbig[nxt],
                       m.cnt + level + 1, m.cnt_big + level + 1, ro, N, level);
}

// Workgroups of the output pass: one wave per 64-vertex group, at most 2048.
static uint32_t ms_output_grid(uint32_t N) {
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(2048, ((uint64_t)N + T - 1) / T));
}

void ms_output(const MsState &m, uint32_t N, uint32_t nsrc, uint32_t *out, uint32_t *reach, hipStream_t st) {
    const uint32_t nb = ms_output_grid(N);
    if (m.lv32)
        hipLaunchKernelGGL(k_ms_output<true>, dim3(nb), dim3(T), 0, st, m.ev, m.evn, (const void *)m.lv32, m.seen, N,
                           nsrc, out, m.partial, reach);
    else
        hipLaunchKernelGGL(k_ms_output<false>, dim3(nb), dim3(T), 0, st, m.ev, m.evn, (const void *)m.lv16, m.seen, N,
                           nsrc, out, m.partial, reach);
    hipLaunchKernelGGL(k_ms_reach, dim3(RB), dim3(RT), 0, st, m.partial, nb, nsrc, reach);
}

uint32_t ms_output_blocks(uint32_t N) { return ms_output_grid(N); }

void ms_widen(const MsState &m, uint32_t N, uint32_t *lv32, hipStream_t st) {
    hipLaunchKernelGGL(k_ms_widen, dim3(4096), dim3(256), 0, st, m.lv16, m.seen, N, lv32);
}

}  // namespace gpma
