// Update-path kernels: stage, locate, pending compaction, level-0 leaf
// rebalance, LDS window rebalance, and the device-wide window passes used
// for wide windows, resize and bulk load.
//
// Semantics follow the reference's update_gpma (gpma.cuh:968-1050):
//   locate: binary search over segment first keys, then a scan of the leaf;
//           an existing key (live or tombstoned) takes the update's value in
//           place, a delete never inserts (gpma.cuh:188-240);
//   rebalance at level L (window W = S << L): drop gaps and tombstones, and
//           if lower[L] <= live + pending <= upper[L] merge the pending
//           inserts in and redistribute evenly (gpma.cuh:294-540,636-663);
//           a failing window is left as it was and its inserts move up.
// The per-window merge replaces the reference's per-window radix sort: the
// live items are already sorted and so are the pending inserts.
#include <atomic>

#include "common.h"
#include "kernels.h"

namespace gpma {

namespace {

// Copies a[0..na) and b[0..nb) to dst[1..] in host-coherent memory (and
// zeroes a[0..na) if clear_a), then raises dst[0] to seq (system-scope
// release) for a host spinning on it.
__global__ void k_publish(uint32_t *a, uint32_t na, const uint32_t *b, uint32_t nb, uint32_t *dst, uint32_t seq,
                          uint32_t clear_a) {
    const uint32_t lane = threadIdx.x;  // one wave
    for (uint32_t i = lane; i < na; i += WAVE) {
        dst[1 + i] = a[i];
        if (clear_a) a[i] = 0;
    }
    for (uint32_t i = lane; i < nb; i += WAVE) dst[1 + na + i] = b[i];
    __threadfence_system();
    if (lane == 0) __hip_atomic_store(dst, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

// stage[0..4) = cnt[0], cnt_big[0] (u64 as two words each), stage[4 .. 4 + n)
// = reach[0 .. n): one captured step's read-back in one place.
__global__ void k_step_gather(const unsigned long long *cnt, const unsigned long long *cnt_big, const uint32_t *reach,
                              uint32_t n, uint32_t *stage) {
    const uint32_t lane = threadIdx.x;  // one wave
    if (lane == 0) {
        const unsigned long long a = *cnt, b = *cnt_big;
        stage[0] = (uint32_t)a;
        stage[1] = (uint32_t)(a >> 32);
        stage[2] = (uint32_t)b;
        stage[3] = (uint32_t)(b >> 32);
    }
    for (uint32_t i = lane; i < n; i += WAVE) stage[4 + i] = reach[i];
}

// Host-work flag of a captured step (see launch_step_gate).
__global__ void k_step_gate(const uint32_t *np0, const uint32_t *np_top, const uint32_t *err, uint32_t bulk_threshold,
                            uint32_t *gate) {
    if (threadIdx.x == 0) *gate = (*np0 >= bulk_threshold || *np_top != 0 || *err != 0) ? 1u : 0u;
}

__global__ void k_fill_u32(uint32_t *p, uint32_t v, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}
__global__ void k_fill_u8(uint8_t *p, uint8_t v, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}
__global__ void k_fill_u64(uint64_t *p, uint64_t v, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}
// A step's device inputs into the handle's buffers, and its record count,
// in one launch.
__global__ void k_stage_step(uint64_t *__restrict__ dk, const uint64_t *__restrict__ sk, uint8_t *__restrict__ dv,
                             const uint8_t *__restrict__ sv, uint64_t n, uint32_t *__restrict__ dn) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x, st = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = t; i < n; i += st) {
        dk[i] = sk[i];
        dv[i] = sv[i];
    }
    if (t == 0) *dn = (uint32_t)n;
}

// Zeroes a[0..na) (u32) and b[0..nb), c[0..nc) (u64) in one launch.
__global__ void k_zero3(uint32_t *a, uint64_t na, uint64_t *b, uint64_t nb, uint64_t *c, uint64_t nc) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x, st = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = t; i < na; i += st) a[i] = 0;
    for (uint64_t i = t; i < nb; i += st) b[i] = 0;
    for (uint64_t i = t; i < nc; i += st) c[i] = 0;
}

uint32_t grid_for(uint64_t n, uint32_t threads, uint32_t cap = 4096) {
    uint64_t b = (n + threads - 1) / threads;
    if (b == 0) b = 1;
    return (uint32_t)std::min<uint64_t>(b, cap);
}

// Wave-aggregated increment of a global counter by the number of lanes with p.
__device__ __forceinline__ void wave_count(uint32_t *ctr, bool p) {
    uint64_t m = ballot(p);
    if (m && lane_id() == (uint32_t)__ffsll((unsigned long long)m) - 1) atomicAdd(ctr, (uint32_t)__popcll(m));
}

// A leaf's S keys and values with 16-byte loads (leaves are S-slot aligned,
// so keys + w is 16 * S / 2 byte aligned and values + w is S byte aligned).
template <uint32_t S>
__device__ __forceinline__ void load_leaf_keys(const gkey *__restrict__ keys, uint64_t w, gkey (&k)[S]) {
    const uint4 *q = reinterpret_cast<const uint4 *>(keys + w);
    uint4 t[S / 2];
#pragma unroll
    for (uint32_t j = 0; j < S / 2; j++) t[j] = q[j];
#pragma unroll
    for (uint32_t j = 0; j < S / 2; j++) {
        k[2 * j] = (uint64_t)t[j].x | ((uint64_t)t[j].y << 32);
        k[2 * j + 1] = (uint64_t)t[j].z | ((uint64_t)t[j].w << 32);
    }
}
template <uint32_t S>
__device__ __forceinline__ void load_leaf_values(const gval *__restrict__ values, uint64_t w, gval (&v)[S]) {
    if constexpr (S >= 16) {
        const uint4 *q = reinterpret_cast<const uint4 *>(values + w);
#pragma unroll
        for (uint32_t j = 0; j < S / 16; j++) {
            const uint4 t = q[j];
            const uint32_t wd[4] = {t.x, t.y, t.z, t.w};
#pragma unroll
            for (uint32_t b = 0; b < 16; b++) v[16 * j + b] = (gval)(wd[b / 4] >> (8 * (b % 4)));
        }
    } else if constexpr (S == 8) {
        const uint2 t = *reinterpret_cast<const uint2 *>(values + w);
#pragma unroll
        for (uint32_t b = 0; b < 8; b++) v[b] = (gval)((b < 4 ? t.x : t.y) >> (8 * (b % 4)));
    } else if constexpr (S == 4) {
        const uint32_t t = *reinterpret_cast<const uint32_t *>(values + w);
#pragma unroll
        for (uint32_t b = 0; b < 4; b++) v[b] = (gval)(t >> (8 * b));
    } else {
        const uint16_t t = *reinterpret_cast<const uint16_t *>(values + w);
        v[0] = (gval)t;
        v[1] = (gval)(t >> 8);
    }
}

// Segment index the bitwise descending search of gpma.cuh:207-215 returns
// for key k: bit b of the result (b = nseg/2 .. 1) is set when
// seps[prefix | b] <= k. Path-dependent on non-monotone separators (the first
// key of an empty segment is KEY_NONE), so the exact walk is kept.
__device__ __forceinline__ uint64_t bitwise_search(const gkey *__restrict__ seps, uint64_t nseg, gkey k) {
    uint64_t p = 0;
    for (uint64_t b = nseg >> 1; b > 0; b >>= 1)
        if (seps[p | b] <= k) p |= b;
    return p;
}

// True when the bitwise search for k would return p: every step's
// comparison along p's path agrees with p's bit. The loads are independent,
// so they cost one round trip instead of log2(nseg) dependent ones.
__device__ __forceinline__ bool search_path_matches(const gkey *__restrict__ seps, uint64_t nseg, gkey k,
                                                    uint64_t p) {
    bool ok = p < nseg;
    for (uint64_t b = nseg >> 1; b > 0; b >>= 1) {
        const uint64_t pos = (p & ~((b << 1) - 1)) | b;
        const bool le = seps[pos] <= k;
        ok &= le == ((p & b) != 0);
    }
    return ok;
}

// One update per thread (gpma.cuh:204-226). Sorted input, so the last of
// equal keys is the one applied (earlier duplicates are skipped).
// With row_hint, the segment is first guessed from row_offset (the key lies
// between the slots of guards src-1 and src). Below 16 slots per segment the
// guess is checked against the full bitwise search path and a failed check
// falls back to the walk, so the result is always the reference search's.
template <uint32_t S>
__global__ __launch_bounds__(CMP_BLOCK) void k_locate(const gkey *__restrict__ ks, const gval *__restrict__ vs,
                                                      uint64_t n, const uint32_t *__restrict__ dn,
                                                      gkey *__restrict__ keys, gval *__restrict__ values,
                                                      const gkey *__restrict__ seps, uint64_t nseg,
                                                      const uint32_t *__restrict__ row_offset, uint32_t row_hint,
                                                      uint32_t *__restrict__ win) {
    if (dn) n = *dn;
    const uint64_t i = (uint64_t)blockIdx.x * CMP_BLOCK + threadIdx.x;
    if (i >= n) return;
    const gkey k = ks[i];
    const gval v = vs[i];
    uint32_t res = SIZE_NONE;
    const bool skip = (k == KEY_NONE) || (i + 1 < n && ks[i + 1] == k);
    if (!skip) {
        uint64_t p = 0;
        bool have = false;
        if (row_hint) {
            const uint32_t src = (uint32_t)(k >> 32);
            uint64_t lo = row_offset[src] / S, hi = row_offset[src + 1] / S;
            if (hi >= nseg) hi = nseg - 1;
            if (lo > hi) lo = hi;
            // last segment in [lo, hi] whose first key is <= k
            while (hi - lo > 4) {
                const uint64_t mid = (lo + hi + 1) >> 1;
                if (seps[mid] <= k) lo = mid;
                else hi = mid - 1;
            }
            p = lo;
            for (uint64_t j = lo + 1; j <= hi; j++)
                if (seps[j] <= k) p = j;
        }
        // the guessed leaf is loaded together with the path check; a failed
        // check (or no hint) searches and loads again. With 16+ slots per
        // segment no segment is ever empty (every redistribution keeps the
        // density above 1/16), so separators increase strictly, the bitwise
        // search's answer is the last separator <= k, and it lies between the
        // guards of rows src - 1 and src: the guess is the answer, unchecked.
        gkey lk[S];
        if (row_hint) {
            load_leaf_keys<S>(keys, p * S, lk);
            have = S >= 16 || search_path_matches(seps, nseg, k, p);
        }
        if (!have) {
            p = bitwise_search(seps, nseg, k);
            load_leaf_keys<S>(keys, p * S, lk);
        }
        const uint64_t base = p * S;
        // handle_del_mod (gpma.cuh:188-199): the first slot holding the key
        uint32_t hit = S;
#pragma unroll
        for (uint32_t j = S; j-- > 0;)
            if (lk[j] == k) hit = j;
        if (hit < S) values[base + hit] = v;
        else if (v != 0) res = (uint32_t)base;
    }
    win[i] = res;
}

// Exclusive scan of per-chunk counts (one workgroup). The number of chunks
// is nb, or ceil(*d_items / CMP_BLOCK) when d_items is given.
__global__ __launch_bounds__(1024) void k_scan_blocks(const uint32_t *__restrict__ cnt, uint64_t nb,
                                                      const uint32_t *__restrict__ d_items,
                                                      uint32_t *__restrict__ off, uint32_t *__restrict__ total) {
    __shared__ uint32_t scratch[1024 / WAVE + 1];
    if (d_items) nb = ((uint64_t)*d_items + CMP_BLOCK - 1) / CMP_BLOCK;
    const uint64_t per = (nb + 1023) / 1024;
    const uint64_t a = std::min<uint64_t>(nb, threadIdx.x * per), b = std::min<uint64_t>(nb, a + per);
    uint32_t s = 0;
    for (uint64_t j = a; j < b; j++) s += cnt[j];
    uint32_t tot;
    uint32_t pre = block_exclusive_scan(s, scratch, &tot);
    for (uint64_t j = a; j < b; j++) {
        uint32_t c = cnt[j];
        off[j] = pre;
        pre += c;
    }
    if (threadIdx.x == 0) *total = tot;
}

// Per-chunk counts of pending (done == 0) entries among the first *d_n;
// grid-stride over chunks of CMP_BLOCK entries.
__global__ __launch_bounds__(CMP_BLOCK) void k_count_pending(const uint8_t *__restrict__ done,
                                                             const uint32_t *__restrict__ d_n,
                                                             uint32_t *__restrict__ blkcnt) {
    __shared__ uint32_t wcnt[CMP_BLOCK / WAVE];
    const uint64_t n = *d_n;
    for (uint64_t c = blockIdx.x; c * CMP_BLOCK < n; c += gridDim.x) {
        const uint64_t i = c * CMP_BLOCK + threadIdx.x;
        bool keep = i < n && done[i] == 0;
        uint64_t m = ballot(keep);
        if (lane_id() == 0) wcnt[threadIdx.x / WAVE] = (uint32_t)__popcll(m);
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t t = 0;
            for (uint32_t w = 0; w < CMP_BLOCK / WAVE; w++) t += wcnt[w];
            blkcnt[c] = t;
        }
        __syncthreads();
    }
}

// Stable compaction of entries whose win != SIZE_NONE (idone == nullptr) or
// whose done flag is 0, over the first *d_n entries (n_ub when d_n is null),
// grid-stride over chunks; writes win & mask and clears the output's done flags.
__global__ __launch_bounds__(CMP_BLOCK) void k_compact(const gkey *__restrict__ ik, const gval *__restrict__ iv,
                                                       const uint32_t *__restrict__ iw,
                                                       const uint8_t *__restrict__ idone, uint64_t n_ub,
                                                       const uint32_t *__restrict__ d_n,
                                                       const uint32_t *__restrict__ off, gkey *__restrict__ ok,
                                                       gval *__restrict__ ov, uint32_t *__restrict__ ow,
                                                       uint8_t *__restrict__ odone, uint32_t mask) {
    __shared__ uint32_t scratch[CMP_BLOCK / WAVE + 1];
    const uint64_t n = d_n ? *d_n : n_ub;
    for (uint64_t c = blockIdx.x; c * CMP_BLOCK < n; c += gridDim.x) {
        const uint64_t i = c * CMP_BLOCK + threadIdx.x;
        bool keep = false;
        if (i < n) keep = idone ? (idone[i] == 0) : (iw[i] != SIZE_NONE);
        uint32_t tot;
        uint32_t pre = block_exclusive_scan(keep ? 1u : 0u, scratch, &tot);
        if (keep) {
            uint64_t o = (uint64_t)off[c] + pre;
            ok[o] = ik[i];
            ov[o] = iv[i];
            ow[o] = iw[i] & mask;
            odone[o] = 0;
        }
    }
}

// Single-pass stable compaction (chained scan with decoupled look-back):
// tiles of CC_TILE entries take tickets in launch order, count their kept
// entries with wave ballots, publish the count, add the counts of earlier
// tiles (read 64 at a time by one wave), publish the inclusive prefix and
// scatter. Status words carry the caller's epoch in their high half, so no
// clearing pass is needed between calls; tiles past *d_n exit at once.
// With fix (the output of k_apply), an entry is kept when iw < WIN_SEP, and
// an entry iw = WIN_SEP | leaf writes fix[i] to seps[leaf]: the merged
// leaf's new first key, held back until every search of the batch ran.
constexpr uint32_t CC_THREADS = 256;
constexpr uint32_t CC_ROWS = 8;                        // 64-entry rows per wave
constexpr uint32_t CC_WAVE_ITEMS = WAVE * CC_ROWS;     // 512
constexpr uint32_t CC_TILE = CC_THREADS * CC_ROWS;     // 2048
constexpr uint32_t CC_GRID_MAX = 256;                  // workgroups; each loops over tiles

__global__ __launch_bounds__(CC_THREADS) void k_compact_chained(
    const gkey *__restrict__ ik, const gval *__restrict__ iv, const uint32_t *__restrict__ iw,
    const uint8_t *__restrict__ idone, uint64_t n_ub, const uint32_t *__restrict__ d_n, gkey *__restrict__ ok,
    gval *__restrict__ ov, uint32_t *__restrict__ ow, uint8_t *__restrict__ odone, uint32_t mask,
    unsigned long long *__restrict__ status, uint32_t epoch, uint32_t *__restrict__ ticket,
    uint32_t *__restrict__ d_out, uint32_t *__restrict__ err, const gkey *__restrict__ fix,
    gkey *__restrict__ seps) {
    __shared__ uint32_t s_tile, s_excl;
    __shared__ uint32_t wtot[CC_THREADS / WAVE];
    const uint64_t n = d_n ? std::min<uint64_t>(*d_n, n_ub) : n_ub;
    // Workgroups beyond what the list needs leave before taking a ticket
    // (tickets are one atomic on one word: a few hundred of them queue up).
    // The rest take one ticket each, in the order they start, and the
    // workgroup with ticket q owns tiles q, q + G, q + 2G, ... (G workgroups
    // take part), each in turn: every tile's look-back waits only on lower
    // tiles, owned by workgroups that started earlier or by itself.
    const uint64_t need = std::max<uint64_t>(1, (n + CC_TILE - 1) / CC_TILE);
    if (blockIdx.x >= need) return;
    const uint32_t G = (uint32_t)std::min<uint64_t>(gridDim.x, need);
    if (threadIdx.x == 0) s_tile = atomicAdd(ticket, 1u);
    __syncthreads();
    for (uint32_t tile = s_tile;; tile += G) {
    __syncthreads();  // wtot and s_excl of the previous tile are consumed
    const uint64_t base = (uint64_t)tile * CC_TILE;
    if (base >= n) {
        if (tile == 0 && threadIdx.x == 0) *d_out = 0;
        return;
    }
    const uint32_t wid = threadIdx.x / WAVE, lane = lane_id();
    const uint64_t wbase = base + (uint64_t)wid * CC_WAVE_ITEMS;
    uint64_t km[CC_ROWS];
    uint32_t wcount = 0;
#pragma unroll
    for (uint32_t j = 0; j < CC_ROWS; j++) {
        const uint64_t i = wbase + (uint64_t)j * WAVE + lane;
        bool keep = false;
        if (i < n) {
            if (idone) {
                keep = idone[i] == 0;
            } else {
                const uint32_t w = iw[i];
                keep = fix ? w < WIN_SEP : w != SIZE_NONE;
                if (fix && w != SIZE_NONE && w >= WIN_SEP) seps[w - WIN_SEP] = fix[i];
            }
        }
        km[j] = ballot(keep);
        wcount += (uint32_t)__popcll(km[j]);
    }
    if (lane == 0) wtot[wid] = wcount;
    __syncthreads();
    uint32_t tile_total = 0, wpre = 0;
#pragma unroll
    for (uint32_t w = 0; w < CC_THREADS / WAVE; w++) {
        if (w < wid) wpre += wtot[w];
        tile_total += wtot[w];
    }
    if (wid == 0) {
        const uint32_t excl = wave_lookback(status, tile, epoch, tile_total, err);
        if (lane == 0) {
            s_excl = excl;
            if (base + CC_TILE >= n) *d_out = excl + tile_total;
        }
    }
    __syncthreads();
    uint32_t o = s_excl + wpre;
#pragma unroll
    for (uint32_t j = 0; j < CC_ROWS; j++) {
        if ((km[j] >> lane) & 1ull) {
            const uint64_t i = wbase + (uint64_t)j * WAVE + lane;
            const uint64_t d = (uint64_t)o + (uint64_t)__popcll(km[j] & lanemask_lt());
            ok[d] = ik[i];
            ov[d] = iv[i];
            ow[d] = iw[i] & mask;
            odone[d] = 0;
        }
        o += (uint32_t)__popcll(km[j]);
    }
    }
}

// Level 0 (window = one leaf of S slots), one thread per run of pending
// inserts that share a leaf. With f = W/S = 1 the redistribution packs the
// merged items at the front of the leaf (gpma.cuh:636-654 with f = 1).
// The leaf is read with 16-byte loads, merged into a per-thread LDS row and
// written back with 16-byte stores.
template <uint32_t S>
__global__ __launch_bounds__(256) void k_leaf(const gkey *__restrict__ pk, const gval *__restrict__ pv,
                                              const uint32_t *__restrict__ pw, uint8_t *__restrict__ pd,
                                              const uint32_t *__restrict__ d_np, gkey *__restrict__ keys,
                                              gval *__restrict__ values, gkey *__restrict__ seps,
                                              uint32_t *__restrict__ row_offset, uint32_t lower, uint32_t upper,
                                              uint32_t bulk_threshold) {
    constexpr uint32_t RK = S + 1;  // padded LDS row (spreads the rows over the banks)
    __shared__ gkey rowk[256 * RK];
    __shared__ gval rowv[256 * S];
    const uint64_t np = *d_np;
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    // a batch with this many pending inserts takes the bulk path instead
    if (i >= np || np >= bulk_threshold) return;
    const uint32_t w = pw[i];
    if (i > 0 && pw[i - 1] == w) return;
    uint32_t r = 1;
    while (r <= upper && i + r < np && pw[i + r] == w) r++;
    if (r > upper) return;
    gkey k[S];
    gval v[S];
    load_leaf_keys<S>(keys, w, k);
    load_leaf_values<S>(values, w, v);
    bool live[S];
    uint32_t shift[S];
    uint32_t c = 0;
#pragma unroll
    for (uint32_t j = 0; j < S; j++) {
        live[j] = is_live(k[j], v[j]);
        c += live[j] ? 1u : 0u;
        shift[j] = 0;
    }
    const uint32_t m = c + r;
    if (m < lower || m > upper) return;
    gkey *ok = rowk + threadIdx.x * RK;
    gval *ov = rowv + threadIdx.x * S;
#pragma unroll
    for (uint32_t j = 0; j < S; j++) {
        ok[j] = KEY_NONE;
        ov[j] = 0;
    }
    for (uint32_t u = 0; u < r; u++) {
        const gkey ik = pk[i + u];
        const gval iv = pv[i + u];
        uint32_t lb = 0;
#pragma unroll
        for (uint32_t j = 0; j < S; j++) {
            if (live[j]) {
                if (k[j] < ik) lb++;
                else shift[j]++;
            }
        }
        const uint32_t pos = u + lb;
        ok[pos] = ik;
        ov[pos] = iv;
        if (is_guard(ik)) row_offset[(ik >> 32) + 1] = w + pos;
        pd[i + u] = 1;
    }
    uint32_t lr = 0;
#pragma unroll
    for (uint32_t j = 0; j < S; j++) {
        if (live[j]) {
            const uint32_t pos = lr + shift[j];
            ok[pos] = k[j];
            ov[pos] = v[j];
            if (is_guard(k[j])) row_offset[(k[j] >> 32) + 1] = w + pos;
            lr++;
        }
    }
    // the row is private to this thread: no barrier needed before reading it back
    uint4 *qk = reinterpret_cast<uint4 *>(keys + w);
#pragma unroll
    for (uint32_t j = 0; j < S / 2; j++) {
        const gkey a = ok[2 * j], b = ok[2 * j + 1];
        qk[j] = make_uint4((uint32_t)a, (uint32_t)(a >> 32), (uint32_t)b, (uint32_t)(b >> 32));
    }
    if constexpr (S >= 16) {
        uint4 *qv = reinterpret_cast<uint4 *>(values + w);
#pragma unroll
        for (uint32_t j = 0; j < S / 16; j++) {
            uint32_t wd[4] = {0, 0, 0, 0};
#pragma unroll
            for (uint32_t b = 0; b < 16; b++) wd[b / 4] |= (uint32_t)ov[16 * j + b] << (8 * (b % 4));
            qv[j] = make_uint4(wd[0], wd[1], wd[2], wd[3]);
        }
    } else {
#pragma unroll
        for (uint32_t j = 0; j < S; j++) values[(uint64_t)w + j] = ov[j];
    }
    seps[w / S] = ok[0];
}

// Locate and level 0 in one pass (S >= 16, row hints): k_locate's search,
// then k_leaf's merge, by the record that heads each leaf's run.
//
// Every thread finds the leaf p of its record as k_locate does. With 16+
// slots per segment the separators increase strictly, so the records of
// leaf p are exactly the sorted records in [seps[p], seps[p + 1]); the
// first of them (the previous record lies below seps[p]) is the run head
// and owns the leaf: it loads the leaf once, applies the run's deletes and
// modifications to it (handle_del_mod, gpma.cuh:188-199; the last of equal
// keys applies), counts the run's pending inserts and, when level 0's
// bounds allow (gpma.cuh:636-654 with f = 1: the merged items packed at the
// front of the leaf), merges them into the leaf. A run may continue past
// the workgroup's tile: the head reads those records from global memory,
// and their own threads, which are not heads, leave them alone.
//
// Separators: a merge only raises its leaf's first key (deletes drop keys,
// inserts are >= the old separator; leaf 0's is never compared), so a
// concurrent search for a key of any other leaf compares the same way with
// the old or the new value. What could change is the head test of a record
// in the same leaf (its predecessor may now lie below the raised separator)
// and the run end of the previous leaf's head (a record of this leaf may
// now lie below it). A merged run whose records all sit in the workgroup's
// tile, after another head of the tile (so the previous leaf's head is here
// too), has those reads done before any merge and writes its separator at
// once; any other run (the tile's first, or one continuing into the next
// tile) passes the new first key on in fix[head] with win[head] = WIN_SEP |
// p, for the compaction that follows. A merge moves guards only inside
// their leaf, so the concurrent row hints (slot / S) read the same segment
// before and after it.
//
// The head keeps the leaf's keys in registers, its values as packed bytes
// and its live slots as a bit mask; the merge stores every key straight to
// its slot (pending insert u, in key order, goes to u + the live keys below
// it; the live keys fill the remaining slots in order) and gathers the
// values in a per-thread LDS row, written back with 16-byte stores (single
// byte stores to the leaf cost more than the whole merge).
template <uint32_t S>
__device__ __forceinline__ uint32_t byte_of(const uint32_t (&vw)[S / 4], uint32_t s) {
    return (vw[s / 4] >> (8 * (s % 4))) & 0xFFu;
}

template <uint32_t S>
__global__ __launch_bounds__(256) void k_apply(const gkey *__restrict__ ks, const gval *__restrict__ vs, uint64_t n,
                                               const uint32_t *__restrict__ dn, gkey *__restrict__ keys,
                                               gval *__restrict__ values, gkey *__restrict__ seps,
                                               uint64_t nseg, uint32_t *__restrict__ row_offset, uint32_t lower,
                                               uint32_t upper, uint32_t *__restrict__ win, gkey *__restrict__ fix,
                                               uint32_t *__restrict__ np0) {
    static_assert(S == 16 || S == 32, "one leaf's live slots fit a 32-bit mask");
    constexpr uint32_t T = 256;
    constexpr uint32_t SMASK = S == 32 ? 0xFFFFFFFFu : (1u << S) - 1;
    __shared__ gkey tk[T];
    __shared__ gval tv[T];
    __shared__ uint32_t wsum[T / WAVE];
    __shared__ uint64_t hmask[T / WAVE];  // run heads of the tile
    __shared__ uint32_t rowv[T * S / 4];  // each head's merged values, packed
    if (dn) n = *dn;
    const uint64_t base = (uint64_t)blockIdx.x * T;
    if (base >= n) return;
    const uint32_t tid = threadIdx.x;
    const uint64_t i = base + tid;
    const uint64_t tend = std::min<uint64_t>(n, base + T);
    gkey k = KEY_NONE, kprev = 0;
    gval v = 0;
    if (i < n) {
        k = ks[i];
        v = vs[i];
        if (i > 0) kprev = ks[i - 1];
    }
    tk[tid] = k;
    tv[tid] = v;
    bool head = false;
    uint64_t p = 0;
    gkey nsep = KEY_NONE;  // first key of leaf p + 1
    if (k != KEY_NONE) {
        const uint32_t src = (uint32_t)(k >> 32);
        uint64_t lo = row_offset[src] / S, hi = row_offset[src + 1] / S;
        if (hi >= nseg) hi = nseg - 1;
        if (lo > hi) lo = hi;
        while (hi - lo > 4) {
            const uint64_t mid = (lo + hi + 1) >> 1;
            if (seps[mid] <= k) lo = mid;
            else hi = mid - 1;
        }
        // the last separator <= k in [lo, hi], with its own and the next
        // one, from one round of loads
        gkey sv[6];
#pragma unroll
        for (uint32_t t = 0; t < 6; t++) sv[t] = (lo + t <= hi + 1 && lo + t < nseg) ? seps[lo + t] : KEY_NONE;
        p = lo;
        gkey sp = sv[0];
        nsep = sv[1];
#pragma unroll
        for (uint32_t t = 1; t < 5; t++) {
            if (lo + t <= hi && sv[t] <= k) {
                p = lo + t;
                sp = sv[t];
                nsep = sv[t + 1];
            }
        }
        // the bitwise search never compares seps[0]: every key below seps[1] is in leaf 0
        head = i == 0 || (p > 0 && kprev < sp);
    } else if (i < n) {
        win[i] = SIZE_NONE;
    }
    const uint64_t hb = ballot(head);
    if (lane_id() == 0) hmask[tid / WAVE] = hb;
    __syncthreads();  // tk / tv / hmask staged
    uint32_t r = 0;
    if (head) {
        const uint64_t w = p * S;
        gkey lk[S];
        uint32_t vw[S / 4];
        load_leaf_keys<S>(keys, w, lk);
        {
            const uint4 *q = reinterpret_cast<const uint4 *>(values + w);
#pragma unroll
            for (uint32_t j = 0; j < S / 16; j++) {
                const uint4 t = q[j];
                vw[4 * j] = t.x;
                vw[4 * j + 1] = t.y;
                vw[4 * j + 2] = t.z;
                vw[4 * j + 3] = t.w;
            }
        }
        // record j + 1 of the run (KEY_NONE past the batch)
        auto next = [&](uint64_t j, gkey &kn, gval &vn) {
            kn = KEY_NONE;
            vn = 0;
            if (j + 1 < tend) {
                kn = tk[j + 1 - base];
                vn = tv[j + 1 - base];
            } else if (j + 1 < n) {
                kn = ks[j + 1];
                vn = vs[j + 1];
            }
        };
        auto hits = [&](gkey x) {
            uint32_t h = 0;
#pragma unroll
            for (uint32_t s = 0; s < S; s++) h |= (lk[s] == x ? 1u : 0u) << s;
            return h;
        };
        // pass 1: deletes and modifications in place, count the pending inserts
        bool dirty = false;
        uint64_t jend = i;
        {
            gkey kj = k, kn;
            gval vj = v, vn;
            for (;;) {
                next(jend, kn, vn);
                if (kn != kj) {
                    const uint32_t h = hits(kj);
                    if (h) {
#pragma unroll
                        for (uint32_t s = 0; s < S; s++) {
                            if ((h >> s) & 1u) {
                                const uint32_t sh = 8 * (s % 4);
                                vw[s / 4] = (vw[s / 4] & ~(0xFFu << sh)) | ((uint32_t)vj << sh);
                            }
                        }
                        dirty = true;
                    } else if (vj != 0) {
                        r++;
                    }
                }
                jend++;
                if (kn >= nsep || kn == KEY_NONE) break;
                kj = kn;
                vj = vn;
            }
        }
        uint32_t livem = 0;
#pragma unroll
        for (uint32_t s = 0; s < S; s++) livem |= (lk[s] != KEY_NONE && byte_of<S>(vw, s) != 0 ? 1u : 0u) << s;
        const uint32_t m = (uint32_t)__popc(livem) + r;
        if (r > 0 && r <= upper && m >= lower && m <= upper) {
            // pass 2: pending inserts to their slots
            gkey kj = k, kn;
            gval vj = v, vn;
            uint32_t u = 0, occ = 0;
            gkey first = KEY_NONE;
            uint32_t *ov = rowv + tid * (S / 4);
            gval *ovb = reinterpret_cast<gval *>(ov);
#pragma unroll
            for (uint32_t q = 0; q < S / 4; q++) ov[q] = 0;
            for (uint64_t j = i; j < jend; j++) {
                next(j, kn, vn);
                if (kn != kj && vj != 0 && hits(kj) == 0) {
                    uint32_t below = 0;
#pragma unroll
                    for (uint32_t s = 0; s < S; s++) below |= (lk[s] < kj ? 1u : 0u) << s;
                    const uint32_t pos = u + (uint32_t)__popc(livem & below);
                    keys[w + pos] = kj;
                    ovb[pos] = vj;
                    if (is_guard(kj)) row_offset[(kj >> 32) + 1] = (uint32_t)(w + pos);
                    if (pos == 0) first = kj;
                    occ |= 1u << pos;
                    u++;
                }
                if (j != i) win[j] = SIZE_NONE;
                kj = kn;
                vj = vn;
            }
            // live keys into the free slots, in order; then the empty tail
            uint32_t fr = ~occ & SMASK;
#pragma unroll
            for (uint32_t s = 0; s < S; s++) {
                if ((livem >> s) & 1u) {
                    const uint32_t pos = (uint32_t)__builtin_ctz(fr);
                    fr &= fr - 1;
                    keys[w + pos] = lk[s];
                    ovb[pos] = (gval)byte_of<S>(vw, s);
                    if (is_guard(lk[s])) row_offset[(lk[s] >> 32) + 1] = (uint32_t)(w + pos);
                    if (pos == 0) first = lk[s];
                }
            }
            for (uint32_t pos = m; pos < S; pos++) keys[w + pos] = KEY_NONE;
            // the row is private to this thread: no barrier needed before reading it back
            uint4 *qv = reinterpret_cast<uint4 *>(values + w);
#pragma unroll
            for (uint32_t j = 0; j < S / 16; j++)
                qv[j] = make_uint4(ov[4 * j], ov[4 * j + 1], ov[4 * j + 2], ov[4 * j + 3]);
            // an earlier head in this tile: the previous leaf's head is here too
            bool earlier = (hmask[tid / WAVE] & lanemask_lt()) != 0;
            for (uint32_t q = 0; q < tid / WAVE; q++) earlier |= hmask[q] != 0;
            if (jend <= tend && (i == 0 || earlier)) {
                // the leaf's records and the previous leaf's head are all in
                // this tile: no other workgroup reads this separator
                seps[p] = first;
                win[i] = SIZE_NONE;
            } else {
                fix[i] = first;
                win[i] = WIN_SEP | (uint32_t)p;
            }
        } else {
            if (dirty) {
                uint4 *q = reinterpret_cast<uint4 *>(values + w);
#pragma unroll
                for (uint32_t j = 0; j < S / 16; j++)
                    q[j] = make_uint4(vw[4 * j], vw[4 * j + 1], vw[4 * j + 2], vw[4 * j + 3]);
            }
            // the run's pending inserts wait for level 1
            gkey kj = k, kn;
            gval vj = v, vn;
            for (uint64_t j = i; j < jend; j++) {
                next(j, kn, vn);
                const bool pend = kn != kj && vj != 0 && hits(kj) == 0;
                win[j] = pend ? (uint32_t)w : SIZE_NONE;
                kj = kn;
                vj = vn;
            }
        }
    }
    // level-0 pending inserts of the tile's runs
#pragma unroll
    for (uint32_t off = 32; off > 0; off >>= 1) r += __shfl_xor(r, off, WAVE);
    if (lane_id() == 0) wsum[tid / WAVE] = r;
    __syncthreads();
    if (tid == 0) {
        uint32_t t = 0;
#pragma unroll
        for (uint32_t q = 0; q < T / WAVE; q++) t += wsum[q];
        if (t) atomicAdd(np0, t);
    }
}

// number of entries of a sorted array a[0..n) strictly below x
__device__ __forceinline__ uint64_t lower_bound_g(const gkey *a, uint64_t n, gkey x) {
    uint64_t lo = 0, hi = n;
    while (lo < hi) {
        uint64_t mid = (lo + hi) >> 1;
        if (a[mid] < x) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// One window run for S < W <= WIN_MAX, by one workgroup: compact the
// window's live items into LDS, check the level's bounds, and write merged
// items straight to their redistributed slots (run = entries i.. with base w).
__device__ void window_one(uint32_t S, uint32_t W, const gkey *__restrict__ pk, const gval *__restrict__ pv,
                           const uint32_t *__restrict__ pw, uint8_t *__restrict__ pd, uint64_t np, uint64_t i,
                           uint32_t w, gkey *__restrict__ keys, gval *__restrict__ values, gkey *__restrict__ seps,
                           uint32_t *__restrict__ row_offset, uint32_t lower, uint32_t upper, gkey *lk, gval *lv,
                           uint32_t *scratch, uint32_t *s_r_ptr) {
    uint32_t &s_r = *s_r_ptr;
    const uint32_t tid = threadIdx.x;
    __syncthreads();  // the previous run of this workgroup is done with LDS
    if (tid == 0) s_r = upper + 1;
    __syncthreads();
    for (uint32_t t = tid; t <= upper; t += 256) {
        bool in = (i + t < np) && pw[i + t] == w;
        if (!in) atomicMin(&s_r, t);
    }
    __syncthreads();
    const uint32_t r = s_r;
    if (r > upper) return;
    uint32_t c = 0;
    for (uint32_t b = 0; b < W; b += 256) {
        const uint32_t p = b + tid;
        bool live = false;
        gkey k = KEY_NONE;
        gval v = 0;
        if (p < W) {
            k = keys[(uint64_t)w + p];
            v = values[(uint64_t)w + p];
            live = is_live(k, v);
        }
        uint32_t tot;
        uint32_t pre = block_exclusive_scan(live ? 1u : 0u, scratch, &tot);
        if (live) {
            lk[c + pre] = k;
            lv[c + pre] = v;
        }
        c += tot;
    }
    __syncthreads();
    const uint32_t m = c + r;
    if (m < lower || m > upper) return;
    const uint64_t f = W / S;
    for (uint32_t t = tid; t < c; t += 256) {
        const gkey k = lk[t];
        const uint64_t mi = t + lower_bound_g(pk + i, r, k);
        const uint32_t slot = redistribute_slot(f, S, m, mi);
        keys[(uint64_t)w + slot] = k;
        values[(uint64_t)w + slot] = lv[t];
        if (is_guard(k)) row_offset[(k >> 32) + 1] = w + slot;
        if (slot % S == 0) seps[(w + slot) / S] = k;
    }
    for (uint32_t u = tid; u < r; u += 256) {
        const gkey k = pk[i + u];
        uint32_t lo = 0, hi = c;
        while (lo < hi) {
            uint32_t mid = (lo + hi) >> 1;
            if (lk[mid] < k) lo = mid + 1;
            else hi = mid;
        }
        const uint64_t mi = u + lo;
        const uint32_t slot = redistribute_slot(f, S, m, mi);
        keys[(uint64_t)w + slot] = k;
        values[(uint64_t)w + slot] = pv[i + u];
        if (is_guard(k)) row_offset[(k >> 32) + 1] = w + slot;
        if (slot % S == 0) seps[(w + slot) / S] = k;
        pd[i + u] = 1;
    }
    for (uint32_t p = tid; p < W; p += 256) {
        const uint64_t s = p / S, l = p % S;
        const uint64_t a = seg_first_item(f, m, s);
        const uint64_t cnt = seg_first_item(f, m, s + 1) - a;
        if (l >= cnt) {
            keys[(uint64_t)w + p] = KEY_NONE;
            values[(uint64_t)w + p] = 0;
            if (l == 0) seps[(w + p) / S] = KEY_NONE;
        }
    }
}

// window_one for W <= WAVE_WIN_MAX by one wave, without workgroup barriers:
// the run length and the live-item compaction by ballots, lk/lv the wave's
// own LDS rows. Same results as window_one. The caller has loaded, in one
// round of loads, the window ids of entries i - 1 + lane (wv) and the first
// 64 pending entries of the run (lpk / lpv, the wave's LDS rows, valid when
// the run is at most 64 long); the window's slots come in one more round.
constexpr uint32_t WAVE_WIN_MAX = 256;
static_assert(WAVE_WIN_MAX == LEVEL_WIN_MAX, "device-driven levels use the wave windows");
__device__ void window_wave(uint32_t S, uint32_t W, const gkey *__restrict__ pk, const gval *__restrict__ pv,
                            const uint32_t *__restrict__ pw, uint8_t *__restrict__ pd, uint64_t np, uint64_t i,
                            uint32_t w, uint32_t wv, gkey *__restrict__ keys, gval *__restrict__ values,
                            gkey *__restrict__ seps, uint32_t *__restrict__ row_offset, uint32_t lower,
                            uint32_t upper, gkey *lk, gval *lv, const gkey *lpk, const gval *lpv) {
    const uint32_t lane = lane_id();
    // run length: the first t <= upper whose entry i + t is not in window w
    // (lane 1 + t holds entry i + t's window for t < 63)
    uint32_t r = upper + 1;
    const uint64_t out0 = ballot(lane >= 1 && wv != w);
    if (out0) {
        r = (uint32_t)__ffsll((unsigned long long)out0) - 2;
    } else {
        for (uint32_t t0 = WAVE - 1; t0 <= upper; t0 += WAVE) {
            const uint32_t t = t0 + lane;
            const bool out = t <= upper && !((i + t < np) && pw[i + t] == w);
            const uint64_t b = ballot(out);
            if (b) {
                r = t0 + (uint32_t)__ffsll((unsigned long long)b) - 1;
                break;
            }
        }
    }
    if (r > upper) return;
    const bool lds_run = r <= WAVE;
    // the window's slots, every load issued before the first ballot
    gkey kq[WAVE_WIN_MAX / WAVE];
    gval vq[WAVE_WIN_MAX / WAVE];
#pragma unroll
    for (uint32_t q = 0; q < WAVE_WIN_MAX / WAVE; q++) {
        const uint32_t p = q * WAVE + lane;
        kq[q] = KEY_NONE;
        vq[q] = 0;
        if (p < W) {
            kq[q] = keys[(uint64_t)w + p];
            vq[q] = values[(uint64_t)w + p];
        }
    }
    uint32_t c = 0;
#pragma unroll
    for (uint32_t q = 0; q < WAVE_WIN_MAX / WAVE; q++) {
        const bool live = is_live(kq[q], vq[q]);
        const uint64_t m = ballot(live);
        if (live) {
            const uint32_t pos = c + (uint32_t)__popcll(m & lanemask_lt());
            lk[pos] = kq[q];
            lv[pos] = vq[q];
        }
        c += (uint32_t)__popcll(m);
    }
    wave_sync();
    const uint32_t m = c + r;
    if (m >= lower && m <= upper) {
        const uint64_t f = W / S;
        for (uint32_t t = lane; t < c; t += WAVE) {
            const gkey k = lk[t];
            uint64_t below;
            if (lds_run) {
                uint32_t lo = 0, hi = r;
                while (lo < hi) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (lpk[mid] < k) lo = mid + 1;
                    else hi = mid;
                }
                below = lo;
            } else {
                below = lower_bound_g(pk + i, r, k);
            }
            const uint32_t slot = redistribute_slot(f, S, m, t + below);
            keys[(uint64_t)w + slot] = k;
            values[(uint64_t)w + slot] = lv[t];
            if (is_guard(k)) row_offset[(k >> 32) + 1] = w + slot;
            if (slot % S == 0) seps[(w + slot) / S] = k;
        }
        for (uint32_t u = lane; u < r; u += WAVE) {
            const gkey k = lds_run ? lpk[u] : pk[i + u];
            uint32_t lo = 0, hi = c;
            while (lo < hi) {
                uint32_t mid = (lo + hi) >> 1;
                if (lk[mid] < k) lo = mid + 1;
                else hi = mid;
            }
            const uint64_t mi = u + lo;
            const uint32_t slot = redistribute_slot(f, S, m, mi);
            keys[(uint64_t)w + slot] = k;
            values[(uint64_t)w + slot] = lds_run ? lpv[u] : pv[i + u];
            if (is_guard(k)) row_offset[(k >> 32) + 1] = w + slot;
            if (slot % S == 0) seps[(w + slot) / S] = k;
            pd[i + u] = 1;
        }
        for (uint32_t p = lane; p < W; p += WAVE) {
            const uint64_t s = p / S, l = p % S;
            const uint64_t a = seg_first_item(f, m, s);
            const uint64_t cnt = seg_first_item(f, m, s + 1) - a;
            if (l >= cnt) {
                keys[(uint64_t)w + p] = KEY_NONE;
                values[(uint64_t)w + p] = 0;
                if (l == 0) seps[(w + p) / S] = KEY_NONE;
            }
        }
    }
}

// Workgroups stride over the pending list; each run head is one window.
constexpr uint64_t WINDOW_GRID = 1024;

// The same for W <= WAVE_WIN_MAX with one wave per window.
__global__ __launch_bounds__(256) void k_window_wave(uint32_t S, uint32_t W, const gkey *__restrict__ pk,
                                                     const gval *__restrict__ pv, const uint32_t *__restrict__ pw,
                                                     uint8_t *__restrict__ pd, const uint32_t *__restrict__ d_np,
                                                     gkey *__restrict__ keys, gval *__restrict__ values,
                                                     gkey *__restrict__ seps, uint32_t *__restrict__ row_offset,
                                                     uint32_t lower, uint32_t upper,
                                                     const uint32_t *__restrict__ d_gate, uint32_t gate) {
    constexpr uint32_t NW = 256 / WAVE;
    __shared__ gkey lk[NW][WAVE_WIN_MAX];
    __shared__ gval lv[NW][WAVE_WIN_MAX];
    __shared__ gkey lpk[NW][WAVE];
    __shared__ gval lpv[NW][WAVE];
    if (d_gate && *d_gate >= gate) return;
    const uint64_t np = *d_np;
    const uint32_t wid = threadIdx.x / WAVE, lane = lane_id();
    const uint64_t nw = (uint64_t)gridDim.x * NW;
    for (uint64_t i = (uint64_t)blockIdx.x * NW + wid; i < np; i += nw) {
        // one round of loads: the window ids of entries i - 1 .. i + 62 and
        // the pending entries i .. i + 63
        uint32_t wv = SIZE_NONE;
        if (lane > 0 ? i - 1 + lane < np : i > 0) wv = pw[i - 1 + lane];
        gkey k = KEY_NONE;
        gval v = 0;
        if (i + lane < np) {
            k = pk[i + lane];
            v = pv[i + lane];
        }
        const uint32_t w = __shfl(wv, 1, WAVE);
        if (i > 0 && __shfl(wv, 0, WAVE) == w) continue;  // not a run head
        wave_sync();  // the wave's previous window is done with its LDS rows
        lpk[wid][lane] = k;
        lpv[wid][lane] = v;
        window_wave(S, W, pk, pv, pw, pd, np, i, w, wv, keys, values, seps, row_offset, lower, upper, lk[wid],
                    lv[wid], lpk[wid], lpv[wid]);
    }
}
// Levels 1..nl of one batch in one launch, for W = S << L <= WAVE_WIN_MAX
// at every L <= nl. Each window of those levels lies inside one aligned
// region of R = S << nl slots, and the windows of different regions never
// meet, so one wave per region runs the levels in the reference's order
// (gpma.cuh:1012-1029: level L rebalances the windows that hold inserts level
// L - 1 could not place) on an LDS copy of the region's slots, and writes the
// windows it rebalanced back once at the end. Results equal level-by-level
// launches exactly.
//
// The level-1 list (*d_lvl[1] entries, sorted) gives each entry its level-1
// window (pw) and a done flag (pd, all clear on entry). The first entry of a
// region makes its wave the region's owner; the entries of one level-1
// window are consecutive, and they are placed or not together, so the wave
// keeps one count per level-1 window (lane l: window l of the region, at
// most R / 2S <= 64) and gathers a window's pending entries only for a merge
// (the region's first 64 entries come from an LDS copy, later ones from the
// list). Inserts that level L cannot place are added to d_lvl[L + 1] (the
// inserts entering level L + 1). Nothing runs when no insert entered level 1
// or the batch takes the bulk path (*d_gate >= gate).
template <uint32_t RM>
__global__ __launch_bounds__(256) void k_window_region(uint32_t S, uint32_t nl, const gkey *__restrict__ pk,
                                                       const gval *__restrict__ pv,
                                                       const uint32_t *__restrict__ pw, uint8_t *__restrict__ pd,
                                                       uint32_t *__restrict__ d_lvl, gkey *__restrict__ keys,
                                                       gval *__restrict__ values, gkey *__restrict__ seps,
                                                       uint32_t *__restrict__ row_offset, LevelBounds lb,
                                                       const uint32_t *__restrict__ d_gate, uint32_t gate) {
    constexpr uint32_t NW = 256 / WAVE;
    static_assert(RM == 128 || RM == WAVE_WIN_MAX, "4 slots per lane cover a region");
    __shared__ gkey rk[NW][RM];  // the region's slots (R = S << nl <= RM)
    __shared__ __align__(16) gval rv[NW][RM];  // read and written as 4-byte words
    __shared__ gkey lk[NW][RM];  // live items of the window being merged
    __shared__ gval lv[NW][RM];
    __shared__ gkey lpk[NW][RM];  // its pending inserts
    __shared__ gval lpv[NW][RM];
    __shared__ gkey ek[NW][WAVE];  // the region's first 64 entries
    __shared__ gval ev[NW][WAVE];
    __shared__ uint32_t wlo[NW][WAVE];  // first entry of each level-1 window, relative
    if (d_gate && *d_gate >= gate) return;
    const uint64_t n1 = d_lvl[1];
    if (n1 == 0) return;
    const uint32_t R = S << nl, W1 = S << 1, nw1 = R / W1;
    const uint32_t sh1 = (uint32_t)__builtin_ctz(W1);
    const uint32_t wid = threadIdx.x / WAVE, lane = lane_id();
    const uint64_t nw = (uint64_t)gridDim.x * NW;
    for (uint64_t i = (uint64_t)blockIdx.x * NW + wid; i < n1; i += nw) {
        const uint32_t a = i + lane < n1 ? pw[i + lane] : SIZE_NONE;
        const uint32_t prev = i > 0 ? pw[i - 1] : SIZE_NONE;
        const uint32_t wr = __shfl(a, 0, WAVE) & ~(R - 1);
        if (i > 0 && (prev & ~(R - 1)) == wr) continue;  // not the region's first entry
        wave_sync();  // the wave's previous region is done with its LDS rows
        // one round of loads: the region's slots (4 per lane) and the first
        // 64 entries' keys and values
        const uint32_t p4 = 4 * lane;
        uint4 ka = make_uint4(0, 0, 0, 0), kb = ka;
        uint32_t va = 0;
        if (p4 < R) {
            const uint4 *src = reinterpret_cast<const uint4 *>(keys + wr + p4);
            ka = src[0];
            kb = src[1];
            va = *reinterpret_cast<const uint32_t *>(values + wr + p4);
        }
        gkey ekj = 0;
        gval evj = 0;
        if (i + lane < n1) {
            ekj = pk[i + lane];
            evj = pv[i + lane];
        }
        if (lane < nw1) wlo[wid][lane] = SIZE_NONE;
        // the region's entries: first entry of each level-1 window, and the count e
        uint32_t e = 0;
        {
            bool in = (a & ~(R - 1)) == wr && a != SIZE_NONE;
            uint32_t carry = SIZE_NONE, x = a;
            for (uint64_t j0 = i;; j0 += WAVE) {
                const uint32_t up = __shfl_up(x, 1, WAVE);
                const uint32_t before = lane == 0 ? carry : up;
                wave_sync();
                if (in && x != before) wlo[wid][(x & (R - 1)) >> sh1] = (uint32_t)(j0 - i) + lane;
                const uint64_t out = ballot(!in);
                if (out) {
                    e = (uint32_t)(j0 - i) + (uint32_t)__builtin_ctzll(out);
                    break;
                }
                carry = __shfl(x, WAVE - 1, WAVE);
                const uint64_t j = j0 + WAVE + lane;
                x = j < n1 ? pw[j] : SIZE_NONE;
                in = (x & ~(R - 1)) == wr && x != SIZE_NONE;
            }
        }
        if (lane < e) {
            ek[wid][lane] = ekj;
            ev[wid][lane] = evj;
        }
        if (p4 < R) {
            rk[wid][p4 + 0] = (gkey)ka.x | ((gkey)ka.y << 32);
            rk[wid][p4 + 1] = (gkey)ka.z | ((gkey)ka.w << 32);
            rk[wid][p4 + 2] = (gkey)kb.x | ((gkey)kb.y << 32);
            rk[wid][p4 + 3] = (gkey)kb.z | ((gkey)kb.w << 32);
            *reinterpret_cast<uint32_t *>(&rv[wid][p4]) = va;
        }
        wave_sync();
        // lane l < nw1: level-1 window l's entries [lo, lo + cnt)
        const uint32_t lo = lane < nw1 ? wlo[wid][lane] : SIZE_NONE;
        const uint64_t present = ballot(lo != SIZE_NONE);
        const uint64_t above = present & ~((2ull << lane) - 1);  // lane 63: 2 << 63 wraps to 0, so none
        const uint32_t nxt = __shfl(lo, above ? (uint32_t)__builtin_ctzll(above) : 0, WAVE);
        const uint32_t cnt0 = lo != SIZE_NONE ? (above ? nxt : e) - lo : 0;
        uint32_t cnt = cnt0;    // pending (not placed yet)
        bool dirty = false;     // window rebalanced: write it back
        for (uint32_t L = 1; L <= nl; L++) {
            const uint32_t g = 1u << (L - 1), W = S << L;
            uint32_t P = cnt;  // pending inserts of this lane's level-L window
            for (uint32_t o = 1; o < g; o <<= 1) P += __shfl_xor(P, o, WAVE);
            uint64_t heads = ballot((lane & (g - 1)) == 0 && P > 0);
            uint32_t fails = 0;
            while (heads) {
                const uint32_t h = (uint32_t)__builtin_ctzll(heads);
                heads &= heads - 1;
                const uint32_t r = __shfl(P, h, WAVE);
                if (r > lb.upper[L - 1] || r > W) {  // cannot fit whatever the window holds
                    fails += r;
                    continue;
                }
                const bool mine = lane >= h && lane < h + g;
                // its pending inserts, in order, from its level-1 windows not placed yet
                uint64_t sub = ballot(mine && cnt > 0);
                uint32_t off = 0;
                while (sub) {
                    const uint32_t s = (uint32_t)__builtin_ctzll(sub);
                    sub &= sub - 1;
                    const uint32_t cs = __shfl(cnt, s, WAVE), ls = __shfl(lo, s, WAVE);
                    for (uint32_t t = lane; t < cs; t += WAVE) {
                        const uint32_t rel = ls + t;
                        gkey k;
                        gval v;
                        if (rel < WAVE) {
                            k = ek[wid][rel];
                            v = ev[wid][rel];
                        } else {
                            k = pk[i + rel];
                            v = pv[i + rel];
                        }
                        lpk[wid][off + t] = k;
                        lpv[wid][off + t] = v;
                    }
                    off += cs;
                }
                // the window's live items, in order
                const uint32_t xo = h << sh1;
                uint32_t c = 0;
                for (uint32_t q = 0; q * WAVE < W; q++) {
                    const uint32_t p = q * WAVE + lane;
                    bool live = false;
                    gkey k = 0;
                    gval v = 0;
                    if (p < W) {
                        k = rk[wid][xo + p];
                        v = rv[wid][xo + p];
                        live = is_live(k, v);
                    }
                    const uint64_t mb = ballot(live);
                    if (live) {
                        const uint32_t pos = c + (uint32_t)__popcll(mb & lanemask_lt());
                        lk[wid][pos] = k;
                        lv[wid][pos] = v;
                    }
                    c += (uint32_t)__popcll(mb);
                }
                const uint32_t m = c + r;
                if (m < lb.lower[L - 1] || m > lb.upper[L - 1]) {
                    fails += r;
                    continue;
                }
                wave_sync();
                for (uint32_t p = lane; p < W; p += WAVE) {
                    rk[wid][xo + p] = KEY_NONE;
                    rv[wid][xo + p] = 0;
                }
                wave_sync();
                const uint64_t f = W / S;
                for (uint32_t t = lane; t < c; t += WAVE) {
                    const gkey k = lk[wid][t];
                    uint32_t b0 = 0, b1 = r;
                    while (b0 < b1) {
                        const uint32_t mid = (b0 + b1) >> 1;
                        if (lpk[wid][mid] < k) b0 = mid + 1;
                        else b1 = mid;
                    }
                    const uint32_t slot = redistribute_slot(f, S, m, t + b0);
                    rk[wid][xo + slot] = k;
                    rv[wid][xo + slot] = lv[wid][t];
                }
                for (uint32_t u = lane; u < r; u += WAVE) {
                    const gkey k = lpk[wid][u];
                    uint32_t b0 = 0, b1 = c;
                    while (b0 < b1) {
                        const uint32_t mid = (b0 + b1) >> 1;
                        if (lk[wid][mid] < k) b0 = mid + 1;
                        else b1 = mid;
                    }
                    const uint32_t slot = redistribute_slot(f, S, m, u + b0);
                    rk[wid][xo + slot] = k;
                    rv[wid][xo + slot] = lpv[wid][u];
                }
                wave_sync();
                if (mine) {
                    cnt = 0;
                    dirty = true;
                }
            }
            if (fails && lane == 0) atomicAdd(d_lvl + L + 1, fails);
        }
        // the rebalanced windows back (4 slots per lane: a level-1 window
        // holds whole groups of 4), their separators and moved guards' rows
        const uint64_t dmask = ballot(dirty);
        if (dmask) {
            if (p4 < R && ((dmask >> (p4 >> sh1)) & 1)) {
                gkey k4[4];
#pragma unroll
                for (uint32_t q = 0; q < 4; q++) k4[q] = rk[wid][p4 + q];
                uint4 *dst = reinterpret_cast<uint4 *>(keys + wr + p4);
                dst[0] = make_uint4((uint32_t)k4[0], (uint32_t)(k4[0] >> 32), (uint32_t)k4[1],
                                    (uint32_t)(k4[1] >> 32));
                dst[1] = make_uint4((uint32_t)k4[2], (uint32_t)(k4[2] >> 32), (uint32_t)k4[3],
                                    (uint32_t)(k4[3] >> 32));
                *reinterpret_cast<uint32_t *>(values + wr + p4) = *reinterpret_cast<const uint32_t *>(&rv[wid][p4]);
#pragma unroll
                for (uint32_t q = 0; q < 4; q++) {
                    const uint32_t p = wr + p4 + q;
                    if (p % S == 0) seps[p / S] = k4[q];
                    if (k4[q] != KEY_NONE && is_guard(k4[q])) row_offset[(k4[q] >> 32) + 1] = p;
                }
            }
            // done flags of the placed entries
            uint64_t placed = ballot(dirty && cnt0 > 0);
            while (placed) {
                const uint32_t s = (uint32_t)__builtin_ctzll(placed);
                placed &= placed - 1;
                const uint32_t cs = __shfl(cnt0, s, WAVE), ls = __shfl(lo, s, WAVE);
                for (uint32_t t = lane; t < cs; t += WAVE) pd[i + ls + t] = 1;
            }
        }
    }
}

__global__ __launch_bounds__(256) void k_window(uint32_t S, uint32_t W, const gkey *__restrict__ pk,
                                                const gval *__restrict__ pv, const uint32_t *__restrict__ pw,
                                                uint8_t *__restrict__ pd, const uint32_t *__restrict__ d_np,
                                                gkey *__restrict__ keys, gval *__restrict__ values,
                                                gkey *__restrict__ seps, uint32_t *__restrict__ row_offset,
                                                uint32_t lower, uint32_t upper, const uint32_t *__restrict__ d_gate,
                                                uint32_t gate) {
    __shared__ gkey lk[WIN_MAX];
    __shared__ gval lv[WIN_MAX];
    __shared__ uint32_t scratch[256 / WAVE + 1];
    __shared__ uint32_t s_r;
    // a batch whose level-0 pending count reached the bulk threshold takes the bulk path
    if (d_gate && *d_gate >= gate) return;
    const uint64_t np = *d_np;
    for (uint64_t i = blockIdx.x; i < np; i += gridDim.x) {
        const uint32_t w = pw[i];
        if (i > 0 && pw[i - 1] == w) continue;
        window_one(S, W, pk, pv, pw, pd, np, i, w, keys, values, seps, row_offset, lower, upper, lk, lv, scratch,
                   &s_r);
    }
}

__global__ __launch_bounds__(CMP_BLOCK) void k_head_count(const uint32_t *__restrict__ pw, uint64_t n_ub,
                                                          const uint32_t *__restrict__ d_np,
                                                          uint32_t *__restrict__ blkcnt) {
    __shared__ uint32_t wcnt[CMP_BLOCK / WAVE];
    const uint64_t n = *d_np;
    const uint64_t i = (uint64_t)blockIdx.x * CMP_BLOCK + threadIdx.x;
    bool head = i < n && (i == 0 || pw[i] != pw[i - 1]);
    uint64_t m = ballot(head);
    if (lane_id() == 0) wcnt[threadIdx.x / WAVE] = (uint32_t)__popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t s = 0;
        for (uint32_t w = 0; w < CMP_BLOCK / WAVE; w++) s += wcnt[w];
        blkcnt[blockIdx.x] = s;
    }
    (void)n_ub;
}

__global__ __launch_bounds__(CMP_BLOCK) void k_head_scatter(const uint32_t *__restrict__ pw, uint64_t n_ub,
                                                            const uint32_t *__restrict__ d_np,
                                                            const uint32_t *__restrict__ off,
                                                            uint32_t *__restrict__ heads) {
    __shared__ uint32_t scratch[CMP_BLOCK / WAVE + 1];
    const uint64_t n = *d_np;
    const uint64_t i = (uint64_t)blockIdx.x * CMP_BLOCK + threadIdx.x;
    bool head = i < n && (i == 0 || pw[i] != pw[i - 1]);
    uint32_t tot;
    uint32_t pre = block_exclusive_scan(head ? 1u : 0u, scratch, &tot);
    if (head) heads[off[blockIdx.x] + pre] = (uint32_t)i;
    (void)n_ub;
}

__global__ __launch_bounds__(256) void k_count_live(const gkey *__restrict__ keys, const gval *__restrict__ values,
                                                    uint64_t w, uint64_t W, uint32_t *__restrict__ chunkcnt) {
    __shared__ uint32_t wcnt[256 / WAVE];
    const uint64_t base = (uint64_t)blockIdx.x * CHUNK;
    uint32_t c = 0;
    for (uint32_t j = threadIdx.x; j < CHUNK; j += 256) {
        uint64_t p = base + j;
        if (p < W) c += is_live(keys[w + p], values[w + p]) ? 1u : 0u;
    }
    // wave reduce
    for (int d = WAVE / 2; d > 0; d >>= 1) c += __shfl_down(c, d, WAVE);
    if (lane_id() == 0) wcnt[threadIdx.x / WAVE] = c;
    __syncthreads();
    if (threadIdx.x == 0) chunkcnt[blockIdx.x] = wcnt[0] + wcnt[1] + wcnt[2] + wcnt[3];
}

__global__ __launch_bounds__(256) void k_compact_live(const gkey *__restrict__ keys, const gval *__restrict__ values,
                                                      uint64_t w, uint64_t W, const uint32_t *__restrict__ chunkoff,
                                                      gkey *__restrict__ tk, gval *__restrict__ tv) {
    __shared__ uint32_t scratch[256 / WAVE + 1];
    const uint64_t base = (uint64_t)blockIdx.x * CHUNK;
    uint64_t o = chunkoff[blockIdx.x];
    for (uint32_t j = 0; j < CHUNK; j += 256) {
        uint64_t p = base + j + threadIdx.x;
        bool live = false;
        gkey k = KEY_NONE;
        gval v = 0;
        if (p < W) {
            k = keys[w + p];
            v = values[w + p];
            live = is_live(k, v);
        }
        uint32_t tot;
        uint32_t pre = block_exclusive_scan(live ? 1u : 0u, scratch, &tot);
        if (live) {
            tk[o + pre] = k;
            tv[o + pre] = v;
        }
        o += tot;
    }
}

__global__ __launch_bounds__(256) void k_merge_scatter(const gkey *__restrict__ tk, const gval *__restrict__ tv,
                                                       uint64_t c, const gkey *__restrict__ ik,
                                                       const gval *__restrict__ iv, uint64_t r,
                                                       gkey *__restrict__ keys, gval *__restrict__ values,
                                                       gkey *__restrict__ seps, uint32_t *__restrict__ row_offset,
                                                       uint64_t w, uint64_t W, uint32_t S) {
    const uint64_t m = c + r;
    const uint64_t f = W / S;
    for (uint64_t idx = (uint64_t)blockIdx.x * 256 + threadIdx.x; idx < m; idx += (uint64_t)gridDim.x * 256) {
        gkey k;
        gval v;
        uint64_t mi;
        if (idx < c) {
            k = tk[idx];
            v = tv[idx];
            mi = idx + lower_bound_g(ik, r, k);
        } else {
            const uint64_t u = idx - c;
            k = ik[u];
            v = iv[u];
            mi = u + lower_bound_g(tk, c, k);
        }
        const uint32_t slot = redistribute_slot(f, S, m, mi);
        keys[w + slot] = k;
        values[w + slot] = v;
        if (is_guard(k)) row_offset[(k >> 32) + 1] = (uint32_t)(w + slot);
        if (slot % S == 0) seps[(w + slot) / S] = k;
    }
}

__global__ __launch_bounds__(256) void k_fill_gaps(uint64_t m, gkey *__restrict__ keys, gval *__restrict__ values,
                                                   gkey *__restrict__ seps, uint64_t w, uint64_t W, uint32_t S) {
    const uint64_t f = W / S;
    for (uint64_t p = (uint64_t)blockIdx.x * 256 + threadIdx.x; p < W; p += (uint64_t)gridDim.x * 256) {
        const uint64_t s = p / S, l = p % S;
        const uint64_t a = seg_first_item(f, m, s);
        const uint64_t cnt = seg_first_item(f, m, s + 1) - a;
        if (l >= cnt) {
            keys[w + p] = KEY_NONE;
            values[w + p] = 0;
            if (l == 0) seps[(w + p) / S] = KEY_NONE;
        }
    }
}

// ---- one rebalance level of windows wider than one workgroup (W > WIN_MAX)
//
// All windows of the level are handled together (rebalancing_kernel's level
// pass, gpma.cuh:727-794): the pending list's runs (heads[0 .. *d_nruns), one
// per window) name the windows; every window is CPW = W / CHUNK chunks. The
// passes: live count per chunk (k_wide_count), one planning workgroup
// (k_wide_plan: chunk offsets, each window's live count c, merge size m =
// c + pending run length, the bounds check, and the prefix of m over the
// windows that pass), compaction of the passing windows' live items into tmp
// (k_wide_compact), the merge of compacted items and pending inserts at the
// reference's slot formula (k_wide_merge), the gaps (k_wide_fill) and the
// done flags of the placed inserts (k_wide_done). Windows are disjoint, so
// the result equals handling them one after another.

__device__ __forceinline__ uint64_t wide_run_len(const uint32_t *heads, uint32_t nruns, uint64_t np, uint32_t j) {
    return (j + 1 < nruns ? (uint64_t)heads[j + 1] : np) - heads[j];
}

// last j in [0, n) with a[j] <= x (a[0] <= x)
template <typename T>
__device__ __forceinline__ uint32_t last_le(const T *a, uint32_t n, uint64_t x) {
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if ((uint64_t)a[mid] <= x) lo = mid;
        else hi = mid;
    }
    return lo;
}

__global__ __launch_bounds__(256) void k_wide_count(const gkey *__restrict__ keys, const gval *__restrict__ values,
                                                    const uint32_t *__restrict__ pw,
                                                    const uint32_t *__restrict__ heads,
                                                    const uint32_t *__restrict__ d_nruns, uint32_t cpw,
                                                    uint32_t *__restrict__ chunkcnt) {
    __shared__ uint32_t wcnt[256 / WAVE];
    const uint64_t nb = (uint64_t)*d_nruns * cpw;
    for (uint64_t g = blockIdx.x; g < nb; g += gridDim.x) {
        const uint64_t base = (uint64_t)pw[heads[g / cpw]] + (g % cpw) * CHUNK;
        uint32_t c = 0;
        for (uint32_t j = threadIdx.x; j < CHUNK; j += 256) c += is_live(keys[base + j], values[base + j]) ? 1u : 0u;
        for (int d = WAVE / 2; d > 0; d >>= 1) c += __shfl_down(c, d, WAVE);
        if (lane_id() == 0) wcnt[threadIdx.x / WAVE] = c;
        __syncthreads();
        if (threadIdx.x == 0) chunkcnt[g] = wcnt[0] + wcnt[1] + wcnt[2] + wcnt[3];
        __syncthreads();
    }
}

// One workgroup: chunkoff[0 .. nb] = exclusive scan of chunkcnt (nb =
// nruns * cpw; chunkoff[nb] = total), live[j] and merge[j] per window,
// ok[j] = lower <= merge <= upper, mpre[0 .. nruns] = exclusive scan of
// merge over the windows that pass.
__global__ __launch_bounds__(1024) void k_wide_plan(const uint32_t *__restrict__ chunkcnt,
                                                   const uint32_t *__restrict__ heads,
                                                   const uint32_t *__restrict__ d_nruns,
                                                   const uint32_t *__restrict__ d_np, uint32_t cpw, uint32_t lower,
                                                   uint32_t upper, uint32_t *__restrict__ chunkoff,
                                                   uint32_t *__restrict__ live, uint32_t *__restrict__ merge,
                                                   uint8_t *__restrict__ ok, unsigned long long *__restrict__ mpre) {
    __shared__ uint32_t scratch[1024 / WAVE + 1];
    __shared__ unsigned long long sc64[1024 / WAVE + 1];
    const uint32_t nruns = *d_nruns;
    const uint64_t np = *d_np;
    const uint64_t nb = (uint64_t)nruns * cpw;
    {
        const uint64_t per = (nb + 1023) / 1024;
        const uint64_t a = min<uint64_t>(nb, threadIdx.x * per), b = min<uint64_t>(nb, a + per);
        uint32_t sum = 0;
        for (uint64_t j = a; j < b; j++) sum += chunkcnt[j];
        uint32_t tot;
        uint32_t pre = block_exclusive_scan(sum, scratch, &tot);
        for (uint64_t j = a; j < b; j++) {
            const uint32_t c = chunkcnt[j];
            chunkoff[j] = pre;
            pre += c;
        }
        if (threadIdx.x == 0) chunkoff[nb] = tot;
    }
    __syncthreads();
    const uint32_t per = (nruns + 1023) / 1024;
    const uint32_t a = min(nruns, threadIdx.x * per), b = min(nruns, a + per);
    unsigned long long sum = 0;
    for (uint32_t j = a; j < b; j++) {
        const uint32_t c = chunkoff[(uint64_t)(j + 1) * cpw] - chunkoff[(uint64_t)j * cpw];
        const uint64_t m = c + wide_run_len(heads, nruns, np, j);
        const bool pass = lower <= m && m <= upper;
        live[j] = c;
        merge[j] = (uint32_t)m;
        ok[j] = pass ? 1 : 0;
        sum += pass ? m : 0;
    }
    // block exclusive scan of the 64-bit sums
    const uint32_t lane = lane_id(), wid = threadIdx.x / WAVE;
    unsigned long long v = sum;
    for (int d = 1; d < WAVE; d <<= 1) {
        const unsigned long long y = __shfl_up(v, d, WAVE);
        if (lane >= (uint32_t)d) v += y;
    }
    if (lane == WAVE - 1) sc64[wid] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long run = 0;
        for (uint32_t w = 0; w < 1024 / WAVE; w++) {
            const unsigned long long t = sc64[w];
            sc64[w] = run;
            run += t;
        }
        sc64[1024 / WAVE] = run;
    }
    __syncthreads();
    unsigned long long pre = sc64[wid] + v - sum;
    for (uint32_t j = a; j < b; j++) {
        mpre[j] = pre;
        pre += ok[j] ? merge[j] : 0;
    }
    if (threadIdx.x == 0) mpre[nruns] = sc64[1024 / WAVE];
}

// Live items of every chunk of a passing window into tk/tv at chunkoff.
__global__ __launch_bounds__(256) void k_wide_compact(const gkey *__restrict__ keys, const gval *__restrict__ values,
                                                      const uint32_t *__restrict__ pw,
                                                      const uint32_t *__restrict__ heads,
                                                      const uint32_t *__restrict__ d_nruns, uint32_t cpw,
                                                      const uint32_t *__restrict__ chunkoff,
                                                      const uint8_t *__restrict__ ok, gkey *__restrict__ tk,
                                                      gval *__restrict__ tv) {
    __shared__ uint32_t scratch[256 / WAVE + 1];
    const uint64_t nb = (uint64_t)*d_nruns * cpw;
    for (uint64_t g = blockIdx.x; g < nb; g += gridDim.x) {
        const uint32_t j = (uint32_t)(g / cpw);
        if (!ok[j]) continue;  // uniform over the workgroup
        const uint64_t base = (uint64_t)pw[heads[j]] + (g % cpw) * CHUNK;
        uint64_t o = chunkoff[g];
        for (uint32_t c = 0; c < CHUNK; c += 256) {
            const gkey k = keys[base + c + threadIdx.x];
            const gval v = values[base + c + threadIdx.x];
            const bool lv = is_live(k, v);
            uint32_t tot;
            const uint32_t pre = block_exclusive_scan(lv ? 1u : 0u, scratch, &tot);
            if (lv) {
                tk[o + pre] = k;
                tv[o + pre] = v;
            }
            o += tot;
        }
    }
}

// Merged item t of the passing windows (t < mpre[nruns]): its window j,
// then the reference's slot formula as in k_merge_scatter.
__global__ __launch_bounds__(256) void k_wide_merge(const gkey *__restrict__ tk, const gval *__restrict__ tv,
                                                    const gkey *__restrict__ pk, const gval *__restrict__ pv,
                                                    const uint32_t *__restrict__ pw,
                                                    const uint32_t *__restrict__ heads,
                                                    const uint32_t *__restrict__ d_nruns, uint32_t cpw,
                                                    const uint32_t *__restrict__ chunkoff,
                                                    const uint32_t *__restrict__ live,
                                                    const uint32_t *__restrict__ merge,
                                                    const unsigned long long *__restrict__ mpre, uint64_t W,
                                                    uint32_t S, gkey *__restrict__ keys, gval *__restrict__ values,
                                                    gkey *__restrict__ seps, uint32_t *__restrict__ row_offset) {
    const uint32_t nruns = *d_nruns;
    const uint64_t total = mpre[nruns];
    const uint64_t f = W / S;
    for (uint64_t t = (uint64_t)blockIdx.x * 256 + threadIdx.x; t < total; t += (uint64_t)gridDim.x * 256) {
        // a window that failed adds nothing to the prefix, so the last j
        // with mpre[j] <= t is the passing window that holds item t
        const uint32_t j = last_le(mpre, nruns, t);
        const uint64_t u = t - mpre[j];
        const uint64_t c = live[j], m = merge[j], r = m - c;
        const gkey *tw = tk + chunkoff[(uint64_t)j * cpw];
        const gval *tvw = tv + chunkoff[(uint64_t)j * cpw];
        const uint64_t h = heads[j];
        gkey k;
        gval v;
        uint64_t mi;
        if (u < c) {
            k = tw[u];
            v = tvw[u];
            mi = u + lower_bound_g(pk + h, r, k);
        } else {
            const uint64_t u2 = u - c;
            k = pk[h + u2];
            v = pv[h + u2];
            mi = u2 + lower_bound_g(tw, c, k);
        }
        const uint64_t wb = pw[h];
        const uint32_t slot = redistribute_slot(f, S, m, mi);
        keys[wb + slot] = k;
        values[wb + slot] = v;
        if (is_guard(k)) row_offset[(k >> 32) + 1] = (uint32_t)(wb + slot);
        if (slot % S == 0) seps[(wb + slot) / S] = k;
    }
}

// Gap slots of the passing windows (the slots the merge left free).
__global__ __launch_bounds__(256) void k_wide_fill(const uint32_t *__restrict__ pw, const uint32_t *__restrict__ heads,
                                                   const uint32_t *__restrict__ d_nruns,
                                                   const uint32_t *__restrict__ merge, const uint8_t *__restrict__ ok,
                                                   uint32_t wshift, uint32_t S, gkey *__restrict__ keys,
                                                   gval *__restrict__ values, gkey *__restrict__ seps) {
    const uint64_t W = 1ull << wshift;
    const uint64_t n = (uint64_t)*d_nruns << wshift;
    const uint64_t f = W / S;
    for (uint64_t p = (uint64_t)blockIdx.x * 256 + threadIdx.x; p < n; p += (uint64_t)gridDim.x * 256) {
        const uint32_t j = (uint32_t)(p >> wshift);
        if (!ok[j]) continue;
        const uint64_t l = p & (W - 1), s = l / S, within = l % S, m = merge[j];
        if (within >= seg_first_item(f, m, s + 1) - seg_first_item(f, m, s)) {
            const uint64_t q = (uint64_t)pw[heads[j]] + l;
            keys[q] = KEY_NONE;
            values[q] = 0;
            if (within == 0) seps[q / S] = KEY_NONE;
        }
    }
}

// Done flags of the pending entries of the passing windows.
__global__ __launch_bounds__(256) void k_wide_done(const uint32_t *__restrict__ heads,
                                                   const uint32_t *__restrict__ d_nruns,
                                                   const uint32_t *__restrict__ d_np, const uint8_t *__restrict__ ok,
                                                   uint8_t *__restrict__ pd) {
    const uint32_t nruns = *d_nruns;
    const uint64_t np = *d_np;
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < np; i += (uint64_t)gridDim.x * 256)
        if (ok[last_le(heads, nruns, i)]) pd[i] = 1;
}

// (count, wrapping sum, xor) of the live keys of keys[0 .. C): a
// size-independent digest of the live set. out[0..3) must be zero.
__global__ __launch_bounds__(256) void k_live_digest(const gkey *__restrict__ keys, const gval *__restrict__ values,
                                                     uint64_t C, unsigned long long *__restrict__ out) {
    unsigned long long c = 0, sm = 0, x = 0;
    for (uint64_t p = (uint64_t)blockIdx.x * 256 + threadIdx.x; p < C; p += (uint64_t)gridDim.x * 256) {
        const gkey k = keys[p];
        if (is_live(k, values[p])) {
            c++;
            sm += k;
            x ^= k;
        }
    }
    for (int d = WAVE / 2; d > 0; d >>= 1) {
        c += __shfl_xor(c, d, WAVE);
        sm += __shfl_xor(sm, d, WAVE);
        x ^= __shfl_xor(x, d, WAVE);
    }
    if (lane_id() == 0) {
        if (c) atomicAdd(out, c);
        if (sm) atomicAdd(out + 1, sm);
        if (x) atomicXor(out + 2, x);
    }
}

// The reference keeps row_offset stale here (gpma.cuh:906-934 compacts
// without touching it); pointing row_offset at the compacted guards instead
// keeps every later BFS range inside the array.
__global__ __launch_bounds__(256) void k_write_compacted(const gkey *__restrict__ tk, const gval *__restrict__ tv,
                                                         uint64_t c, gkey *__restrict__ keys,
                                                         gval *__restrict__ values, uint64_t C,
                                                         gkey *__restrict__ seps, uint32_t *__restrict__ row_offset,
                                                         uint32_t S) {
    for (uint64_t p = (uint64_t)blockIdx.x * 256 + threadIdx.x; p < C; p += (uint64_t)gridDim.x * 256) {
        gkey k = p < c ? tk[p] : KEY_NONE;
        keys[p] = k;
        values[p] = p < c ? tv[p] : 0;
        if (p % S == 0) seps[p / S] = k;
        if (p < c && is_guard(k)) row_offset[(k >> 32) + 1] = (uint32_t)p;
    }
}

}  // namespace

void fill_u32(uint32_t *p, uint32_t v, uint64_t n, hipStream_t st) {
    if (!n) return;
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n, 256, 2048)), dim3(256), 0, st, p, v, n);
}
void launch_publish(const uint32_t *a, uint32_t na, const uint32_t *b, uint32_t nb, uint32_t *dst, uint32_t seq,
                    bool clear_a, hipStream_t st) {
    hipLaunchKernelGGL(k_publish, dim3(1), dim3(WAVE), 0, st, const_cast<uint32_t *>(a), na, b, nb, dst, seq,
                       clear_a ? 1u : 0u);
}
void launch_step_gather(const unsigned long long *cnt, const unsigned long long *cnt_big, const uint32_t *reach,
                        uint32_t n, uint32_t *stage, hipStream_t st) {
    hipLaunchKernelGGL(k_step_gather, dim3(1), dim3(WAVE), 0, st, cnt, cnt_big, reach, n, stage);
}
void launch_step_gate(const uint32_t *np0, const uint32_t *np_top, const uint32_t *err, uint32_t bulk_threshold,
                      uint32_t *gate, hipStream_t st) {
    hipLaunchKernelGGL(k_step_gate, dim3(1), dim3(64), 0, st, np0, np_top, err, bulk_threshold, gate);
}
void fill_u8(uint8_t *p, uint8_t v, uint64_t n, hipStream_t st) {
    if (!n) return;
    hipLaunchKernelGGL(k_fill_u8, dim3(grid_for(n, 256, 2048)), dim3(256), 0, st, p, v, n);
}
void fill_u64(uint64_t *p, uint64_t v, uint64_t n, hipStream_t st) {
    if (!n) return;
    hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(n, 256, 2048)), dim3(256), 0, st, p, v, n);
}

void stage_step(uint64_t *dk, const uint64_t *sk, uint8_t *dv, const uint8_t *sv, uint64_t n, uint32_t *dn,
                hipStream_t st) {
    hipLaunchKernelGGL(k_stage_step, dim3(grid_for(n, 256, 2048)), dim3(256), 0, st, dk, sk, dv, sv, n, dn);
}

void zero3(uint32_t *a, uint64_t na, uint64_t *b, uint64_t nb, uint64_t *c, uint64_t nc, hipStream_t st) {
    const uint64_t n = std::max(na, std::max(nb, nc));
    if (!n) return;
    hipLaunchKernelGGL(k_zero3, dim3(grid_for(n, 256, 2048)), dim3(256), 0, st, a, na, b, nb, c, nc);
}

void launch_locate(const gkey *ks, const gval *vs, uint64_t n, const uint32_t *dn, gkey *keys, gval *values,
                   const gkey *seps, uint64_t nseg, uint32_t S, const uint32_t *row_offset, bool row_hint,
                   uint32_t *win, hipStream_t st) {
    if (!n) return;
    dim3 g((uint32_t)((n + CMP_BLOCK - 1) / CMP_BLOCK)), b(CMP_BLOCK);
    const uint32_t rh = row_hint ? 1u : 0u;
#define GPMA_LOCATE(SS) \
    hipLaunchKernelGGL(k_locate<SS>, g, b, 0, st, ks, vs, n, dn, keys, values, seps, nseg, row_offset, rh, win)
    switch (S) {
    case 2: GPMA_LOCATE(2); break;
    case 4: GPMA_LOCATE(4); break;
    case 8: GPMA_LOCATE(8); break;
    case 16: GPMA_LOCATE(16); break;
    default: GPMA_LOCATE(32); break;
    }
#undef GPMA_LOCATE
}

void launch_scan_blocks(const uint32_t *cnt, uint64_t nb, uint32_t *off, uint32_t *total, hipStream_t st,
                        const uint32_t *d_items) {
    hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(1024), 0, st, cnt, nb, d_items, off, total);
}

void launch_compact(const gkey *ik, const gval *iv, const uint32_t *iw, const uint8_t *idone, uint64_t n_ub,
                    const uint32_t *d_n, const uint32_t *off, gkey *ok, gval *ov, uint32_t *ow, uint8_t *odone,
                    uint32_t mask, hipStream_t st) {
    if (!n_ub) return;
    uint64_t nb = (n_ub + CMP_BLOCK - 1) / CMP_BLOCK;
    hipLaunchKernelGGL(k_compact, dim3((uint32_t)std::min<uint64_t>(nb, 2048)), dim3(CMP_BLOCK), 0, st, ik, iv, iw,
                       idone, n_ub, d_n, off, ok, ov, ow, odone, mask);
}

uint64_t compact_status_words(uint64_t n_ub) { return (n_ub + CC_TILE - 1) / CC_TILE + 1; }

void launch_compact_chained(const gkey *ik, const gval *iv, const uint32_t *iw, const uint8_t *idone, uint64_t n_ub,
                            const uint32_t *d_n, gkey *ok, gval *ov, uint32_t *ow, uint8_t *odone, uint32_t mask,
                            unsigned long long *status, uint32_t epoch, uint32_t *ticket, uint32_t *d_out,
                            uint32_t *err, hipStream_t st, const gkey *fix, gkey *seps) {
    const uint64_t nt = std::max<uint64_t>(1, std::min<uint64_t>((n_ub + CC_TILE - 1) / CC_TILE, CC_GRID_MAX));
    hipLaunchKernelGGL(k_compact_chained, dim3((uint32_t)nt), dim3(CC_THREADS), 0, st, ik, iv, iw, idone, n_ub, d_n,
                       ok, ov, ow, odone, mask, status, epoch, ticket, d_out, err, fix, seps);
}

void launch_count_pending(const uint8_t *done, uint64_t n_ub, const uint32_t *d_n, uint32_t *blkcnt,
                          hipStream_t st) {
    if (!n_ub) return;
    uint64_t nb = (n_ub + CMP_BLOCK - 1) / CMP_BLOCK;
    hipLaunchKernelGGL(k_count_pending, dim3((uint32_t)std::min<uint64_t>(nb, 2048)), dim3(CMP_BLOCK), 0, st, done,
                       d_n, blkcnt);
}

void launch_apply(uint32_t S, const gkey *ks, const gval *vs, uint64_t n_ub, const uint32_t *d_n, gkey *keys,
                  gval *values, gkey *seps, uint64_t nseg, uint32_t *row_offset, uint32_t lower, uint32_t upper,
                  uint32_t *win, gkey *fix, uint32_t *np0, hipStream_t st) {
    if (!n_ub) return;
    dim3 g((uint32_t)((n_ub + 255) / 256)), b(256);
    if (S == 16)
        hipLaunchKernelGGL(k_apply<16>, g, b, 0, st, ks, vs, n_ub, d_n, keys, values, seps, nseg, row_offset, lower,
                           upper, win, fix, np0);
    else
        hipLaunchKernelGGL(k_apply<32>, g, b, 0, st, ks, vs, n_ub, d_n, keys, values, seps, nseg, row_offset, lower,
                           upper, win, fix, np0);
}

void launch_leaf(uint32_t S, const gkey *pk, const gval *pv, const uint32_t *pw, uint8_t *pd, uint64_t n_ub,
                 const uint32_t *d_np, gkey *keys, gval *values, gkey *seps, uint32_t *row_offset, uint32_t lower,
                 uint32_t upper, uint32_t bulk_threshold, hipStream_t st) {
    if (!n_ub) return;
    dim3 g((uint32_t)((n_ub + 255) / 256)), b(256);
#define GPMA_LEAF(SS)                                                                                            \
    hipLaunchKernelGGL(k_leaf<SS>, g, b, 0, st, pk, pv, pw, pd, d_np, keys, values, seps, row_offset, lower, upper, \
                       bulk_threshold)
    switch (S) {
    case 2: GPMA_LEAF(2); break;
    case 4: GPMA_LEAF(4); break;
    case 8: GPMA_LEAF(8); break;
    case 16: GPMA_LEAF(16); break;
    default: GPMA_LEAF(32); break;
    }
#undef GPMA_LEAF
}

// Workgroups of k_window_region<RM> the current device holds at once (at
// least WINDOW_GRID), cached per device.
template <uint32_t RM>
static uint64_t region_grid() {
    static std::atomic<uint64_t> cache[64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return WINDOW_GRID;
    uint64_t g = cache[dev].load(std::memory_order_relaxed);
    if (!g) {
        int per_cu = 0, cus = 0;
        g = WINDOW_GRID;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_window_region<RM>, 256, 0) == hipSuccess &&
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && per_cu > 0 &&
            cus > 0)
            g = std::max<uint64_t>(WINDOW_GRID, (uint64_t)per_cu * (uint64_t)cus);
        cache[dev].store(g, std::memory_order_relaxed);
    }
    return g;
}

void window_region_prepare() {
    (void)region_grid<128>();
    (void)region_grid<WAVE_WIN_MAX>();
}

void launch_window_region(uint32_t S, uint32_t nl, const gkey *pk, const gval *pv, const uint32_t *pw, uint8_t *pd,
                          uint64_t n_ub, uint32_t *d_lvl, gkey *keys, gval *values, gkey *seps, uint32_t *row_offset,
                          const LevelBounds &lb, const uint32_t *d_gate, uint32_t gate, hipStream_t st) {
    if (!n_ub || nl == 0) return;
    // LDS rows sized for the region (128 slots at three device levels of 16-slot
    // segments), and as many workgroups as the GPU holds at once: each wave walks
    // its share of the regions one after another
    const bool narrow = ((uint64_t)S << nl) <= 128;
    const dim3 grid((uint32_t)std::min<uint64_t>(n_ub, narrow ? region_grid<128>() : region_grid<WAVE_WIN_MAX>()));
    if (narrow)
        hipLaunchKernelGGL(k_window_region<128>, grid, dim3(256), 0, st, S, nl, pk, pv, pw, pd, d_lvl, keys, values,
                           seps, row_offset, lb, d_gate, gate);
    else
        hipLaunchKernelGGL(k_window_region<WAVE_WIN_MAX>, grid, dim3(256), 0, st, S, nl, pk, pv, pw, pd, d_lvl, keys,
                           values, seps, row_offset, lb, d_gate, gate);
}

void launch_window(uint32_t S, uint32_t W, const gkey *pk, const gval *pv, const uint32_t *pw, uint8_t *pd,
                   uint64_t n_ub, const uint32_t *d_np, gkey *keys, gval *values, gkey *seps,
                   uint32_t *row_offset, uint32_t lower, uint32_t upper, const uint32_t *d_gate, uint32_t gate,
                   hipStream_t st) {
    if (!n_ub) return;
    const dim3 grid((uint32_t)std::min<uint64_t>(n_ub, WINDOW_GRID));
    if (W <= WAVE_WIN_MAX)
        hipLaunchKernelGGL(k_window_wave, grid, dim3(256), 0, st, S, W, pk, pv, pw, pd, d_np, keys, values, seps,
                           row_offset, lower, upper, d_gate, gate);
    else
        hipLaunchKernelGGL(k_window, grid, dim3(256), 0, st, S, W, pk, pv, pw, pd, d_np, keys, values, seps,
                           row_offset, lower, upper, d_gate, gate);
}

void launch_run_heads(const uint32_t *pw, uint64_t n_ub, const uint32_t *d_np, uint32_t *blkcnt, uint32_t *off,
                      uint32_t *heads, uint32_t *d_nruns, hipStream_t st) {
    if (!n_ub) return;
    uint64_t nb = (n_ub + CMP_BLOCK - 1) / CMP_BLOCK;
    hipLaunchKernelGGL(k_head_count, dim3((uint32_t)nb), dim3(CMP_BLOCK), 0, st, pw, n_ub, d_np, blkcnt);
    launch_scan_blocks(blkcnt, nb, off, d_nruns, st, nullptr);
    hipLaunchKernelGGL(k_head_scatter, dim3((uint32_t)nb), dim3(CMP_BLOCK), 0, st, pw, n_ub, d_np, off, heads);
}

void launch_count_live(const gkey *keys, const gval *values, uint64_t w, uint64_t W, uint32_t *chunkcnt,
                       hipStream_t st) {
    uint64_t nb = (W + CHUNK - 1) / CHUNK;
    hipLaunchKernelGGL(k_count_live, dim3((uint32_t)nb), dim3(256), 0, st, keys, values, w, W, chunkcnt);
}

void launch_compact_live(const gkey *keys, const gval *values, uint64_t w, uint64_t W, const uint32_t *chunkoff,
                         gkey *tk, gval *tv, hipStream_t st) {
    uint64_t nb = (W + CHUNK - 1) / CHUNK;
    hipLaunchKernelGGL(k_compact_live, dim3((uint32_t)nb), dim3(256), 0, st, keys, values, w, W, chunkoff, tk, tv);
}

void launch_merge_redistribute(const gkey *tk, const gval *tv, uint64_t c, const gkey *ik, const gval *iv,
                               uint64_t r, gkey *keys, gval *values, gkey *seps, uint32_t *row_offset,
                               uint64_t w, uint64_t W, uint32_t S, hipStream_t st) {
    uint64_t m = c + r;
    if (m) hipLaunchKernelGGL(k_merge_scatter, dim3(grid_for(m, 256, 8192)), dim3(256), 0, st, tk, tv, c, ik, iv, r,
                              keys, values, seps, row_offset, w, W, S);
    hipLaunchKernelGGL(k_fill_gaps, dim3(grid_for(W, 256, 8192)), dim3(256), 0, st, m, keys, values, seps, w, W, S);
}

void launch_wide_level(const gkey *pk, const gval *pv, const uint32_t *pw, uint8_t *pd, uint64_t n_ub,
                       const uint32_t *d_np, uint64_t W, uint32_t S, uint32_t lower, uint32_t upper, gkey *keys,
                       gval *values, gkey *seps, uint32_t *row_offset, const WideWork &w, hipStream_t st) {
    if (!n_ub) return;
    const uint32_t cpw = (uint32_t)(W / CHUNK);
    uint32_t wshift = 0;
    while ((1ull << wshift) < W) wshift++;
    launch_run_heads(pw, n_ub, d_np, w.blkcnt, w.blkoff, w.heads, w.d_nruns, st);
    const dim3 grid(4096), blk(256);
    hipLaunchKernelGGL(k_wide_count, grid, blk, 0, st, keys, values, pw, w.heads, w.d_nruns, cpw, w.chunkcnt);
    hipLaunchKernelGGL(k_wide_plan, dim3(1), dim3(1024), 0, st, w.chunkcnt, w.heads, w.d_nruns, d_np, cpw, lower,
                       upper, w.chunkoff, w.live, w.merge, w.ok, w.mpre);
    hipLaunchKernelGGL(k_wide_compact, grid, blk, 0, st, keys, values, pw, w.heads, w.d_nruns, cpw, w.chunkoff, w.ok,
                       w.tk, w.tv);
    hipLaunchKernelGGL(k_wide_merge, grid, blk, 0, st, w.tk, w.tv, pk, pv, pw, w.heads, w.d_nruns, cpw, w.chunkoff,
                       w.live, w.merge, w.mpre, W, S, keys, values, seps, row_offset);
    hipLaunchKernelGGL(k_wide_fill, grid, blk, 0, st, pw, w.heads, w.d_nruns, w.merge, w.ok, wshift, S, keys, values,
                       seps);
    hipLaunchKernelGGL(k_wide_done, grid, blk, 0, st, w.heads, w.d_nruns, d_np, w.ok, pd);
}

void launch_live_digest(const gkey *keys, const gval *values, uint64_t C, unsigned long long *out, hipStream_t st) {
    fill_u64((uint64_t *)out, 0, 3, st);
    hipLaunchKernelGGL(k_live_digest, dim3(grid_for(C, 256, 4096)), dim3(256), 0, st, keys, values, C, out);
}

void launch_mark_done(uint8_t *pd, uint64_t a, uint64_t b, hipStream_t st) {
    if (b > a) fill_u8(pd + a, 1, b - a, st);
}

void launch_write_compacted(const gkey *tk, const gval *tv, uint64_t c, gkey *keys, gval *values, uint64_t C,
                            gkey *seps, uint32_t *row_offset, uint32_t S, hipStream_t st) {
    hipLaunchKernelGGL(k_write_compacted, dim3(grid_for(C, 256, 8192)), dim3(256), 0, st, tk, tv, c, keys, values, C,
                       seps, row_offset, S);
}

}  // namespace gpma
