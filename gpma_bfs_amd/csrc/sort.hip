// Stable LSD radix sort of (u64 key, u8 value) pairs for update batches.
//
// Reference counterpart: thrust::sort_by_key(update_keys, update_values)
// (gpma.cuh:977), which sorts the batch by the full 64-bit key. Here only
// the 8-bit digits that can differ between valid keys are sorted: a valid
// key has src, dst < row_num, so bits above the row-number width are zero in
// both 32-bit halves; the caller passes that digit list.
//
// Pass 0 (k_stage_hist) copies the batch into the sort buffer, replaces
// out-of-range keys by KEY_NONE (counted; it sorts after every valid key),
// and builds the digit histograms of every pass from one read.
// Then one launch per digit pass ("onesweep", Merrill & Garland / Adinets &
// Merrill 2022): each 4096-item tile ranks its items per digit with wave64
// ballots, publishes its per-digit counts, looks back over earlier tiles'
// published counts to find its global offsets, and scatters. Tiles take
// tickets in launch order and publish their own counts before they wait on
// anyone, so every wait is on a tile that is already running and will publish
// without waiting; the look-back reads 16 predecessors per round trip, and is
// bounded by a spin limit that raises an error flag instead of hanging.
#include "common.h"
#include "kernels.h"

namespace gpma {

namespace {

constexpr uint32_t SORT_THREADS = 256;
constexpr uint32_t SORT_WAVES = SORT_THREADS / WAVE;
constexpr uint32_t SORT_IPT = 16;                      // items per lane
constexpr uint32_t WAVE_ITEMS = WAVE * SORT_IPT;       // 1024
constexpr uint32_t TILE = SORT_THREADS * SORT_IPT;     // 4096
constexpr uint32_t LOOKBACK = 32;                      // predecessors read per round trip
constexpr uint32_t FLAG_AGG = 1u << 30;
constexpr uint32_t FLAG_INC = 2u << 30;
constexpr uint32_t VAL_MASK = (1u << 30) - 1;
constexpr uint32_t SPIN_LIMIT = 1u << 22;
constexpr uint32_t HIST_GRID_MAX = 128;

// Validate + copy the batch and build every pass's digit histogram at once;
// also clears the look-back status words of the passes. Each thread takes
// HIST_IPT items per round with all loads issued before any is used; each
// wave counts into its own LDS histograms (less same-address contention on
// skewed digits), merged into the global histograms at the end.
constexpr uint32_t HIST_IPT = 8;
__global__ __launch_bounds__(SORT_THREADS) void k_stage_hist(const gkey *__restrict__ ik, const gval *__restrict__ iv,
                                                             gkey *__restrict__ ok, gval *__restrict__ ov,
                                                             uint64_t n, uint32_t N, uint32_t guards_ok,
                                                             uint32_t *__restrict__ invalid, SortShifts shifts,
                                                             uint32_t npass, uint32_t *__restrict__ hist,
                                                             uint32_t *__restrict__ status, uint64_t nstatus) {
    __shared__ uint32_t h[SORT_WAVES][8][256];
    for (uint32_t i = threadIdx.x; i < SORT_WAVES * 8 * 256; i += SORT_THREADS) (&h[0][0][0])[i] = 0;
    __syncthreads();
    const uint32_t wid = threadIdx.x / WAVE;
    const uint64_t stride = (uint64_t)gridDim.x * SORT_THREADS;
    uint32_t nbad = 0;
    for (uint64_t r0 = (uint64_t)blockIdx.x * SORT_THREADS * HIST_IPT; r0 < n; r0 += stride * HIST_IPT) {
        gkey k[HIST_IPT];
        gval v[HIST_IPT];
#pragma unroll
        for (uint32_t j = 0; j < HIST_IPT; j++) {
            const uint64_t i = r0 + (uint64_t)j * SORT_THREADS + threadIdx.x;
            k[j] = i < n ? ik[i] : 0;
            v[j] = i < n ? iv[i] : 0;
        }
#pragma unroll
        for (uint32_t j = 0; j < HIST_IPT; j++) {
            const uint64_t i = r0 + (uint64_t)j * SORT_THREADS + threadIdx.x;
            bool bad = false;
            if (i < n) {
                gkey kk = k[j];
                const uint32_t dst = (uint32_t)kk;
                bad = (uint32_t)(kk >> 32) >= N || (dst >= N && !(guards_ok && dst == COL_NONE));
                if (bad) kk = KEY_NONE;
                ok[i] = kk;
                ov[i] = v[j];
                for (uint32_t p = 0; p < npass; p++) atomicAdd(&h[wid][p][(uint32_t)(kk >> shifts.s[p]) & 0xFF], 1u);
            }
            nbad += (uint32_t)__popcll(ballot(bad));
        }
    }
    if (lane_id() == 0 && nbad) atomicAdd(invalid, nbad);
    for (uint64_t i = (uint64_t)blockIdx.x * SORT_THREADS + threadIdx.x; i < nstatus; i += stride) status[i] = 0;
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < npass * 256; i += SORT_THREADS) {
        uint32_t c = 0;
#pragma unroll
        for (uint32_t w = 0; w < SORT_WAVES; w++) c += (&h[w][0][0])[i];
        if (c) atomicAdd(&hist[i], c);
    }
}

__global__ __launch_bounds__(SORT_THREADS) void k_sort_pass(const gkey *__restrict__ kin, const gval *__restrict__ vin,
                                                            gkey *__restrict__ kout, gval *__restrict__ vout,
                                                            uint64_t n, uint32_t shift,
                                                            const uint32_t *__restrict__ hist,  // [256] of this pass
                                                            uint32_t *__restrict__ status,      // [ntiles][256]
                                                            uint32_t *__restrict__ ticket, uint32_t *__restrict__ err) {
    __shared__ uint32_t cnt[SORT_WAVES][256];
    __shared__ uint32_t gofs[256];
    __shared__ uint32_t scan_tmp[SORT_WAVES + 1];
    __shared__ uint32_t s_tile;

    const uint32_t tid = threadIdx.x;
    const uint32_t wid = tid / WAVE;
    const uint32_t lane = lane_id();
    for (uint32_t i = tid; i < SORT_WAVES * 256; i += SORT_THREADS) (&cnt[0][0])[i] = 0;
    if (tid == 0) s_tile = atomicAdd(ticket, 1u);
    const uint32_t hist_d = hist[tid];  // SORT_THREADS == 256: thread d owns digit d
    __syncthreads();
    const uint32_t tile = s_tile;
    const uint64_t wbase = (uint64_t)tile * TILE + (uint64_t)wid * WAVE_ITEMS;

    gkey k[SORT_IPT];
    gval v[SORT_IPT];
    uint32_t rank[SORT_IPT];
    uint32_t dig[SORT_IPT];
#pragma unroll
    for (uint32_t j = 0; j < SORT_IPT; j++) {
        uint64_t idx = wbase + (uint64_t)j * WAVE + lane;
        bool ok = idx < n;
        k[j] = ok ? kin[idx] : 0;
        v[j] = ok ? vin[idx] : 0;
        dig[j] = ok ? ((uint32_t)(k[j] >> shift) & 0xFF) : 256u;
    }
    // Stable per-wave ranking in striped order (item j of all lanes, then j+1).
#pragma unroll
    for (uint32_t j = 0; j < SORT_IPT; j++) {
        const uint32_t d = dig[j];
        const bool ok = d < 256;
        uint64_t match = ballot(ok);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            uint64_t bb = ballot(ok && ((d >> b) & 1));
            match &= ((d >> b) & 1) ? bb : ~bb;
        }
        uint32_t leader = ok ? (uint32_t)__ffsll((unsigned long long)match) - 1 : 0;
        uint32_t base = 0;
        if (ok && lane == leader) {
            base = cnt[wid][d];
            cnt[wid][d] = base + (uint32_t)__popcll(match);
        }
        base = __shfl(base, leader, WAVE);
        rank[j] = base + (uint32_t)__popcll(match & lanemask_lt());
    }
    __syncthreads();
    // Thread d: tile count for digit d, and per-wave prefixes.
    const uint32_t d = tid;  // SORT_THREADS == 256
    uint32_t tot = 0;
#pragma unroll
    for (uint32_t w = 0; w < SORT_WAVES; w++) {
        uint32_t c = cnt[w][d];
        cnt[w][d] = tot;
        tot += c;
    }
    // Global digit base: exclusive scan of the pass histogram (before the
    // look-back, so its latency overlaps the wait).
    uint32_t htot;
    const uint32_t hb = block_exclusive_scan(hist_d, scan_tmp, &htot);
    // Publish this tile's aggregate, then look back LOOKBACK tiles per round.
    uint32_t *my = status + (uint64_t)tile * 256 + d;
    __hip_atomic_store(my, (tile == 0 ? FLAG_INC : FLAG_AGG) | tot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    uint32_t excl = 0;
    if (tile > 0) {
        int64_t t = (int64_t)tile - 1;
        uint32_t spins = 0;
        while (t >= 0) {
            uint32_t s[LOOKBACK];
#pragma unroll
            for (uint32_t j = 0; j < LOOKBACK; j++)
                s[j] = (t - (int64_t)j >= 0)
                           ? __hip_atomic_load(status + (uint64_t)(t - j) * 256 + d, __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT)
                           : FLAG_INC;
            uint32_t used = 0;
            bool stop = false;
#pragma unroll
            for (uint32_t j = 0; j < LOOKBACK; j++) {
                if (stop || used != j) break;
                const uint32_t flag = s[j] & ~VAL_MASK;
                if (flag == 0) break;  // not published yet: poll again from here
                excl += s[j] & VAL_MASK;
                used++;
                if (flag == FLAG_INC) stop = true;
            }
            if (stop) break;
            if (used == 0) {
                if (++spins > SPIN_LIMIT) {
                    atomicOr(err, 1u);
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
            t -= used;
        }
        __hip_atomic_store(my, FLAG_INC | (excl + tot), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    gofs[d] = hb + excl;
    __syncthreads();
#pragma unroll
    for (uint32_t j = 0; j < SORT_IPT; j++) {
        if (dig[j] < 256) {
            uint64_t pos = (uint64_t)gofs[dig[j]] + cnt[wid][dig[j]] + rank[j];
            kout[pos] = k[j];
            vout[pos] = v[j];
        }
    }
}

}  // namespace

uint64_t sort_status_words(uint64_t n, uint32_t npass) {
    uint64_t tiles = (n + TILE - 1) / TILE;
    return tiles * 256 * npass;
}

// Stages (ik, iv) into (k0, v0) and sorts them by the byte digits at
// `shifts` (LSD first), ping-ponging with (k1, v1). Returns 0 if the result
// is in (k0, v0), 1 if in (k1, v1).
// work: hist [npass*256] + tickets [npass] + status words.
int stage_and_sort(const gkey *ik, const gval *iv, gkey *k0, gval *v0, gkey *k1, gval *v1, uint64_t n, uint32_t N,
                   bool guards_ok, uint32_t *invalid, const SortShifts &shifts, uint32_t npass, uint32_t *work,
                   uint32_t *d_err, hipStream_t st) {
    if (n == 0) return 0;
    uint32_t *hist = work;
    uint32_t *tickets = hist + npass * 256;
    uint32_t *status = tickets + npass;
    const uint64_t tiles = (n + TILE - 1) / TILE;
    const uint64_t nstatus = tiles * 256 * npass;
    fill_u32(hist, 0, npass * 256 + npass, st);
    uint64_t want = (std::max<uint64_t>(n, nstatus) + SORT_THREADS * HIST_IPT - 1) / (SORT_THREADS * HIST_IPT);
    uint32_t hb = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(HIST_GRID_MAX, want));
    hipLaunchKernelGGL(k_stage_hist, dim3(hb), dim3(SORT_THREADS), 0, st, ik, iv, k0, v0, n, N, guards_ok ? 1u : 0u,
                       invalid, shifts, npass, hist, status, nstatus);
    int cur = 0;
    for (uint32_t p = 0; p < npass; p++) {
        gkey *ki = cur ? k1 : k0;
        gval *vi = cur ? v1 : v0;
        gkey *ko = cur ? k0 : k1;
        gval *vo = cur ? v0 : v1;
        hipLaunchKernelGGL(k_sort_pass, dim3((uint32_t)tiles), dim3(SORT_THREADS), 0, st, ki, vi, ko, vo, n,
                           shifts.s[p], hist + p * 256, status + (uint64_t)p * tiles * 256, tickets + p, d_err);
        cur ^= 1;
    }
    return cur;
}

}  // namespace gpma
