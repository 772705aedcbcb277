// Bucket (MSD) sort of an update batch whose keys pack into 64 bits.
//
// Reference counterpart: thrust::sort_by_key(update_keys, update_values)
// (gpma.cuh:977). Same result as the LSD sort in sort.hip (stable order by
// the full key), by a different route that suits batches of 10^4..10^6:
//
//   pack      a valid key (src, dst) with value v becomes
//             src << (bs + 8) | dst << 8 | v, bs = bit width of row_num, so
//             one u64 carries key and value; a guard's dst (0xFFFFFFFF)
//             packs as row_num and an out-of-range key as all ones, which
//             keeps the order of the unpacked keys.
//   rank      (k_msd_rank) each tile (1024..4096 items) ranks its items by bucket
//             (wave ballots) = the row's slot position scaled to D bits, read
//             from a table of row splitters in LDS (k_msd_splitters), or the
//             top D bits of the packed key for the guard insertion; writes the
//             packed keys, their rank inside (tile, bucket) and the tile's
//             bucket counts;
//   scan      (k_msd_scan) exclusive scan of the counts in (bucket, tile)
//             order, chained across blocks of 64 buckets;
//   scatter   (k_msd_scatter) every item to its bucket's range, stable;
//   local     (k_msd_local) one workgroup per bucket sorts the bucket by the
//             remaining bits: runs sorted in registers and merged in LDS up to 1024 items,
//             else LSD passes with 8-bit digits over the bucket's range, and
//             writes the unpacked keys and values.
//
// No pass waits on a chain of tiles longer than the 64 scan blocks, so the
// whole sort is four launches whatever the key width.
#include "common.h"
#include "kernels.h"

namespace gpma {

namespace {

constexpr uint32_t MT = 256;                 // threads per workgroup
constexpr uint32_t MW = MT / WAVE;           // waves per workgroup
constexpr uint32_t MIPT = 4;                 // items per lane (the local pass's LSD passes over a large bucket)
constexpr uint32_t MWI = WAVE * MIPT;        // items per wave
constexpr uint32_t MTILE = MT * MIPT;        // items per tile
constexpr uint32_t MAX_D = 12;               // bucket bits
constexpr uint32_t MAX_NB = 1u << MAX_D;
static_assert(SPL_BITS == MAX_D, "the splitter table serves every bucket width up to MAX_D");
constexpr uint32_t TINY = 128;               // buckets ranked by plain comparison in LDS
constexpr uint32_t SMALL = 1024;             // buckets sorted in LDS (runs of 64, then a merge rank)
constexpr uint32_t RUN = WAVE;               // items per register-sorted run
constexpr uint32_t SCAN_B = WAVE;            // buckets per scan block
constexpr uint32_t SCAN_T = 1024;            // threads per scan block
constexpr uint32_t SCAN_G = SCAN_T / SCAN_B; // tile groups per scan block
constexpr uint32_t SCAN_R = 32;              // tiles per scan thread held in registers (else read twice)

struct PackInfo {
    uint32_t N;          // row_num
    uint32_t bs;         // bits per packed src / dst field
    uint32_t guards_ok;  // dst == 0xFFFFFFFF accepted (guard keys)
};

__device__ __forceinline__ uint64_t pack_key(gkey k, gval v, const PackInfo &pi, bool &bad) {
    const uint32_t src = (uint32_t)(k >> 32), dst = (uint32_t)k;
    const bool guard = pi.guards_ok && dst == COL_NONE;
    bad = src >= pi.N || (dst >= pi.N && !guard);
    if (bad) return ~0ull >> (56 - 2 * pi.bs);
    return ((uint64_t)src << (pi.bs + 8)) | ((uint64_t)(guard ? pi.N : dst) << 8) | v;
}

__device__ __forceinline__ void unpack_key(uint64_t p, const PackInfo &pi, gkey &k, gval &v) {
    const uint32_t src = (uint32_t)(p >> (pi.bs + 8));
    const uint32_t dst = (uint32_t)((p >> 8) & ((1ull << pi.bs) - 1));
    v = (gval)p;
    k = src >= pi.N ? KEY_NONE : (((uint64_t)src << 32) | (dst == pi.N ? COL_NONE : dst));
}

// Tiles of MT * IPT items (IPT items per lane): the plan picks IPT from the
// batch size so that small batches still spread over every CU. The bucket of
// an item is the top D bits of its packed key, or, with spl (row buckets), the
// last bucket whose splitter row is <= the item's row: the splitters cut the
// rows where their slot positions cross multiples of C / 2^D (k_msd_splitters),
// so the buckets follow the key order and stay balanced when the sources are
// skewed (R-MAT). The table sits in LDS: no per-item global read.
// rk[i] = bucket << 16 | rank inside (tile, bucket).
template <uint32_t IPT>
__global__ __launch_bounds__(MT) void k_msd_rank(const gkey *__restrict__ ik, const gval *__restrict__ iv, uint64_t n,
                                                 const uint32_t *__restrict__ dn, PackInfo pi, uint32_t bshift,
                                                 uint32_t D, const uint32_t *__restrict__ spl,
                                                 uint64_t *__restrict__ kp, uint32_t *__restrict__ rk,
                                                 uint16_t *__restrict__ counts, uint32_t *__restrict__ invalid) {
    constexpr uint32_t TILE = MT * IPT, WI = WAVE * IPT;
    __shared__ uint16_t cnt[MW][MAX_NB];
    __shared__ uint32_t s_spl[MAX_NB];
    if (dn) n = *dn;
    if ((uint64_t)blockIdx.x * TILE >= n) return;  // tiles past the batch (grid sized for a capacity)
    const uint32_t NB = 1u << D;
    if (spl) {
        // splitters 1..NB-1 (every (2^(12 - D))-th of the 12-bit table) in
        // breadth-first tree order: splitter b sits at depth D - 1 - ctz(b),
        // so the lanes of one search round read neighbouring words (no bank
        // conflicts), where a sorted table would put them in one bank
        for (uint32_t b = 1 + threadIdx.x; b < NB; b += MT) {
            const uint32_t t = (uint32_t)__builtin_ctz(b);
            s_spl[(1u << (D - 1 - t)) + (b >> (t + 1))] = spl[b << (SPL_BITS - D)];
        }
    }
    {  // zero the first NB counters of every wave's row (NB >= 2)
        const uint32_t hs = D - 1;  // log2 of the u32 words in use per row
        for (uint32_t i = threadIdx.x; i < (MW << hs); i += MT)
            reinterpret_cast<uint32_t *>(&cnt[i >> hs][0])[i & ((1u << hs) - 1)] = 0;
    }
    const uint32_t wid = threadIdx.x / WAVE, lane = lane_id();
    const uint64_t wbase = (uint64_t)blockIdx.x * TILE + (uint64_t)wid * WI;
    gkey k[IPT];
    gval v[IPT];
#pragma unroll
    for (uint32_t j = 0; j < IPT; j++) {
        const uint64_t idx = wbase + (uint64_t)j * WAVE + lane;
        k[j] = idx < n ? ik[idx] : 0;
        v[j] = idx < n ? iv[idx] : 0;
    }
    uint64_t p[IPT];
    uint32_t b[IPT];
    uint32_t nbad = 0;
#pragma unroll
    for (uint32_t j = 0; j < IPT; j++) {
        const uint64_t idx = wbase + (uint64_t)j * WAVE + lane;
        bool bad = false;
        p[j] = 0;
        b[j] = NB;  // no item
        if (idx < n) {
            p[j] = pack_key(k[j], v[j], pi, bad);
            if (!spl) b[j] = (uint32_t)(p[j] >> bshift) & (NB - 1);
            else if (bad) b[j] = NB - 1;  // packs as all ones: after every valid key
        }
        nbad += (uint32_t)__popcll(ballot(bad));
    }
    if (lane == 0 && nbad) atomicAdd(invalid, nbad);
    __syncthreads();  // counters zeroed, splitters staged
    if (spl) {
        // the bucket = the number of splitters 1..NB-1 <= src (splitter 0
        // is row 0): D rounds down the tree, the IPT searches interleaved
        uint32_t x[IPT], src[IPT];
#pragma unroll
        for (uint32_t j = 0; j < IPT; j++) {
            x[j] = 1;
            src[j] = (uint32_t)(k[j] >> 32);
        }
        for (uint32_t d = 0; d < D; d++) {
#pragma unroll
            for (uint32_t j = 0; j < IPT; j++) x[j] = 2 * x[j] + (s_spl[x[j]] <= src[j] ? 1u : 0u);
        }
#pragma unroll
        for (uint32_t j = 0; j < IPT; j++)
            if (b[j] == NB && wbase + (uint64_t)j * WAVE + lane < n) b[j] = x[j] - NB;
    }
    uint32_t r[IPT];
#pragma unroll
    for (uint32_t j = 0; j < IPT; j++) r[j] = wave_rank<MAX_D>(b[j], b[j] < NB, cnt[wid]);
    __syncthreads();
    for (uint32_t d = threadIdx.x; d < NB; d += MT) {
        uint32_t run = 0;
#pragma unroll
        for (uint32_t w = 0; w < MW; w++) {
            const uint32_t c = cnt[w][d];
            cnt[w][d] = (uint16_t)run;
            run += c;
        }
        counts[(uint64_t)blockIdx.x * NB + d] = (uint16_t)run;  // a tile holds <= 4096 items
    }
    __syncthreads();
#pragma unroll
    for (uint32_t j = 0; j < IPT; j++) {
        const uint64_t idx = wbase + (uint64_t)j * WAVE + lane;
        if (idx < n) {
            kp[idx] = p[j];
            rk[idx] = (b[j] << 16) | (cnt[wid][b[j]] + r[j]);
        }
    }
}

// Row splitters at SPL_BITS resolution: spl[c] = the first row r whose slot
// position ro[r] scaled to SPL_BITS bits (ro[r] >> (cs - SPL_BITS), cs =
// log2 C) is >= c, or N when no row reaches c. Row r writes the c in
// (scale(ro[r - 1]), scale(ro[r])] (row N stands for every c past the last
// row), so each entry is written exactly once; a thread takes 4 rows with one
// 16-byte load. The bucket of src at D bits is then the last b with
// spl[b << (SPL_BITS - D)] <= src, which is ro[src] scaled to D bits. Any
// non-decreasing table sorts correctly; a stale one only balances the buckets
// less well.
__global__ void k_msd_splitters(const uint32_t *__restrict__ ro, uint32_t N, int32_t cs, uint32_t *__restrict__ spl) {
    const uint64_t r0 = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) * 4;
    if (r0 > N) return;
    constexpr uint32_t TOP = (1u << SPL_BITS) - 1;
    auto scale = [&](uint32_t pos) {
        const uint64_t c = cs >= (int32_t)SPL_BITS ? (uint64_t)pos >> (cs - (int32_t)SPL_BITS)
                                                   : (uint64_t)pos << ((int32_t)SPL_BITS - cs);
        return (uint32_t)(c < TOP ? c : TOP);
    };
    uint32_t p[4];
    if (r0 + 4 <= N) {  // ro is 256-byte aligned: rows 4t.. sit in one 16-byte word
        const uint4 x = *reinterpret_cast<const uint4 *>(ro + r0);
        p[0] = x.x;
        p[1] = x.y;
        p[2] = x.z;
        p[3] = x.w;
    } else {
#pragma unroll
        for (uint32_t j = 0; j < 4; j++) p[j] = r0 + j < N ? ro[r0 + j] : 0u;
    }
    uint32_t lo = r0 > 0 ? scale(ro[r0 - 1]) + 1 : 0;
#pragma unroll
    for (uint32_t j = 0; j < 4; j++) {
        const uint64_t r = r0 + j;
        if (r > N) break;
        const uint32_t hi = r < N ? scale(p[j]) : TOP;
        for (uint32_t c = lo; c <= hi; c++) spl[c] = (uint32_t)r;
        lo = hi + 1;
    }
}

// offs[t][b] = items in buckets < b + items of bucket b in tiles < t;
// bstart[b] = offs[0][b], bstart[NB] = n. A block covers SCAN_B buckets with
// SCAN_G tile groups (each thread walks ntiles / SCAN_G tiles of one
// bucket's column); blocks chain their bucket totals by look-back. With dn,
// n = *dn and only its ceil(n / tile) tiles take part.
__global__ __launch_bounds__(SCAN_T) void k_msd_scan(const uint16_t *__restrict__ counts, uint32_t ntiles, uint32_t tile,
                                                 uint32_t NB, uint64_t n, const uint32_t *__restrict__ dn,
                                                 uint32_t *__restrict__ offs, uint32_t *__restrict__ bstart,
                                                 unsigned long long *status, uint32_t epoch,
                                                 uint32_t *__restrict__ ticket, uint32_t *__restrict__ err) {
    __shared__ uint32_t gs[SCAN_G][SCAN_B];
    __shared__ uint32_t base_s[SCAN_B];
    __shared__ uint32_t s_blk;
    if (dn) {
        n = *dn;
        ntiles = min(ntiles, (uint32_t)((n + tile - 1) / tile));
    }
    if (threadIdx.x == 0) s_blk = atomicAdd(ticket, 1u);
    __syncthreads();
    const uint32_t blk = s_blk;
    const uint32_t g = threadIdx.x / SCAN_B, bl = threadIdx.x % SCAN_B;
    const uint32_t bkt = blk * SCAN_B + bl;
    const uint32_t tg = (ntiles + SCAN_G - 1) / SCAN_G;
    const uint32_t t0 = min(ntiles, g * tg), t1 = min(ntiles, t0 + tg);
    // up to SCAN_R tiles per thread: every count load in flight at once and
    // kept in registers for the second pass
    const bool inreg = tg <= SCAN_R;
    uint32_t cr[SCAN_R];
    uint32_t sum = 0;
    if (inreg) {
#pragma unroll
        for (uint32_t r = 0; r < SCAN_R; r++) {
            const uint32_t t = t0 + r;
            cr[r] = bkt < NB && t < t1 ? counts[(uint64_t)t * NB + bkt] : 0u;
        }
#pragma unroll
        for (uint32_t r = 0; r < SCAN_R; r++) sum += cr[r];
    } else if (bkt < NB) {
#pragma unroll 8
        for (uint32_t t = t0; t < t1; t++) sum += counts[(uint64_t)t * NB + bkt];
    }
    gs[g][bl] = sum;
    __syncthreads();
    uint32_t gpre = 0, col = 0;
#pragma unroll
    for (uint32_t q = 0; q < SCAN_G; q++) {
        if (q < g) gpre += gs[q][bl];
        col += gs[q][bl];
    }
    if (g == 0) {
        // wave 0: bucket-exclusive prefix inside the block, block total, look-back
        uint32_t v = col;
#pragma unroll
        for (uint32_t d = 1; d < WAVE; d <<= 1) {
            const uint32_t y = __shfl_up(v, d, WAVE);
            if (bl >= d) v += y;
        }
        const uint32_t agg = __shfl(v, WAVE - 1, WAVE);
        const uint32_t excl = wave_lookback(status, blk, epoch, agg, err);
        base_s[bl] = excl + v - col;
    }
    __syncthreads();
    if (bkt < NB && inreg) {
        uint32_t run = base_s[bl] + gpre;
        if (g == 0) bstart[bkt] = base_s[bl];
#pragma unroll
        for (uint32_t r = 0; r < SCAN_R; r++) {
            const uint32_t t = t0 + r;
            if (t < t1) offs[(uint64_t)t * NB + bkt] = run;
            run += cr[r];
        }
    } else if (bkt < NB) {
        uint32_t run = base_s[bl] + gpre;
        if (g == 0) bstart[bkt] = base_s[bl];
#pragma unroll 8
        for (uint32_t t = t0; t < t1; t++) {
            const uint64_t at = (uint64_t)t * NB + bkt;
            const uint32_t c = counts[at];
            offs[at] = run;
            run += c;
        }
    }
    if (blk == 0 && threadIdx.x == 0) bstart[NB] = (uint32_t)n;
}

template <uint32_t IPT>
__global__ __launch_bounds__(MT) void k_msd_scatter(const uint64_t *__restrict__ kp, const uint32_t *__restrict__ rk,
                                                    uint64_t n, const uint32_t *__restrict__ dn, uint32_t NB,
                                                    const uint32_t *__restrict__ offs, uint64_t *__restrict__ kb) {
    if (dn) n = *dn;
    const uint64_t base = (uint64_t)blockIdx.x * (MT * IPT);
    if (base >= n) return;
    const uint32_t *o = offs + (uint64_t)blockIdx.x * NB;
    uint64_t p[IPT];
    uint32_t r[IPT];
#pragma unroll
    for (uint32_t j = 0; j < IPT; j++) {
        const uint64_t idx = base + (uint64_t)j * MT + threadIdx.x;
        p[j] = idx < n ? kp[idx] : 0;
        r[j] = idx < n ? rk[idx] : 0;
    }
#pragma unroll
    for (uint32_t j = 0; j < IPT; j++) {
        const uint64_t idx = base + (uint64_t)j * MT + threadIdx.x;
        if (idx < n) kb[(uint64_t)o[r[j] >> 16] + (r[j] & 0xFFFFu)] = p[j];
    }
}

// (key, position) order of two bucket items: the packed key without its
// value byte, then the position in the bucket (which makes the sort stable).
__device__ __forceinline__ bool item_less(uint64_t pa, uint32_t ia, uint64_t pb, uint32_t ib) {
    const uint64_t a = pa >> 8, b = pb >> 8;
    return a < b || (a == b && ia < ib);
}

// One wave sorts its 64 (item, position) pairs across lanes, ascending:
// bitonic network over lane exchanges.
__device__ __forceinline__ void wave_bitonic(uint64_t &p, uint32_t &id) {
    const uint32_t lane = lane_id();
#pragma unroll
    for (uint32_t k2 = 2; k2 <= RUN; k2 <<= 1) {
#pragma unroll
        for (uint32_t j = k2 >> 1; j > 0; j >>= 1) {
            const uint64_t q = __shfl_xor(p, (int)j, WAVE);
            const uint32_t qi = __shfl_xor(id, (int)j, WAVE);
            const bool want_min = ((lane & k2) == 0) == ((lane & j) == 0);
            const bool q_less = item_less(q, qi, p, id);
            if (want_min ? q_less : item_less(p, id, q, qi)) {
                p = q;
                id = qi;
            }
        }
    }
}

// One workgroup per bucket: stable sort of kb[bstart[b], bstart[b+1]) by the
// packed key's bits [8, 8 + R) (the bucket bits above them are equal), then
// unpack into (ok, ov) at the same positions. tmp is scratch of the same
// range. Buckets above SMALL items raise *maxb to their size.
//
// Up to TINY items: each item's rank is the count of smaller (key, position)
// pairs, by comparison with all of them. Up to SMALL: each wave sorts runs
// of 64 in registers (bitonic), and an item's rank is its place in its run
// plus, for every other run, the count of smaller pairs there (binary search
// in LDS). Larger buckets: LSD passes with 8-bit digits over the bucket's
// range in global memory.
__global__ __launch_bounds__(MT) void k_msd_local(uint64_t *__restrict__ kb, uint64_t *__restrict__ tmp,
                                                  const uint32_t *__restrict__ bstart, uint32_t R, PackInfo pi,
                                                  gkey *__restrict__ ok, gval *__restrict__ ov,
                                                  uint32_t *__restrict__ maxb) {
    __shared__ uint64_t lk[SMALL];
    __shared__ uint16_t lid[SMALL];
    __shared__ uint32_t cnt[MW][256];
    __shared__ uint32_t run[256];
    __shared__ uint32_t tot[256];
    __shared__ uint32_t scan_tmp[MW + 1];
    const uint32_t tid = threadIdx.x;
    const uint32_t start = bstart[blockIdx.x], end = bstart[blockIdx.x + 1];
    const uint32_t c = end - start;
    if (c == 0) return;
    if (c > TINY && c <= SMALL) {
        const uint32_t nrun = (c + RUN - 1) / RUN;
        const uint32_t lane = lane_id();
        for (uint32_t r = tid / WAVE; r < nrun; r += MW) {
            const uint32_t i = r * RUN + lane;
            uint64_t p = i < c ? kb[(uint64_t)start + i] : ~0ull;  // padding sorts last
            uint32_t id = i < c ? i : 0xFFFFu;
            wave_bitonic(p, id);
            lk[i] = p;
            lid[i] = (uint16_t)id;
        }
        __syncthreads();
        for (uint32_t x = tid; x < nrun * RUN; x += MT) {
            const uint32_t id = lid[x];
            if (id == 0xFFFFu) continue;  // padding
            const uint64_t p = lk[x];
            const uint32_t r = x / RUN;
            uint32_t rank = x % RUN;
            for (uint32_t q = 0; q < nrun; q++) {
                if (q == r) continue;
                const uint32_t b = q * RUN;
                uint32_t lo = 0;
#pragma unroll
                for (uint32_t step = RUN; step > 0; step >>= 1)
                    if (lo + step <= RUN && item_less(lk[b + lo + step - 1], lid[b + lo + step - 1], p, id)) lo += step;
                rank += lo;
            }
            gkey k;
            gval v;
            unpack_key(p, pi, k, v);
            ok[(uint64_t)start + rank] = k;
            ov[(uint64_t)start + rank] = v;
        }
        return;
    }
    if (c <= TINY) {
        for (uint32_t i = tid; i < c; i += MT) lk[i] = kb[(uint64_t)start + i];
        __syncthreads();
        for (uint32_t i = tid; i < c; i += MT) {
            const uint64_t me = lk[i] >> 8;
            uint32_t rank = 0;
            for (uint32_t j = 0; j < c; j++) {
                const uint64_t o = lk[j] >> 8;
                rank += (o < me || (o == me && j < i)) ? 1u : 0u;
            }
            gkey k;
            gval v;
            unpack_key(lk[i], pi, k, v);
            ok[(uint64_t)start + rank] = k;
            ov[(uint64_t)start + rank] = v;
        }
        return;
    }
    if (tid == 0) atomicMax(maxb, c);
    const uint32_t wid = tid / WAVE, lane = lane_id();
    const uint32_t npass = (R + 7) / 8;
    if (npass == 0) {
        for (uint32_t i = tid; i < c; i += MT) {
            gkey k;
            gval v;
            unpack_key(kb[(uint64_t)start + i], pi, k, v);
            ok[(uint64_t)start + i] = k;
            ov[(uint64_t)start + i] = v;
        }
        return;
    }
    uint64_t *src = kb, *dst = tmp;
    for (uint32_t q = 0; q < npass; q++) {
        const uint32_t shift = 8 + 8 * q;
        const bool last = q + 1 == npass;
        run[tid] = 0;
        __syncthreads();
        for (uint32_t i = tid; i < c; i += MT) atomicAdd(&run[(uint32_t)(src[(uint64_t)start + i] >> shift) & 0xFF], 1u);
        __syncthreads();
        const uint32_t h = run[tid];
        uint32_t total;
        const uint32_t ex = block_exclusive_scan(h, scan_tmp, &total);
        run[tid] = start + ex;
        __syncthreads();
        for (uint32_t c0 = 0; c0 < c; c0 += MTILE) {
            for (uint32_t i = tid; i < MW * 256; i += MT) (&cnt[0][0])[i] = 0;
            __syncthreads();
            const uint32_t wb = c0 + wid * MWI;
            uint64_t p[MIPT];
            uint32_t d[MIPT], r[MIPT];
#pragma unroll
            for (uint32_t j = 0; j < MIPT; j++) {
                const uint32_t idx = wb + j * WAVE + lane;
                const bool in = idx < c;
                p[j] = in ? src[(uint64_t)start + idx] : 0;
                d[j] = in ? ((uint32_t)(p[j] >> shift) & 0xFF) : 256u;
            }
#pragma unroll
            for (uint32_t j = 0; j < MIPT; j++) r[j] = wave_rank<8>(d[j], d[j] < 256, cnt[wid]);
            __syncthreads();
            {
                uint32_t rw = 0;
#pragma unroll
                for (uint32_t w = 0; w < MW; w++) {
                    const uint32_t cc = cnt[w][tid];
                    cnt[w][tid] = rw;
                    rw += cc;
                }
                tot[tid] = rw;
            }
            __syncthreads();
#pragma unroll
            for (uint32_t j = 0; j < MIPT; j++) {
                if (d[j] < 256) {
                    const uint64_t pos = (uint64_t)run[d[j]] + cnt[wid][d[j]] + r[j];
                    if (last) {
                        gkey k;
                        gval v;
                        unpack_key(p[j], pi, k, v);
                        ok[pos] = k;
                        ov[pos] = v;
                    } else {
                        dst[pos] = p[j];
                    }
                }
            }
            __syncthreads();
            run[tid] += tot[tid];
            __syncthreads();
        }
        uint64_t *t = src;
        src = dst;
        dst = t;
    }
}

}  // namespace

bool msd_sort_fits(uint32_t N) { return N > 0 && N < (1u << 28); }

MsdPlan msd_plan(uint64_t n, uint32_t N, uint32_t bucket_bits, bool by_row, uint32_t items_per_lane) {
    MsdPlan pl{};
    uint32_t bs = 0;
    while (bs < 32 && (N >> bs) != 0) bs++;  // bit width of N itself
    pl.bs = bs;
    const uint32_t B = 2 * bs;
    uint32_t D = 1;
    while (D < MAX_D && (n >> (D + 8)) != 0) D++;  // about 256 items per bucket
    if (bucket_bits) D = std::min(bucket_bits, MAX_D);
    if (D > B) D = B;
    pl.D = D;
    pl.by_row = by_row;
    pl.R = by_row ? B : B - D;  // row buckets do not fix the key's top bits
    pl.bshift = 8 + B - D;
    // items per lane: enough tiles (about 400) to occupy every CU
    pl.ipt = n >= 400ull * MT * 16 ? 16 : n >= 400ull * MT * 8 ? 8 : 4;
    if (items_per_lane == 4 || items_per_lane == 8 || items_per_lane == 16) pl.ipt = items_per_lane;
    pl.ntiles = (uint32_t)((n + MT * pl.ipt - 1) / (MT * pl.ipt));
    pl.NB = 1u << D;
    pl.nscan = (pl.NB + SCAN_B - 1) / SCAN_B;
    return pl;
}

void msd_splitters(const uint32_t *row_offset, uint32_t N, uint64_t C, uint32_t *spl, hipStream_t st) {
    int32_t cs = 0;
    while ((1ull << (cs + 1)) <= C) cs++;  // C is a power of two
    const uint64_t threads = ((uint64_t)N + 1 + 3) / 4;  // 4 rows each, rows 0..N
    hipLaunchKernelGGL(k_msd_splitters, dim3((uint32_t)((threads + 255) / 256)), dim3(256), 0, st, row_offset, N, cs,
                       spl);
}

void msd_sort(const gkey *ik, const gval *iv, gkey *ok, gval *ov, const MsdPlan &pl, const MsdWork &w, uint64_t n,
              const uint32_t *dn, uint32_t N, bool guards_ok, const uint32_t *spl, uint32_t *invalid, uint32_t epoch,
              uint32_t *ticket, uint32_t *err, uint32_t *maxb, hipStream_t st) {
    if (n == 0) return;
    const uint32_t *sp = pl.by_row ? spl : nullptr;
    const PackInfo pi{N, pl.bs, guards_ok ? 1u : 0u};
    uint16_t *counts = reinterpret_cast<uint16_t *>(w.mat);                        // [ntiles][NB] u16
    uint32_t *offs = w.mat + ((uint64_t)pl.ntiles * pl.NB + 1) / 2;             // [ntiles][NB] u32
#define GPMA_MSD_TILE(IPT)                                                                                    \
    hipLaunchKernelGGL(k_msd_rank<IPT>, dim3(pl.ntiles), dim3(MT), 0, st, ik, iv, n, dn, pi, pl.bshift, pl.D, sp,     \
                       w.kp, w.rk, counts, invalid);                                                              \
    hipLaunchKernelGGL(k_msd_scan, dim3(pl.nscan), dim3(SCAN_T), 0, st, counts, pl.ntiles, MT * IPT, pl.NB, n, dn, \
                       offs, w.bstart, w.status, epoch, ticket, err);                                              \
    hipLaunchKernelGGL(k_msd_scatter<IPT>, dim3(pl.ntiles), dim3(MT), 0, st, w.kp, w.rk, n, dn, pl.NB, offs, w.kb)
    switch (pl.ipt) {
    case 4: GPMA_MSD_TILE(4); break;
    case 8: GPMA_MSD_TILE(8); break;
    default: GPMA_MSD_TILE(16); break;
    }
#undef GPMA_MSD_TILE
    hipLaunchKernelGGL(k_msd_local, dim3(pl.NB), dim3(MT), 0, st, w.kb, w.kp, w.bstart, pl.R, pi, ok, ov, maxb);
}

}  // namespace gpma
