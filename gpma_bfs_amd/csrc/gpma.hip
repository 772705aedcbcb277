// C ABI and host orchestration of the gfx950 GPMA.
//
// Device state (structure of arrays in HBM, one allocation each):
//   keys[C] u64, values[C] u8       the gapped packed array (C = S << h)
//   seps[C/S] u64                   first key of every segment (== keys[j*S]),
//                                   the search array of locate
//   row_offset[N+1] u32             row_offset[v+1] = slot of guard v
// The host keeps S, h and the per-level bounds (gpma.cuh:37-46).
//
// update_gpma (gpma.cuh:968-1050) becomes: stage + radix sort (sort.hip),
// locate, compaction of the pending inserts, one leaf-level kernel, then
// per higher level either one LDS window kernel or the device-wide window
// passes, a compaction, and one host read of the pending count; a resize or
// a bulk load rebuilds the array with the same device-wide passes. The
// common case (every insert placed at level 0 or 1) reads two counters back
// per batch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/gpma.h"
#include "common.h"
#include "kernels.h"

using namespace gpma;

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

#define HIP_TRY(expr)                                                                                     \
    do {                                                                                                  \
        hipError_t e_ = (expr);                                                                           \
        if (e_ != hipSuccess)                                                                             \
            return fail(e_ == hipErrorOutOfMemory ? GPMA_ERR_OOM : GPMA_ERR_HIP,                          \
                        std::string(#expr) + ": " + hipGetErrorString(e_));                               \
    } while (0)

#define TRY(expr)                 \
    do {                          \
        int r_ = (expr);          \
        if (r_ != GPMA_OK) return r_; \
    } while (0)

// Entry points on a handle whose array was lost to a failed allocation.
#define CHECK_ALIVE(g, what)                                                                              \
    do {                                                                                                  \
        if ((g)->failed)                                                                                  \
            return fail(GPMA_ERR_OOM, std::string(what) + ": the handle lost its array to a failed allocation"); \
    } while (0)

// gpma.cuh:104-126
uint32_t fls32(uint32_t x) { return x ? 32u - (uint32_t)__builtin_clz(x) : 0u; }

// Device memory held by one handle, and the allocation-failure injection of
// gpma_test_fail_allocs (after `skip` more successful allocations, the next
// `fail` ones report out of memory).
struct Alloc {
    uint64_t bytes = 0;
    uint32_t skip = 0, fail = 0;
    bool injected() {
        if (!fail) return false;
        if (skip) {
            skip--;
            return false;
        }
        fail--;
        return true;
    }
};

// Device buffer that only grows.
struct Buf {
    void *p = nullptr;
    size_t bytes = 0;
    int ensure(size_t need, Alloc *a) {
        if (need <= bytes && p) return GPMA_OK;
        size_t nb = std::max(need, bytes + bytes / 2);
        if (nb < 256) nb = 256;
        if (p) {
            (void)hipFree(p);
            a->bytes -= bytes;
            p = nullptr;
            bytes = 0;
        }
        if (a->injected()) return fail(GPMA_ERR_OOM, "hipMalloc(" + std::to_string(nb) + "): injected failure");
        hipError_t e = hipMalloc(&p, nb);
        if (e != hipSuccess) {
            p = nullptr;
            return fail(GPMA_ERR_OOM, "hipMalloc(" + std::to_string(nb) + "): " + hipGetErrorString(e));
        }
        bytes = nb;
        a->bytes += nb;
        return GPMA_OK;
    }
    void release(Alloc *a) {
        if (p) {
            (void)hipFree(p);
            a->bytes -= bytes;
        }
        p = nullptr;
        bytes = 0;
    }
    void swap(Buf &o) {
        std::swap(p, o.p);
        std::swap(bytes, o.bytes);
    }
    template <typename T>
    T *as() const {
        return (T *)p;
    }
};

constexpr uint32_t BFS_MAX_CHUNK = 64;

enum Ctr : uint32_t {
    C_INVALID = 0,
    C_NP0 = 1,
    C_NP1 = 2,
    C_SORTERR = 3,
    C_LIVE = 4,
    C_NEXT = 5,
    C_NRUNS = 6,
    C_GATE = 7,         // a captured step's update left work for the host
    C_MSD_TICKET = 8,   // scan-block tickets of the bucket sort
    C_MSD_MAXB = 9,     // largest bucket above 1024 items in the bucket sort
    C_TICKET = 16,  // C_TICKET + k: tile tickets of the k-th compaction of one update
    C_LVL = 64,     // C_LVL + L: pending inserts entering level L of the device-driven prefix
    C_COUNT = 96
};
constexpr uint32_t N_TICKETS = C_LVL - C_TICKET;
constexpr uint32_t C_LVL_SLOTS = C_COUNT - C_LVL;
// Bucket sort for batches up to this size (its count and offset matrices
// take 6 to 24 bytes per record); larger batches and row counts >= 2^28 use
// the LSD passes.
constexpr uint64_t MSD_MAX_N = 4ull << 20;
constexpr uint32_t MSD_BIG_BUCKET = 16384;
constexpr uint32_t MSD_COOL_BATCHES = 64;
// Updates (or captured steps) between refreshes of the row splitters of the
// bucket sort. Between refreshes the guards move only inside the windows the
// updates rebalance, so a stale table shifts a bucket boundary by little; any
// table sorts correctly, a stale one only balances the buckets less well.
constexpr uint32_t SPL_EVERY = 8;
// Levels 1..spec_levels (default SPEC_LEVELS) run without a host round trip;
// the host reads the counters once after them and drives any higher level
// itself. Each device level costs a launch on every batch (one that no insert
// reaches exits at once), each host level a compaction and a round trip on
// the batches that reach it; over the 100 slides of the power-law 1.63M-node
// protocol (rebalance depth grows from 1-2 to 4-5) four device levels give
// the best protocol mean and tail (three: 3% faster middle slides, 2% slower
// protocol, p99 0.20 instead of 0.17 ms). device_levels tunes it.
constexpr uint32_t SPEC_LEVELS = 4;

}  // namespace

namespace gpma {
// Sets gpma_last_error() for this thread (host code outside this file).
void set_last_error(const std::string &msg) { g_last_error = msg; }
}  // namespace gpma

struct gpma_handle {
    int device = 0;
    hipStream_t st = nullptr;
    uint32_t N = 0;
    uint64_t C = 0;
    uint32_t S = 0, h = 0;
    std::vector<uint32_t> lower, upper;
    uint32_t bulk_threshold = 5u * 1000u * 1000u;
    uint32_t seg_override = 0;  // segment length after a resize (0: the reference formula)
    uint32_t spec_levels = SPEC_LEVELS;
    uint32_t sort_bits = 0;     // bucket bits of the batch sort (0: from the batch size)
    uint32_t sort_ipt = 0;      // items per lane of the sort's tiles (0: from the batch size)
    uint32_t push_tile = 64;    // most frontier vertices per BFS push tile
    uint32_t ms_tile = 64;      // the same for the multi-source BFS
    // density thresholds (gpma.cuh:40-43): lower / upper at the leaves and at the root
    double lower_leaf = 0.08, lower_root = 0.42, upper_leaf = 0.92, upper_root = 0.84;
    Alloc mem;          // device bytes held, allocation-failure injection
    bool failed = false;  // the array was lost to an allocation failure: every call returns GPMA_ERR_OOM
    // structure
    Buf keys, values, seps, row_offset;
    // update scratch
    Buf in_k, in_v, sk0, sv0, sk1, sv1, win, blkcnt, blkoff, sortwork;
    Buf pk0, pv0, pw0, pd0, pk1, pv1, pw1, pd1;
    Buf tmpk, tmpv, chunkcnt, chunkoff;
    Buf wide;                // per-window scratch of the wide rebalance levels
    Buf ccstat;              // look-back status words of the chained compaction
    Buf msd_kp, msd_rk, msd_mat, msd_bstart, msd_status;  // bucket sort scratch
    Buf msd_spl;             // row splitters of the bucket sort (SPL_TABLE u32)
    uint32_t spl_age = SPL_EVERY;  // updates since the splitters were computed (SPL_EVERY: refresh first)
    uint32_t msd_cool = 0;   // batches left on the LSD sort after a skewed batch
    uint32_t cc_epoch = 0;   // tag of the latest compaction call (never 0)
    uint32_t cc_ticket = 0;  // next ticket slot of this update
    // bfs scratch
    Buf lev, fr0, fr1, fb0, fb1, bcnt, bcntb;
    // multi-source bfs scratch
    Buf ms_seen, ms_fr0, ms_fr1, ms_nx, ms_lv16, ms_q0, ms_q1, ms_qb0, ms_qb1, ms_touched, ms_cnt, ms_cntb, ms_tcnt,
        ms_partial, ms_starts, ms_out, ms_reach;
    Buf ms_lv32;          // u32 level staging, allocated when a traversal reaches level 65,535
    Buf ms_ev, ms_evn;    // level events per vertex, and their counts
    bool ms_wide = false; // the running multi-source traversal stages u32 levels
    unsigned long long *ms_hcnt = nullptr;
    unsigned long long *hcnt = nullptr;  // pinned mirror of one chunk of bcnt and bcntb
    uint32_t bfs_depth = 0, ms_depth = 0;  // depth of the last traversal (sizes the next first chunk)
    double pull_fraction = 0.25;  // pull when the frontier's slots exceed this share of C
    Buf ctr;
    Buf digest;                // 3 u64: live-set digest
    uint32_t *hctr = nullptr;  // pinned mirror of ctr
    bool ctr_clean = false;    // ctr is all zero on the device (the last read cleared it)
    // completion signal (host-coherent pinned): [0] sequence, [1..] payload
    uint32_t *sig = nullptr;
    uint32_t sig_seq = 0;
    // stats
    uint64_t applied = 0, invalid = 0, dropped = 0, resizes = 0, bulk = 0;
    uint32_t last_levels = 0;
    // phase timing (HIP events on st): events come from a pool; every
    // begin/end pair is kept until the next stream synchronisation
    bool prof = false;
    uint32_t prof_mask = (1u << GPMA_PHASES) - 1;
    struct EvPair {
        int phase;
        hipEvent_t a, b;
    };
    std::vector<hipEvent_t> ev_pool, ev_all;
    std::vector<EvPair> ev_done;
    hipEvent_t ev_open[GPMA_PHASES] = {};
    double ph_ms[GPMA_PHASES] = {};
    uint64_t ph_calls[GPMA_PHASES] = {};
    // captured steps (gpma_step): one executable graph per (size class,
    // sources, output) and layout; sig lists every pointer and size the
    // graph's launches bake in
    struct StepGraph {
        std::vector<uint64_t> sig;
        hipGraphExec_t exec = nullptr;
        uint64_t used = 0;
        uint32_t spec = 0;  // device-driven update levels it holds
    };
    std::vector<StepGraph> graphs;
    bool capturing = false;
    uint32_t step_blevels = 8;  // BFS levels a new capture holds (grows to the depth seen)
    uint32_t step_ulevels = 4;  // rebalance levels a new capture holds (grows when a replay needed more)
    uint64_t step_clock = 0;
    Buf step_n;                    // device u32: records of the step being replayed
    Buf step_stage;                // device: BFS counts and reachable counts gathered for the read-back
    uint32_t step_rb[4 + 64] = {};  // a replay's read-back: frontier counts (2 x u64), reachable counts
    gpma_step_stats sst{};
};

namespace {

// gpma.cuh:147-168, with the handle's thresholds (the reference's
// 0.08 / 0.42 / 0.92 / 0.84 unless configured)
void recalc_density(gpma_handle *g) {
    const double lower_leaf = g->lower_leaf, lower_root = g->lower_root, upper_leaf = g->upper_leaf,
                 upper_root = g->upper_root;
    g->lower.assign(g->h + 1, 0);
    g->upper.assign(g->h + 1, 0);
    uint64_t len = g->S;
    for (uint32_t i = 0; i <= g->h; i++) {
        double dl = lower_root + (lower_leaf - lower_root) * (g->h - i) / g->h;
        double du = upper_root + (upper_leaf - upper_root) * (g->h - i) / g->h;
        g->lower[i] = (uint32_t)std::ceil(dl * (double)len);
        g->upper[i] = (uint32_t)std::floor(du * (double)len);
        if (i > 0) {
            g->lower[i] = std::max(g->lower[i], 2 * g->lower[i - 1]);
            g->upper[i] = std::min(g->upper[i], 2 * g->upper[i - 1]);
        }
        len <<= 1;
    }
}

enum Phase : int { P_STAGE = 0, P_SORT, P_LOCATE, P_COMPACT, P_LEAF, P_UPPER, P_REBUILD, P_BFS };

// Events skip the system-scope fence: they only time work on the stream.
hipEvent_t pool_event(gpma_handle *g) {
    if (g->ev_pool.empty()) {
        hipEvent_t e = nullptr;
        if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) return nullptr;
        g->ev_all.push_back(e);
        return e;
    }
    hipEvent_t e = g->ev_pool.back();
    g->ev_pool.pop_back();
    return e;
}
bool phase_on(const gpma_handle *g, int p) { return g->prof && !g->capturing && ((g->prof_mask >> p) & 1u); }
void phase_begin(gpma_handle *g, int p) {
    if (!phase_on(g, p)) return;
    if (g->ev_open[p]) g->ev_pool.push_back(g->ev_open[p]);
    hipEvent_t e = pool_event(g);
    if (e && hipEventRecord(e, g->st) == hipSuccess) g->ev_open[p] = e;
    else {
        if (e) g->ev_pool.push_back(e);
        g->ev_open[p] = nullptr;
    }
}
void phase_end(gpma_handle *g, int p, uint64_t launches) {
    if (!phase_on(g, p) || !g->ev_open[p]) return;
    hipEvent_t e = pool_event(g);
    if (!e || hipEventRecord(e, g->st) != hipSuccess) {
        if (e) g->ev_pool.push_back(e);
        return;
    }
    g->ev_done.push_back({p, g->ev_open[p], e});
    g->ev_open[p] = nullptr;
    g->ph_calls[p] += launches;
}
// After a stream synchronisation: fold finished event pairs into the totals.
// Also called right after a completion signal, when the events have run on
// the GPU but the runtime may not have seen them yet: then wait for them.
void phase_collect(gpma_handle *g) {
    for (const auto &r : g->ev_done) {
        float ms = 0;
        hipError_t e = hipEventElapsedTime(&ms, r.a, r.b);
        if (e == hipErrorNotReady && hipEventSynchronize(r.b) == hipSuccess) e = hipEventElapsedTime(&ms, r.a, r.b);
        if (e == hipSuccess) g->ph_ms[r.phase] += ms;
        g->ev_pool.push_back(r.a);
        g->ev_pool.push_back(r.b);
    }
    g->ev_done.clear();
}

int stream_sync(gpma_handle *g) {
    HIP_TRY(hipStreamSynchronize(g->st));
    phase_collect(g);
    return GPMA_OK;
}

constexpr uint32_t SIG_WORDS = 1024;

// Spin until the signal holds seq. The stream is polled now and then, so an
// error, or a drained stream whose signal never arrived, ends the wait.
int wait_signal(gpma_handle *g, uint32_t seq) {
    const volatile uint32_t *f = g->sig;
    uint32_t drained = 0;  // spins since the stream was seen idle
    for (uint32_t spin = 1; f[0] != seq; spin++) {
        if (drained) {
            if (++drained > (1u << 24))
                return fail(GPMA_ERR_HIP, "completion signal missing after the stream drained");
        } else if ((spin & 1023) == 0) {
            const hipError_t e = hipStreamQuery(g->st);
            if (e == hipSuccess) drained = 1;
            else if (e != hipErrorNotReady)
                return fail(GPMA_ERR_HIP, std::string("hipStreamQuery: ") + hipGetErrorString(e));
        }
    }
    std::atomic_thread_fence(std::memory_order_acquire);
    return GPMA_OK;
}

// out[0 .. na + nb) = a[0 .. na) ++ b[0 .. nb) (device u32 words, na + nb <
// SIG_WORDS), read once everything queued before on the stream is done: a
// one-wave kernel writes them to host-coherent memory and raises a sequence
// number the host spins on, which takes less time than a stream
// synchronisation after a device-to-host copy.
// A captured step's own publish writes this sequence word; the host clears
// the word before each replay.
constexpr uint32_t SIG_STEP = 0xFFFFFFFFu;

int publish_and_wait(gpma_handle *g, const uint32_t *a, uint32_t na, const uint32_t *b, uint32_t nb, uint32_t *out,
                     bool clear_a = false) {
    if (++g->sig_seq == 0 || g->sig_seq == SIG_STEP) g->sig_seq = 1;
    launch_publish(a, na, b, nb, g->sig, g->sig_seq, clear_a, g->st);
    HIP_TRY(hipGetLastError());
    TRY(wait_signal(g, g->sig_seq));
    std::memcpy(out, g->sig + 1, (size_t)(na + nb) * 4);
    return GPMA_OK;
}

// Counters into hctr once the queued work is done, leaving them zero on the
// device (the next update then needs no clearing launch).
int read_ctr(gpma_handle *g) {
    TRY(publish_and_wait(g, g->ctr.as<uint32_t>(), C_COUNT, nullptr, 0, g->hctr, true));
    g->ctr_clean = true;
    if (g->prof) phase_collect(g);
    return GPMA_OK;
}

int ensure_tmp(gpma_handle *g, uint64_t n) {
    TRY(g->tmpk.ensure(n * 8, &g->mem));
    TRY(g->tmpv.ensure(n, &g->mem));
    uint64_t nch = (n + CHUNK - 1) / CHUNK + 1;
    TRY(g->chunkcnt.ensure(nch * 4, &g->mem));
    TRY(g->chunkoff.ensure(nch * 4, &g->mem));
    return GPMA_OK;
}

// S, h and the level bounds for a T-slot array: gpma.cuh:917-920 (S =
// 2^(fls(fls T) - 1), h = fls(T / S) - 1), or the configured segment length
// (capped at T / 2).
void size_tree(gpma_handle *g, uint64_t T) {
    g->S = 1u << (fls32(fls32((uint32_t)T)) - 1);
    if (g->seg_override) g->S = (uint32_t)std::min<uint64_t>(g->seg_override, T / 2);
    g->h = fls32((uint32_t)(T / g->S)) - 1;
    recalc_density(g);
}

// Live count of window [w, w+W) (one host sync), leaving chunk offsets for
// a following compact_live.
int count_live(gpma_handle *g, uint64_t w, uint64_t W, uint64_t *c) {
    TRY(ensure_tmp(g, std::max<uint64_t>(W, 1)));
    uint64_t nch = (W + CHUNK - 1) / CHUNK;
    launch_count_live(g->keys.as<gkey>(), g->values.as<gval>(), w, W, g->chunkcnt.as<uint32_t>(), g->st);
    launch_scan_blocks(g->chunkcnt.as<uint32_t>(), nch, g->chunkoff.as<uint32_t>(), g->ctr.as<uint32_t>() + C_LIVE,
                       g->st);
    g->ctr_clean = false;
    TRY(read_ctr(g));
    if (g->hctr[C_SORTERR]) return fail(GPMA_ERR_HIP, "compaction look-back exceeded its spin limit");
    *c = g->hctr[C_LIVE];
    return GPMA_OK;
}

// Array shape: capacity, segment length, height and level bounds.
struct Shape {
    uint64_t C;
    uint32_t S, h;
    std::vector<uint32_t> lower, upper;
};
Shape shape_of(const gpma_handle *g) { return {g->C, g->S, g->h, g->lower, g->upper}; }
void set_shape(gpma_handle *g, const Shape &s) {
    g->C = s.C;
    g->S = s.S;
    g->h = s.h;
    g->lower = s.lower;
    g->upper = s.upper;
}

// New keys / values / seps arrays of T slots (segment length g->S) for a
// rebuild whose c live items wait in tmpk / tmpv. The new arrays are
// allocated while the old ones still exist; if that fails, the old ones go
// first (their items are in tmp) and the allocation is tried again. If that
// fails too, the live items are put back into arrays of the old shape
// (pending inserts are not applied) and GPMA_ERR_OOM is returned; if even
// that fails, the handle is marked failed and every later call returns
// GPMA_ERR_OOM without touching the device.
int replace_arrays(gpma_handle *g, uint64_t T, uint64_t c, const Shape &old) {
    auto alloc3 = [&](Buf &k, Buf &v, Buf &sp, uint64_t slots, uint32_t S) {
        int rc = k.ensure(slots * 8, &g->mem);
        if (rc == GPMA_OK) rc = v.ensure(slots, &g->mem);
        if (rc == GPMA_OK) rc = sp.ensure((slots / S) * 8, &g->mem);
        if (rc != GPMA_OK) {
            k.release(&g->mem);
            v.release(&g->mem);
            sp.release(&g->mem);
        }
        return rc;
    };
    Buf nk, nv, ns;
    int rc = alloc3(nk, nv, ns, T, g->S);
    g->keys.release(&g->mem);
    g->values.release(&g->mem);
    g->seps.release(&g->mem);
    if (rc != GPMA_OK) rc = alloc3(nk, nv, ns, T, g->S);
    if (rc == GPMA_OK) {
        g->keys.swap(nk);
        g->values.swap(nv);
        g->seps.swap(ns);
        return GPMA_OK;
    }
    const std::string why = g_last_error;
    set_shape(g, old);
    if (alloc3(g->keys, g->values, g->seps, old.C, old.S) != GPMA_OK) {
        g->failed = true;
        return fail(GPMA_ERR_OOM, "rebuild to " + std::to_string(T) + " slots failed (" + why +
                                      ") and the array could not be restored; the handle is unusable");
    }
    launch_merge_redistribute(g->tmpk.as<gkey>(), g->tmpv.as<gval>(), c, nullptr, nullptr, 0, g->keys.as<gkey>(),
                              g->values.as<gval>(), g->seps.as<gkey>(), g->row_offset.as<uint32_t>(), 0, old.C, old.S,
                              g->st);
    (void)stream_sync(g);
    return fail(GPMA_ERR_OOM, "rebuild to " + std::to_string(T) + " slots failed (" + why +
                                  "); the live items were kept at the old capacity and the pending inserts dropped");
}

// resize_gpma + root rebalance (gpma.cuh:906-934,1034-1044) when
// check_bounds, or significant_insert (gpma.cuh:936-966) when not: compact
// every live item, size the array to the smallest power of two >= 4 with
// floor(upper_root * C') >= live + pending (0.84 in the reference), and
// redistribute the merge of the live
// items and the pending inserts over the whole new array.
int rebuild(gpma_handle *g, const gkey *pk, const gval *pv, uint64_t np, bool check_bounds) {
    const Shape old = shape_of(g);
    uint64_t c = 0;
    phase_begin(g, P_REBUILD);
    TRY(count_live(g, 0, g->C, &c));
    phase_begin(g, P_REBUILD);
    TRY(ensure_tmp(g, g->C));
    launch_compact_live(g->keys.as<gkey>(), g->values.as<gval>(), 0, g->C, g->chunkoff.as<uint32_t>(),
                        g->tmpk.as<gkey>(), g->tmpv.as<gval>(), g->st);
    uint64_t merge = c + np;
    uint64_t T = 4;
    while (std::floor(g->upper_root * (double)T) < (double)merge) T *= 2;
    if (T > (1ull << 31)) return fail(GPMA_ERR_CAPACITY, "packed array would exceed 2^31 slots");
    size_tree(g, T);
    if (check_bounds) {
        // The reference checks nothing here and loses the pending inserts when
        // the trimmed root bound is below live + pending (gpma.cuh:1034-1044);
        // grow until they fit instead, and place them even below the lower bound.
        while (merge > g->upper[g->h]) {
            T *= 2;
            if (T > (1ull << 31)) return fail(GPMA_ERR_CAPACITY, "packed array would exceed 2^31 slots");
            size_tree(g, T);
        }
    }
    if (T != g->C) {
        phase_end(g, P_REBUILD, 1);
        TRY(stream_sync(g));
        phase_begin(g, P_REBUILD);
        TRY(replace_arrays(g, T, c, old));
    }
    g->C = T;
    g->spl_age = SPL_EVERY;  // the rows moved: new splitters before the next sort
    if (check_bounds) g->resizes++;
    else g->bulk++;
    launch_merge_redistribute(g->tmpk.as<gkey>(), g->tmpv.as<gval>(), c, pk, pv, np, g->keys.as<gkey>(),
                              g->values.as<gval>(), g->seps.as<gkey>(), g->row_offset.as<uint32_t>(), 0, T, g->S,
                              g->st);
    phase_end(g, P_REBUILD, 2);
    return GPMA_OK;
}

int init_structure(gpma_handle *g) {
    // gpma.cuh:54-69: keys [MAX, NONE, MAX, NONE], values [1, 0, 1, 0], S = 2, h = 1
    g->C = 4;
    g->S = 2;
    g->h = 1;
    recalc_density(g);
    TRY(g->keys.ensure(4 * 8, &g->mem));
    TRY(g->values.ensure(4, &g->mem));
    TRY(g->seps.ensure(2 * 8, &g->mem));
    TRY(g->row_offset.ensure(((uint64_t)g->N + 1) * 4, &g->mem));
    gkey hk[4] = {KEY_MAX, KEY_NONE, KEY_MAX, KEY_NONE};
    gval hv[4] = {1, 0, 1, 0};
    gkey hs[2] = {KEY_MAX, KEY_MAX};
    HIP_TRY(hipMemcpyAsync(g->keys.p, hk, sizeof(hk), hipMemcpyHostToDevice, g->st));
    HIP_TRY(hipMemcpyAsync(g->values.p, hv, sizeof(hv), hipMemcpyHostToDevice, g->st));
    HIP_TRY(hipMemcpyAsync(g->seps.p, hs, sizeof(hs), hipMemcpyHostToDevice, g->st));
    fill_u32(g->row_offset.as<uint32_t>(), 0, (uint64_t)g->N + 1, g->st);
    HIP_TRY(hipStreamSynchronize(g->st));
    return GPMA_OK;
}

int ensure_update_buffers(gpma_handle *g, uint64_t n) {
    Alloc *t = &g->mem;
    TRY(g->in_k.ensure(n * 8, t));
    TRY(g->in_v.ensure(n, t));
    TRY(g->sk0.ensure(n * 8, t));
    TRY(g->sv0.ensure(n, t));
    TRY(g->sk1.ensure(n * 8, t));
    TRY(g->sv1.ensure(n, t));
    TRY(g->win.ensure(n * 4, t));
    uint64_t nb = (n + CMP_BLOCK - 1) / CMP_BLOCK + 1;
    TRY(g->blkcnt.ensure(nb * 4, t));
    TRY(g->blkoff.ensure(nb * 4, t));
    TRY(g->sortwork.ensure((8 * 256 + 8 + 8 + sort_status_words(n, 8)) * 4, t));
    TRY(g->pk0.ensure(n * 8, t));
    TRY(g->pv0.ensure(n, t));
    TRY(g->pw0.ensure(n * 4, t));
    TRY(g->pd0.ensure(n, t));
    TRY(g->pk1.ensure(n * 8, t));
    TRY(g->pv1.ensure(n, t));
    TRY(g->pw1.ensure(n * 4, t));
    TRY(g->pd1.ensure(n, t));
    void *before = g->ccstat.p;
    const uint64_t nstat = compact_status_words(n);
    TRY(g->ccstat.ensure(nstat * 8, t));
    if (g->ccstat.p != before) fill_u64(g->ccstat.as<uint64_t>(), 0, g->ccstat.bytes / 8, g->st);
    return GPMA_OK;
}

// Next look-back tag; on wrap-around every status word is cleared so no
// stale tag can match.
void drop_graphs(gpma_handle *g);

uint32_t next_epoch(gpma_handle *g) {
    if (++g->cc_epoch == 0) {
        drop_graphs(g);  // their tags could now repeat
        fill_u64(g->ccstat.as<uint64_t>(), 0, g->ccstat.bytes / 8, g->st);
        if (g->msd_status.p) fill_u64(g->msd_status.as<uint64_t>(), 0, g->msd_status.bytes / 8, g->st);
        g->cc_epoch = 1;
    }
    return g->cc_epoch;
}

// Stable compaction through the chained single-pass kernel; see
// launch_compact_chained for the arguments.
void compact_step(gpma_handle *g, const gkey *ik, const gval *iv, const uint32_t *iw, const uint8_t *idone,
                  uint64_t n_ub, const uint32_t *d_n, gkey *ok, gval *ov, uint32_t *ow, uint8_t *odone, uint32_t mask,
                  uint32_t *d_out, const gkey *fix = nullptr, gkey *seps = nullptr) {
    uint32_t *ctr = g->ctr.as<uint32_t>();
    if (g->cc_ticket >= N_TICKETS) {
        fill_u32(ctr + C_TICKET, 0, N_TICKETS, g->st);
        g->cc_ticket = 0;
    }
    const uint32_t epoch = next_epoch(g);
    launch_compact_chained(ik, iv, iw, idone, n_ub, d_n, ok, ov, ow, odone, mask,
                           g->ccstat.as<unsigned long long>(), epoch, ctr + C_TICKET + g->cc_ticket, d_out,
                           ctr + C_SORTERR, g->st, fix, seps);
    g->cc_ticket++;
}

int ensure_msd(gpma_handle *g, uint64_t n, const MsdPlan &pl) {
    Alloc *t = &g->mem;
    TRY(g->msd_kp.ensure(n * 8, t));
    TRY(g->msd_rk.ensure(n * 4, t));
    TRY(g->msd_mat.ensure(2ull * pl.ntiles * pl.NB * 4, t));
    TRY(g->msd_bstart.ensure(((uint64_t)pl.NB + 1) * 4, t));
    TRY(g->msd_spl.ensure((uint64_t)SPL_TABLE * 4, t));
    void *before = g->msd_status.p;
    TRY(g->msd_status.ensure((uint64_t)pl.nscan * 8, t));
    if (g->msd_status.p != before) fill_u64(g->msd_status.as<uint64_t>(), 0, g->msd_status.bytes / 8, g->st);
    return GPMA_OK;
}

// Device part of one update: sort, locate, the level-0 compaction, the leaf
// pass and levels 1..dev_levels, each reading its own count on the device,
// so nothing here waits for the host. The record count is n_ub, or *d_n
// (<= n_ub) when d_n is given: launch shapes then cover n_ub and the kernels
// past the batch exit at once, which lets a captured graph serve batches of
// any size up to n_ub. pl: bucket sort plan for n_ub (nullptr: LSD passes).
// Returns the number of device-driven levels above the leaves.
uint32_t enqueue_update(gpma_handle *g, const gkey *src_k, const gval *src_v, uint64_t n_ub, const uint32_t *d_n,
                        bool guards_ok, const MsdPlan *pl, uint32_t dev_levels) {
    hipStream_t st = g->st;
    uint32_t *ctr = g->ctr.as<uint32_t>();
    phase_begin(g, P_SORT);
    int cur = 0;
    if (pl) {
        MsdWork w{g->msd_kp.as<uint64_t>(), g->msd_rk.as<uint32_t>(), g->sk1.as<uint64_t>(), g->msd_mat.as<uint32_t>(),
                  g->msd_bstart.as<uint32_t>(), g->msd_status.as<unsigned long long>()};
        // the row splitters: refreshed every SPL_EVERY updates and after a
        // rebuild (a captured step's graph holds none: run_step refreshes
        // them before a replay on the same schedule)
        const bool refresh = pl->by_row && !g->capturing && g->spl_age >= SPL_EVERY;
        if (refresh) {
            msd_splitters(g->row_offset.as<uint32_t>(), g->N, g->C, g->msd_spl.as<uint32_t>(), st);
            g->spl_age = 0;
        }
        if (pl->by_row && !g->capturing) g->spl_age++;
        msd_sort(src_k, src_v, g->sk0.as<gkey>(), g->sv0.as<gval>(), *pl, w, n_ub, d_n, g->N, guards_ok,
                 g->msd_spl.as<uint32_t>(), ctr + C_INVALID, next_epoch(g), ctr + C_MSD_TICKET, ctr + C_SORTERR,
                 ctr + C_MSD_MAXB, st);
        phase_end(g, P_SORT, refresh ? 5 : 4);
    } else {
        // digits: bits of N-1 plus one (so KEY_NONE sorts after every valid key)
        const uint32_t nb = std::max<uint32_t>(1, fls32(g->N ? g->N - 1 : 0)) + 1;
        const uint32_t dpw = std::min<uint32_t>(4, (nb + 7) / 8);
        SortShifts sh{};
        uint32_t npass = 0;
        for (uint32_t d = 0; d < dpw; d++) sh.s[npass++] = 8 * d;
        for (uint32_t d = 0; d < dpw; d++) sh.s[npass++] = 32 + 8 * d;
        cur = stage_and_sort(src_k, src_v, g->sk0.as<gkey>(), g->sv0.as<gval>(), g->sk1.as<gkey>(), g->sv1.as<gval>(),
                             n_ub, g->N, guards_ok, ctr + C_INVALID, sh, npass, g->sortwork.as<uint32_t>(),
                             ctr + C_SORTERR, st);
        phase_end(g, P_SORT, npass + 2);
    }
    const gkey *ks = cur ? g->sk1.as<gkey>() : g->sk0.as<gkey>();
    const gval *vs = cur ? g->sv1.as<gval>() : g->sv0.as<gval>();
    Buf *PK[2] = {&g->pk0, &g->pk1}, *PV[2] = {&g->pv0, &g->pv1}, *PW[2] = {&g->pw0, &g->pw1},
        *PD[2] = {&g->pd0, &g->pd1};
    // Drop placed inserts of list `pc` (count in *d_in), move the rest up to
    // level+1 into the other list, count in *d_out.
    auto compact_pending = [&](int pc, uint32_t level, const uint32_t *d_in, uint32_t *d_out) {
        const uint32_t mask = ~(uint32_t)(((uint64_t)g->S << (level + 1)) - 1);
        phase_begin(g, P_COMPACT);
        compact_step(g, PK[pc]->as<gkey>(), PV[pc]->as<gval>(), PW[pc]->as<uint32_t>(), PD[pc]->as<uint8_t>(), n_ub,
                     d_in, PK[pc ^ 1]->as<gkey>(), PV[pc ^ 1]->as<gval>(), PW[pc ^ 1]->as<uint32_t>(),
                     PD[pc ^ 1]->as<uint8_t>(), mask, d_out);
        phase_end(g, P_COMPACT, 1);
    };
    if (!guards_ok && (g->S == 16 || g->S == 32)) {
        // locate and level 0 in one pass, then the pending inserts of the
        // leaves that could not take them straight into the level-1 list
        // (the compaction also publishes the merged leaves' separators)
        phase_begin(g, P_LOCATE);
        launch_apply(g->S, ks, vs, n_ub, d_n, g->keys.as<gkey>(), g->values.as<gval>(), g->seps.as<gkey>(),
                     g->C / g->S, g->row_offset.as<uint32_t>(), g->lower[0], g->upper[0], g->win.as<uint32_t>(),
                     g->pk0.as<gkey>(), ctr + C_LVL, st);
        phase_end(g, P_LOCATE, 1);
        phase_begin(g, P_COMPACT);
        compact_step(g, ks, vs, g->win.as<uint32_t>(), nullptr, n_ub, d_n, g->pk1.as<gkey>(), g->pv1.as<gval>(),
                     g->pw1.as<uint32_t>(), g->pd1.as<uint8_t>(), ~(uint32_t)(((uint64_t)g->S << 1) - 1),
                     ctr + C_LVL + 1, g->pk0.as<gkey>(), g->seps.as<gkey>());
        phase_end(g, P_COMPACT, 1);
    } else {
        // locate + in-place delete / modify
        phase_begin(g, P_LOCATE);
        launch_locate(ks, vs, n_ub, d_n, g->keys.as<gkey>(), g->values.as<gval>(), g->seps.as<gkey>(), g->C / g->S,
                      g->S, g->row_offset.as<uint32_t>(), !guards_ok, g->win.as<uint32_t>(), st);
        phase_end(g, P_LOCATE, 1);
        phase_begin(g, P_COMPACT);
        compact_step(g, ks, vs, g->win.as<uint32_t>(), nullptr, n_ub, d_n, g->pk0.as<gkey>(), g->pv0.as<gval>(),
                     g->pw0.as<uint32_t>(), g->pd0.as<uint8_t>(), 0xFFFFFFFFu, ctr + C_LVL);
        phase_end(g, P_COMPACT, 1);
        // level 0; it and every level above stand down if the batch turns
        // out to be a bulk load
        phase_begin(g, P_LEAF);
        launch_leaf(g->S, g->pk0.as<gkey>(), g->pv0.as<gval>(), g->pw0.as<uint32_t>(), g->pd0.as<uint8_t>(), n_ub,
                    ctr + C_LVL, g->keys.as<gkey>(), g->values.as<gval>(), g->seps.as<gkey>(),
                    g->row_offset.as<uint32_t>(), g->lower[0], g->upper[0], g->bulk_threshold, st);
        phase_end(g, P_LEAF, 1);
        compact_pending(0, 0, ctr + C_LVL, ctr + C_LVL + 1);
    }
    // Levels 1..spec on the device. Windows of one wave (W <= LEVEL_WIN_MAX)
    // work straight from the level-1 list, all such levels in one launch: its
    // done flags mark the placed entries, and each level counts the inserts
    // entering the next one.
    uint32_t spec = 0;
    LevelBounds lb{};
    while (spec < dev_levels && spec + 1 <= g->h && ((uint64_t)g->S << (spec + 1)) <= LEVEL_WIN_MAX &&
           spec + 2 < C_LVL_SLOTS) {
        lb.lower[spec] = g->lower[spec + 1];
        lb.upper[spec] = g->upper[spec + 1];
        spec++;
    }
    if (spec) {
        phase_begin(g, P_UPPER);
        launch_window_region(g->S, spec, g->pk1.as<gkey>(), g->pv1.as<gval>(), g->pw1.as<uint32_t>(),
                             g->pd1.as<uint8_t>(), n_ub, ctr + C_LVL, g->keys.as<gkey>(), g->values.as<gval>(),
                             g->seps.as<gkey>(), g->row_offset.as<uint32_t>(), lb, ctr + C_LVL, g->bulk_threshold,
                             st);
        phase_end(g, P_UPPER, 1);
    }
    // Wider windows (one workgroup, W <= WIN_MAX) take compacted lists: the
    // entries not placed yet, with their windows at the level, into list 0,
    // then a compaction after each level alternating lists 0 and 1 (level L's
    // list is list (L - waves - 1) & 1).
    const uint32_t waves = spec;
    int pc = 0;
    while (spec < dev_levels && spec + 1 <= g->h && ((uint64_t)g->S << (spec + 1)) <= WIN_MAX &&
           spec + 2 < C_LVL_SLOTS) {
        const uint32_t L = spec + 1;
        if (L == waves + 1) {
            phase_begin(g, P_COMPACT);
            compact_step(g, g->pk1.as<gkey>(), g->pv1.as<gval>(), g->pw1.as<uint32_t>(), g->pd1.as<uint8_t>(), n_ub,
                         ctr + C_LVL + 1, g->pk0.as<gkey>(), g->pv0.as<gval>(), g->pw0.as<uint32_t>(),
                         g->pd0.as<uint8_t>(), ~(uint32_t)(((uint64_t)g->S << L) - 1), ctr + C_LVL + L);
            phase_end(g, P_COMPACT, 1);
        }
        phase_begin(g, P_UPPER);
        launch_window(g->S, g->S << L, PK[pc]->as<gkey>(), PV[pc]->as<gval>(), PW[pc]->as<uint32_t>(),
                      PD[pc]->as<uint8_t>(), n_ub, ctr + C_LVL + L, g->keys.as<gkey>(), g->values.as<gval>(),
                      g->seps.as<gkey>(), g->row_offset.as<uint32_t>(), g->lower[L], g->upper[L], ctr + C_LVL,
                      g->bulk_threshold, st);
        phase_end(g, P_UPPER, 1);
        compact_pending(pc, L, ctr + C_LVL + L, ctr + C_LVL + L + 1);
        pc ^= 1;
        spec++;
    }
    return spec;
}

// Host part of one update of n records, after the counters of its device
// part (enqueue_update, `spec` device levels) are in hctr: errors, counts,
// the bulk path, and any level above the device-driven ones.
int finish_update(gpma_handle *g, uint64_t n, uint32_t spec, bool used_msd) {
    hipStream_t st = g->st;
    uint32_t *ctr = g->ctr.as<uint32_t>();
    if (g->hctr[C_SORTERR]) return fail(GPMA_ERR_HIP, "sort or compaction look-back exceeded its spin limit");
    // a batch with a bucket this large sorts faster with the LSD passes: use
    // them for the next batches of this handle
    if (used_msd && g->hctr[C_MSD_MAXB] > MSD_BIG_BUCKET) g->msd_cool = MSD_COOL_BATCHES;
    const uint64_t inval = g->hctr[C_INVALID];
    g->invalid += inval;
    g->applied += n - inval;
    const uint64_t np0 = g->hctr[C_LVL];
    if (np0 >= g->bulk_threshold) {
        // The level kernels stood down: the level-1 list holds every pending
        // insert the level-0 pass did not place. Any leaf merge of the fused
        // level-0 pass is kept: the rebuild lays out the live set (array
        // items plus this list) from scratch, whichever of them a leaf holds.
        g->last_levels = 0;
        TRY(rebuild(g, g->pk1.as<gkey>(), g->pv1.as<gval>(), g->hctr[C_LVL + 1], false));
        TRY(stream_sync(g));
        HIP_TRY(hipGetLastError());
        return inval ? fail(GPMA_ERR_KEY_RANGE, std::to_string(inval) + " keys with src/dst >= row_num skipped")
                     : GPMA_OK;
    }
    g->last_levels = 0;
    for (uint32_t L = 0; L <= spec; L++)
        if (g->hctr[C_LVL + L]) g->last_levels = L + 1;
    Buf *PK[2] = {&g->pk0, &g->pk1}, *PV[2] = {&g->pv0, &g->pv1}, *PW[2] = {&g->pw0, &g->pw1},
        *PD[2] = {&g->pd0, &g->pd1};
    // host-driven levels above the prefix (rare: wide windows, growth)
    uint32_t level = spec + 1;
    uint64_t np = g->hctr[C_LVL + level];
    const uint32_t npslot[2] = {C_NP0, C_NP1};
    // the device prefix's one-wave levels (enqueue_update): after wider ones
    // the next level's entries are a compacted list already
    uint32_t waves = 0;
    while (waves < spec && ((uint64_t)g->S << (waves + 1)) <= LEVEL_WIN_MAX) waves++;
    int pc = spec > waves ? (int)((spec - waves) & 1) : 0;
    if (np > 0 && spec == waves) {
        // the level-1 list's entries not placed yet, with their windows at
        // `level`, into list 0
        fill_u32(ctr + C_NP1, (uint32_t)g->hctr[C_LVL + 1], 1, st);
        compact_step(g, g->pk1.as<gkey>(), g->pv1.as<gval>(), g->pw1.as<uint32_t>(), g->pd1.as<uint8_t>(), n,
                     ctr + C_NP1, g->pk0.as<gkey>(), g->pv0.as<gval>(), g->pw0.as<uint32_t>(), g->pd0.as<uint8_t>(),
                     ~(uint32_t)(((uint64_t)g->S << level) - 1), ctr + C_NP0);
        g->ctr_clean = false;
    }
    while (np > 0 && level <= g->h) {
        const uint64_t W = (uint64_t)g->S << level;
        // the list's count into its device slot (reading the counters cleared them)
        fill_u32(ctr + npslot[pc], (uint32_t)np, 1, st);
        g->ctr_clean = false;
        gkey *pk = PK[pc]->as<gkey>();
        gval *pv = PV[pc]->as<gval>();
        uint32_t *pw = PW[pc]->as<uint32_t>();
        uint8_t *pd = PD[pc]->as<uint8_t>();
        uint32_t *d_np = ctr + npslot[pc];
        g->last_levels = level + 1;
        phase_begin(g, P_UPPER);
        if (W <= WIN_MAX) {
            launch_window(g->S, (uint32_t)W, pk, pv, pw, pd, np, d_np, g->keys.as<gkey>(), g->values.as<gval>(),
                          g->seps.as<gkey>(), g->row_offset.as<uint32_t>(), g->lower[level], g->upper[level],
                          nullptr, 0, st);
            phase_end(g, P_UPPER, 1);
        } else {
            // windows wider than a workgroup: every window of the level in
            // one set of device-wide passes (no host round trip per window)
            TRY(ensure_tmp(g, g->C));
            TRY(g->wide.ensure((np + 1) * 24, &g->mem));
            WideWork w;
            w.blkcnt = g->blkcnt.as<uint32_t>();
            w.blkoff = g->blkoff.as<uint32_t>();
            w.heads = g->win.as<uint32_t>();  // free once the level-0 compaction ran
            w.d_nruns = ctr + C_NRUNS;
            w.chunkcnt = g->chunkcnt.as<uint32_t>();
            w.chunkoff = g->chunkoff.as<uint32_t>();
            w.mpre = g->wide.as<unsigned long long>();
            w.live = (uint32_t *)(w.mpre + np + 1);
            w.merge = w.live + np + 1;
            w.ok = (uint8_t *)(w.merge + np + 1);
            w.tk = g->tmpk.as<gkey>();
            w.tv = g->tmpv.as<gval>();
            launch_wide_level(pk, pv, pw, pd, np, d_np, W, g->S, g->lower[level], g->upper[level],
                              g->keys.as<gkey>(), g->values.as<gval>(), g->seps.as<gkey>(),
                              g->row_offset.as<uint32_t>(), w, st);
            phase_end(g, P_UPPER, 9);
        }
        const uint32_t mask = ~(uint32_t)(((uint64_t)g->S << (level + 1)) - 1);
        g->ctr_clean = false;
        phase_begin(g, P_COMPACT);
        compact_step(g, pk, pv, pw, pd, np, d_np, PK[pc ^ 1]->as<gkey>(), PV[pc ^ 1]->as<gval>(),
                     PW[pc ^ 1]->as<uint32_t>(), PD[pc ^ 1]->as<uint8_t>(), mask, ctr + npslot[pc ^ 1]);
        phase_end(g, P_COMPACT, 1);
        TRY(read_ctr(g));
        if (g->hctr[C_SORTERR]) return fail(GPMA_ERR_HIP, "compaction look-back exceeded its spin limit");
        pc ^= 1;
        np = g->hctr[npslot[pc]];
        level++;
    }
    if (np > 0) {
        // gpma.cuh:1034-1044: grow (or shrink) and rebalance the root
        TRY(rebuild(g, PK[pc]->as<gkey>(), PV[pc]->as<gval>(), np, true));
        TRY(stream_sync(g));
    }
    HIP_TRY(hipGetLastError());
    if (inval) return fail(GPMA_ERR_KEY_RANGE, std::to_string(inval) + " keys with src/dst >= row_num skipped");
    return GPMA_OK;
}

bool msd_usable(const gpma_handle *g, uint64_t n) {
    return msd_sort_fits(g->N) && n <= MSD_MAX_N && g->msd_cool == 0;
}

int do_update(gpma_handle *g, const uint64_t *keys, const uint8_t *values, uint64_t n, int on_device,
              bool guards_ok = false) {
    if (n == 0) return GPMA_OK;
    if (!keys || !values) return fail(GPMA_ERR_ARG, "gpma_update: null keys or values");
    if (n >= (1ull << 30)) return fail(GPMA_ERR_CAPACITY, "gpma_update: batch larger than 2^30 records");
    HIP_TRY(hipSetDevice(g->device));
    TRY(ensure_update_buffers(g, n));
    hipStream_t st = g->st;
    // host-side preparation before the first launch, so the launches that
    // follow reach the GPU back to back
    const bool use_msd = msd_usable(g, n);
    if (g->msd_cool) g->msd_cool--;
    MsdPlan pl{};
    if (use_msd) {
        pl = msd_plan(n, g->N, g->sort_bits, !guards_ok, g->sort_ipt);
        TRY(ensure_msd(g, n, pl));
    }
    if (!g->ctr_clean) fill_u32(g->ctr.as<uint32_t>(), 0, C_COUNT, st);
    g->ctr_clean = false;
    g->cc_ticket = 0;
    const gkey *src_k = keys;
    const gval *src_v = values;
    if (!on_device) {
        HIP_TRY(hipMemcpyAsync(g->in_k.p, keys, n * 8, hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(g->in_v.p, values, n, hipMemcpyHostToDevice, st));
        src_k = g->in_k.as<gkey>();
        src_v = g->in_v.as<gval>();
    }
    const uint32_t spec = enqueue_update(g, src_k, src_v, n, nullptr, guards_ok, use_msd ? &pl : nullptr,
                                         g->spec_levels);
    TRY(read_ctr(g));
    return finish_update(g, n, spec, use_msd);
}

int ensure_bfs(gpma_handle *g) {
    uint64_t n = (uint64_t)g->N + 1;
    Alloc *t = &g->mem;
    TRY(g->lev.ensure(n * 4, t));
    TRY(g->fr0.ensure(n * 4, t));
    TRY(g->fr1.ensure(n * 4, t));
    TRY(g->fb0.ensure(n * 4, t));
    TRY(g->fb1.ensure(n * 4, t));
    TRY(g->bcnt.ensure((n + BFS_MAX_CHUNK + 2) * 8, t));
    TRY(g->bcntb.ensure((n + BFS_MAX_CHUNK + 2) * 8, t));
    if (!g->hcnt) HIP_TRY(hipHostMalloc((void **)&g->hcnt, 2 * (BFS_MAX_CHUNK + 1) * 8, hipHostMallocDefault));
    return GPMA_OK;
}

BfsState bfs_state(gpma_handle *g) {
    BfsState b;
    b.tile_cap = g->push_tile;
    b.levels = g->lev.as<uint32_t>();
    b.q[0] = g->fr0.as<uint32_t>();
    b.q[1] = g->fr1.as<uint32_t>();
    b.qbig[0] = g->fb0.as<uint32_t>();
    b.qbig[1] = g->fb1.as<uint32_t>();
    b.cnt = g->bcnt.as<unsigned long long>();
    b.cnt_big = g->bcntb.as<unsigned long long>();
    return b;
}

// The level loop of gpma_bfs_multidev_single (gpma_bfs.cuh:446-509) on one
// device, without a host round trip per level: levels are launched in
// chunks of growing size and the host reads the chunk's frontier sizes once.
// hcnt: pinned host buffer of 2 * (BFS_MAX_CHUNK + 1) u64 words.
int run_bfs(const gkey *keys, const gval *values, uint64_t C, const uint32_t *row_offset, uint32_t N, uint32_t start,
            const BfsState &b, unsigned long long *hcnt, hipStream_t st, uint32_t *reach_out, double pull_fraction,
            gpma_handle *prof = nullptr) {
    launch_bfs_init(b, N, start, row_offset, st);
    const unsigned long long thr =
        C ? (unsigned long long)(pull_fraction * (double)C) : ~0ull;  // no pull without the array size
    uint64_t reach = 1;
    // 4 levels first (a start with a short reach ends there); a traversal
    // still alive then takes the rest of the last one's depth + 2 at once,
    // so one like it ends after two round trips; then chunks of 4, doubling
    uint32_t level = 1, chunk = 4;
    bool done = false, hinted = false;
    while (!done) {
        if ((uint64_t)level + chunk + 1 > (uint64_t)N + BFS_MAX_CHUNK) chunk = 1;
        launch_bfs_clear(b, level + 1, level + chunk + 1, st);
        if (prof) phase_begin(prof, P_BFS);
        for (uint32_t j = 0; j < chunk; j++) launch_bfs_level(keys, values, C, N, row_offset, b, level + j, thr, st);
        if (prof) phase_end(prof, P_BFS, chunk);
        if (prof) {
            uint32_t w[4 * BFS_MAX_CHUNK];
            TRY(publish_and_wait(prof, (const uint32_t *)(b.cnt + level + 1), 2 * chunk,
                                 (const uint32_t *)(b.cnt_big + level + 1), 2 * chunk, w));
            std::memcpy(hcnt, w, chunk * 8);
            std::memcpy(hcnt + BFS_MAX_CHUNK + 1, w + 2 * chunk, chunk * 8);
            if (prof->prof) phase_collect(prof);
        } else {
            HIP_TRY(hipMemcpyAsync(hcnt, b.cnt + level + 1, chunk * 8, hipMemcpyDeviceToHost, st));
            HIP_TRY(hipMemcpyAsync(hcnt + BFS_MAX_CHUNK + 1, b.cnt_big + level + 1, chunk * 8, hipMemcpyDeviceToHost,
                                   st));
            HIP_TRY(hipStreamSynchronize(st));
        }
        for (uint32_t j = 0; j < chunk && !done; j++) {
            const uint64_t c = (uint64_t)(uint32_t)hcnt[j] + (uint32_t)hcnt[BFS_MAX_CHUNK + 1 + j];
            if (c == 0) {
                done = true;
                if (prof) prof->bfs_depth = level + j;  // the last level with a frontier
            }
            reach += c;
        }
        level += chunk;
        chunk = hinted ? 4 : std::min<uint32_t>(chunk * 2, BFS_MAX_CHUNK);
        hinted = false;
        if (level == 5 && prof && prof->bfs_depth + 2 > level + chunk) {
            chunk = std::min(BFS_MAX_CHUNK, prof->bfs_depth + 2 - 4);
            hinted = true;
        }
        if (level >= N + 1) done = true;
    }
    HIP_TRY(hipGetLastError());
    if (reach_out) *reach_out = (uint32_t)reach;
    return GPMA_OK;
}

int ensure_msbfs(gpma_handle *g) {
    const uint64_t n = (uint64_t)g->N + 1;
    Alloc *t = &g->mem;
    TRY(g->ms_seen.ensure(n * 8, t));
    TRY(g->ms_fr0.ensure(n * 8, t));
    TRY(g->ms_fr1.ensure(n * 8, t));
    TRY(g->ms_nx.ensure(n * 8, t));
    TRY(g->ms_lv16.ensure(n * 128, t));
    TRY(g->ms_ev.ensure(n * 128, t));
    TRY(g->ms_evn.ensure((n + 3) / 4 * 4, t));
    TRY(g->ms_q0.ensure(n * 4, t));
    TRY(g->ms_q1.ensure(n * 4, t));
    TRY(g->ms_qb0.ensure(n * 4, t));
    TRY(g->ms_qb1.ensure(n * 4, t));
    TRY(g->ms_touched.ensure(n * 4, t));
    TRY(g->ms_cnt.ensure((n + BFS_MAX_CHUNK + 2) * 8, t));
    TRY(g->ms_cntb.ensure((n + BFS_MAX_CHUNK + 2) * 8, t));
    TRY(g->ms_tcnt.ensure((n + BFS_MAX_CHUNK + 2) * 4, t));
    TRY(g->ms_partial.ensure((uint64_t)ms_output_blocks(g->N) * 64 * 4, t));
    TRY(g->ms_starts.ensure(64 * 4, t));
    TRY(g->ms_reach.ensure(64 * 4, t));
    if (!g->ms_hcnt)
        HIP_TRY(hipHostMalloc((void **)&g->ms_hcnt, 2 * (BFS_MAX_CHUNK + 1) * 8, hipHostMallocDefault));
    return GPMA_OK;
}

MsState ms_state(gpma_handle *g) {
    MsState m;
    m.tile_cap = g->ms_tile;
    m.seen = g->ms_seen.as<uint64_t>();
    m.fr[0] = g->ms_fr0.as<uint64_t>();
    m.fr[1] = g->ms_fr1.as<uint64_t>();
    m.nx = g->ms_nx.as<uint64_t>();
    m.lv16 = g->ms_lv16.as<uint16_t>();
    m.ev = g->ms_ev.as<uint4>();
    m.evn = g->ms_evn.as<uint8_t>();
    m.lv32 = g->ms_wide ? g->ms_lv32.as<uint32_t>() : nullptr;
    m.q[0] = g->ms_q0.as<uint32_t>();
    m.q[1] = g->ms_q1.as<uint32_t>();
    m.qbig[0] = g->ms_qb0.as<uint32_t>();
    m.qbig[1] = g->ms_qb1.as<uint32_t>();
    m.touched = g->ms_touched.as<uint32_t>();
    m.cnt = g->ms_cnt.as<unsigned long long>();
    m.cnt_big = g->ms_cntb.as<unsigned long long>();
    m.tcnt = g->ms_tcnt.as<uint32_t>();
    m.partial = g->ms_partial.as<uint32_t>();
    return m;
}

// Pull threshold of the multi-source traversal, in frontier slots: twice the
// single-source share of C. Its streaming step ORs a 64-bit mask per slot,
// so pushing stays cheaper for longer: on the power-law 1.63M-node graph
// with 64 random sources, 2.22 ms with the single-source 25%, 1.90 ms at 50%
// or above, 3.45 ms when every level pulls.
unsigned long long ms_pull_threshold(const gpma_handle *g) {
    return (unsigned long long)(std::min(1.0, 2.0 * g->pull_fraction) * (double)g->C);
}

// Multi-source levels from `level` on, in chunks without a host round trip
// per level (like run_bfs), until a level adds no vertex. cnt[level] must
// hold the level's frontier counts and the touched count of `level` must be
// clear. *depth: the last level run. Widens the level staging to u32 before
// the first level that stages a level above 65,535 (g->ms_wide).
int msbfs_levels(gpma_handle *g, uint32_t level, uint32_t *depth) {
    hipStream_t st = g->st;
    MsState m = ms_state(g);
    const uint32_t N = g->N;
    const unsigned long long thr = ms_pull_threshold(g);
    // a traversal from level 1 first covers the last one's depth + 2, then
    // takes chunks of 4, doubling
    const bool hinted = level == 1 && g->ms_depth + 2 > 4;
    uint32_t chunk = hinted ? std::min(BFS_MAX_CHUNK, g->ms_depth + 2) : 4;
    bool first = true;
    bool done = false;
    *depth = level;
    while (!done) {
        if ((uint64_t)level + chunk + 1 > (uint64_t)N + BFS_MAX_CHUNK) chunk = 1;
        // counts of the chunk's levels (touched counts are indexed by the
        // expanding level; the first level's is cleared by ms_init or the
        // previous chunk)
        ms_clear(m, level + 1, level + chunk + 1, st);
        phase_begin(g, P_BFS);
        for (uint32_t j = 0; j < chunk; j++) {
            // level L stages level L + 1: past 65,535 the staging moves to u32 rows
            if (!g->ms_wide && level + j + 1 > 0xFFFFu) {
                TRY(g->ms_lv32.ensure((uint64_t)N * 256, &g->mem));
                ms_widen(m, N, g->ms_lv32.as<uint32_t>(), st);
                g->ms_wide = true;
                m = ms_state(g);
            }
            ms_level(g->keys.as<gkey>(), g->values.as<gval>(), g->C, N, g->row_offset.as<uint32_t>(), m, level + j,
                     thr, st);
        }
        phase_end(g, P_BFS, 2 * chunk);
        {
            uint32_t w[4 * BFS_MAX_CHUNK];
            TRY(publish_and_wait(g, (const uint32_t *)(m.cnt + level + 1), 2 * chunk,
                                 (const uint32_t *)(m.cnt_big + level + 1), 2 * chunk, w));
            std::memcpy(g->ms_hcnt, w, chunk * 8);
            std::memcpy(g->ms_hcnt + BFS_MAX_CHUNK + 1, w + 2 * chunk, chunk * 8);
            if (g->prof) phase_collect(g);
        }
        for (uint32_t j = 0; j < chunk && !done; j++) {
            *depth = level + j;
            if ((uint32_t)g->ms_hcnt[j] == 0 && (uint32_t)g->ms_hcnt[BFS_MAX_CHUNK + 1 + j] == 0) done = true;
        }
        level += chunk;
        chunk = first && hinted ? 4 : std::min<uint32_t>(chunk * 2, BFS_MAX_CHUNK);
        first = false;
        if (level >= N + 1) done = true;
    }
    g->ms_depth = *depth;
    return GPMA_OK;
}

// Multi-source BFS: levels land in out[nsrc][N] (device), counts in
// reach[nsrc] (device).
int run_msbfs(gpma_handle *g, const uint32_t *d_starts, uint32_t nsrc, uint32_t *out, uint32_t *reach,
              uint32_t *depth = nullptr) {
    g->ms_wide = false;
    ms_init(ms_state(g), g->N, d_starts, nsrc, g->row_offset.as<uint32_t>(), g->st);
    uint32_t d = 0;
    TRY(msbfs_levels(g, 1, &d));
    if (depth) *depth = d;
    phase_begin(g, P_BFS);
    ms_output(ms_state(g), g->N, nsrc, out, reach, g->st);
    phase_end(g, P_BFS, 2);
    if (g->prof) TRY(stream_sync(g));  // callers synchronise before reading results
    HIP_TRY(hipGetLastError());
    return GPMA_OK;
}

// ---- captured steps

constexpr uint32_t STEP_MIN_CLASS = 16;   // smallest size class: 2^16 records
constexpr uint32_t STEP_MAX_GRAPHS = 16;  // graphs kept per handle (least recently used go first)
constexpr uint32_t STEP_MAX_BLEVELS = 64; // BFS levels a graph may hold

void drop_graphs(gpma_handle *g) {
    for (auto &sg : g->graphs)
        if (sg.exec) (void)hipGraphExecDestroy(sg.exec);
    g->graphs.clear();
}

// Everything a captured step's launches bake in: buffers, layout, limits.
std::vector<uint64_t> step_signature(const gpma_handle *g, uint32_t cls, uint32_t nsrc, const uint32_t *out,
                                     uint32_t ulevels, uint32_t blevels) {
    std::vector<uint64_t> v = {cls, nsrc, (uint64_t)(uintptr_t)out, ulevels, blevels, g->N, g->C, g->S, g->h,
                               g->bulk_threshold, g->sort_bits, (uint64_t)(g->pull_fraction * 1e9), g->sort_ipt,
                               g->ms_tile};
    const Buf *bufs[] = {&g->keys, &g->values, &g->seps, &g->row_offset, &g->ctr, &g->step_n, &g->step_stage,
                         &g->in_k, &g->in_v,
                         &g->sk0, &g->sv0, &g->sk1, &g->win, &g->pk0, &g->pv0, &g->pw0, &g->pd0, &g->pk1, &g->pv1,
                         &g->pw1, &g->pd1, &g->ccstat, &g->msd_kp, &g->msd_rk, &g->msd_mat, &g->msd_bstart,
                         &g->msd_status, &g->msd_spl, &g->ms_seen, &g->ms_fr0, &g->ms_fr1, &g->ms_nx, &g->ms_lv16, &g->ms_q0,
                         &g->ms_q1, &g->ms_qb0, &g->ms_qb1, &g->ms_touched, &g->ms_cnt, &g->ms_cntb, &g->ms_tcnt,
                         &g->ms_partial, &g->ms_starts, &g->ms_reach, &g->ms_ev, &g->ms_evn};
    for (const Buf *b : bufs) {
        v.push_back((uint64_t)(uintptr_t)b->p);
        v.push_back(b->bytes);
    }
    return v;
}

// Capture one step for batches of up to 2^cls records: counters and look-back
// status cleared, the update's device part with `ulevels` levels, the gate,
// and (nsrc > 0) the multi-source BFS with `blevels` levels into out.
int capture_step(gpma_handle *g, uint32_t cls, const MsdPlan &pl, uint32_t ulevels, uint32_t nsrc, uint32_t *out,
                 uint32_t blevels, hipGraphExec_t *exec, uint32_t *spec_out) {
    hipStream_t st = g->st;
    uint32_t *ctr = g->ctr.as<uint32_t>();
    // the graph's look-back tags must stay clear of a wrap-around while it lives
    if (g->cc_epoch > 0xFFFF0000u) {
        fill_u64(g->ccstat.as<uint64_t>(), 0, g->ccstat.bytes / 8, st);
        if (g->msd_status.p) fill_u64(g->msd_status.as<uint64_t>(), 0, g->msd_status.bytes / 8, st);
        g->cc_epoch = 0;
        drop_graphs(g);
    }
    g->ms_wide = false;  // graphs hold at most 64 BFS levels: u16 staging only
    HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    g->capturing = true;
    // the counters, and the look-back status words (each replay reuses the
    // same tags, so they start clear), in one launch
    zero3(ctr, C_COUNT, g->ccstat.as<uint64_t>(), g->ccstat.bytes / 8, g->msd_status.as<uint64_t>(),
          g->msd_status.bytes / 8, st);
    g->cc_ticket = 0;
    const uint32_t spec = enqueue_update(g, g->in_k.as<gkey>(), g->in_v.as<gval>(), 1ull << cls,
                                         g->step_n.as<uint32_t>(), false, &pl, ulevels);
    *spec_out = spec;
    if (!nsrc)
        launch_step_gate(ctr + C_LVL, ctr + C_LVL + spec + 1, ctr + C_SORTERR, g->bulk_threshold, ctr + C_GATE, st);
    if (nsrc) {
        const MsState m = ms_state(g);
        const unsigned long long thr = ms_pull_threshold(g);
        MsStepGate sg;
        sg.np0 = ctr + C_LVL;
        sg.np_top = ctr + C_LVL + spec + 1;
        sg.err = ctr + C_SORTERR;
        sg.bulk = g->bulk_threshold;
        ms_init_step(m, g->N, g->ms_starts.as<uint32_t>(), nsrc, g->row_offset.as<uint32_t>(), sg, ctr + C_GATE, 2,
                     blevels + 2, st);
        for (uint32_t L = 1; L <= blevels; L++)
            ms_level(g->keys.as<gkey>(), g->values.as<gval>(), g->C, g->N, g->row_offset.as<uint32_t>(), m, L, thr,
                     st);
        ms_output(m, g->N, nsrc, out, g->ms_reach.as<uint32_t>(), st);
        launch_step_gather(m.cnt + blevels + 1, m.cnt_big + blevels + 1, g->ms_reach.as<uint32_t>(), nsrc,
                           g->step_stage.as<uint32_t>(), st);
    }
    // the read-back: counters (cleared on the device) and the BFS words, then the signal
    launch_publish(ctr, C_COUNT, g->step_stage.as<uint32_t>(), nsrc ? 4 + nsrc : 0, g->sig, SIG_STEP, true, st);
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(st, &graph);
    g->capturing = false;
    const hipError_t el = hipGetLastError();
    if (ec != hipSuccess || el != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        return fail(GPMA_ERR_HIP, std::string("step capture: ") + hipGetErrorString(ec != hipSuccess ? ec : el));
    }
    const hipError_t ei = hipGraphInstantiate(exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess) return fail(GPMA_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
    g->sst.captures++;
    return GPMA_OK;
}

int step_host_path(gpma_handle *g, const uint64_t *keys, const uint8_t *values, uint64_t n, int on_device,
                   const uint32_t *starts, uint32_t nsrc, uint32_t *levels, int levels_on_device,
                   uint32_t *reachable);

// The graph for size class cls (batches up to 2^cls records), nsrc sources
// and levels destination dlev (device; nullptr: the handle's own buffer),
// captured now if the cache has none for the current layout. *out: where
// its BFS writes levels; *blevels: BFS levels it holds.
int step_graph(gpma_handle *g, uint32_t cls, uint32_t nsrc, uint32_t *dlev, gpma_handle::StepGraph **res,
               uint32_t **out, uint32_t *blevels) {
    const uint64_t cap = 1ull << cls;
    TRY(ensure_update_buffers(g, cap));
    const MsdPlan pl = msd_plan(cap, g->N, g->sort_bits, true, g->sort_ipt);
    TRY(ensure_msd(g, cap, pl));
    TRY(g->step_n.ensure(256, &g->mem));
    TRY(g->step_stage.ensure((4 + 64) * 4, &g->mem));
    *out = nullptr;
    if (nsrc) {
        TRY(ensure_msbfs(g));
        if (dlev) {
            *out = dlev;
        } else {
            TRY(g->ms_out.ensure((uint64_t)nsrc * g->N * 4, &g->mem));
            *out = g->ms_out.as<uint32_t>();
        }
    }
    // levels whose window fits one workgroup, at most as many as steps have
    // needed (an empty captured level still costs two launches)
    uint32_t ulevels = 0;
    while (ulevels + 1 <= g->h && ((uint64_t)g->S << (ulevels + 1)) <= WIN_MAX && ulevels + 2 < C_LVL_SLOTS &&
           ulevels < g->step_ulevels)
        ulevels++;
    *blevels = nsrc ? std::min<uint32_t>(g->step_blevels, g->N) : 0;
    const std::vector<uint64_t> sig = step_signature(g, cls, nsrc, *out, ulevels, *blevels);
    for (auto &x : g->graphs)
        if (x.sig == sig) {
            *res = &x;
            return GPMA_OK;
        }
    if (g->graphs.size() >= STEP_MAX_GRAPHS) {
        auto lru = std::min_element(g->graphs.begin(), g->graphs.end(),
                                    [](const auto &a, const auto &b) { return a.used < b.used; });
        (void)hipGraphExecDestroy(lru->exec);
        g->graphs.erase(lru);
    }
    hipGraphExec_t exec = nullptr;
    uint32_t spec = 0;
    TRY(capture_step(g, cls, pl, ulevels, nsrc, *out, *blevels, &exec, &spec));
    g->graphs.push_back({sig, exec, ++g->step_clock, spec});
    *res = &g->graphs.back();
    return GPMA_OK;
}

int run_step(gpma_handle *g, const uint64_t *keys, const uint8_t *values, uint64_t n, int on_device,
             const uint32_t *starts, uint32_t nsrc, uint32_t *levels, int levels_on_device, uint32_t *reachable) {
    g->sst.steps++;
    if (n == 0 || !msd_usable(g, n) || g->prof)
        return step_host_path(g, keys, values, n, on_device, starts, nsrc, levels, levels_on_device, reachable);
    hipStream_t st = g->st;
    uint32_t cls = STEP_MIN_CLASS;
    while ((1ull << cls) < n) cls++;
    uint32_t *out = nullptr;
    uint32_t blevels = 0;
    gpma_handle::StepGraph *sg = nullptr;
    TRY(step_graph(g, cls, nsrc, levels && levels_on_device ? levels : nullptr, &sg, &out, &blevels));
    sg->used = ++g->step_clock;
    // stage the step's inputs where the graph reads them
    if (on_device) {
        stage_step(g->in_k.as<uint64_t>(), keys, g->in_v.as<uint8_t>(), values, n, g->step_n.as<uint32_t>(), st);
    } else {
        HIP_TRY(hipMemcpyAsync(g->in_k.p, keys, n * 8, hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(g->in_v.p, values, n, hipMemcpyHostToDevice, st));
        fill_u32(g->step_n.as<uint32_t>(), (uint32_t)n, 1, st);
    }
    if (nsrc) HIP_TRY(hipMemcpyAsync(g->ms_starts.p, starts, nsrc * 4, hipMemcpyHostToDevice, st));
    const uint32_t spec = sg->spec;
    // the graph ends with its own publish: counters (left cleared on the
    // device), the frontier counts after the last captured level and the
    // reachable counts, then the step's sequence word
    *(volatile uint32_t *)g->sig = 0;
    std::atomic_thread_fence(std::memory_order_seq_cst);
    g->ms_wide = false;  // the replay's traversal stages u16 levels
    // the bucket sort's row splitters, every SPL_EVERY steps, ahead of the graph
    if (g->spl_age >= SPL_EVERY) {
        msd_splitters(g->row_offset.as<uint32_t>(), g->N, g->C, g->msd_spl.as<uint32_t>(), st);
        g->spl_age = 0;
    }
    g->spl_age++;
    HIP_TRY(hipGraphLaunch(sg->exec, st));
    g->sst.replays++;
    TRY(wait_signal(g, SIG_STEP));
    g->ctr_clean = true;
    HIP_TRY(hipGetLastError());
    uint32_t *rb = g->step_rb;  // [0..2) count words, [2..4) long-range count words, then reachable
    std::memcpy(g->hctr, g->sig + 1, C_COUNT * 4);
    if (nsrc) std::memcpy(rb, g->sig + 1 + C_COUNT, (4 + nsrc) * 4);
    const bool gated = g->hctr[C_GATE] != 0;
    const int urc = finish_update(g, n, spec, true);
    if (urc != GPMA_OK && urc != GPMA_ERR_KEY_RANGE) return urc;
    if (gated) g->sst.update_tails++;
    // inserts climbed above the captured levels: later captures hold one more
    // than the highest level they reached
    const uint32_t need = g->last_levels ? g->last_levels - 1 : 0;
    if (need > spec && need >= g->step_ulevels) g->step_ulevels = need + 1;
    if (!nsrc) return urc;
    bool reach_ready = true;
    if (gated) {
        // the captured BFS stood down: traverse the finished graph now
        uint32_t depth = 0;
        TRY(run_msbfs(g, g->ms_starts.as<uint32_t>(), nsrc, out, g->ms_reach.as<uint32_t>(), &depth));
        if (depth + 2 > g->step_blevels) g->step_blevels = std::min(depth + 2, STEP_MAX_BLEVELS);
        reach_ready = false;
    } else if (rb[0] != 0 || rb[2] != 0) {  // vertices in the frontier after the last captured level
        // deeper than the captured levels: continue from the next one
        g->sst.bfs_tails++;
        uint32_t depth = 0;
        TRY(msbfs_levels(g, blevels + 1, &depth));
        ms_output(ms_state(g), g->N, nsrc, out, g->ms_reach.as<uint32_t>(), st);
        if (depth + 2 > g->step_blevels) g->step_blevels = std::min(depth + 2, STEP_MAX_BLEVELS);
        reach_ready = false;
    }
    if (reachable) {
        if (reach_ready) {
            std::memcpy(reachable, rb + 4, nsrc * 4);
        } else {
            HIP_TRY(hipMemcpyAsync(reachable, g->ms_reach.p, nsrc * 4, hipMemcpyDeviceToHost, st));
        }
    }
    if (levels && !levels_on_device)
        HIP_TRY(hipMemcpyAsync(levels, out, (uint64_t)nsrc * g->N * 4, hipMemcpyDeviceToHost, st));
    if (!reach_ready || (levels && !levels_on_device)) TRY(stream_sync(g));
    HIP_TRY(hipGetLastError());
    return urc;
}

// gpma_bfs_multi after argument checks.
int bfs_multi_impl(gpma_handle *g, const uint32_t *starts, uint32_t nsrc, uint32_t *levels, int on_device,
                   uint32_t *reachable) {
    HIP_TRY(hipSetDevice(g->device));
    TRY(ensure_msbfs(g));
    HIP_TRY(hipMemcpyAsync(g->ms_starts.p, starts, nsrc * 4, hipMemcpyHostToDevice, g->st));
    uint32_t *out = levels;
    if (!on_device || !levels) {
        TRY(g->ms_out.ensure((uint64_t)nsrc * g->N * 4, &g->mem));
        out = g->ms_out.as<uint32_t>();
    }
    TRY(run_msbfs(g, g->ms_starts.as<uint32_t>(), nsrc, out, g->ms_reach.as<uint32_t>()));
    if (levels && !on_device)
        HIP_TRY(hipMemcpyAsync(levels, out, (uint64_t)nsrc * g->N * 4, hipMemcpyDeviceToHost, g->st));
    if (reachable) HIP_TRY(hipMemcpyAsync(reachable, g->ms_reach.p, nsrc * 4, hipMemcpyDeviceToHost, g->st));
    HIP_TRY(hipStreamSynchronize(g->st));
    return GPMA_OK;
}

int step_host_path(gpma_handle *g, const uint64_t *keys, const uint8_t *values, uint64_t n, int on_device,
                   const uint32_t *starts, uint32_t nsrc, uint32_t *levels, int levels_on_device,
                   uint32_t *reachable) {
    g->sst.host_steps++;
    const int urc = do_update(g, keys, values, n, on_device);
    if (urc != GPMA_OK && urc != GPMA_ERR_KEY_RANGE) return urc;
    if (nsrc) TRY(bfs_multi_impl(g, starts, nsrc, levels, levels_on_device, reachable));
    return urc;
}

// ---- BFS without allocation (gpma_csr_bfs_noalloc)
//
// The frontier queues and per-level counts of one traversal live in the
// file-scope device array `scratch`, sized for node_size <= NOALLOC_MAX_NODES:
// two count arrays of NA_CNT u64 (indexed by level), then four frontier
// queues of NA_Q u32. Every GPU the library runs on holds its own copy (64 MiB)
// from module load to process end. The traversal clears what it reads before
// reading it (k_bfs_init: levels and the first counts; k_bfs_clear: each
// chunk's counts), so the contents a previous call left do not matter. Calls
// that overlap on one GPU share the array and race.
constexpr uint32_t NOALLOC_MAX_NODES = 1u << 21;
constexpr uint64_t NA_CNT = (uint64_t)NOALLOC_MAX_NODES + 1 + BFS_MAX_CHUNK + 2;
constexpr uint64_t NA_Q = (uint64_t)NOALLOC_MAX_NODES + 1;
constexpr uint64_t NA_WORDS64 = 2 * NA_CNT + (4 * NA_Q + 1) / 2;

}  // namespace

static __device__ __align__(256) unsigned long long scratch[NA_WORDS64];

namespace {

constexpr int NA_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void *g_scratch_dev[NA_MAX_DEVICES] = {};  // device address of `scratch` per device, once looked up

// Device address of `scratch` on the current device, looked up once per
// device with that device current. Host code uses only this address: the
// array's own name is a host-side stand-in.
int noalloc_scratch(void **out) {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= NA_MAX_DEVICES)
        return fail(GPMA_ERR_NO_DEVICE, "gpma_csr_bfs_noalloc: device ordinal " + std::to_string(dev) + " too large");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(GPMA_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_dev[dev] = p;
    }
    *out = g_scratch_dev[dev];
    return GPMA_OK;
}

}  // namespace

extern "C" {

const char *gpma_last_error(void) { return g_last_error.c_str(); }

int gpma_profile(gpma_t g, int enable) {
    if (!g) return fail(GPMA_ERR_ARG, "gpma_profile: null handle");
    HIP_TRY(hipSetDevice(g->device));
    if (enable == 2) {
        HIP_TRY(hipStreamSynchronize(g->st));
        phase_collect(g);
        for (int p = 0; p < GPMA_PHASES; p++) {
            g->ph_ms[p] = 0;
            g->ph_calls[p] = 0;
        }
    }
    g->prof = enable != 0;
    return GPMA_OK;
}

int gpma_profile_phases(gpma_t g, uint32_t mask) {
    if (!g) return fail(GPMA_ERR_ARG, "gpma_profile_phases: null handle");
    g->prof_mask = mask & ((1u << GPMA_PHASES) - 1);
    return GPMA_OK;
}

int gpma_get_profile(gpma_t g, double *ms, uint64_t *calls, uint32_t n) {
    if (!g) return fail(GPMA_ERR_ARG, "gpma_get_profile: null handle");
    for (uint32_t p = 0; p < n && p < GPMA_PHASES; p++) {
        if (ms) ms[p] = g->ph_ms[p];
        if (calls) calls[p] = g->ph_calls[p];
    }
    return GPMA_OK;
}

int gpma_device_malloc(void **ptr, uint64_t bytes) {
    if (!ptr) return fail(GPMA_ERR_ARG, "gpma_device_malloc: null ptr");
    HIP_TRY(hipMalloc(ptr, bytes ? bytes : 1));
    return GPMA_OK;
}

int gpma_mem_info(int32_t device, uint64_t *free_bytes, uint64_t *total_bytes) {
    if (!free_bytes || !total_bytes) return fail(GPMA_ERR_ARG, "gpma_mem_info: null output");
    int old = 0;
    HIP_TRY(hipGetDevice(&old));
    if (device >= 0) HIP_TRY(hipSetDevice(device));
    size_t f = 0, t = 0;
    hipError_t e = hipMemGetInfo(&f, &t);
    if (device >= 0) (void)hipSetDevice(old);
    HIP_TRY(e);
    *free_bytes = f;
    *total_bytes = t;
    return GPMA_OK;
}

int gpma_device_free(void *ptr) {
    if (ptr) HIP_TRY(hipFree(ptr));
    return GPMA_OK;
}

int gpma_memcpy(void *dst, const void *src, uint64_t bytes, int kind) {
    hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
    HIP_TRY(hipMemcpy(dst, src, bytes, k));
    return GPMA_OK;
}

int gpma_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int gpma_create(uint32_t row_num, const gpma_config *cfg, gpma_t *out) {
    if (!out) return fail(GPMA_ERR_ARG, "gpma_create: null out");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(GPMA_ERR_NO_DEVICE, "no HIP device");
    if (row_num >= 0xFFFFFFFFu) return fail(GPMA_ERR_ARG, "gpma_create: row_num must be < 2^32-1");
    gpma_handle *g = new gpma_handle();
    g->N = row_num;
    // callers built against the first 32-byte layout (up to `reserved`) keep working
    if (cfg && cfg->struct_size >= GPMA_CONFIG_V1_SIZE) {
        if (cfg->device >= 0) g->device = cfg->device;
        else hipGetDevice(&g->device);
        if (cfg->bulk_threshold) g->bulk_threshold = cfg->bulk_threshold;
        if (cfg->pull_permille) g->pull_fraction = cfg->pull_permille / 1000.0;
        const uint32_t sl = cfg->segment_length;
        if (sl && (sl < 4 || sl > 32 || (sl & (sl - 1)))) {
            delete g;
            return fail(GPMA_ERR_ARG, "gpma_create: segment_length must be 4, 8, 16 or 32");
        }
        if (cfg->device_levels > 8 || (cfg->sort_bucket_bits && (cfg->sort_bucket_bits < 4 ||
                                                                    cfg->sort_bucket_bits > 12))) {
            delete g;
            return fail(GPMA_ERR_ARG, "gpma_create: device_levels must be <= 8, sort_bucket_bits 4..12");
        }
        g->seg_override = sl;
        if (cfg->device_levels) g->spec_levels = cfg->device_levels;
        g->sort_bits = cfg->sort_bucket_bits;
        if (cfg->struct_size >= sizeof(gpma_config)) {
            const uint16_t *dp = cfg->density_permille;
            if (dp[0]) g->lower_leaf = dp[0] / 1000.0;
            if (dp[1]) g->lower_root = dp[1] / 1000.0;
            if (dp[2]) g->upper_leaf = dp[2] / 1000.0;
            if (dp[3]) g->upper_root = dp[3] / 1000.0;
            // every redistributed window keeps at least one item per segment
            // (lower_leaf * 16 >= 1), bounds nest, and a resized root (density
            // above upper_root / 2) never starts below its lower bound
            if (g->lower_leaf < 0.0625 || g->lower_leaf > g->lower_root || g->lower_root * 2 > g->upper_root ||
                g->upper_root > g->upper_leaf || g->upper_leaf > 1.0) {
                delete g;
                return fail(GPMA_ERR_ARG, "gpma_create: density thresholds need 0.0625 <= lower_leaf <= lower_root, "
                                          "2 * lower_root <= upper_root <= upper_leaf <= 1");
            }
            const uint32_t ipt = cfg->sort_items_per_lane, pt = cfg->push_tile, mt = cfg->ms_tile;
            if ((ipt && ipt != 4 && ipt != 8 && ipt != 16) || (pt && (pt < 16 || pt > 64)) || (mt && (mt < 16 || mt > 64))) {
                delete g;
                return fail(GPMA_ERR_ARG, "gpma_create: sort_items_per_lane must be 4, 8 or 16; push_tile and "
                                          "ms_tile 16..64");
            }
            g->sort_ipt = ipt;
            if (pt) g->push_tile = pt;
            if (mt) g->ms_tile = mt;
        }
    } else {
        hipGetDevice(&g->device);
    }
    int rc = GPMA_OK;
    do {
        if (hipSetDevice(g->device) != hipSuccess) {
            rc = fail(GPMA_ERR_ARG, "gpma_create: bad device");
            break;
        }
        if (hipStreamCreateWithFlags(&g->st, hipStreamNonBlocking) != hipSuccess) {
            rc = fail(GPMA_ERR_HIP, "hipStreamCreate failed");
            break;
        }
        window_region_prepare();  // its launch shape, before any graph capture
        if ((rc = g->ctr.ensure(C_COUNT * 4, &g->mem)) != GPMA_OK) break;
        if (hipHostMalloc((void **)&g->hctr, C_COUNT * 4, hipHostMallocDefault) != hipSuccess ||
            hipHostMalloc((void **)&g->sig, SIG_WORDS * 4, hipHostMallocCoherent) != hipSuccess) {
            rc = fail(GPMA_ERR_OOM, "hipHostMalloc failed");
            break;
        }
        std::memset(g->sig, 0, SIG_WORDS * 4);
        if ((rc = init_structure(g)) != GPMA_OK) break;
        // gpma.cuh:70-78: insert every row guard (v << 32 | 0xFFFFFFFF) with value 1
        if (row_num) {
            std::vector<uint64_t> guards(row_num);
            std::vector<uint8_t> ones(row_num, 1);
            for (uint32_t v = 0; v < row_num; v++) guards[v] = ((uint64_t)v << 32) | 0xFFFFFFFFull;
            rc = do_update(g, guards.data(), ones.data(), row_num, 0, /*guards_ok=*/true);
            g->invalid = 0;
            g->applied = 0;
        }
    } while (0);
    if (rc != GPMA_OK) {
        gpma_destroy(g);
        return rc;
    }
    *out = g;
    return GPMA_OK;
}

int gpma_destroy(gpma_t g) {
    if (!g) return GPMA_OK;
    hipSetDevice(g->device);
    if (g->st) hipStreamSynchronize(g->st);
    Buf *all[] = {&g->keys, &g->values, &g->seps, &g->row_offset, &g->in_k, &g->in_v, &g->sk0, &g->sv0,
                  &g->sk1, &g->sv1, &g->win, &g->blkcnt, &g->blkoff, &g->sortwork, &g->msd_kp, &g->msd_rk,
                  &g->msd_mat, &g->msd_bstart, &g->msd_status, &g->msd_spl, &g->ccstat, &g->pk0, &g->pv0,
                  &g->pw0, &g->pd0, &g->pk1, &g->pv1, &g->pw1, &g->pd1, &g->tmpk, &g->tmpv, &g->chunkcnt,
                  &g->chunkoff, &g->lev, &g->fr0, &g->fr1, &g->fb0, &g->fb1, &g->bcnt, &g->bcntb,
                  &g->ms_seen, &g->ms_fr0, &g->ms_fr1, &g->ms_nx, &g->ms_lv16, &g->ms_q0, &g->ms_q1, &g->ms_qb0,
                  &g->ms_qb1, &g->ms_touched, &g->ms_cnt, &g->ms_cntb, &g->ms_tcnt, &g->ms_partial, &g->ms_starts,
                  &g->ms_out, &g->ms_reach, &g->ms_lv32, &g->ms_ev, &g->ms_evn, &g->ctr, &g->digest, &g->step_n,
                  &g->step_stage,
                  &g->wide};
    drop_graphs(g);
    for (Buf *b : all) b->release(&g->mem);
    if (g->hctr) (void)hipHostFree(g->hctr);
    if (g->sig) (void)hipHostFree(g->sig);
    if (g->hcnt) (void)hipHostFree(g->hcnt);
    if (g->ms_hcnt) (void)hipHostFree(g->ms_hcnt);
    if (g->st) (void)hipStreamSynchronize(g->st);
    for (hipEvent_t e : g->ev_all) (void)hipEventDestroy(e);
    if (g->st) hipStreamDestroy(g->st);
    delete g;
    return GPMA_OK;
}

int gpma_update(gpma_t g, const uint64_t *keys, const uint8_t *values, uint64_t n, int on_device) {
    if (!g) return fail(GPMA_ERR_ARG, "gpma_update: null handle");
    CHECK_ALIVE(g, "gpma_update");
    return do_update(g, keys, values, n, on_device);
}

int gpma_update_batches(gpma_t g, const uint64_t *keys, const uint8_t *values, const uint64_t *offsets,
                        uint32_t nbatches, int on_device) {
    if (!g || !offsets) return fail(GPMA_ERR_ARG, "gpma_update_batches: null argument");
    CHECK_ALIVE(g, "gpma_update_batches");
    if (nbatches == 0) return GPMA_OK;
    if (offsets[0] != 0) return fail(GPMA_ERR_ARG, "gpma_update_batches: offsets[0] must be 0");
    for (uint32_t b = 0; b < nbatches; b++)
        if (offsets[b + 1] < offsets[b]) return fail(GPMA_ERR_ARG, "gpma_update_batches: offsets must not decrease");
    // Batches are contiguous in (batch, position) order, which the stable sort
    // keeps within each key: the last record of a key is its last update.
    return do_update(g, keys, values, offsets[nbatches], on_device);
}

int gpma_bfs(gpma_t g, uint32_t start, uint32_t *levels, int on_device, uint32_t *reachable) {
    if (!g) return fail(GPMA_ERR_ARG, "gpma_bfs: null handle");
    CHECK_ALIVE(g, "gpma_bfs");
    if (start >= g->N) return fail(GPMA_ERR_ARG, "gpma_bfs: start node >= row_num");
    HIP_TRY(hipSetDevice(g->device));
    TRY(ensure_bfs(g));
    uint32_t reach = 0;
    TRY(run_bfs(g->keys.as<gkey>(), g->values.as<gval>(), g->C, g->row_offset.as<uint32_t>(), g->N, start,
                bfs_state(g), g->hcnt, g->st, &reach, g->pull_fraction, g));
    if (levels) {
        HIP_TRY(hipMemcpyAsync(levels, g->lev.p, (uint64_t)g->N * 4,
                               on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, g->st));
        HIP_TRY(hipStreamSynchronize(g->st));
    }
    if (reachable) *reachable = reach;
    return GPMA_OK;
}

int gpma_bfs_multi(gpma_t g, const uint32_t *starts, uint32_t nsrc, uint32_t *levels, int on_device,
                   uint32_t *reachable) {
    if (!g || !starts) return fail(GPMA_ERR_ARG, "gpma_bfs_multi: null argument");
    CHECK_ALIVE(g, "gpma_bfs_multi");
    if (nsrc == 0 || nsrc > 64) return fail(GPMA_ERR_ARG, "gpma_bfs_multi: nsrc must be in [1, 64]");
    for (uint32_t i = 0; i < nsrc; i++)
        if (starts[i] >= g->N) return fail(GPMA_ERR_ARG, "gpma_bfs_multi: start node >= row_num");
    return bfs_multi_impl(g, starts, nsrc, levels, on_device, reachable);
}

int gpma_step(gpma_t g, const uint64_t *keys, const uint8_t *values, const uint64_t *offsets, uint32_t nbatches,
              int on_device, const uint32_t *starts, uint32_t nsrc, uint32_t *levels, int levels_on_device,
              uint32_t *reachable) {
    if (!g || !offsets) return fail(GPMA_ERR_ARG, "gpma_step: null argument");
    CHECK_ALIVE(g, "gpma_step");
    if (offsets[0] != 0) return fail(GPMA_ERR_ARG, "gpma_step: offsets[0] must be 0");
    for (uint32_t b = 0; b < nbatches; b++)
        if (offsets[b + 1] < offsets[b]) return fail(GPMA_ERR_ARG, "gpma_step: offsets must not decrease");
    const uint64_t n = nbatches ? offsets[nbatches] : 0;
    if (n && (!keys || !values)) return fail(GPMA_ERR_ARG, "gpma_step: null keys or values");
    if (n >= (1ull << 30)) return fail(GPMA_ERR_CAPACITY, "gpma_step: step larger than 2^30 records");
    if (nsrc > 64) return fail(GPMA_ERR_ARG, "gpma_step: nsrc must be in [0, 64]");
    if (nsrc && !starts) return fail(GPMA_ERR_ARG, "gpma_step: null starts");
    for (uint32_t i = 0; i < nsrc; i++)
        if (starts[i] >= g->N) return fail(GPMA_ERR_ARG, "gpma_step: start node >= row_num");
    HIP_TRY(hipSetDevice(g->device));
    return run_step(g, keys, values, n, on_device, starts, nsrc, levels, levels_on_device, reachable);
}

int gpma_step_prepare(gpma_t g, uint64_t max_records, uint32_t nsrc, uint32_t *levels, uint32_t bfs_levels) {
    if (!g) return fail(GPMA_ERR_ARG, "gpma_step_prepare: null handle");
    CHECK_ALIVE(g, "gpma_step_prepare");
    if (nsrc > 64) return fail(GPMA_ERR_ARG, "gpma_step_prepare: nsrc must be in [0, 64]");
    if (bfs_levels > STEP_MAX_BLEVELS) return fail(GPMA_ERR_ARG, "gpma_step_prepare: bfs_levels above 64");
    HIP_TRY(hipSetDevice(g->device));
    if (bfs_levels) g->step_blevels = bfs_levels;
    if (!msd_sort_fits(g->N)) return GPMA_OK;  // such steps take the host path
    const uint64_t top = std::min<uint64_t>(std::max<uint64_t>(max_records, 1), MSD_MAX_N);
    // size every buffer for the largest class first: growing one later would
    // move it and leave the graphs captured before unusable
    uint32_t top_cls = STEP_MIN_CLASS;
    while ((1ull << top_cls) < top) top_cls++;
    TRY(ensure_update_buffers(g, 1ull << top_cls));
    TRY(ensure_msd(g, 1ull << top_cls, msd_plan(1ull << top_cls, g->N, g->sort_bits, true, g->sort_ipt)));
    if (nsrc) {
        TRY(ensure_msbfs(g));
        if (!levels) TRY(g->ms_out.ensure((uint64_t)nsrc * g->N * 4, &g->mem));
    }
    for (uint32_t cls = STEP_MIN_CLASS;; cls++) {
        uint32_t *out = nullptr;
        uint32_t bl = 0;
        gpma_handle::StepGraph *sg = nullptr;
        TRY(step_graph(g, cls, nsrc, levels, &sg, &out, &bl));
        if ((1ull << cls) >= top) break;
    }
    return GPMA_OK;
}

int gpma_get_step_stats(gpma_t g, gpma_step_stats *out) {
    if (!g || !out) return fail(GPMA_ERR_ARG, "gpma_get_step_stats: null argument");
    *out = g->sst;
    out->graphs = (uint32_t)g->graphs.size();
    out->bfs_levels = g->step_blevels;
    return GPMA_OK;
}

int gpma_csr_bfs(const uint64_t *keys, const uint8_t *values, const uint32_t *row_offsets, uint32_t node_size,
                 uint32_t edge_size, uint32_t start_node, uint32_t *results) {
    (void)edge_size;
    if (!keys || !values || !row_offsets || !results) return fail(GPMA_ERR_ARG, "gpma_csr_bfs: null pointer");
    if (start_node >= node_size) return fail(GPMA_ERR_ARG, "gpma_csr_bfs: start_node >= node_size");
    // The caller's array length is unknown here, so only the push kernel runs.
    hipStream_t st = nullptr;
    HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    const uint64_t n = (uint64_t)node_size + 1;
    const uint64_t cw = n + BFS_MAX_CHUNK + 2;
    const uint64_t words = 5 * n + 4 * cw + 2;
    uint32_t *buf = nullptr;
    unsigned long long *hc = nullptr;
    int rc = GPMA_OK;
    if (hipMalloc((void **)&buf, words * 4) != hipSuccess) rc = fail(GPMA_ERR_OOM, "gpma_csr_bfs: hipMalloc failed");
    if (rc == GPMA_OK && hipHostMalloc((void **)&hc, 2 * (BFS_MAX_CHUNK + 1) * 8, hipHostMallocDefault) != hipSuccess)
        rc = fail(GPMA_ERR_OOM, "gpma_csr_bfs: hipHostMalloc failed");
    if (rc == GPMA_OK) {
        uint32_t *p = buf;
        BfsState b;
        b.levels = p;
        p += n;
        b.q[0] = p;
        p += n;
        b.q[1] = p;
        p += n;
        b.qbig[0] = p;
        p += n;
        b.qbig[1] = p;
        p += n;
        if ((uintptr_t)p & 7) p++;
        b.cnt = (unsigned long long *)p;
        p += 2 * cw;
        b.cnt_big = (unsigned long long *)p;
        rc = run_bfs(keys, values, 0, row_offsets, node_size, start_node, b, hc, st, nullptr, 0.0);
        if (rc == GPMA_OK && (hipMemcpyAsync(results, b.levels, (uint64_t)node_size * 4, hipMemcpyDeviceToHost, st) !=
                                  hipSuccess ||
                              hipStreamSynchronize(st) != hipSuccess))
            rc = fail(GPMA_ERR_HIP, "gpma_csr_bfs: copy-out failed");
    }
    if (buf) (void)hipFree(buf);
    if (hc) (void)hipHostFree(hc);
    (void)hipStreamDestroy(st);
    return rc;
}

int gpma_csr_bfs_noalloc(const uint64_t *keys, const uint8_t *values, const uint32_t *row_offsets, uint32_t node_size,
                         uint32_t edge_size, uint32_t start_node, uint32_t *levels, uint32_t *reachable,
                         void *stream) {
    (void)edge_size;
    if (!keys || !values || !row_offsets || !levels)
        return fail(GPMA_ERR_ARG, "gpma_csr_bfs_noalloc: null pointer");
    if (start_node >= node_size) return fail(GPMA_ERR_ARG, "gpma_csr_bfs_noalloc: start_node >= node_size");
    if (node_size > NOALLOC_MAX_NODES)
        return fail(GPMA_ERR_CAPACITY, "gpma_csr_bfs_noalloc: node_size " + std::to_string(node_size) +
                                           " exceeds the preallocated maximum of " +
                                           std::to_string(NOALLOC_MAX_NODES) + " nodes");
    void *base = nullptr;
    TRY(noalloc_scratch(&base));
    unsigned long long *cnt = (unsigned long long *)base;
    uint32_t *q = (uint32_t *)(cnt + 2 * NA_CNT);
    BfsState b;
    b.levels = levels;  // the caller's output doubles as the claim words
    b.cnt = cnt;
    b.cnt_big = cnt + NA_CNT;
    b.q[0] = q;
    b.q[1] = q + NA_Q;
    b.qbig[0] = q + 2 * NA_Q;
    b.qbig[1] = q + 3 * NA_Q;
    unsigned long long hc[2 * (BFS_MAX_CHUNK + 1)];  // per-chunk counts read back into host stack memory
    uint32_t reach = 0;
    // the caller's array length is unknown here, so only the push kernel runs
    TRY(run_bfs(keys, values, 0, row_offsets, node_size, start_node, b, hc, (hipStream_t)stream, &reach, 0.0));
    if (reachable) *reachable = reach;
    return GPMA_OK;
}

int gpma_get_csr(gpma_t g, gpma_csr *out) {
    if (!g || !out) return fail(GPMA_ERR_ARG, "gpma_get_csr: null argument");
    CHECK_ALIVE(g, "gpma_get_csr");
    out->keys = g->keys.as<uint64_t>();
    out->values = g->values.as<uint8_t>();
    out->row_offset = g->row_offset.as<uint32_t>();
    out->capacity = g->C;
    out->segment_length = g->S;
    out->tree_height = g->h;
    out->row_num = g->N;
    out->reserved = 0;
    return GPMA_OK;
}

int gpma_copy_layout(gpma_t g, uint64_t *keys, uint8_t *values, uint32_t *row_offset) {
    if (!g) return fail(GPMA_ERR_ARG, "gpma_copy_layout: null handle");
    CHECK_ALIVE(g, "gpma_copy_layout");
    HIP_TRY(hipSetDevice(g->device));
    if (keys) HIP_TRY(hipMemcpyAsync(keys, g->keys.p, g->C * 8, hipMemcpyDeviceToHost, g->st));
    if (values) HIP_TRY(hipMemcpyAsync(values, g->values.p, g->C, hipMemcpyDeviceToHost, g->st));
    if (row_offset)
        HIP_TRY(hipMemcpyAsync(row_offset, g->row_offset.p, ((uint64_t)g->N + 1) * 4, hipMemcpyDeviceToHost, g->st));
    HIP_TRY(hipStreamSynchronize(g->st));
    return GPMA_OK;
}

int gpma_export_live(gpma_t g, uint64_t *out, uint64_t cap, uint64_t *n) {
    if (!g || !n) return fail(GPMA_ERR_ARG, "gpma_export_live: null argument");
    CHECK_ALIVE(g, "gpma_export_live");
    HIP_TRY(hipSetDevice(g->device));
    uint64_t c = 0;
    TRY(count_live(g, 0, g->C, &c));
    *n = c;
    if (!out) return GPMA_OK;
    if (cap < c) return fail(GPMA_ERR_CAPACITY, "gpma_export_live: buffer holds " + std::to_string(cap) +
                                                     " keys, " + std::to_string(c) + " live");
    launch_compact_live(g->keys.as<gkey>(), g->values.as<gval>(), 0, g->C, g->chunkoff.as<uint32_t>(),
                        g->tmpk.as<gkey>(), g->tmpv.as<gval>(), g->st);
    HIP_TRY(hipMemcpyAsync(out, g->tmpk.p, c * 8, hipMemcpyDeviceToHost, g->st));
    HIP_TRY(hipStreamSynchronize(g->st));
    HIP_TRY(hipGetLastError());
    return GPMA_OK;
}

int gpma_live_digest(gpma_t g, uint64_t *count, uint64_t *sum, uint64_t *xr) {
    if (!g) return fail(GPMA_ERR_ARG, "gpma_live_digest: null handle");
    CHECK_ALIVE(g, "gpma_live_digest");
    HIP_TRY(hipSetDevice(g->device));
    TRY(g->digest.ensure(3 * 8, &g->mem));
    launch_live_digest(g->keys.as<gkey>(), g->values.as<gval>(), g->C, g->digest.as<unsigned long long>(), g->st);
    uint64_t h[3] = {0, 0, 0};
    HIP_TRY(hipMemcpyAsync(h, g->digest.p, sizeof(h), hipMemcpyDeviceToHost, g->st));
    HIP_TRY(hipStreamSynchronize(g->st));
    HIP_TRY(hipGetLastError());
    if (count) *count = h[0];
    if (sum) *sum = h[1];
    if (xr) *xr = h[2];
    return GPMA_OK;
}

int gpma_test_fail_allocs(gpma_t g, uint32_t skip, uint32_t count) {
    if (!g) return fail(GPMA_ERR_ARG, "gpma_test_fail_allocs: null handle");
    g->mem.skip = skip;
    g->mem.fail = count;
    return GPMA_OK;
}

int gpma_get_stats(gpma_t g, gpma_stats *out) {
    if (!g || !out) return fail(GPMA_ERR_ARG, "gpma_get_stats: null argument");
    out->capacity = g->C;
    out->segment_length = g->S;
    out->tree_height = g->h;
    out->row_num = g->N;
    out->last_levels = g->last_levels;
    out->updates_applied = g->applied;
    out->invalid_skipped = g->invalid;
    out->dropped = g->dropped;
    out->resizes = g->resizes;
    out->bulk_loads = g->bulk;
    out->device_bytes = g->mem.bytes;
    return GPMA_OK;
}

int gpma_sync(gpma_t g) {
    if (!g) return fail(GPMA_ERR_ARG, "gpma_sync: null handle");
    HIP_TRY(hipStreamSynchronize(g->st));
    return GPMA_OK;
}

void *gpma_stream(gpma_t g) { return g ? (void *)g->st : nullptr; }

int gpma_copy_async(gpma_t g, void *dst, const void *src, uint64_t bytes, int kind) {
    if (!g || (!dst && bytes) || (!src && bytes)) return fail(GPMA_ERR_ARG, "gpma_copy_async: null argument");
    if (kind < 0 || kind > 2) return fail(GPMA_ERR_ARG, "gpma_copy_async: kind must be 0, 1 or 2");
    if (!bytes) return GPMA_OK;
    HIP_TRY(hipSetDevice(g->device));
    const hipMemcpyKind k[3] = {hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice};
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, k[kind], g->st));
    return GPMA_OK;
}

int gpma_device_of(gpma_t g) { return g ? g->device : -1; }

}  // extern "C"
