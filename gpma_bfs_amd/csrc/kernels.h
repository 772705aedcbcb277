// Host-side launchers for the gfx950 kernels (all asynchronous on `st`).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "common.h"

namespace gpma {

// Digit shifts (bit positions of 8-bit digits, LSD first) for one sort.
struct SortShifts {
    uint32_t s[8];
};

void fill_u32(uint32_t *p, uint32_t v, uint64_t n, hipStream_t st);
void fill_u8(uint8_t *p, uint8_t v, uint64_t n, hipStream_t st);
void fill_u64(uint64_t *p, uint64_t v, uint64_t n, hipStream_t st);
// dk[0..n) = sk[0..n), dv[0..n) = sv[0..n) and *dn = n by one launch (device buffers)
void stage_step(uint64_t *dk, const uint64_t *sk, uint8_t *dv, const uint8_t *sv, uint64_t n, uint32_t *dn,
                hipStream_t st);
// a[0..na) (u32), b[0..nb) and c[0..nc) (u64) set to zero by one launch
void zero3(uint32_t *a, uint64_t na, uint64_t *b, uint64_t nb, uint64_t *c, uint64_t nc, hipStream_t st);

// ---- sort.hip
uint64_t sort_status_words(uint64_t n, uint32_t npass);
// Validate + copy (ik, iv) into (k0, v0) (keys with src/dst >= N become
// KEY_NONE and are counted in *invalid; with guards_ok dst == 0xFFFFFFFF is
// accepted), then stable-sort by the digits at `shifts`. Returns which
// buffer pair holds the result (0: k0/v0, 1: k1/v1).
int stage_and_sort(const gkey *ik, const gval *iv, gkey *k0, gval *v0, gkey *k1, gval *v1, uint64_t n, uint32_t N,
                   bool guards_ok, uint32_t *invalid, const SortShifts &shifts, uint32_t npass, uint32_t *work,
                   uint32_t *d_err, hipStream_t st);

// ---- msd.hip
// Bucket sort for row_num < 2^28 (keys and values packed into one u64).
bool msd_sort_fits(uint32_t N);
struct MsdPlan {
    uint32_t bs;      // bits per packed src / dst field
    uint32_t D;       // bucket bits
    bool by_row;      // buckets from row_offset (the row's slot position), not the key's top bits
    uint32_t R;       // bits sorted inside a bucket
    uint32_t bshift;  // shift of the bucket bits in a packed key
    uint32_t ipt;     // items per lane of the rank and scatter tiles (4, 8 or 16)
    uint32_t ntiles;  // tiles of 256 * ipt items in the batch
    uint32_t NB;      // buckets
    uint32_t nscan;   // scan blocks (status words used)
};
// bucket_bits 0: from n. by_row: buckets by the row's slot position
// (row_offset valid, capacity C), balanced for skewed sources; else by the
// key's top bits (the guard insertion, before row_offset exists).
// items_per_lane: 4, 8 or 16 items per lane of the rank and scatter tiles (0: from n).
MsdPlan msd_plan(uint64_t n, uint32_t N, uint32_t bucket_bits, bool by_row, uint32_t items_per_lane = 0);
struct MsdWork {
    uint64_t *kp;               // n packed keys
    uint32_t *rk;               // n (bucket << 16 | rank in (tile, bucket))
    uint64_t *kb;               // n bucketed packed keys
    uint32_t *mat;              // ntiles * NB u16 counts, then ntiles * NB u32 offsets (2 * ntiles * NB words)
    uint32_t *bstart;           // NB + 1 bucket starts
    unsigned long long *status; // nscan look-back words (zeroed once)
};
// Row splitters of the row buckets: SPL_TABLE u32 entries from row_offset
// (N + 1 entries) for an array of C slots (k_msd_splitters).
constexpr uint32_t SPL_BITS = 12;
constexpr uint32_t SPL_TABLE = 1u << SPL_BITS;
void msd_splitters(const uint32_t *row_offset, uint32_t N, uint64_t C, uint32_t *spl, hipStream_t st);
// Sorted (ok, ov) from (ik, iv); same contract as stage_and_sort. spl: the
// splitter table (by_row plans). epoch:
// non-zero tag distinct from earlier calls on the same status words;
// *ticket zero before the call; *maxb raised to the largest bucket above 1024
// items. dn (optional): the batch size is read from *dn on the device, and
// n and the plan are an upper bound of it (launch shapes fixed in advance).
void msd_sort(const gkey *ik, const gval *iv, gkey *ok, gval *ov, const MsdPlan &pl, const MsdWork &w, uint64_t n,
              const uint32_t *dn, uint32_t N, bool guards_ok, const uint32_t *spl, uint32_t *invalid, uint32_t epoch,
              uint32_t *ticket, uint32_t *err, uint32_t *maxb, hipStream_t st);

// ---- update.hip
constexpr uint32_t CMP_BLOCK = 256;  // items per compaction block
// win of a record that heads a leaf merged by launch_apply: WIN_SEP | leaf
// index (leaf indices stay below 2^27, slot indices below 2^31).
constexpr uint32_t WIN_SEP = 0x80000000u;
constexpr uint32_t WIN_MAX = 4096;   // widest window rebalanced by one workgroup in LDS
constexpr uint32_t CHUNK = 4096;     // slots per workgroup in device-wide window passes

// Locate + in-place delete/modify. win[i] = leaf base or SIZE_NONE. With
// row_hint the search starts from row_offset (valid guards in the array).
// dn (optional): the record count is *dn (<= n) on the device.
void launch_locate(const gkey *ks, const gval *vs, uint64_t n, const uint32_t *dn, gkey *keys, gval *values,
                   const gkey *seps, uint64_t nseg, uint32_t S, const uint32_t *row_offset, bool row_hint,
                   uint32_t *win, hipStream_t st);
// Completion signal: copies a[0..na) then b[0..nb) (u32 words) to dst[1..]
// (host-coherent pinned memory), zeroes a[0..na) when clear_a (a must then
// be writable device memory), and then stores seq to dst[0] with
// system-scope release, after everything queued before it on the stream.
void launch_publish(const uint32_t *a, uint32_t na, const uint32_t *b, uint32_t nb, uint32_t *dst, uint32_t seq,
                    bool clear_a, hipStream_t st);
// stage[0..4) = *cnt, *cnt_big as u32 pairs; stage[4 .. 4 + n) = reach[0 .. n).
void launch_step_gather(const unsigned long long *cnt, const unsigned long long *cnt_big, const uint32_t *reach,
                        uint32_t n, uint32_t *stage, hipStream_t st);
// *gate = 1 when the update before it left work for the host: a bulk batch
// (*np0 >= bulk_threshold), inserts pending above the device-driven levels
// (*np_top != 0) or a look-back error (*err != 0); else 0.
void launch_step_gate(const uint32_t *np0, const uint32_t *np_top, const uint32_t *err, uint32_t bulk_threshold,
                      uint32_t *gate, hipStream_t st);
// Exclusive scan of per-chunk counts (one workgroup); *total gets the sum.
// The chunk count is nb, or ceil(*d_items / CMP_BLOCK) when d_items is given.
void launch_scan_blocks(const uint32_t *cnt, uint64_t nb, uint32_t *off, uint32_t *total, hipStream_t st,
                        const uint32_t *d_items = nullptr);
// Stable compaction of entries whose win != SIZE_NONE (idone == nullptr) or
// whose done flag is 0, writing win & mask and clearing odone. Entries at or
// beyond *d_n (or n_ub when d_n is null) are ignored; grid-stride over chunks.
void launch_compact(const gkey *ik, const gval *iv, const uint32_t *iw, const uint8_t *idone, uint64_t n_ub,
                    const uint32_t *d_n, const uint32_t *off, gkey *ok, gval *ov, uint32_t *ow, uint8_t *odone,
                    uint32_t mask, hipStream_t st);
// Per-chunk counts of entries with done == 0 among the first *d_n.
// Single-pass stable compaction of the first *d_n entries (n_ub when d_n is
// null) keeping done == 0 (idone given) or win != SIZE_NONE; writes win & mask,
// clears the output done flags and stores the kept count in *d_out. status:
// compact_status_words(n_ub) words, zeroed once at allocation; epoch: non-zero
// and distinct per call; *ticket zero before the call; *err set if the
// look-back spin limit is exceeded. With fix (after launch_apply): keeps
// win < WIN_SEP and writes seps[leaf] = fix[i] for win = WIN_SEP | leaf.
uint64_t compact_status_words(uint64_t n_ub);
void launch_compact_chained(const gkey *ik, const gval *iv, const uint32_t *iw, const uint8_t *idone, uint64_t n_ub,
                            const uint32_t *d_n, gkey *ok, gval *ov, uint32_t *ow, uint8_t *odone, uint32_t mask,
                            unsigned long long *status, uint32_t epoch, uint32_t *ticket, uint32_t *d_out,
                            uint32_t *err, hipStream_t st, const gkey *fix = nullptr, gkey *seps = nullptr);
void launch_count_pending(const uint8_t *done, uint64_t n_ub, const uint32_t *d_n, uint32_t *blkcnt,
                          hipStream_t st);
// Level-0 rebalance: one thread per leaf run; does nothing when *d_np >=
// bulk_threshold (that batch takes the bulk path).
// Locate and level 0 in one pass, for S = 16 or 32 with row hints: every
// sorted record finds its leaf; the first record of each leaf's run applies
// the run's deletes and modifications and merges its pending inserts into
// the leaf when level 0's bounds allow (a run at a tile edge leaves its new
// first key in fix[head], win[head] = WIN_SEP | leaf, for the compaction),
// else leaves them pending (win = leaf base);
// other records get SIZE_NONE. Adds the level-0 pending count to *np0.
void launch_apply(uint32_t S, const gkey *ks, const gval *vs, uint64_t n_ub, const uint32_t *d_n, gkey *keys,
                  gval *values, gkey *seps, uint64_t nseg, uint32_t *row_offset, uint32_t lower, uint32_t upper,
                  uint32_t *win, gkey *fix, uint32_t *np0, hipStream_t st);
void launch_leaf(uint32_t S, const gkey *pk, const gval *pv, const uint32_t *pw, uint8_t *pd, uint64_t n_ub,
                 const uint32_t *d_np, gkey *keys, gval *values, gkey *seps, uint32_t *row_offset, uint32_t lower,
                 uint32_t upper, uint32_t bulk_threshold, hipStream_t st);
// Device-driven levels: windows of at most this many slots, one wave each.
constexpr uint32_t LEVEL_WIN_MAX = 256;
// Density bounds of levels 1..8 (index L - 1), in items per window.
struct LevelBounds {
    uint32_t lower[8], upper[8];
};
// Device-driven levels 1..nl (S << nl <= LEVEL_WIN_MAX) in one launch over
// the level-1 list itself (d_lvl[1] entries, level-1 windows in pw, done
// flags in pd): one wave per region of S << nl slots runs the levels in
// order; placed entries get their done flag, and the inserts level L could
// not place are added to d_lvl[L + 1]. Nothing runs when d_lvl[1] == 0 or
// d_gate is set and *d_gate >= gate.
// Looks up (once per device) how many workgroups of the region kernel the
// current device holds at once; the region launches use it.
void window_region_prepare();
void launch_window_region(uint32_t S, uint32_t nl, const gkey *pk, const gval *pv, const uint32_t *pw, uint8_t *pd,
                          uint64_t n_ub, uint32_t *d_lvl, gkey *keys, gval *values, gkey *seps, uint32_t *row_offset,
                          const LevelBounds &lb, const uint32_t *d_gate, uint32_t gate, hipStream_t st);
// Levels whose window fits in LDS: workgroups stride over the first *d_np
// pending entries, one window per run head. Does nothing when d_gate is set
// and *d_gate >= gate (that batch takes the bulk path).
void launch_window(uint32_t S, uint32_t W, const gkey *pk, const gval *pv, const uint32_t *pw, uint8_t *pd,
                   uint64_t n_ub, const uint32_t *d_np, gkey *keys, gval *values, gkey *seps,
                   uint32_t *row_offset, uint32_t lower, uint32_t upper, const uint32_t *d_gate, uint32_t gate,
                   hipStream_t st);
// Run heads of the pending list: heads[j] = index of the j-th run start.
void launch_run_heads(const uint32_t *pw, uint64_t n_ub, const uint32_t *d_np, uint32_t *flags_cnt,
                      uint32_t *off, uint32_t *heads, uint32_t *d_nruns, hipStream_t st);
// Device-wide window passes (window [w, w+W) of the array).
void launch_count_live(const gkey *keys, const gval *values, uint64_t w, uint64_t W, uint32_t *chunkcnt,
                       hipStream_t st);
void launch_compact_live(const gkey *keys, const gval *values, uint64_t w, uint64_t W, const uint32_t *chunkoff,
                         gkey *tk, gval *tv, hipStream_t st);
// Merge c compacted items with r sorted inserts and redistribute them into
// window [w, w+W) of (keys, values), writing gaps, row_offset and seps.
void launch_merge_redistribute(const gkey *tk, const gval *tv, uint64_t c, const gkey *ik, const gval *iv,
                               uint64_t r, gkey *keys, gval *values, gkey *seps, uint32_t *row_offset,
                               uint64_t w, uint64_t W, uint32_t S, hipStream_t st);
void launch_mark_done(uint8_t *pd, uint64_t a, uint64_t b, hipStream_t st);
// Scratch of one wide rebalance level (launch_wide_level): heads and counts
// of the pending list's runs, per-chunk live counts and offsets (up to
// C / CHUNK + 1 entries), per-window live count, merge size, pass flag and
// merge prefix (up to the pending count + 1 entries), the compaction buffer
// (C items).
struct WideWork {
    uint32_t *blkcnt, *blkoff;  // run-head counting (pending count / CMP_BLOCK + 1 words each)
    uint32_t *heads, *d_nruns;
    uint32_t *chunkcnt, *chunkoff;
    uint32_t *live, *merge;
    uint8_t *ok;
    unsigned long long *mpre;
    gkey *tk;
    gval *tv;
};
// One rebalance level whose windows (W = S << level > WIN_MAX slots) are
// wider than a workgroup, over every window the pending list names, with no
// host round trip: windows whose merge size lies in [lower, upper] are
// compacted, merged with their pending inserts and redistributed, and those
// inserts marked done (pd). The pending count is *d_np (<= n_ub).
void launch_wide_level(const gkey *pk, const gval *pv, const uint32_t *pw, uint8_t *pd, uint64_t n_ub,
                       const uint32_t *d_np, uint64_t W, uint32_t S, uint32_t lower, uint32_t upper, gkey *keys,
                       gval *values, gkey *seps, uint32_t *row_offset, const WideWork &w, hipStream_t st);
// out[0..3) = (count, wrapping sum, xor) of the live keys of keys[0 .. C).
void launch_live_digest(const gkey *keys, const gval *values, uint64_t C, unsigned long long *out, hipStream_t st);
// Write c compacted items to keys[0..c), KEY_NONE after; rebuild seps and
// point row_offset at the compacted guards.
void launch_write_compacted(const gkey *tk, const gval *tv, uint64_t c, gkey *keys, gval *values, uint64_t C,
                            gkey *seps, uint32_t *row_offset, uint32_t S, hipStream_t st);

// ---- bfs.hip
// Device state of one BFS: levels[N] (0 = not reached; also the claim
// words), two frontier queues per parity (short ranges / long ranges) and
// per-level counts and slot work.
struct BfsState {
    uint32_t *levels;
    uint32_t *q[2], *qbig[2];
    // per level: low 32 bits = queued vertices, high 32 bits = their slot count
    unsigned long long *cnt, *cnt_big;
    uint32_t tile_cap = 64;  // most frontier vertices per push tile (16..64)
};
// levels = 0 except start (1); start queued for level 1.
void launch_bfs_init(const BfsState &b, uint32_t N, uint32_t start, const uint32_t *row_offset, hipStream_t st);
void launch_bfs_clear(const BfsState &b, uint32_t a, uint32_t e, hipStream_t st);
// Expand level `level` into level+1: push when the level's slot count is at
// most pull_threshold, else stream all C slots (pull; C = 0 disables it).
void launch_bfs_level(const gkey *keys, const gval *values, uint64_t C, uint32_t N, const uint32_t *row_offset,
                      const BfsState &b, uint32_t level, unsigned long long pull_threshold, hipStream_t st);

// ---- msbfs.hip
// Device state of one multi-source BFS (<= 64 sources, bit i = starts[i]).
struct MsState {
    uint64_t *seen, *fr[2], *nx;
    uint4 *ev;                 // [8][N] level events, slot-major: fresh source mask (x, y), level (z)
    uint8_t *evn;              // [N] events recorded (flag 0x80: more levels in the overflow row)
    uint16_t *lv16;            // [N][64] overflow levels, vertex-major
    uint32_t *lv32;            // [N][64] u32 levels once the traversal passed level 65,534 (else null)
    uint32_t *q[2], *qbig[2];  // frontier lists by parity (short / long ranges)
    uint32_t *touched;         // vertices whose nx went non-zero this level
    unsigned long long *cnt, *cnt_big;  // per level: low = vertices, high = slots
    uint32_t *tcnt;            // per level touched count
    uint32_t *partial;         // [ms_output_blocks(N)][64] reachable partials
    uint32_t tile_cap = 64;    // most frontier vertices per push tile (16..64)
};
// gate (optional): when *gate != 0 no source is queued, so every level finds
// an empty frontier.
void ms_init(const MsState &m, uint32_t N, const uint32_t *starts, uint32_t nsrc, const uint32_t *ro, hipStream_t st,
             const uint32_t *gate = nullptr);
// A captured step's form of ms_init: the source setup also computes the
// step's host-work flag (*gate = 1 when *np0 >= bulk, *np_top != 0 or
// *err != 0; then no source is queued) and clears the counts of levels
// clear_a..clear_b-1 (what launch_step_gate and ms_clear did as launches of
// their own).
struct MsStepGate {
    const uint32_t *np0 = nullptr, *np_top = nullptr, *err = nullptr;
    uint32_t bulk = 0;
};
void ms_init_step(const MsState &m, uint32_t N, const uint32_t *starts, uint32_t nsrc, const uint32_t *ro,
                  const MsStepGate &sg, uint32_t *gate, uint32_t clear_a, uint32_t clear_b, hipStream_t st);
void ms_clear(const MsState &m, uint32_t a, uint32_t b, hipStream_t st);
void ms_level(const gkey *keys, const gval *values, uint64_t C, uint32_t N, const uint32_t *ro, const MsState &m,
              uint32_t level, unsigned long long pull_threshold, hipStream_t st);
// Source-major u32 levels out[nsrc][N] and reachable counts reach[nsrc].
void ms_output(const MsState &m, uint32_t N, uint32_t nsrc, uint32_t *out, uint32_t *reach, hipStream_t st);
uint32_t ms_output_blocks(uint32_t N);
// Copy every staged level (where seen) into the u32 rows lv32[N][64]; the
// levels after it then go to lv32 (MsState::lv32).
void ms_widen(const MsState &m, uint32_t N, uint32_t *lv32, hipStream_t st);

}  // namespace gpma
