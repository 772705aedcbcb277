/*
 * gpma.h - C ABI of the MI355X (gfx950) GPMA: a gapped, sorted packed array
 * of 64-bit edge keys (src << 32 | dst) with per-row guard keys and a CSR
 * row_offset view, updated in batches and traversed by BFS on the GPU.
 *
 * Every entry point is extern "C", takes plain pointers and sizes, returns
 * an int status (GPMA_OK = 0) and never throws. gpma_last_error() returns a
 * thread-local message for the last failure. One handle is used by one host
 * thread at a time; each handle owns one HIP stream and all calls are
 * synchronous with respect to the caller unless stated otherwise.
 *
 * Key and value encodings follow the reference (utils.cuh:11,19-21;
 * gpma.cuh:15-19): key = src << 32 | dst; value 0 = delete / tombstone,
 * any other value = insert / live. KEY_NONE marks a gap, KEY_MAX is the
 * sentinel, a key whose low word is 0xFFFFFFFF is row v's guard.
 */
#ifndef GPMA_AMD_GPMA_H
#define GPMA_AMD_GPMA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GPMA_OK 0
#define GPMA_ERR_ARG 1       /* bad argument (null pointer, size mismatch, start >= N, ...) */
#define GPMA_ERR_HIP 2       /* a HIP runtime call failed */
#define GPMA_ERR_OOM 3       /* device allocation failed */
#define GPMA_ERR_CAPACITY 4  /* caller buffer too small, or u32 slot index would overflow */
#define GPMA_ERR_KEY_RANGE 5 /* some update keys had src or dst >= row_num; they were skipped */
#define GPMA_ERR_NO_DEVICE 6 /* no HIP device visible */
#define GPMA_ERR_TIMEOUT 7   /* a wait ended before the request completed (gpma_serve.h) */
#define GPMA_ERR_STOPPED 8   /* the server is shutting down (gpma_serve.h) */

#define GPMA_KEY_NONE 0xFFFFFFFFFFFFFFFFull
#define GPMA_KEY_MAX 0xFFFFFFFFFFFFFFFEull
#define GPMA_COL_IDX_NONE 0xFFFFFFFFu

typedef struct gpma_handle *gpma_t;

/* Optional construction parameters; pass NULL for the reference's values. */
typedef struct gpma_config {
    uint32_t struct_size;      /* sizeof(gpma_config) */
    int32_t device;            /* HIP device ordinal, -1 = current device */
    uint32_t bulk_threshold;   /* pending inserts that take the bulk path (reference: 5,000,000) */
    uint32_t pull_permille;    /* BFS streams the whole array when the frontier's slots exceed
                                  this many thousandths of the capacity (0 = 250); the
                                  multi-source BFS uses twice this share (at most all of it) */
    /* Tuning knobs (0 = default). None changes the live-key set or BFS results;
     * segment_length changes the physical layout (the reference's formula is
     * S = 2^(fls(fls C) - 1)). */
    uint32_t segment_length;   /* slots per leaf after a resize / bulk load: 4, 8, 16 or 32 */
    uint32_t device_levels;    /* rebalance levels 1..k queued without a host round trip (1..8, 0 = 3) */
    uint32_t sort_bucket_bits; /* bucket bits of the batch sort (4..12, 0 = from the batch size) */
    uint32_t reserved;
    /* Fields from here on are read when struct_size >= sizeof(gpma_config);
     * a caller built against the first layout (struct_size 32) gets their
     * defaults. None changes the live-key set or BFS results. */
    uint16_t density_permille[4]; /* density thresholds in thousandths (reference gpma.cuh:40-43;
                                     0 = the reference's value): [0] lower bound at the leaves (80),
                                     [1] lower at the root (420), [2] upper at the leaves (920),
                                     [3] upper at the root (840, also the resize target). Needs
                                     62.5 <= [0] <= [1], 2 * [1] <= [3] <= [2] <= 1000 */
    uint32_t sort_items_per_lane; /* items per lane of the batch sort's tiles: 4, 8 or 16 (0 = from the batch size) */
    uint32_t push_tile;           /* most frontier vertices per BFS push tile, 16..64 (0 = 64) */
    uint32_t ms_tile;             /* the same for the multi-source BFS, 16..64 (0 = 64) */
    uint32_t reserved2;
} gpma_config;
#define GPMA_CONFIG_V1_SIZE 32

/* State summary. Pointers are device pointers owned by the handle and are
 * valid until the next gpma_update / gpma_destroy. */
typedef struct gpma_csr {
    const uint64_t *keys;        /* [capacity] */
    const uint8_t *values;       /* [capacity] */
    const uint32_t *row_offset;  /* [row_num + 1] */
    uint64_t capacity;           /* slots, a power of two: segment_length << tree_height */
    uint32_t segment_length;
    uint32_t tree_height;
    uint32_t row_num;
    uint32_t reserved;
} gpma_csr;

typedef struct gpma_stats {
    uint64_t capacity;
    uint32_t segment_length;
    uint32_t tree_height;
    uint32_t row_num;
    uint32_t last_levels;        /* rebalance levels run by the last update */
    uint64_t updates_applied;    /* cumulative update records accepted */
    uint64_t invalid_skipped;    /* cumulative keys skipped for src/dst >= row_num */
    uint64_t dropped;            /* cumulative inserts lost to a failed post-resize root
                                    rebalance (reference gpma.cuh:1034-1044 drops them too) */
    uint64_t resizes;
    uint64_t bulk_loads;
    uint64_t device_bytes;       /* device memory held by the handle */
} gpma_stats;

/* GPMA(row_num) + init_csr_gpma (reference gpma.cuh:54-82): a 4-slot seed
 * tree, then one update inserting the guard key of every row v < row_num. */
int gpma_create(uint32_t row_num, const gpma_config *cfg, gpma_t *out);
int gpma_destroy(gpma_t g);

/* update_gpma(gpma, keys, values) (reference gpma.cuh:968-1050): apply one
 * batch of n updates. keys/values are host pointers (on_device = 0) or
 * device pointers (on_device = 1); they are read only (the reference sorts
 * and compacts the caller's vectors in place; this ABI does not). Keys must
 * have src < row_num and dst < row_num; other keys are skipped and reported
 * with GPMA_ERR_KEY_RANGE after the rest of the batch is applied. Duplicate
 * keys within one batch: the last record in input order wins. */
int gpma_update(gpma_t g, const uint64_t *keys, const uint8_t *values, uint64_t n, int on_device);

/* Several update batches in one call (no reference counterpart; the
 * batched form): batch b is records [offsets[b], offsets[b+1]) of
 * keys/values, offsets has nbatches + 1 entries (host), offsets[0] = 0.
 * The result equals applying the batches one after another with
 * gpma_update (same live keys): all records go through one sort, locate and
 * rebalance pass in which the last record of each key in (batch, position)
 * order decides. */
int gpma_update_batches(gpma_t g, const uint64_t *keys, const uint8_t *values, const uint64_t *offsets,
                        uint32_t nbatches, int on_device);

/* gpma_bfs(gpma, node_size, edge_size, start_node, results) (reference
 * gpma_bfs.cuh:517-533): levels[v] = 1 + hop distance from start over live
 * edges, 0 if unreachable; *reachable = #{v : levels[v] != 0}
 * (gpma_bfs_demo.cu:80). levels is a host (on_device = 0) or device
 * (on_device = 1) array of row_num entries; reachable may be NULL. */
int gpma_bfs(gpma_t g, uint32_t start, uint32_t *levels, int on_device, uint32_t *reachable);

/* Multi-source BFS (no reference counterpart; the batched form's query):
 * nsrc <= 64 sources (host array starts[nsrc], duplicates allowed) share one
 * traversal with a 64-bit source mask per vertex. levels is nsrc x row_num
 * uint32, source-major (levels[i * row_num + v] is what gpma_bfs(starts[i])
 * writes for v), on the host (on_device = 0) or the device (1); levels may
 * be NULL. reachable (host, nsrc entries, may be NULL) gets each source's
 * reachable count. */
int gpma_bfs_multi(gpma_t g, const uint32_t *starts, uint32_t nsrc, uint32_t *levels, int on_device,
                   uint32_t *reachable);

/* One batched step through a captured HIP graph (no reference counterpart;
 * the batched form of the update -> rebalance -> BFS step that the
 * reference's driver runs as update_gpma then gpma_bfs,
 * gpma_bfs_demo.cu:93-104): gpma_update_batches(keys, values, offsets,
 * nbatches, on_device) followed by gpma_bfs_multi(starts, nsrc, levels,
 * levels_on_device, reachable), with the same results and error codes.
 * nsrc = 0 runs the update alone.
 *
 * The device part is one graph launch and one host synchronisation. Graphs
 * are captured on first use and kept per size class (the step's record count
 * rounded up to a power of two, at least 2^16), source count and device
 * levels pointer, for the current array layout (a resize captures anew).
 * The kernels read the step's record count on the device and launches past
 * it exit at once. A graph holds the rebalance levels of every window up to
 * 4096 slots and a fixed number of BFS levels (grown to the deepest BFS seen);
 * what lies beyond them (a bulk load, wider windows, a root resize, a deeper
 * BFS) continues on the host path after the replay. Steps above 4M records,
 * steps while a skewed batch keeps the bucket sort off, and steps while phase
 * profiling is on run the host path only. */
int gpma_step(gpma_t g, const uint64_t *keys, const uint8_t *values, const uint64_t *offsets, uint32_t nbatches,
              int on_device, const uint32_t *starts, uint32_t nsrc, uint32_t *levels, int levels_on_device,
              uint32_t *reachable);

typedef struct gpma_step_stats {
    uint64_t steps;         /* gpma_step calls */
    uint64_t replays;       /* steps run as a graph replay */
    uint64_t captures;      /* graphs captured */
    uint64_t host_steps;    /* steps run on the host path only */
    uint64_t update_tails;  /* replays whose update continued on the host path */
    uint64_t bfs_tails;     /* replays whose BFS continued past the captured levels */
    uint32_t graphs;        /* graphs held now */
    uint32_t bfs_levels;    /* BFS levels the next capture holds */
} gpma_step_stats;

int gpma_get_step_stats(gpma_t g, gpma_step_stats *out);

/* Capture ahead of time the graphs gpma_step would capture on first use, for
 * every size class up to max_records (<= 4M), nsrc sources and the given
 * levels destination (device pointer, or NULL for levels copied to the host
 * or not kept), so no step pays a capture later. bfs_levels > 0 sets the BFS
 * levels a graph holds (at most 64; a BFS deeper than that continues on the
 * host path) for this and later captures. */
int gpma_step_prepare(gpma_t g, uint64_t max_records, uint32_t nsrc, uint32_t *levels, uint32_t bfs_levels);

/* gpma_bfs(keys, values, row_offsets, node_size, edge_size, start_node,
 * results) (reference gpma_bfs.cuh:511-515): BFS over a caller-provided
 * device CSR view in the GPMA encoding. edge_size is accepted for signature
 * compatibility and only used as a sizing hint. results is a host array of
 * node_size entries. */
int gpma_csr_bfs(const uint64_t *keys, const uint8_t *values, const uint32_t *row_offsets, uint32_t node_size,
                 uint32_t edge_size, uint32_t start_node, uint32_t *results);

/* gpma_csr_bfs without any allocation (no reference counterpart: for
 * callers that cannot allocate device or pinned memory, such as tight
 * serving loops and harnesses that forbid hipMalloc). Same traversal and
 * results as gpma_csr_bfs, but levels is a DEVICE array of node_size entries
 * written on `stream` (a hipStream_t; NULL = the default stream), and
 * *reachable (host, may be NULL) gets the reachable count. The frontier
 * queues and level counts live in a static per-device array sized for at
 * most 2,097,152 nodes (64 MiB, held by every GPU the library runs on until
 * the process ends); larger node_size returns GPMA_ERR_CAPACITY before any
 * launch. The call returns after the traversal finished. Calls that overlap
 * on one GPU share that array and race: serialise them. */
int gpma_csr_bfs_noalloc(const uint64_t *keys, const uint8_t *values, const uint32_t *row_offsets,
                         uint32_t node_size, uint32_t edge_size, uint32_t start_node, uint32_t *levels,
                         uint32_t *reachable, void *stream);

/* Public members of the reference's GPMA (gpma.cuh:31-52): device pointers
 * to keys, values, row_offset plus segment_length, tree_height, row_num. */
int gpma_get_csr(gpma_t g, gpma_csr *out);

/* Host copies of the whole layout (GPMA::mirror(), gpma.cuh:90-101): keys and
 * values [capacity], row_offset [row_num + 1]. Any pointer may be NULL. */
int gpma_copy_layout(gpma_t g, uint64_t *keys, uint8_t *values, uint32_t *row_offset);

/* Sorted live keys (key != KEY_NONE and value != 0, guards and the two
 * KEY_MAX sentinels included) into a host buffer of cap entries; *n gets the
 * live count. With out == NULL only the count is returned. */
int gpma_export_live(gpma_t g, uint64_t *out, uint64_t cap, uint64_t *n);

int gpma_get_stats(gpma_t g, gpma_stats *out);

/* Digest of the live-key set (the keys gpma_export_live returns): their
 * count, their sum modulo 2^64 and their xor, computed on the device. A
 * size-independent parity check against a closed-form key set (no
 * reference counterpart). Any output pointer may be NULL. */
int gpma_live_digest(gpma_t g, uint64_t *count, uint64_t *sum, uint64_t *xr);

/* Testing hook for the out-of-memory paths (no reference counterpart):
 * after `skip` more successful device allocations of this handle, the next
 * `count` ones fail as if the device were out of memory (count = 0: off).
 * A rebuild (resize or bulk load) that cannot allocate its new arrays keeps
 * the live items at the old capacity and returns GPMA_ERR_OOM without the
 * pending inserts; if even that fails, every later call on the handle
 * returns GPMA_ERR_OOM. */
int gpma_test_fail_allocs(gpma_t g, uint32_t skip, uint32_t count);

/* Wait for all work queued on the handle's stream. */
int gpma_sync(gpma_t g);

/* Queue a copy on the handle's stream (kind as gpma_memcpy); ordered before
 * the handle's next launches. Pageable host memory makes it synchronous. */
int gpma_copy_async(gpma_t g, void *dst, const void *src, uint64_t bytes, int kind);

/* HIP device ordinal the handle lives on (-1 for a null handle). */
int gpma_device_of(gpma_t g);

/* The HIP stream the handle launches on (as void*, a hipStream_t). */
void *gpma_stream(gpma_t g);

/* Per-phase timing with HIP events on the handle's stream (off by default).
 * Phases: 0 stage, 1 sort, 2 locate, 3 compact, 4 leaf rebalance,
 * 5 upper-level rebalance, 6 rebuild (resize / bulk load), 7 BFS expand.
 * With 16 or 32 slots per segment (after the guard load) the locate and
 * the leaf rebalance are one kernel, counted under 2; 4 then stays empty.
 * gpma_get_profile fills ms[i] (cumulative milliseconds) and calls[i]
 * (kernel launches timed) for i < n; gpma_profile(g, 0) clears nothing,
 * gpma_profile(g, 2) resets the counters and enables. */
#define GPMA_PHASES 8
int gpma_profile(gpma_t g, int enable);
/* Restrict phase timing to the phases whose bit (1 << phase) is set in mask
 * (default: all). Each timed phase adds two event records to the stream, so
 * a caller timing one kernel group masks the others out. */
int gpma_profile_phases(gpma_t g, uint32_t mask);
int gpma_get_profile(gpma_t g, double *ms, uint64_t *calls, uint32_t n);

/* Device-memory plumbing for callers without their own allocator (tests,
 * benchmarks): plain hipMalloc / hipFree / hipMemcpy on the current device.
 * kind: 0 host->device, 1 device->host, 2 device->device. */
int gpma_device_malloc(void **ptr, uint64_t bytes);
/* Free and total bytes of device memory on `device` (-1 = current). */
int gpma_mem_info(int32_t device, uint64_t *free_bytes, uint64_t *total_bytes);
int gpma_device_free(void *ptr);
int gpma_memcpy(void *dst, const void *src, uint64_t bytes, int kind);

/* Number of visible HIP devices (0 when none, or when the runtime fails). */
int gpma_device_count(void);

/* Message of the last failed call on this host thread ("" if none). */
const char *gpma_last_error(void);

/* Seeded generators (gen.cpp). R-MAT: unique, self-loop-free, shuffled;
 * *n_out = number of unique edges (call with src = dst = NULL to size). */
int gpma_gen_rmat(uint32_t scale, uint64_t ndraws, double a, double b, double c, uint64_t seed, uint32_t *src,
                  uint32_t *dst, uint64_t cap, uint64_t *n_out);
/* Power-law out-degree graph with exactly E unique edges (see gen.cpp). */
int gpma_gen_powerlaw(uint32_t N, uint64_t E, double alpha, double soft_cap, double p1, uint64_t seed, uint32_t *src,
                      uint32_t *dst);

/* The reference's graph text format (dataset/prepare_dataset.sh:34-37, read
 * by gpma_bfs_demo.cu:6-30): a header "N E", then E lines "src dst" with
 * 0-based node ids. gpma_read_edge_text stores N and E, and with src and dst
 * non-NULL (cap >= E entries each) the edges in file order;
 * gpma_write_edge_text writes that format. */
int gpma_read_edge_text(const char *path, uint32_t *N, uint64_t *E, uint32_t *src, uint32_t *dst, uint64_t cap);
int gpma_write_edge_text(const char *path, uint32_t N, uint64_t E, const uint32_t *src, const uint32_t *dst);

#ifdef __cplusplus
}
#endif

#endif /* GPMA_AMD_GPMA_H */
