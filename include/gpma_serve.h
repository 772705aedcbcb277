/*
 * gpma_serve.h - request server over one GPMA handle: continuous batching of
 * update and BFS requests with device-memory admission control.
 *
 * No counterpart in the reference, which has no serving layer: its drivers
 * call update_gpma and gpma_bfs back to back (gpma_bfs_demo.cu:69-111). The
 * server takes a stream of such requests from many threads; each submission
 * gets a ticket, and the ticket order is the order the requests take effect
 * in (exactly, or for BFS at least, depending on the order chosen below).
 *
 * Requests are served in rounds by a two-stage pipeline (continuous
 * batching): a planner thread takes the next round from the queue and copies
 * its update records into a device staging area on the server's own stream
 * while an executor thread runs the round before on the handle's stream; two
 * staging areas alternate. Requests that arrive while a round runs join a
 * later one. Updates are always applied in ticket order, fused into one
 * gpma_update_batches launch per round; BFS requests of a round share one
 * multi-source gpma_bfs_multi traversal (up to 64 sources). Two orders:
 *
 *   GPMA_SERVE_STRICT (default): every request sees exactly the requests
 *     ticketed before it. A round serves the longest run of queued requests
 *     of one kind, so BFS requests separated by updates do not share a
 *     traversal.
 *   GPMA_SERVE_SNAPSHOT: a BFS sees every update ticketed before it and
 *     possibly later ones (those applied before its traversal starts). A
 *     round applies every queued update that fits, then serves up to 64
 *     queued BFS requests together.
 *
 * Each result carries state_ticket, the ticket of the last update applied
 * before the request was served, so a caller can replay exactly what it saw.
 *
 * Admission: the server plans within a device-memory budget (default: half
 * the device's free memory at creation). Out of it come a reserve for a
 * rebuild of the packed array (its compaction copy and new arrays at twice
 * the current capacity, which exist beside the old arrays during a resize;
 * 28 bytes per slot), the BFS group's level buffer and traversal scratch,
 * and two staging areas; the rest, at 146 bytes per record (library scratch
 * and staging), sets how many update records one round may hold. The limit
 * is recomputed after every round as the array grows. A round always takes
 * the first queued update; later ones that do not fit wait for a later
 * round (counted as deferred). An update request larger than the limit at
 * creation is rejected with GPMA_ERR_CAPACITY at submission; submitters
 * block while more than max_queued_records update records wait.
 *
 * While a server exists, the handle must not be used directly.
 */
#ifndef GPMA_AMD_GPMA_SERVE_H
#define GPMA_AMD_GPMA_SERVE_H

#include "gpma.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gpma_server *gpma_server_t;

#define GPMA_SERVE_STRICT 0
#define GPMA_SERVE_SNAPSHOT 1

typedef struct gpma_server_config {
    uint32_t struct_size;        /* sizeof(gpma_server_config) */
    uint32_t max_bfs_group;      /* BFS requests per traversal, 1..64 (0 = 64) */
    uint64_t hbm_budget_bytes;   /* device memory the server may plan for (0 = free / 2) */
    uint64_t max_fused_records;  /* update records per fused launch (0 = from the budget) */
    uint64_t max_queued_records; /* update records queued before submitters block (0 = 2^26) */
    uint32_t order;              /* GPMA_SERVE_STRICT or GPMA_SERVE_SNAPSHOT */
    uint32_t reserved[3];
} gpma_server_config;

typedef struct gpma_request_result {
    int32_t status;      /* GPMA_OK, GPMA_ERR_KEY_RANGE (the request's valid keys were applied), ... */
    uint32_t kind;       /* 0 update, 1 BFS */
    uint32_t reachable;  /* BFS: vertices with a non-zero level */
    uint32_t group;      /* requests served by the same launch */
    double queue_us;     /* submission -> start of its launch */
    double service_us;   /* start -> end of its launch */
    double latency_us;   /* submission -> completion */
    uint64_t state_ticket; /* last update applied before this request was served (0: none) */
} gpma_request_result;

typedef struct gpma_server_stats {
    uint64_t submitted;
    uint64_t completed;
    uint64_t update_launches;
    uint64_t update_records;
    uint64_t bfs_launches;
    uint64_t bfs_requests;
    uint64_t rejected;
    uint64_t max_fused_records;    /* update records one round may hold now */
    uint32_t max_bfs_group;
    uint32_t reserved;
    uint64_t deferred;             /* update requests left for a later round by the records limit */
    uint64_t staged_ahead;         /* rounds whose records were staged while the round before ran */
    uint64_t budget_bytes;         /* device memory the server plans within */
    uint64_t growth_reserve_bytes; /* part of it held for a rebuild of the array */
} gpma_server_stats;

/* Start a server (and its worker thread) over handle g. */
int gpma_server_create(gpma_t g, const gpma_server_config *cfg, gpma_server_t *out);

/* Serve every request already submitted, then stop and free the server.
 * The handle stays valid and may be used directly again. */
int gpma_server_destroy(gpma_server_t s);

/* Queue an update request (same encoding as gpma_update). Host arrays are
 * copied before the call returns; device arrays (on_device = 1) must stay
 * valid and unchanged until the request completes. */
int gpma_server_submit_update(gpma_server_t s, const uint64_t *keys, const uint8_t *values, uint64_t n,
                              int on_device, uint64_t *ticket);

/* Queue a BFS request from `start`. levels: host uint32[row_num] written
 * before completion, or NULL for the reachable count only. */
int gpma_server_submit_bfs(gpma_server_t s, uint32_t start, uint32_t *levels, uint64_t *ticket);

/* Wait for a request (timeout_us < 0: no limit). GPMA_ERR_TIMEOUT leaves it
 * waitable; any other return consumes the ticket. The request's own status
 * is in result->status. */
int gpma_server_wait(gpma_server_t s, uint64_t ticket, int64_t timeout_us, gpma_request_result *result);

int gpma_server_get_stats(gpma_server_t s, gpma_server_stats *out);

#ifdef __cplusplus
}
#endif

#endif
