// Request server over one GPMA handle (include/gpma_serve.h): continuous
// batching in a two-stage pipeline.
//
//   planner   admits queued requests into the open round as they arrive
//             (strict or snapshot order, within the round's records limit)
//             and copies each admitted update's records into the round's
//             device staging area on the server's copy stream (host arrays:
//             H2D, device arrays: D2D) at once. When the executor asks for
//             work and the open round holds requests, the planner closes it
//             (records the area's event), hands it over and opens the next.
//   executor  makes the handle's stream wait for the round's event, runs
//             the fused update (gpma_update_batches over the staged area)
//             and the round's BFS group (one multi-source traversal), and
//             completes the requests.
//
// So requests arriving while a round runs are staged while it runs, and a
// round closes only when the device is ready for it: an idle server serves a
// request at once, a busy one batches everything that arrived meanwhile. Two
// staging areas alternate; an area is free again once its round's update
// ran. Admission: a round takes queued update requests in ticket order while
// their records fit the current per-round limit, which comes from the
// device-memory budget minus the BFS group, the staging areas, the library's
// scratch and a reserve for a rebuild of the array (the compaction copy and
// the new arrays at twice the capacity, which exist beside the old arrays
// during a resize). The limit is recomputed after every round, as the array
// grows. Requests left for a later round because the limit was reached count
// as deferred.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/gpma_serve.h"

namespace {

using clk = std::chrono::steady_clock;

double us_between(clk::time_point a, clk::time_point b) {
    return std::chrono::duration<double, std::micro>(b - a).count();
}

enum Kind : uint32_t { K_UPDATE = 0, K_BFS = 1 };

// Planning figures for admission: device bytes the library's update path
// holds per record of a fused launch (sort, locate and pending-list scratch),
// per row for a multi-source traversal, and per slot of the array for a
// rebuild (9 B compaction copy + 2 x 9 B new slots; separators 16 B per
// segment of the new array, below 1 B per slot).
constexpr uint64_t UPDATE_BYTES_PER_RECORD = 128;
constexpr uint64_t STAGE_BYTES_PER_RECORD = 9;
constexpr uint64_t MSBFS_BYTES_PER_ROW = 384;
constexpr uint64_t REBUILD_BYTES_PER_SLOT = 28;
constexpr uint64_t MIN_FUSED = 1024;

struct Request {
    uint64_t ticket = 0;
    Kind kind = K_UPDATE;
    // update
    std::vector<uint64_t> hk;
    std::vector<uint8_t> hv;
    const uint64_t *dk = nullptr;
    const uint8_t *dv = nullptr;
    uint64_t n = 0;
    bool on_device = false;
    bool deferred = false;
    uint64_t nbad = 0;
    // BFS
    uint32_t start = 0;
    uint32_t *levels = nullptr;
    // outcome
    gpma_request_result res{};
    bool done = false;
    clk::time_point t_submit;
};

using ReqPtr = std::shared_ptr<Request>;

struct Area {  // device staging of one round's update records
    void *dk = nullptr, *dv = nullptr;
    uint64_t cap = 0;
    hipEvent_t ev = nullptr;
    bool free = true;
};

struct Round {
    std::vector<ReqPtr> upd, bfs;
    std::vector<uint64_t> offs{0};
    int area = -1;
    int stage_rc = GPMA_OK;
    uint64_t state_ticket = 0;  // last update ticket applied once this round's updates ran
    bool stop = false;
    // admission state of the open round
    Kind kind = K_UPDATE;     // strict order: the round's one kind
    bool closed = false;      // strict: no further request joins
    bool upd_closed = false;  // snapshot: an update was left out, so later requests wait
    size_t staged = 0;        // updates whose copies are issued
    bool empty() const { return upd.empty() && bfs.empty(); }
};

}  // namespace

struct gpma_server {
    gpma_t g = nullptr;
    int device = 0;
    uint32_t N = 0;
    uint32_t max_group = 64;
    uint32_t order = GPMA_SERVE_STRICT;
    uint64_t budget = 0, reserve = 0;
    uint64_t max_fused = 0, fused_cap = 0, max_queued = 0;
    bool fused_fixed = false;
    std::mutex mu;
    std::condition_variable cv_work, cv_done, cv_space, cv_area, cv_round;
    std::deque<ReqPtr> q;
    std::unordered_map<uint64_t, ReqPtr> live;
    uint64_t next_ticket = 1, queued_records = 0;
    uint64_t q_updates = 0;  // update requests in q
    uint64_t planned_update_ticket = 0;  // planner: last update ticket of the rounds planned so far
    bool stopping = false;
    bool executor_busy = false;
    bool executor_wants = false;  // the executor waits for a round
    gpma_server_stats st{};
    std::thread planner, executor;
    hipStream_t copy_st = nullptr;
    Area area[2];
    int next_area = 0;
    std::deque<Round> rounds;  // planned, not yet executed (at most one waits)
    void *dlev = nullptr;      // BFS levels of one group
};

namespace {

int hip_rc(hipError_t e) { return e == hipSuccess ? GPMA_OK : e == hipErrorOutOfMemory ? GPMA_ERR_OOM : GPMA_ERR_HIP; }

int ensure_area(Area &a, uint64_t n) {
    if (n <= a.cap) return GPMA_OK;
    const uint64_t c = std::max<uint64_t>(n, a.cap + a.cap / 2);
    if (a.dk) (void)hipFree(a.dk);
    if (a.dv) (void)hipFree(a.dv);
    a.dk = a.dv = nullptr;
    a.cap = 0;
    int rc = hip_rc(hipMalloc(&a.dk, c * 8));
    if (rc == GPMA_OK) rc = hip_rc(hipMalloc(&a.dv, c));
    if (rc != GPMA_OK) return rc;
    a.cap = c;
    return GPMA_OK;
}

// The per-round records limit from the budget (s->mu held). Everything the
// server plans for beyond the handle's own arrays must fit the budget: the
// rebuild reserve for the current capacity, the BFS group, two staging
// areas and the library's scratch for a fused launch.
void plan_limits(gpma_server *s) {
    gpma_stats gs{};
    if (gpma_get_stats(s->g, &gs) != GPMA_OK) return;
    s->reserve = REBUILD_BYTES_PER_SLOT * gs.capacity;
    const uint64_t bfs = (uint64_t)s->max_group * s->N * 4 + (uint64_t)s->N * MSBFS_BYTES_PER_ROW;
    const uint64_t fixed = s->reserve + bfs;
    const uint64_t rest = s->budget > fixed ? s->budget - fixed : 0;
    if (!s->fused_fixed) {
        const uint64_t f = rest / (UPDATE_BYTES_PER_RECORD + 2 * STAGE_BYTES_PER_RECORD);
        s->max_fused = std::min(std::max(f, MIN_FUSED), s->fused_cap);
    }
    s->st.max_fused_records = s->max_fused;
    s->st.budget_bytes = s->budget;
    s->st.growth_reserve_bytes = s->reserve;
}

// Whether an update of n records may join the open round (s->mu held): the
// first update of a round always does; later ones while the round's records
// stay within the limit and the staging area's size.
bool fits(const gpma_server *s, const Round &rd, uint64_t n) {
    if (rd.upd.empty()) return true;
    const uint64_t t = rd.offs.back() + n;
    return t <= s->max_fused && t <= s->area[rd.area].cap;
}

void defer(gpma_server *s, Request &r) {
    if (!r.deferred) {
        r.deferred = true;
        s->st.deferred++;
    }
}

void admit_update(gpma_server *s, Round &rd, ReqPtr r) {
    s->q_updates--;
    // the round's first update may be larger than its area: no copy into the
    // area is in flight yet, so it can grow
    if (rd.upd.empty() && rd.stage_rc == GPMA_OK && r->n > s->area[rd.area].cap)
        rd.stage_rc = ensure_area(s->area[rd.area], r->n);
    rd.offs.push_back(rd.offs.back() + r->n);
    rd.upd.push_back(std::move(r));
}

// Strict order: the open round is a run of one kind from the queue's head.
void admit_strict(gpma_server *s, Round &rd) {
    while (!s->q.empty() && !rd.closed) {
        ReqPtr &r = s->q.front();
        if (rd.empty()) rd.kind = r->kind;
        if (r->kind != rd.kind) {
            rd.closed = true;
            break;
        }
        if (r->kind == K_UPDATE) {
            if (!fits(s, rd, r->n)) {
                defer(s, *r);
                rd.closed = true;
                break;
            }
            admit_update(s, rd, std::move(r));
        } else {
            if (rd.bfs.size() >= s->max_group) {
                rd.closed = true;
                break;
            }
            rd.bfs.push_back(std::move(r));
        }
        s->q.pop_front();
    }
}

// Snapshot order: updates in ticket order while they fit, and BFS requests
// (up to the group size) ticketed before the first update left out.
void admit_snapshot(gpma_server *s, Round &rd) {
    std::deque<ReqPtr> rest;
    while (!s->q.empty()) {
        ReqPtr r = std::move(s->q.front());
        s->q.pop_front();
        if (!rd.upd_closed) {
            if (r->kind == K_UPDATE) {
                if (fits(s, rd, r->n)) {
                    admit_update(s, rd, std::move(r));
                    continue;
                }
                defer(s, *r);
                rd.upd_closed = true;
            } else if (rd.bfs.size() < s->max_group) {
                rd.bfs.push_back(std::move(r));
                continue;
            }
        }
        rest.push_back(std::move(r));
    }
    s->q.swap(rest);
}

// Whether a queued request could join the open round now (s->mu held):
// admission then either takes it or closes the round to later requests.
bool admissible(const gpma_server *s, const Round &rd) {
    if (s->q.empty()) return false;
    if (s->order == GPMA_SERVE_STRICT) return !rd.closed;
    return !rd.upd_closed && (s->q_updates > 0 || rd.bfs.size() < s->max_group);
}

// Issue the copies of the round's updates admitted since the last call
// into its staging area on the copy stream.
int stage(gpma_server *s, Round &rd, size_t upto) {
    Area &a = s->area[rd.area];
    int rc = rd.stage_rc;
    for (size_t i = rd.staged; i < upto && rc == GPMA_OK; i++) {
        Request &r = *rd.upd[i];
        if (!r.n) continue;
        const hipMemcpyKind kind = r.on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
        const void *sk = r.on_device ? (const void *)r.dk : (const void *)r.hk.data();
        const void *sv = r.on_device ? (const void *)r.dv : (const void *)r.hv.data();
        rc = hip_rc(hipMemcpyAsync((uint8_t *)a.dk + rd.offs[i] * 8, sk, r.n * 8, kind, s->copy_st));
        if (rc == GPMA_OK) rc = hip_rc(hipMemcpyAsync((uint8_t *)a.dv + rd.offs[i], sv, r.n, kind, s->copy_st));
    }
    rd.staged = upto;
    return rc;
}

// Open a round on staging area `ai` (free: its last round's update ran),
// sized for the records waiting now (s->mu held; may allocate).
void open_round(gpma_server *s, Round &rd, int ai) {
    rd = Round();
    rd.area = ai;
    const uint64_t want = std::min(s->max_fused, std::max<uint64_t>(1ull << 20, s->queued_records + (1ull << 20)));
    if (s->area[ai].cap < want) rd.stage_rc = ensure_area(s->area[ai], want);
    if (rd.stage_rc != GPMA_OK) s->area[ai].cap = 0;
}

void planner_loop(gpma_server *s) {
    (void)hipSetDevice(s->device);
    Round rd;
    std::unique_lock<std::mutex> lk(s->mu);
    open_round(s, rd, s->next_area);
    for (;;) {
        s->cv_work.wait(lk, [&] {
            return admissible(s, rd) || (s->executor_wants && !rd.empty()) ||
                   (s->stopping && s->q.empty() && rd.empty());
        });
        if (rd.empty() && s->q.empty() && s->stopping) {
            Round stop;
            stop.stop = true;
            s->rounds.push_back(std::move(stop));
            s->cv_round.notify_all();
            return;
        }
        if (!(s->executor_wants && !rd.empty())) {
            if (s->order == GPMA_SERVE_SNAPSHOT) admit_snapshot(s, rd);
            else admit_strict(s, rd);
            const size_t upto = rd.upd.size();
            if (upto > rd.staged) {
                lk.unlock();
                rd.stage_rc = stage(s, rd, upto);
                lk.lock();
                if (s->executor_busy) s->st.staged_ahead++;
            }
            continue;
        }
        // the executor waits and the open round holds requests: close it
        if (!rd.upd.empty()) {
            if (rd.stage_rc == GPMA_OK) rd.stage_rc = hip_rc(hipEventRecord(s->area[rd.area].ev, s->copy_st));
            s->planned_update_ticket = rd.upd.back()->ticket;
        }
        rd.state_ticket = s->planned_update_ticket;
        const int used = rd.upd.empty() ? -1 : rd.area;
        if (used >= 0) s->area[used].free = false;
        s->executor_wants = false;
        s->rounds.push_back(std::move(rd));
        s->cv_round.notify_all();
        // the next round takes the other area if this one was used; that one
        // is free: its round completed before the executor asked again
        if (used >= 0) s->next_area ^= 1;
        s->cv_area.wait(lk, [&] { return s->area[s->next_area].free; });
        open_round(s, rd, s->next_area);
    }
}

void finish(gpma_server *s, std::vector<ReqPtr> &batch, Kind kind, clk::time_point t0, clk::time_point t1,
            uint64_t state_ticket) {
    std::lock_guard<std::mutex> lk(s->mu);
    for (auto &r : batch) {
        r->res.kind = kind;
        r->res.group = (uint32_t)batch.size();
        r->res.queue_us = us_between(r->t_submit, t0);
        r->res.service_us = us_between(t0, t1);
        r->res.latency_us = us_between(r->t_submit, t1);
        r->res.state_ticket = state_ticket;
        r->done = true;
        if (kind == K_UPDATE) s->queued_records -= r->n;
    }
    s->st.completed += batch.size();
}

// One fused launch over a staged round; fills each request's status.
void run_updates(gpma_server *s, Round &rd) {
    Area &a = s->area[rd.area];
    int rc = rd.stage_rc;
    if (rc == GPMA_OK) rc = hip_rc(hipStreamWaitEvent((hipStream_t)gpma_stream(s->g), a.ev, 0));
    if (rc == GPMA_OK)
        rc = gpma_update_batches(s->g, (const uint64_t *)a.dk, (const uint8_t *)a.dv, rd.offs.data(),
                                 (uint32_t)rd.upd.size(), 1);
    for (auto &r : rd.upd) {
        if (rc == GPMA_ERR_KEY_RANGE)
            // a host request knows its own invalid keys; a device one cannot tell
            r->res.status = (r->on_device || r->nbad) ? GPMA_ERR_KEY_RANGE : GPMA_OK;
        else
            r->res.status = rc;
    }
    s->st.update_launches++;
    s->st.update_records += rd.offs.back();
}

// One traversal for a round's BFS requests.
void run_bfs(gpma_server *s, std::vector<ReqPtr> &batch) {
    const uint32_t nb = (uint32_t)batch.size();
    std::vector<uint32_t> starts(nb), reach(nb, 0);
    for (uint32_t i = 0; i < nb; i++) starts[i] = batch[i]->start;
    int rc;
    if (nb == 1) rc = gpma_bfs(s->g, starts[0], (uint32_t *)s->dlev, 1, reach.data());
    else rc = gpma_bfs_multi(s->g, starts.data(), nb, (uint32_t *)s->dlev, 1, reach.data());
    for (uint32_t i = 0; i < nb; i++) {
        Request &r = *batch[i];
        int ri = rc;
        if (ri == GPMA_OK && r.levels)
            ri = gpma_memcpy(r.levels, (const uint8_t *)s->dlev + (uint64_t)i * s->N * 4, (uint64_t)s->N * 4, 1);
        r.res.status = ri;
        r.res.reachable = rc == GPMA_OK ? reach[i] : 0;
    }
    s->st.bfs_launches++;
    s->st.bfs_requests += nb;
}

void executor_loop(gpma_server *s) {
    (void)hipSetDevice(s->device);
    uint64_t applied_ticket = 0;
    for (;;) {
        Round rd;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->executor_wants = true;
            s->cv_work.notify_all();
            s->cv_round.wait(lk, [&] { return !s->rounds.empty(); });
            rd = std::move(s->rounds.front());
            s->rounds.pop_front();
            s->executor_busy = !rd.stop;
        }
        if (rd.stop) return;
        if (!rd.upd.empty()) {
            const uint64_t before = applied_ticket;
            const clk::time_point t0 = clk::now();
            run_updates(s, rd);
            applied_ticket = rd.upd.back()->ticket;
            {
                std::lock_guard<std::mutex> lk(s->mu);
                s->area[rd.area].free = true;
            }
            s->cv_area.notify_all();
            finish(s, rd.upd, K_UPDATE, t0, clk::now(), before);
            s->cv_done.notify_all();
            s->cv_space.notify_all();
        }
        if (!rd.bfs.empty()) {
            const clk::time_point t0 = clk::now();
            run_bfs(s, rd.bfs);
            finish(s, rd.bfs, K_BFS, t0, clk::now(), rd.state_ticket);
            s->cv_done.notify_all();
        }
        {
            std::lock_guard<std::mutex> lk(s->mu);
            plan_limits(s);  // the array may have grown
            s->executor_busy = false;
        }
    }
}

int enqueue(gpma_server *s, ReqPtr r, uint64_t *ticket) {
    std::unique_lock<std::mutex> lk(s->mu);
    if (r->kind == K_UPDATE)
        s->cv_space.wait(lk, [&] {
            return s->stopping || s->queued_records == 0 || s->queued_records + r->n <= s->max_queued;
        });
    if (s->stopping) return GPMA_ERR_STOPPED;
    r->ticket = s->next_ticket++;
    r->t_submit = clk::now();
    if (r->kind == K_UPDATE) {
        s->queued_records += r->n;
        s->q_updates++;
    }
    s->live[r->ticket] = r;
    s->q.push_back(r);
    s->st.submitted++;
    *ticket = r->ticket;
    lk.unlock();
    s->cv_work.notify_one();
    return GPMA_OK;
}

void release(gpma_server *s) {
    for (Area &a : s->area) {
        if (a.dk) (void)hipFree(a.dk);
        if (a.dv) (void)hipFree(a.dv);
        if (a.ev) (void)hipEventDestroy(a.ev);
    }
    if (s->dlev) gpma_device_free(s->dlev);
    if (s->copy_st) (void)hipStreamDestroy(s->copy_st);
    delete s;
}

}  // namespace

extern "C" {

int gpma_server_create(gpma_t g, const gpma_server_config *cfg, gpma_server_t *out) {
    if (!g || !out) return GPMA_ERR_ARG;
    if (cfg && cfg->struct_size != sizeof(gpma_server_config)) return GPMA_ERR_ARG;
    gpma_stats gs{};
    int rc = gpma_get_stats(g, &gs);
    if (rc != GPMA_OK) return rc;
    auto *s = new gpma_server();
    s->g = g;
    s->device = gpma_device_of(g);
    s->N = gs.row_num;
    s->budget = cfg ? cfg->hbm_budget_bytes : 0;
    if (!s->budget) {
        uint64_t fr = 0, tot = 0;
        rc = gpma_mem_info(s->device, &fr, &tot);
        if (rc != GPMA_OK) {
            delete s;
            return rc;
        }
        s->budget = fr / 2;
    }
    s->max_group = (cfg && cfg->max_bfs_group) ? std::min<uint32_t>(cfg->max_bfs_group, 64) : 64;
    s->order = cfg ? cfg->order : GPMA_SERVE_STRICT;
    if (s->order > GPMA_SERVE_SNAPSHOT) {
        delete s;
        return GPMA_ERR_ARG;
    }
    // the level buffer of a full group and the traversal scratch take at most half the budget
    auto bfs_bytes = [&](uint32_t grp) { return (uint64_t)grp * s->N * 4 + (uint64_t)s->N * MSBFS_BYTES_PER_ROW; };
    while (s->max_group > 1 && bfs_bytes(s->max_group) > s->budget / 2) s->max_group /= 2;
    s->fused_fixed = cfg && cfg->max_fused_records;
    s->fused_cap = (1ull << 30) - 1;
    if (s->fused_fixed) s->max_fused = std::min<uint64_t>(cfg->max_fused_records, s->fused_cap);
    s->max_queued = (cfg && cfg->max_queued_records) ? cfg->max_queued_records : (1ull << 26);
    s->st.max_bfs_group = s->max_group;
    plan_limits(s);
    // largest update request ever admitted: the limit at creation
    s->fused_cap = s->max_fused;
    if (hipSetDevice(s->device) != hipSuccess ||
        hipStreamCreateWithFlags(&s->copy_st, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&s->area[0].ev, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&s->area[1].ev, hipEventDisableTiming) != hipSuccess) {
        release(s);
        return GPMA_ERR_HIP;
    }
    rc = gpma_device_malloc(&s->dlev, (uint64_t)s->max_group * std::max<uint32_t>(s->N, 1) * 4);
    if (rc != GPMA_OK) {
        release(s);
        return rc;
    }
    s->planner = std::thread(planner_loop, s);
    s->executor = std::thread(executor_loop, s);
    *out = s;
    return GPMA_OK;
}

int gpma_server_destroy(gpma_server_t s) {
    if (!s) return GPMA_ERR_ARG;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stopping = true;
    }
    s->cv_work.notify_all();
    s->cv_space.notify_all();
    s->planner.join();
    s->executor.join();
    s->cv_done.notify_all();
    (void)hipStreamSynchronize(s->copy_st);
    release(s);
    return GPMA_OK;
}

int gpma_server_submit_update(gpma_server_t s, const uint64_t *keys, const uint8_t *values, uint64_t n,
                              int on_device, uint64_t *ticket) {
    if (!s || !ticket || (n && (!keys || !values))) return GPMA_ERR_ARG;
    if (n > s->fused_cap) {
        std::lock_guard<std::mutex> lk(s->mu);
        s->st.rejected++;
        return GPMA_ERR_CAPACITY;
    }
    auto r = std::make_shared<Request>();
    r->kind = K_UPDATE;
    r->n = n;
    r->on_device = on_device != 0;
    if (r->on_device) {
        r->dk = keys;
        r->dv = values;
    } else if (n) {
        r->hk.assign(keys, keys + n);
        r->hv.assign(values, values + n);
        for (uint64_t i = 0; i < n; i++) {
            const uint64_t k = keys[i];
            r->nbad += ((uint32_t)(k >> 32) >= s->N || (uint32_t)k >= s->N) ? 1 : 0;
        }
    }
    return enqueue(s, r, ticket);
}

int gpma_server_submit_bfs(gpma_server_t s, uint32_t start, uint32_t *levels, uint64_t *ticket) {
    if (!s || !ticket || start >= s->N) return GPMA_ERR_ARG;
    auto r = std::make_shared<Request>();
    r->kind = K_BFS;
    r->start = start;
    r->levels = levels;
    return enqueue(s, r, ticket);
}

int gpma_server_wait(gpma_server_t s, uint64_t ticket, int64_t timeout_us, gpma_request_result *result) {
    if (!s) return GPMA_ERR_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->live.find(ticket);
    if (it == s->live.end()) return GPMA_ERR_ARG;
    ReqPtr r = it->second;
    if (timeout_us < 0) {
        s->cv_done.wait(lk, [&] { return r->done; });
    } else if (!s->cv_done.wait_for(lk, std::chrono::microseconds(timeout_us), [&] { return r->done; })) {
        return GPMA_ERR_TIMEOUT;
    }
    if (result) *result = r->res;
    s->live.erase(ticket);
    return GPMA_OK;
}

int gpma_server_get_stats(gpma_server_t s, gpma_server_stats *out) {
    if (!s || !out) return GPMA_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->st;
    return GPMA_OK;
}

}  // extern "C"
