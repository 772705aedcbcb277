// Request server over one GPMA handle (include/gpma_serve.h): a worker
// thread serves the queue in rounds of at most one fused update launch and
// one multi-source BFS traversal, in strict or snapshot ticket order.
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
// holds per record of a fused launch (sort, locate and pending-list scratch
// plus the server's staging), and per row for a multi-source traversal.
constexpr uint64_t UPDATE_BYTES_PER_RECORD = 128;
constexpr uint64_t MSBFS_BYTES_PER_ROW = 256;

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

}  // namespace

struct gpma_server {
    gpma_t g = nullptr;
    uint32_t N = 0;
    uint32_t max_group = 64;
    uint32_t order = GPMA_SERVE_STRICT;
    uint64_t last_update_ticket = 0;  // worker only
    uint64_t max_fused = 0, max_queued = 0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done, cv_space;
    std::deque<ReqPtr> q;
    std::unordered_map<uint64_t, ReqPtr> live;
    uint64_t next_ticket = 1, queued_records = 0;
    bool stopping = false;
    gpma_server_stats st{};
    std::thread worker;
    // device staging of fused update batches and of BFS levels
    void *dk = nullptr, *dv = nullptr, *dlev = nullptr;
    uint64_t cap = 0;
    std::vector<uint64_t> hk_cat;
    std::vector<uint8_t> hv_cat;
};

namespace {

int ensure_staging(gpma_server *s, uint64_t n) {
    if (n <= s->cap) return GPMA_OK;
    uint64_t c = std::max<uint64_t>(n, s->cap + s->cap / 2);
    if (s->dk) gpma_device_free(s->dk);
    if (s->dv) gpma_device_free(s->dv);
    s->dk = s->dv = nullptr;
    s->cap = 0;
    int rc = gpma_device_malloc(&s->dk, c * 8);
    if (rc == GPMA_OK) rc = gpma_device_malloc(&s->dv, c);
    if (rc != GPMA_OK) return rc;
    s->cap = c;
    return GPMA_OK;
}

// One fused launch for a run of update requests; fills each request's status.
void serve_updates(gpma_server *s, std::vector<ReqPtr> &batch) {
    const uint32_t nb = (uint32_t)batch.size();
    std::vector<uint64_t> offs(nb + 1, 0);
    bool any_device = false;
    for (uint32_t i = 0; i < nb; i++) {
        offs[i + 1] = offs[i] + batch[i]->n;
        any_device |= batch[i]->on_device;
    }
    const uint64_t total = offs[nb];
    int rc = GPMA_OK;
    if (!any_device) {
        s->hk_cat.resize(total);
        s->hv_cat.resize(total);
        for (uint32_t i = 0; i < nb; i++) {
            if (!batch[i]->n) continue;
            std::memcpy(s->hk_cat.data() + offs[i], batch[i]->hk.data(), batch[i]->n * 8);
            std::memcpy(s->hv_cat.data() + offs[i], batch[i]->hv.data(), batch[i]->n);
        }
        rc = gpma_update_batches(s->g, s->hk_cat.data(), s->hv_cat.data(), offs.data(), nb, 0);
    } else {
        rc = ensure_staging(s, std::max<uint64_t>(total, 1));
        for (uint32_t i = 0; i < nb && rc == GPMA_OK; i++) {
            Request &r = *batch[i];
            if (!r.n) continue;
            const int kind = r.on_device ? 2 : 0;
            const void *sk = r.on_device ? (const void *)r.dk : (const void *)r.hk.data();
            const void *sv = r.on_device ? (const void *)r.dv : (const void *)r.hv.data();
            rc = gpma_copy_async(s->g, (uint8_t *)s->dk + offs[i] * 8, sk, r.n * 8, kind);
            if (rc == GPMA_OK) rc = gpma_copy_async(s->g, (uint8_t *)s->dv + offs[i], sv, r.n, kind);
        }
        if (rc == GPMA_OK)
            rc = gpma_update_batches(s->g, (const uint64_t *)s->dk, (const uint8_t *)s->dv, offs.data(), nb, 1);
    }
    for (auto &r : batch) {
        if (rc == GPMA_ERR_KEY_RANGE)
            // a host request knows its own invalid keys; a device one cannot tell
            r->res.status = (r->on_device || r->nbad) ? GPMA_ERR_KEY_RANGE : GPMA_OK;
        else
            r->res.status = rc;
    }
    s->st.update_launches++;
    s->st.update_records += total;
}

// One traversal for a run of BFS requests.
void serve_bfs(gpma_server *s, std::vector<ReqPtr> &batch) {
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

// Strict order: the longest run of same-kind requests at the queue's head.
void take_strict(gpma_server *s, std::vector<ReqPtr> &upd, std::vector<ReqPtr> &bfs) {
    const Kind kind = s->q.front()->kind;
    uint64_t total = 0;
    while (!s->q.empty() && s->q.front()->kind == kind) {
        const ReqPtr &r = s->q.front();
        if (kind == K_UPDATE) {
            if (!upd.empty() && total + r->n > s->max_fused) break;
            total += r->n;
            upd.push_back(r);
        } else {
            if (bfs.size() >= s->max_group) break;
            bfs.push_back(r);
        }
        s->q.pop_front();
    }
}

// Snapshot order: every queued update that fits (in ticket order), and up
// to max_group BFS requests ticketed before the first update left queued.
void take_snapshot(gpma_server *s, std::vector<ReqPtr> &upd, std::vector<ReqPtr> &bfs) {
    uint64_t total = 0;
    bool upd_open = true;
    std::deque<ReqPtr> rest;
    while (!s->q.empty()) {
        ReqPtr r = std::move(s->q.front());
        s->q.pop_front();
        if (r->kind == K_UPDATE) {
            if (upd_open && (upd.empty() || total + r->n <= s->max_fused)) {
                total += r->n;
                upd.push_back(std::move(r));
                continue;
            }
            upd_open = false;
        } else if (upd_open && bfs.size() < s->max_group) {
            bfs.push_back(std::move(r));
            continue;
        }
        rest.push_back(std::move(r));
    }
    s->q.swap(rest);
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

void worker_loop(gpma_server *s) {
    for (;;) {
        std::vector<ReqPtr> upd, bfs;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_work.wait(lk, [&] { return !s->q.empty() || s->stopping; });
            if (s->q.empty()) return;  // stopping and drained
            if (s->order == GPMA_SERVE_SNAPSHOT) take_snapshot(s, upd, bfs);
            else take_strict(s, upd, bfs);
        }
        if (!upd.empty()) {
            const uint64_t before = s->last_update_ticket;
            const clk::time_point t0 = clk::now();
            serve_updates(s, upd);
            s->last_update_ticket = upd.back()->ticket;
            finish(s, upd, K_UPDATE, t0, clk::now(), before);
            s->cv_done.notify_all();
            s->cv_space.notify_all();
        }
        if (!bfs.empty()) {
            const clk::time_point t0 = clk::now();
            serve_bfs(s, bfs);
            finish(s, bfs, K_BFS, t0, clk::now(), s->last_update_ticket);
            s->cv_done.notify_all();
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
    if (r->kind == K_UPDATE) s->queued_records += r->n;
    s->live[r->ticket] = r;
    s->q.push_back(r);
    s->st.submitted++;
    *ticket = r->ticket;
    lk.unlock();
    s->cv_work.notify_one();
    return GPMA_OK;
}

}  // namespace

extern "C" {

int gpma_server_create(gpma_t g, const gpma_server_config *cfg, gpma_server_t *out) {
    if (!g || !out) return GPMA_ERR_ARG;
    if (cfg && cfg->struct_size != sizeof(gpma_server_config)) return GPMA_ERR_ARG;
    gpma_stats gs{};
    int rc = gpma_get_stats(g, &gs);
    if (rc != GPMA_OK) return rc;
    uint64_t budget = cfg ? cfg->hbm_budget_bytes : 0;
    if (!budget) {
        uint64_t fr = 0, tot = 0;
        rc = gpma_mem_info(gpma_device_of(g), &fr, &tot);
        if (rc != GPMA_OK) return rc;
        budget = fr / 4;
    }
    auto *s = new gpma_server();
    s->g = g;
    s->N = gs.row_num;
    s->max_group = (cfg && cfg->max_bfs_group) ? std::min<uint32_t>(cfg->max_bfs_group, 64) : 64;
    s->order = cfg ? cfg->order : GPMA_SERVE_STRICT;
    if (s->order > GPMA_SERVE_SNAPSHOT) {
        delete s;
        return GPMA_ERR_ARG;
    }
    // the level buffer of a full group and the traversal scratch take at most half the budget
    auto bfs_bytes = [&](uint32_t grp) { return (uint64_t)grp * s->N * 4 + (uint64_t)s->N * MSBFS_BYTES_PER_ROW; };
    while (s->max_group > 1 && bfs_bytes(s->max_group) > budget / 2) s->max_group /= 2;
    const uint64_t rest = budget > bfs_bytes(s->max_group) ? budget - bfs_bytes(s->max_group) : 0;
    s->max_fused = (cfg && cfg->max_fused_records) ? cfg->max_fused_records : rest / UPDATE_BYTES_PER_RECORD;
    s->max_fused = std::min<uint64_t>(std::max<uint64_t>(s->max_fused, 1024), (1ull << 30) - 1);
    s->max_queued = (cfg && cfg->max_queued_records) ? cfg->max_queued_records : (1ull << 26);
    s->st.max_fused_records = s->max_fused;
    s->st.max_bfs_group = s->max_group;
    rc = gpma_device_malloc(&s->dlev, (uint64_t)s->max_group * std::max<uint32_t>(s->N, 1) * 4);
    if (rc != GPMA_OK) {
        delete s;
        return rc;
    }
    s->worker = std::thread(worker_loop, s);
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
    s->worker.join();
    s->cv_done.notify_all();
    if (s->dk) gpma_device_free(s->dk);
    if (s->dv) gpma_device_free(s->dv);
    if (s->dlev) gpma_device_free(s->dlev);
    delete s;
    return GPMA_OK;
}

int gpma_server_submit_update(gpma_server_t s, const uint64_t *keys, const uint8_t *values, uint64_t n,
                              int on_device, uint64_t *ticket) {
    if (!s || !ticket || (n && (!keys || !values))) return GPMA_ERR_ARG;
    if (n > s->max_fused) {
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
