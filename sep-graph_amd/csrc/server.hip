// Batched query serving over a resident graph (C ABI in include/sepgraph.h).
//
// The reference has no serving layer: each sample answers one query per
// process (samples/hybrid_*/main.cpp). Here one server keeps a graph and an
// engine resident and answers BFS, SSSP and BC-dependency queries:
//   * admission by device memory: every query needs one result row of n
//     32-bit values; the server owns a pool of rows sized by the caller's
//     memory budget, and a query waits in its kind's FIFO until a row is free
//     (rows come back when the client releases a result);
//   * continuous batching: a worker thread forms the next batch the moment
//     the GPU is free, from whatever is queued at that time -- the kind whose
//     oldest request has waited longest, up to that kind's batch limit (<= 64)
//     -- so requests that arrive while a batch runs join the next one;
//   * few launches per batch, each writing every query's result straight into
//     its row of the pool: BC-dependency batches, and BFS / SSSP batches of
//     at least the configured bfs_packed_min / sssp_packed_min (defaults
//     kMsBfsMin, kMsSsspMin) run the bit-packed multi-source kernels (up to
//     64 sources per launch); smaller BFS and SSSP batches run their queries
//     one after another over the whole grid, which is faster per query than a
//     small bit-packed launch.
// Batches are serialised on the engine's stream: a persistent kernel assumes
// all of its blocks are co-resident, which two concurrent launches would not
// be. Host work (queueing, result hand-off) overlaps the running batch.
//
// That is the SG_SERVE_BATCHES form. The default, SG_SERVE_SLOTS, keeps one
// persistent serving kernel (serve_kernel.hip) in which every XCD is a query
// slot: the worker publishes each admitted query into the kernel's ring at
// once, oldest first, and a slot that finishes claims the next, so queries
// that arrive while others run join without waiting for a batch to end.
#include "serve.hpp"
#include "sg_internal.hpp"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <memory>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

using namespace sg;

namespace {

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int kind_to_ms(int kind)
{
    return kind == SG_QUERY_BFS ? kMsBfs : kind == SG_QUERY_SSSP ? kMsSssp : kMsBcDep;
}

enum { kQueued = 0, kRunning = 1, kDone = 2, kFailed = 3 };

// Default smallest SSSP batch run as one bit-packed launch: below it,
// whole-GPU single queries are faster per query (RMAT-22: 2.2 ms alone
// against 1.8 ms per query for 8 sources per launch and 1.2 ms for 16).
constexpr uint32_t kMsSsspMin = 8;
// How long a short SSSP queue may yield to other kinds' batches.
constexpr double kSsspHoldMs = 10.0;
// Default smallest BFS batch run as one bit-packed launch (RMAT-22: a lone
// BFS takes 0.17 ms; bit-packed, 16 sources take 0.19 ms per source, 64 take
// 0.08).
constexpr uint32_t kMsBfsMin = 16;

}  // namespace

struct sg_server {
    sg_graph *g = nullptr;                 // the first graph (the engine's)
    std::vector<sg_graph *> graphs;        // every graph queries may name
    sg_engine *e = nullptr;
    sg_server_config cfg{};
    uint64_t row_bytes = 0;
    uint32_t nrows = 0;
    void *pool = nullptr;
    uint32_t *d_src = nullptr, *d_rows = nullptr;  // 64 entries each
    uint32_t *h_src = nullptr, *h_rows = nullptr;  // pinned staging
    uint32_t h_qsrc[64] = {};                      // the batch's sources, caller's ids
    ServeDev *sd = nullptr;                        // SG_SERVE_SLOTS: the serving kernel
    uint64_t slot_launches = 0;
    double t_launch = 0.0;

    struct Req {
        int graph;
        int kind;
        uint32_t source;     // device vertex id
        uint32_t source_id;  // the caller's id
        int32_t row;
        int state;
        int status;
        bool auto_release;
        double t_arrival, t_start, t_done;
        double *latency_out;  // replay: where to store t_done - t_arrival (ms)
    };
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::unordered_map<uint64_t, Req> reqs;
    std::deque<uint64_t> queue[SG_QUERY_KINDS];
    std::vector<uint32_t> free_rows;
    uint64_t next_ticket = 1;
    bool stop = false;
    std::thread worker;
    // counters
    uint64_t submitted = 0, completed = 0, failed = 0, batches = 0, batched_queries = 0;
    int last_failure = SG_OK;  // status of the last failed query
    double busy_s = 0.0;

    bool has_work() const
    {
        if (free_rows.empty()) return false;
        for (int k = 0; k < SG_QUERY_KINDS; k++)
            if (!queue[k].empty()) return true;
        return false;
    }

    // status: the query's error when !ok (SG_ERR_HIP when none is known)
    void finish_locked(uint64_t t, Req &r, bool ok, double t_done, int status = SG_ERR_HIP)
    {
        r.state = ok ? kDone : kFailed;
        r.status = ok ? SG_OK : status;
        if (!ok) last_failure = status;
        r.t_done = t_done;
        if (ok) completed++;
        else failed++;
        if (r.latency_out) *r.latency_out = ok ? (t_done - r.t_arrival) * 1e3 : -1.0;
        if (r.auto_release) {
            if (r.row >= 0) free_rows.push_back((uint32_t)r.row);
            reqs.erase(t);
        }
    }

    // SG_SERVE_SLOTS worker: publish admitted queries into the serving
    // kernel's ring (oldest first, as result rows allow), keep the kernel
    // running while entries are unclaimed, and hand back finished queries.
    void run_slots()
    {
        std::vector<std::pair<uint32_t, uint64_t>> inflight;  // (ring sequence, ticket)
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
            if (inflight.empty()) {
                // while a launch drains its idle time, poll it every
                // millisecond, so its end (busy time, launch count) is seen
                if (running_)
                    cv_work.wait_for(lk, std::chrono::milliseconds(1), [&] { return stop || has_work(); });
                else
                    cv_work.wait(lk, [&] { return stop || has_work(); });
            }
            if (stop) return;
            const double tp = now_s();
            while (inflight.size() < kServeRing && !free_rows.empty()) {
                int kind = -1;
                double oldest = 0.0;
                for (int k = 0; k < SG_QUERY_KINDS; k++) {
                    if (queue[k].empty()) continue;
                    const double t = reqs[queue[k].front()].t_arrival;
                    if (kind < 0 || t < oldest) {
                        kind = k;
                        oldest = t;
                    }
                }
                if (kind < 0) break;
                const uint64_t t = queue[kind].front();
                queue[kind].pop_front();
                Req &r = reqs[t];
                r.row = (int32_t)free_rows.back();
                free_rows.pop_back();
                r.state = kRunning;
                r.t_start = tp;
                inflight.emplace_back(serve_publish(sd, r.graph, kind, r.source, r.source_id, (uint32_t)r.row), t);
            }
            lk.unlock();
            const bool was_running = running_;
            bool running = false;
            const int rc = serve_pump(sd, &running);
            const double tn = now_s();
            lk.lock();
            if (was_running && !running) busy_s += tn - t_launch;
            if (running && !running_) {
                slot_launches++;
                t_launch = tn;
            }
            running_ = running;
            bool changed = false;
            for (size_t i = 0; i < inflight.size();) {
                int32_t st = SG_OK;
                if (serve_poll(sd, inflight[i].first, &st)) {
                    const uint64_t t = inflight[i].second;
                    batched_queries++;
                    finish_locked(t, reqs[t], st == SG_OK, tn, st);
                    inflight[i] = inflight.back();
                    inflight.pop_back();
                    changed = true;
                } else {
                    i++;
                }
            }
            if (rc != SG_OK && !running) {
                // the launch that ended failed: queries it had claimed and not
                // finished are lost; the others stay in the ring
                const uint32_t head = serve_claimed(sd);
                for (size_t i = 0; i < inflight.size();) {
                    if ((int32_t)(head - inflight[i].first) > 0) {
                        const uint64_t t = inflight[i].second;
                        finish_locked(t, reqs[t], false, tn, rc);
                        inflight[i] = inflight.back();
                        inflight.pop_back();
                        changed = true;
                    } else {
                        i++;
                    }
                }
            }
            if (changed) {
                cv_done.notify_all();
                continue;
            }
            // nothing finished: wait briefly, or until a new query arrives
            if (!inflight.empty()) cv_work.wait_for(lk, std::chrono::microseconds(20));
        }
    }
    bool running_ = false;

    void run()
    {
        if (sd) {
            run_slots();
            return;
        }
        std::vector<uint64_t> tickets;
        for (;;) {
            std::unique_lock<std::mutex> lk(mu);
            cv_work.wait(lk, [&] { return stop || has_work(); });
            if (stop) return;
            // the kind whose oldest query has waited longest; an SSSP queue
            // too short for a bit-packed launch yields to the other kinds for
            // up to kSsspHoldMs, so that under load SSSP batches grow into
            // the cheaper bit-packed form instead of running one by one
            const double tnow = now_s();
            int kind = -1;
            double oldest = 0.0;
            for (int pass = 0; pass < 2 && kind < 0; pass++) {
                for (int k = 0; k < SG_QUERY_KINDS; k++) {
                    if (queue[k].empty()) continue;
                    const double t = reqs[queue[k].front()].t_arrival;
                    if (pass == 0 && k == SG_QUERY_SSSP && queue[k].size() < (size_t)cfg.sssp_packed_min &&
                        (tnow - t) * 1e3 < kSsspHoldMs)
                        continue;
                    if (kind < 0 || t < oldest) {
                        kind = k;
                        oldest = t;
                    }
                }
            }
            const uint32_t limit = (uint32_t)std::max(1, std::min(64, cfg.max_batch[kind] ? cfg.max_batch[kind] : 64));
            const uint32_t b = (uint32_t)std::min<size_t>({queue[kind].size(), free_rows.size(), (size_t)limit});
            tickets.clear();
            const double t0 = now_s();
            for (uint32_t i = 0; i < b; i++) {
                const uint64_t t = queue[kind].front();
                queue[kind].pop_front();
                Req &r = reqs[t];
                r.row = (int32_t)free_rows.back();
                free_rows.pop_back();
                r.state = kRunning;
                r.t_start = t0;
                h_src[i] = r.source;
                h_qsrc[i] = r.source_id;
                h_rows[i] = (uint32_t)r.row;
                tickets.push_back(t);
            }
            lk.unlock();

            // A graph kept in degree order on the device computes into the
            // engine's work rows 0..b-1, then gathers each into its pool row
            // in the caller's vertex order; otherwise kernels write the pool.
            const bool via_work = g->relabeled;
            const size_t n = g->dev.n;
            uint32_t *work = (uint32_t *)e->work;
            bool ok = true;
            if (kind == SG_QUERY_BFS && b < (uint32_t)cfg.bfs_packed_min) {
                // one whole-grid query after another, each writing its row in
                // the caller's order
                for (uint32_t i = 0; ok && i < b; i++) {
                    int32_t *row = (int32_t *)pool + (size_t)h_rows[i] * n;
                    ok = bfs_launch(e, h_qsrc[i], row, false) == SG_OK && engine_run_status(e) == SG_OK;
                }
            } else if (kind == SG_QUERY_SSSP && b < (uint32_t)cfg.sssp_packed_min) {
                // one whole-grid query after another, each into its row
                // (a relabelled graph's kernel gathers into the pool row)
                for (uint32_t i = 0; ok && i < b; i++) {
                    uint32_t *row = (uint32_t *)pool + (size_t)h_rows[i] * n;
                    ok = sssp_single_launch(e, h_src[i], via_work ? work : row, row) == SG_OK &&
                         engine_run_status(e) == SG_OK;
                }
            } else {
                ok = hipMemcpyAsync(d_src, h_src, b * 4, hipMemcpyHostToDevice, e->stream) == hipSuccess &&
                     hipMemcpyAsync(d_rows, h_rows, b * 4, hipMemcpyHostToDevice, e->stream) == hipSuccess &&
                     hipMemsetAsync(e->stats, 0, sizeof(DevStats), e->stream) == hipSuccess;
                ok = ok && ms_launch(e, kind_to_ms(kind), d_src, b, via_work ? (void *)work : pool,
                                     via_work ? nullptr : d_rows, true, true) == SG_OK;
                for (uint32_t j = 0; ok && via_work && j < b; j++)
                    ok = engine_gather_rows(e, work + j * n, (uint32_t *)pool + (size_t)h_rows[j] * n, 1) == SG_OK;
                ok = ok && engine_run_status(e) == SG_OK;
            }
            const double t1 = now_s();

            lk.lock();
            batches++;
            batched_queries += b;
            busy_s += t1 - t0;
            for (uint64_t t : tickets) finish_locked(t, reqs[t], ok, t1);
            lk.unlock();
            cv_done.notify_all();
            cv_work.notify_all();
        }
    }
};

extern "C" void sg_server_config_default(sg_server_config *c)
{
    if (!c) return;
    std::memset(c, 0, sizeof(*c));
    c->memory_budget = 4ull << 30;
    c->max_rows = 4096;
    for (int k = 0; k < SG_QUERY_KINDS; k++) c->max_batch[k] = 64;
}

namespace {

int server_create(sg_graph *const *graphs, int32_t ngraphs, const sg_server_config *cfg, const sg_engine_options *opt,
                  sg_server **out)
{
    if (!graphs || !out) SG_FAIL(SG_ERR_INVALID, "sg_server_create: null argument");
    *out = nullptr;
    if (ngraphs < 1 || ngraphs > kServeGraphs) SG_FAIL(SG_ERR_RANGE, "sg_server_create: 1..4 graphs");
    for (int32_t i = 0; i < ngraphs; i++)
        if (!graphs[i]) SG_FAIL(SG_ERR_INVALID, "sg_server_create: null graph");
    sg_graph *g = graphs[0];
    sg_server_config c;
    if (cfg) c = *cfg;
    else sg_server_config_default(&c);
    for (int k = 0; k < SG_QUERY_KINDS; k++)
        if (c.max_batch[k] < 0 || c.max_batch[k] > 64) SG_FAIL(SG_ERR_RANGE, "sg_server_create: max_batch in [0, 64]");
    if (c.bfs_packed_min < 0 || c.bfs_packed_min > 65 || c.sssp_packed_min < 0 || c.sssp_packed_min > 65)
        SG_FAIL(SG_ERR_RANGE, "sg_server_create: packed_min in [0, 65]");
    if (c.mode != SG_SERVE_SLOTS && c.mode != SG_SERVE_BATCHES) SG_FAIL(SG_ERR_INVALID, "sg_server_create: unknown mode");
    if (c.mode == SG_SERVE_BATCHES && ngraphs > 1)
        SG_FAIL(SG_ERR_UNSUPPORTED, "sg_server_create: several graphs need SG_SERVE_SLOTS");
    if (c.sssp_prio_graphs >> ngraphs) SG_FAIL(SG_ERR_RANGE, "sg_server_create: sssp_prio_graphs names a graph not given");
    if (c.mode == SG_SERVE_BATCHES && c.sssp_prio_graphs)
        SG_FAIL(SG_ERR_UNSUPPORTED, "sg_server_create: per-graph SSSP priority needs SG_SERVE_SLOTS");
    if (!c.bfs_packed_min) c.bfs_packed_min = (int32_t)kMsBfsMin;
    if (!c.sssp_packed_min) c.sssp_packed_min = (int32_t)kMsSsspMin;
    // a result row holds one query's values on the largest graph
    uint64_t n = 0;
    for (int32_t i = 0; i < ngraphs; i++) n = std::max<uint64_t>(n, graphs[i]->dev.n);
    const uint64_t row_bytes = n * 4;
    uint64_t rows = c.memory_budget / row_bytes;
    if (c.max_rows) rows = std::min<uint64_t>(rows, (uint64_t)c.max_rows);
    if (rows < 1) SG_FAIL(SG_ERR_CAPACITY, "sg_server_create: memory budget below one result row");
    std::unique_ptr<sg_server> s(new sg_server());
    s->g = g;
    s->graphs.assign(graphs, graphs + ngraphs);
    s->cfg = c;
    s->row_bytes = row_bytes;
    s->nrows = (uint32_t)std::min<uint64_t>(rows, 0xFFFFFFFFull);
    int rc = sg_engine_create(g, &s->e);
    if (rc) return rc;
    if (opt && (rc = sg_engine_set_options(s->e, opt))) {
        sg_engine_destroy(s->e);
        return rc;
    }
    auto fail = [&](int code) {
        sg_server *p = s.release();
        sg_server_destroy(p);
        return code;
    };
    if (c.mode == SG_SERVE_BATCHES) {
        // working arenas of every kind (the largest wins); small SSSP batches
        // run single-source queries on the engine's own frontier state,
        // everything else the bit-packed kernels
        if ((rc = ms_prepare(s->e, kMsBfs))) return fail(rc);
        if ((rc = ms_prepare(s->e, kMsSssp))) return fail(rc);
        if ((rc = ms_prepare(s->e, kMsBcDep))) return fail(rc);
        if (g->relabeled && (rc = engine_ensure_work(s->e, 64ull * g->dev.n * 4))) return fail(rc);
    }
    if ((rc = engine_alloc(&s->pool, (uint64_t)s->nrows * row_bytes))) return fail(rc);
    if (c.mode == SG_SERVE_SLOTS &&
        (rc = serve_create(graphs, ngraphs, s->e->opt, c.sssp_prio_graphs, s->e->stream, s->pool, row_bytes, &s->sd)))
        return fail(rc);
    if ((rc = engine_alloc((void **)&s->d_src, 256))) return fail(rc);
    if ((rc = engine_alloc((void **)&s->d_rows, 256))) return fail(rc);
    if (hipHostMalloc((void **)&s->h_src, 256, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void **)&s->h_rows, 256, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        sg::set_error("sg_server_create: pinned staging allocation failed");
        return fail(SG_ERR_NOMEM);
    }
    s->free_rows.reserve(s->nrows);
    for (uint32_t r = s->nrows; r-- > 0;) s->free_rows.push_back(r);
    sg_server *p = s.release();
    p->worker = std::thread([p] { p->run(); });
    *out = p;
    return SG_OK;
}

}  // namespace

extern "C" int sg_server_create(sg_graph *g, const sg_server_config *cfg, const sg_engine_options *opt,
                                sg_server **out)
{
    if (!g || !out) SG_FAIL(SG_ERR_INVALID, "sg_server_create: null argument");
    return server_create(&g, 1, cfg, opt, out);
}

extern "C" int sg_server_create_multi(sg_graph *const *graphs, int32_t ngraphs, const sg_server_config *cfg,
                                      const sg_engine_options *opt, sg_server **out)
{
    return server_create(graphs, ngraphs, cfg, opt, out);
}

extern "C" int sg_server_destroy(sg_server *s)
{
    if (!s) return SG_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->worker.joinable()) s->worker.join();
    if (s->sd) serve_destroy(s->sd);  // lets a running launch drain the ring, waits for it
    if (s->e) {
        (void)hipStreamSynchronize(s->e->stream);
        sg_engine_destroy(s->e);
    }
    if (s->pool) (void)hipFree(s->pool);
    if (s->d_src) (void)hipFree(s->d_src);
    if (s->d_rows) (void)hipFree(s->d_rows);
    if (s->h_src) (void)hipHostFree(s->h_src);
    if (s->h_rows) (void)hipHostFree(s->h_rows);
    delete s;
    return SG_OK;
}

namespace {

int submit_locked(sg_server *s, int32_t graph, int32_t kind, uint32_t source, double t_arrival, bool auto_release,
                  double *lat, uint64_t *ticket)
{
    const uint64_t t = s->next_ticket++;
    sg_server::Req r{};
    r.graph = graph;
    r.kind = kind;
    r.source = dev_vertex(s->graphs[graph], source);  // device vertex order
    r.source_id = source;
    r.row = -1;
    r.state = kQueued;
    r.status = SG_OK;
    r.auto_release = auto_release;
    r.t_arrival = t_arrival;
    r.latency_out = lat;
    s->reqs.emplace(t, r);
    s->queue[kind].push_back(t);
    s->submitted++;
    if (ticket) *ticket = t;
    return SG_OK;
}

}  // namespace

extern "C" int sg_server_submit_to(sg_server *s, int32_t graph, int32_t kind, uint32_t source, uint64_t *ticket)
{
    if (!s || !ticket) SG_FAIL(SG_ERR_INVALID, "sg_server_submit: null argument");
    if (graph < 0 || graph >= (int32_t)s->graphs.size()) SG_FAIL(SG_ERR_RANGE, "sg_server_submit: unknown graph");
    if (kind < 0 || kind >= SG_QUERY_KINDS) SG_FAIL(SG_ERR_INVALID, "sg_server_submit: unknown query kind");
    if (source >= s->graphs[graph]->dev.n) SG_FAIL(SG_ERR_RANGE, "sg_server_submit: source >= n");
    {
        std::lock_guard<std::mutex> lk(s->mu);
        submit_locked(s, graph, kind, source, now_s(), false, nullptr, ticket);
    }
    s->cv_work.notify_one();
    return SG_OK;
}

extern "C" int sg_server_submit(sg_server *s, int32_t kind, uint32_t source, uint64_t *ticket)
{
    return sg_server_submit_to(s, 0, kind, source, ticket);
}

extern "C" int sg_server_wait(sg_server *s, uint64_t ticket, int32_t timeout_ms)
{
    if (!s) SG_FAIL(SG_ERR_INVALID, "sg_server_wait: null server");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) SG_FAIL(SG_ERR_INVALID, "sg_server_wait: unknown ticket");
    auto done = [&] {
        auto jt = s->reqs.find(ticket);
        return jt == s->reqs.end() || jt->second.state >= kDone;
    };
    if (timeout_ms < 0) s->cv_done.wait(lk, done);
    else if (!s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), done))
        SG_FAIL(SG_ERR_TIMEOUT, "sg_server_wait: not finished");
    auto jt = s->reqs.find(ticket);
    if (jt != s->reqs.end() && jt->second.state == kFailed)
        SG_FAIL(jt->second.status, "sg_server_wait: query failed with status " + std::to_string(jt->second.status));
    return SG_OK;
}

extern "C" int sg_server_result(sg_server *s, uint64_t ticket, void **dev_ptr, uint64_t *bytes, double *queue_ms,
                                double *total_ms)
{
    if (!s) SG_FAIL(SG_ERR_INVALID, "sg_server_result: null server");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) SG_FAIL(SG_ERR_INVALID, "sg_server_result: unknown ticket");
    const sg_server::Req &r = it->second;
    if (r.state != kDone) SG_FAIL(SG_ERR_STATE, "sg_server_result: query not finished");
    if (dev_ptr) *dev_ptr = (char *)s->pool + (size_t)r.row * s->row_bytes;
    if (bytes) *bytes = (uint64_t)s->graphs[r.graph]->dev.n * 4;
    if (queue_ms) *queue_ms = (r.t_start - r.t_arrival) * 1e3;
    if (total_ms) *total_ms = (r.t_done - r.t_arrival) * 1e3;
    return SG_OK;
}

extern "C" int sg_server_copy_result(sg_server *s, uint64_t ticket, void *host_out, uint64_t bytes)
{
    void *p = nullptr;
    uint64_t nb = 0;
    int rc = sg_server_result(s, ticket, &p, &nb, nullptr, nullptr);
    if (rc) return rc;
    if (!host_out || bytes < nb) SG_FAIL(SG_ERR_CAPACITY, "sg_server_copy_result: host buffer too small");
    SG_HIP(hipMemcpy(host_out, p, nb, hipMemcpyDeviceToHost));
    return SG_OK;
}

extern "C" int sg_server_release(sg_server *s, uint64_t ticket)
{
    if (!s) SG_FAIL(SG_ERR_INVALID, "sg_server_release: null server");
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->reqs.find(ticket);
        if (it == s->reqs.end()) SG_FAIL(SG_ERR_INVALID, "sg_server_release: unknown ticket");
        if (it->second.state < kDone) SG_FAIL(SG_ERR_STATE, "sg_server_release: query still queued or running");
        if (it->second.row >= 0) s->free_rows.push_back((uint32_t)it->second.row);
        s->reqs.erase(it);
    }
    s->cv_work.notify_one();
    return SG_OK;
}

extern "C" int sg_server_get_stats(sg_server *s, sg_server_stats *out)
{
    if (!s || !out) SG_FAIL(SG_ERR_INVALID, "sg_server_get_stats: null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    std::memset(out, 0, sizeof(*out));
    out->submitted = s->submitted;
    out->completed = s->completed;
    out->failed = s->failed;
    // slots: launches of the serving kernel, each serving many queries
    const uint64_t nb = s->sd ? s->slot_launches : s->batches;
    out->batches = nb;
    out->mean_batch = nb ? (double)s->batched_queries / (double)nb : 0.0;
    out->busy_ms = s->busy_s * 1e3;
    out->rows = s->nrows;
    out->free_rows = (uint32_t)s->free_rows.size();
    uint64_t q = 0;
    for (int k = 0; k < SG_QUERY_KINDS; k++) q += s->queue[k].size();
    out->queued = q;
    uint64_t gbytes = 0;
    for (const sg_graph *g : s->graphs) gbytes += g->device_bytes;
    out->device_bytes = gbytes + s->e->aux_bytes + (uint64_t)s->nrows * s->row_bytes + s->e->value_cap +
                        (s->sd ? serve_device_bytes(s->sd) : 0);
    return SG_OK;
}

extern "C" int sg_server_replay_multi(sg_server *s, const int32_t *graphs, const int32_t *kinds, const uint32_t *sources,
                                      const double *arrival_s, int64_t count, double *latency_ms)
{
    if (!s || !kinds || !sources || !arrival_s || !latency_ms) SG_FAIL(SG_ERR_INVALID, "sg_server_replay: null argument");
    for (int64_t i = 0; i < count; i++) {
        const int32_t gi = graphs ? graphs[i] : 0;
        if (gi < 0 || gi >= (int32_t)s->graphs.size()) SG_FAIL(SG_ERR_RANGE, "sg_server_replay: unknown graph");
        if (kinds[i] < 0 || kinds[i] >= SG_QUERY_KINDS) SG_FAIL(SG_ERR_INVALID, "sg_server_replay: unknown query kind");
        if (sources[i] >= s->graphs[gi]->dev.n) SG_FAIL(SG_ERR_RANGE, "sg_server_replay: source >= n");
        if (i && arrival_s[i] < arrival_s[i - 1]) SG_FAIL(SG_ERR_INVALID, "sg_server_replay: arrivals must not decrease");
        latency_ms[i] = -1.0;
    }
    uint64_t base_completed, base_failed;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        base_completed = s->completed;
        base_failed = s->failed;
    }
    // open loop: each query is submitted at its arrival time (measured from
    // the start of the replay); latency counts from that scheduled time
    const double t0 = now_s();
    for (int64_t i = 0; i < count; i++) {
        const double due = t0 + arrival_s[i];
        for (;;) {
            const double t = now_s();
            if (t >= due) break;
            if (due - t > 2e-4) std::this_thread::sleep_for(std::chrono::microseconds((int64_t)((due - t) * 1e6) - 100));
        }
        {
            std::lock_guard<std::mutex> lk(s->mu);
            submit_locked(s, graphs ? graphs[i] : 0, kinds[i], sources[i], due, true, &latency_ms[i], nullptr);
        }
        s->cv_work.notify_one();
    }
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait(lk, [&] { return s->completed + s->failed - base_completed - base_failed >= (uint64_t)count; });
    if (s->failed != base_failed)
        SG_FAIL(s->last_failure, "sg_server_replay: " + std::to_string(s->failed - base_failed) +
                                     " queries failed; the last with status " + std::to_string(s->last_failure));
    return SG_OK;
}

extern "C" int sg_server_replay(sg_server *s, const int32_t *kinds, const uint32_t *sources, const double *arrival_s,
                                int64_t count, double *latency_ms)
{
    return sg_server_replay_multi(s, nullptr, kinds, sources, arrival_s, count, latency_ms);
}
