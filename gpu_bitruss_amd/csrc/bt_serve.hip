// Serving: a request loop with continuous batching on one device.
//
// No counterpart in the reference (one graph per process). Requests carry a
// host graph and an algorithm; the server copies the graph, estimates the
// device memory the request needs, and routes it to one of three lanes:
//   - light: small graphs with few priority wedges (below
//     cfg.light_wedges); their worker takes every pending light request of
//     the oldest request's algorithm that fits the memory budget (up to
//     cfg.max_batch) and runs them as one batch (bt_batch_*); requests that
//     arrive meanwhile form the next batch;
//   - heavy: the other small graphs, batched the same way, a batch holding
//     at most cfg.batch_wedges priority wedges so that it stays short (a
//     batch ends when its slowest graph does);
//   - large: graphs with at least cfg.large_edges edges, batched the same
//     way (a union of large graphs costs about as many phases as its
//     largest graph, so concurrent large requests finish together instead
//     of one after the other).
// The light lane has two workers, each with its own context: while one
// worker's batch runs, requests that arrive are taken by the other, whose
// copy to the device, support count and bloom index overlap the running
// peel (its own peel starts when the CU lease has room).
// Each lane has its own context (HIP stream); the lanes' kernels overlap on
// the device, and their persistent peels run side by side on disjoint CU
// shares (bt_cu_lease; a quarter for the light lane, three eighths for the
// heavy lane and the rest for large requests by default), so a large
// request does not hold up the batches. Admission: a request starts only
// when the estimated bytes of all running requests plus its own stay within
// cfg.budget_bytes; requests wait in arrival order per lane.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <new>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/bitruss_hip.h"
#include "bt_internal.h"

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
    uint64_t ticket = 0;
    int algo = 0;
    bt_host_graph g{};                  // owned copy
    std::vector<uint32_t> arrays;       // storage of g's arrays
    uint64_t est_bytes = 0;
    uint64_t wedges = 0;                // priority wedges
    uint64_t run_bytes = 0;             // the run's fixed share, on one request of it
    std::vector<uint64_t> out;          // per-edge result
    bt_batch_result res{};
    int status = BT_OK;
    bool done = false;
    Clock::time_point t_submit, t_done;
};

using RequestPtr = std::shared_ptr<Request>;

// Device bytes a request may hold: the resident graph and its scratch
// (~137 B per edge, 40 B per vertex) and the bloom index (~44 B per
// priority wedge). Each run (a batch or a large request) adds the peel's
// control buffers once.
uint64_t estimate_bytes(const bt_host_graph& g, uint64_t wedges) {
    return 140ull * g.m + 48ull * g.n + 48ull * wedges;
}
constexpr uint64_t kRunBytes = 64ull << 20;

}  // namespace

struct bt_server {
    bt_serve_config cfg{};
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<RequestPtr> light_q, heavy_q, large_q;
    std::unordered_map<uint64_t, RequestPtr> all;
    uint64_t next_ticket = 1;
    uint64_t in_use = 0;
    bool stopping = false;
    bt_serve_stats stats{};
    bt_ctx* ctx_light = nullptr;
    bt_ctx* ctx_light2 = nullptr;  // the light lane's second worker (may stay null)
    bt_ctx* ctx_heavy = nullptr;
    bt_ctx* ctx_large = nullptr;
    std::thread thr_light, thr_light2, thr_heavy, thr_large;

    // takes the next group of requests from q that fits the budget (caller
    // holds mu); empty when the front does not fit yet
    std::vector<RequestPtr> take(std::deque<RequestPtr>& q, bool batch, uint64_t max_wedges) {
        std::vector<RequestPtr> got;
        uint64_t add = kRunBytes, wedges = 0;
        int const algo = q.empty() ? 0 : q.front()->algo;
        for (auto it = q.begin(); it != q.end() && got.size() < (batch ? cfg.max_batch : 1u);) {
            RequestPtr const& r = *it;
            if (r->algo != algo) {
                ++it;
                continue;
            }
            if (in_use + add + r->est_bytes > cfg.budget_bytes) break;
            if (!got.empty() && wedges + r->wedges > max_wedges) break;
            wedges += r->wedges;
            add += r->est_bytes;
            got.push_back(r);
            it = q.erase(it);
        }
        if (got.empty()) return got;
        in_use += add;
        if (in_use > stats.peak_bytes) stats.peak_bytes = in_use;
        got.front()->run_bytes = kRunBytes;  // given back with the first request
        return got;
    }

    void finish(std::vector<RequestPtr>& rs, int st) {
        std::lock_guard<std::mutex> lk(mu);
        for (RequestPtr& r : rs) {
            r->status = r->status == BT_OK ? st : r->status;
            r->done = true;
            r->t_done = Clock::now();
            in_use -= r->est_bytes + r->run_bytes;
            stats.completed++;
            if (r->status != BT_OK) stats.failed++;
            // the graph copy is no longer needed
            r->arrays.clear();
            r->arrays.shrink_to_fit();
        }
        cv_done.notify_all();
        cv_work.notify_all();
    }

    // waits for work in q (or the end); false when the server stops
    bool next(std::deque<RequestPtr>& q, bool batch, std::vector<RequestPtr>& rs) {
        std::unique_lock<std::mutex> lk(mu);
        uint64_t const max_wedges = &q == &heavy_q ? cfg.batch_wedges : ~0ull;
        for (;;) {
            if (!q.empty()) {
                rs = take(q, batch, max_wedges);
                if (!rs.empty()) break;
            } else if (stopping) {
                return false;
            }
            cv_work.wait(lk);
        }
        if (&q == &large_q) {
            stats.large_runs++;
        } else {
            stats.batches++;
            stats.batched_requests += rs.size();
        }
        return true;
    }

    void run_batches(std::deque<RequestPtr>& q, bt_ctx* ctx) {
        std::vector<RequestPtr> rs;
        while (next(q, true, rs)) {
            std::vector<bt_host_graph> gs(rs.size());
            std::vector<uint64_t*> outs(rs.size());
            std::vector<bt_batch_result> res(rs.size());
            for (size_t i = 0; i < rs.size(); i++) {
                gs[i] = rs[i]->g;
                rs[i]->out.assign(rs[i]->algo == BT_ALGO_BFC ? 0 : rs[i]->g.m, 0);
                outs[i] = rs[i]->out.empty() ? nullptr : rs[i]->out.data();
            }
            bt_batch* b = nullptr;
            int st = bt_batch_upload(ctx, gs.data(), gs.size(), &b);
            if (st == BT_OK) st = bt_batch_run(ctx, b, rs.front()->algo, outs.data(), res.data());
            if (b) bt_batch_free(b);
            if (st == BT_OK)
                for (size_t i = 0; i < rs.size(); i++) rs[i]->res = res[i];
            finish(rs, st);
        }
    }

};

extern "C" {

int bt_serve_start(const bt_serve_config* cfg, bt_server** out) {
    if (!out) return BT_ERR_INVALID;
    *out = nullptr;
    bt_server* s = new (std::nothrow) bt_server();
    if (!s) return BT_ERR_NOMEM;
    if (cfg) s->cfg = *cfg;
    if (!s->cfg.budget_bytes) s->cfg.budget_bytes = 64ull << 30;
    if (!s->cfg.large_edges) s->cfg.large_edges = 1000000;
    if (!s->cfg.max_batch) s->cfg.max_batch = 1024;
    if (!s->cfg.light_wedges) s->cfg.light_wedges = 2000000;
    if (!s->cfg.batch_wedges) s->cfg.batch_wedges = 100000000;
    if (!s->cfg.light_workers) s->cfg.light_workers = 2;
    if (s->cfg.light_workers > 2) {
        delete s;
        return BT_ERR_INVALID;
    }
    int st = bt_ctx_create(s->cfg.device, &s->ctx_light);
    if (st == BT_OK && s->cfg.light_workers == 2) st = bt_ctx_create(s->cfg.device, &s->ctx_light2);
    if (st == BT_OK) st = bt_ctx_create(s->cfg.device, &s->ctx_heavy);
    if (st == BT_OK) st = bt_ctx_create(s->cfg.device, &s->ctx_large);
    if (st == BT_OK) {
        // disjoint CU shares: a quarter for the light lane, three eighths
        // for the heavy lane, the rest for large requests
        // (each a multiple of the 8 XCDs, so that shares adding up to the
        // CU count fit every XCD together)
        int const cus = s->ctx_light->cu_count;
        auto xcd = [](int x) { return std::max(kXcds, x / kXcds * kXcds); };
        int const light = xcd(s->cfg.light_cus ? int(s->cfg.light_cus) : cus / 4);
        int const heavy = xcd(s->cfg.heavy_cus ? int(s->cfg.heavy_cus) : 3 * cus / 8);
        int const large = xcd(s->cfg.large_cus ? int(s->cfg.large_cus) : cus - light - heavy);
        if (light > cus || heavy > cus || large > cus) st = BT_ERR_INVALID;
        // a share of all CUs is a cooperative launch on the whole device
        auto share = [cus](int x) { return x >= cus ? 0 : x; };
        if (st == BT_OK) st = bt_ctx_set_option(s->ctx_light, BT_OPT_PEEL_BLOCKS, share(light));
        if (st == BT_OK && s->ctx_light2) st = bt_ctx_set_option(s->ctx_light2, BT_OPT_PEEL_BLOCKS, share(light));
        if (st == BT_OK) st = bt_ctx_set_option(s->ctx_heavy, BT_OPT_PEEL_BLOCKS, share(heavy));
        if (st == BT_OK) st = bt_ctx_set_option(s->ctx_large, BT_OPT_PEEL_BLOCKS, share(large));
        // a served batch is one union on its lane's CUs
        if (st == BT_OK) st = bt_ctx_set_option(s->ctx_light, BT_OPT_BATCH_LANES, 1);
        if (st == BT_OK && s->ctx_light2) st = bt_ctx_set_option(s->ctx_light2, BT_OPT_BATCH_LANES, 1);
        if (st == BT_OK) st = bt_ctx_set_option(s->ctx_heavy, BT_OPT_BATCH_LANES, 1);
        if (st == BT_OK) st = bt_ctx_set_option(s->ctx_large, BT_OPT_BATCH_LANES, 1);
    }
    auto destroy_ctx = [s] {
        if (s->ctx_light) bt_ctx_destroy(s->ctx_light);
        if (s->ctx_light2) bt_ctx_destroy(s->ctx_light2);
        if (s->ctx_heavy) bt_ctx_destroy(s->ctx_heavy);
        if (s->ctx_large) bt_ctx_destroy(s->ctx_large);
    };
    if (st != BT_OK) {
        destroy_ctx();
        delete s;
        return st;
    }
    try {
        s->thr_light = std::thread([s] { s->run_batches(s->light_q, s->ctx_light); });
        if (s->ctx_light2) s->thr_light2 = std::thread([s] { s->run_batches(s->light_q, s->ctx_light2); });
        s->thr_heavy = std::thread([s] { s->run_batches(s->heavy_q, s->ctx_heavy); });
        s->thr_large = std::thread([s] { s->run_batches(s->large_q, s->ctx_large); });
    } catch (...) {
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->stopping = true;
        }
        s->cv_work.notify_all();
        for (std::thread* t : {&s->thr_light, &s->thr_light2, &s->thr_heavy, &s->thr_large})
            if (t->joinable()) t->join();
        destroy_ctx();
        delete s;
        return BT_ERR_NOMEM;
    }
    *out = s;
    return BT_OK;
}

int bt_serve_submit(bt_server* s, const bt_host_graph* g, int algo, uint64_t* ticket) {
    if (!s || !g || !ticket || algo < BT_ALGO_BFC || algo > BT_ALGO_MSP) return BT_ERR_INVALID;
    BT_TRY(bt_graph_check(g));
    uint64_t wedges = 0;
    BT_TRY(bt_graph_wedges(g, &wedges));
    RequestPtr r;
    try {
        r = std::make_shared<Request>();
        size_t const m2 = 2ull * g->m, n = g->n;
        r->arrays.resize(3 * m2 + 2 * n + 1);
        uint32_t* p = r->arrays.data();
        auto put = [&](const uint32_t* src, size_t cnt) {
            uint32_t* d = p;
            if (cnt) std::memcpy(d, src, sizeof(uint32_t) * cnt);
            p += cnt;
            return d;
        };
        r->g = *g;
        r->g.neighbors = put(g->neighbors, m2);
        r->g.edges = put(g->edges, m2);
        r->g.edge_ids = put(g->edge_ids, m2);
        r->g.degrees = put(g->degrees, n);
        r->g.offsets = put(g->offsets, n + 1);
    } catch (...) {
        return BT_ERR_NOMEM;
    }
    r->algo = algo;
    r->est_bytes = estimate_bytes(*g, wedges);
    r->wedges = wedges;
    if (r->est_bytes + kRunBytes > s->cfg.budget_bytes) return BT_ERR_BUDGET;
    r->t_submit = Clock::now();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stopping) return BT_ERR_STATE;
        r->ticket = s->next_ticket++;
        s->all[r->ticket] = r;
        std::deque<RequestPtr>& q = g->m >= s->cfg.large_edges ? s->large_q
                                    : wedges < s->cfg.light_wedges ? s->light_q
                                                                   : s->heavy_q;
        q.push_back(r);
        s->stats.submitted++;
        *ticket = r->ticket;
    }
    s->cv_work.notify_all();
    return BT_OK;
}

int bt_serve_poll(bt_server* s, uint64_t ticket, int* done) {
    if (!s || !done) return BT_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->all.find(ticket);
    if (it == s->all.end()) return BT_ERR_INVALID;
    *done = it->second->done ? 1 : 0;
    return BT_OK;
}

int bt_serve_wait(bt_server* s, uint64_t ticket, uint64_t* out, bt_batch_result* result, double* latency_ms) {
    if (!s) return BT_ERR_INVALID;
    RequestPtr r;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        auto it = s->all.find(ticket);
        if (it == s->all.end()) return BT_ERR_INVALID;
        // the request itself is held, not the iterator: the wait releases
        // the lock, and a submit meanwhile may rehash the map
        r = it->second;
        s->cv_done.wait(lk, [&] { return r->done; });
        // another waiter on the same ticket may have released it already
        if (s->all.erase(ticket) == 0) return BT_ERR_INVALID;
    }
    if (r->status != BT_OK) return r->status;
    if (out && !r->out.empty()) std::memcpy(out, r->out.data(), sizeof(uint64_t) * r->out.size());
    if (result) *result = r->res;
    if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
    return BT_OK;
}

int bt_serve_stats_get(bt_server* s, bt_serve_stats* st) {
    if (!s || !st) return BT_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    *st = s->stats;
    st->in_use_bytes = s->in_use;
    return BT_OK;
}

int bt_serve_stop(bt_server* s) {
    if (!s) return BT_ERR_INVALID;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stopping = true;
    }
    s->cv_work.notify_all();
    for (std::thread* t : {&s->thr_light, &s->thr_light2, &s->thr_heavy, &s->thr_large})
        if (t->joinable()) t->join();
    bt_ctx_destroy(s->ctx_light);
    if (s->ctx_light2) bt_ctx_destroy(s->ctx_light2);
    bt_ctx_destroy(s->ctx_heavy);
    bt_ctx_destroy(s->ctx_large);
    delete s;
    return BT_OK;
}

}  // extern "C"
