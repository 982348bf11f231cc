// serve.cpp -- a query server over a resident database: continuous batching
// of single queries into batched scans, written on the public C ABI.
//
// Requests are queued on submit.  A worker thread admits queued requests
// into a batch when a device slot is free (at most max_inflight batches on
// the device, on the context's two stream lanes) and the batch's result
// buffers fit the HBM budget: immediately when the device is idle, otherwise
// once max_batch requests wait or the oldest has waited max_wait_us.
// Requests arriving while batches run are admitted as slots free up; two
// batches in flight overlap on the device.  Each finished batch's records
// are split per request (records are ordered by query first), so a request's
// result is exactly what scanning it alone returns.
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

#include "bpx.h"

namespace bpx {
// sets bpx_last_error() of the calling thread (runtime.hip)
void set_last_error(const char* msg);
}

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
    uint64_t id = 0;
    std::vector<uint8_t> q;
    Clock::time_point t_submit, t_launch, t_done;
    bool done = false;
    int status = BPX_OK;
    std::string err;  // bpx_last_error() of the worker when status != BPX_OK
    uint32_t diag = 0;  // diagonal-table slots of the query (bpx_query_footprint)
    uint32_t batch_size = 0;
    std::vector<bpx_init_hsp> recs;
    std::vector<uint32_t> subs;
    bpx_ungapped_stats stats{};
};
using ReqPtr = std::shared_ptr<Request>;

struct Inflight {
    bpx_batch* b = nullptr;
    bpx_result* r = nullptr;
    std::vector<ReqPtr> reqs;
};

double ms_between(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

}  // namespace

struct bpx_server {
    bpx_ctx* ctx = nullptr;
    const bpx_db* db = nullptr;
    bpx_server_config cfg{};
    uint32_t nsub = 0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<ReqPtr> pending;
    std::unordered_map<uint64_t, ReqPtr> reqs;
    uint64_t next_id = 1;
    bool stop = false;
    std::thread worker;
    // totals
    uint64_t batches = 0, completed = 0;

    void run();
    void finish(Inflight& f);
};

void bpx_server::finish(Inflight& f) {
    int rc = bpx_result_wait(f.r);
    std::vector<bpx_init_hsp> recs;
    std::vector<uint32_t> subs, qids;
    std::vector<bpx_ungapped_stats> qst(f.reqs.size());
    uint64_t n = 0;
    if (rc == BPX_OK) rc = bpx_result_summary(f.r, &n, nullptr);
    if (rc == BPX_OK) {
        recs.resize(std::max<uint64_t>(n, 1));
        subs.resize(recs.size());
        qids.resize(recs.size());
        rc = bpx_result_records(f.r, recs.data(), subs.data(), qids.data(), recs.size());
    }
    if (rc == BPX_OK) rc = bpx_result_query_stats(f.r, qst.data(), (uint32_t)qst.size());
    const std::string err = rc == BPX_OK ? std::string() : std::string(bpx_last_error());
    const auto now = Clock::now();
    {
        std::lock_guard<std::mutex> lk(mu);
        uint64_t at = 0;
        for (uint32_t j = 0; j < f.reqs.size(); ++j) {
            Request& q = *f.reqs[j];
            const uint64_t from = at;
            while (rc == BPX_OK && at < n && qids[at] == j) ++at;
            if (rc == BPX_OK) {
                q.recs.assign(recs.begin() + (ptrdiff_t)from, recs.begin() + (ptrdiff_t)at);
                q.subs.assign(subs.begin() + (ptrdiff_t)from, subs.begin() + (ptrdiff_t)at);
                q.stats = qst[j];
            }
            q.status = rc;
            q.err = err;
            q.t_done = now;
            q.done = true;
        }
        completed += f.reqs.size();
    }
    cv_done.notify_all();
    bpx_result_destroy(f.r);
    bpx_batch_destroy(f.b);
}

void bpx_server::run() {
    std::deque<Inflight> fl;
    const size_t max_inflight = (size_t)cfg.max_inflight;
    // result buffers of one query over the database: counts and run
    // locations per subject (16 B) plus record segments (8 B per subject)
    const uint64_t per_query = (uint64_t)nsub * 24 + (1u << 20);
    const size_t budget_queries =
        (size_t)std::max<uint64_t>(1, (uint64_t)cfg.hbm_budget / std::max<uint64_t>(per_query, 1));
    const uint64_t diag_budget = (uint64_t)cfg.diag_budget;
    for (;;) {
        std::vector<ReqPtr> take;
        bool idle_wait = false;
        {
            std::unique_lock<std::mutex> lk(mu);
            if (fl.empty())
                cv_work.wait(lk, [&] { return stop || !pending.empty(); });
            if (stop && pending.empty() && fl.empty()) return;
            if (fl.size() < max_inflight && !pending.empty()) {
                // the longest prefix of the queue within max_batch, the HBM
                // result budget and the diagonal-table budget (a query above
                // the latter goes alone)
                const size_t cap = std::min((size_t)cfg.max_batch, budget_queries);
                size_t n = 0;
                uint64_t diag = 0;
                while (n < pending.size() && n < cap && (n == 0 || diag + pending[n]->diag <= diag_budget))
                    diag += pending[n++]->diag;
                const auto oldest = ms_between(pending.front()->t_submit, Clock::now()) * 1000.0;
                const bool full = n < pending.size();
                const bool launch = fl.empty() || full || oldest >= (double)cfg.max_wait_us;
                if (launch) {
                    take.assign(pending.begin(), pending.begin() + (ptrdiff_t)n);
                    pending.erase(pending.begin(), pending.begin() + (ptrdiff_t)n);
                } else {
                    idle_wait = true;
                }
            }
        }
        if (!take.empty()) {
            std::vector<const uint8_t*> ptrs(take.size());
            std::vector<int32_t> lens(take.size());
            for (size_t i = 0; i < take.size(); ++i) {
                ptrs[i] = take[i]->q.data();
                lens[i] = (int32_t)take[i]->q.size();
            }
            Inflight f;
            f.reqs = take;
            int rc = bpx_batch_prepare(ctx, (uint32_t)take.size(), ptrs.data(), lens.data(), nullptr, &cfg.params,
                                       nullptr, nullptr, &cfg.policy, &f.b);
            if (rc == BPX_OK) rc = bpx_scan_batch(ctx, f.b, db, 0, nsub, &cfg.launch, 0u, &f.r);
            const auto now = Clock::now();
            if (rc != BPX_OK) {
                const std::string err = bpx_last_error();
                bpx_batch_destroy(f.b);
                std::lock_guard<std::mutex> lk(mu);
                for (auto& q : take) {
                    q->status = rc;
                    q->err = err;
                    q->t_launch = q->t_done = now;
                    q->done = true;
                }
                cv_done.notify_all();
                continue;
            }
            {
                // batches is read by bpx_server_counters on other threads
                std::lock_guard<std::mutex> lk(mu);
                for (auto& q : take) {
                    q->t_launch = now;
                    q->batch_size = (uint32_t)take.size();
                }
                ++batches;
            }
            fl.push_back(std::move(f));
            continue;
        }
        if (fl.empty()) continue;
        const int ready = bpx_result_ready(fl.front().r);
        if (ready == 0) {
            // nothing finished yet: wait briefly for an arrival (a slot may
            // be free) or for the device
            std::unique_lock<std::mutex> lk(mu);
            cv_work.wait_for(lk, std::chrono::microseconds(idle_wait ? std::max(50, cfg.max_wait_us / 4) : 200));
            continue;
        }
        finish(fl.front());
        fl.pop_front();
    }
}

extern "C" {

int bpx_server_create(bpx_ctx* ctx, const bpx_db* db, const bpx_server_config* cfg, bpx_server** out) {
    if (!ctx || !db || !out) return BPX_E_ARG;
    *out = nullptr;
    auto* s = new (std::nothrow) bpx_server;
    if (!s) return BPX_E_NOMEM;
    s->ctx = ctx;
    s->db = db;
    if (cfg) s->cfg = *cfg;
    bpx_server_config& c = s->cfg;
    if (c.max_batch <= 0) c.max_batch = 256;
    if (c.max_inflight <= 0) c.max_inflight = 2;
    c.max_inflight = std::min(c.max_inflight, 2);
    if (c.max_wait_us <= 0) c.max_wait_us = 10000;
    if (c.diag_budget <= 0) c.diag_budget = 1 << 20;
    if (c.params.word_length == 0) {
        c.params.word_length = 3;
        c.params.threshold = 11;
        c.params.window = 40;
        c.params.x_dropoff = 16;
        c.params.cutoff_score = 41;
        c.params.diag_mask = 0;
    }
    int32_t cus = 0;
    int64_t lds = 0, hbm = 0;
    int rc = bpx_context_info(ctx, &cus, &lds, &hbm);
    if (rc == BPX_OK) rc = bpx_db_info(db, &s->nsub, nullptr, nullptr);
    if (rc) {
        delete s;
        return rc;
    }
    if (c.hbm_budget <= 0) c.hbm_budget = std::max<int64_t>(hbm / 8, 1 << 28);
    s->worker = std::thread([s] { s->run(); });
    *out = s;
    return BPX_OK;
}

int bpx_server_submit(bpx_server* s, const uint8_t* query, int32_t qlen, uint64_t* ticket) {
    if (!s || !ticket || qlen < 0 || qlen > 65535 || (qlen > 0 && !query)) {
        bpx::set_last_error("bpx_server_submit: NULL argument or query length outside [0, 65535]");
        return BPX_E_ARG;
    }
    for (int32_t i = 0; i < qlen; ++i)
        if (query[i] >= BPX_ALPHABET) {
            bpx::set_last_error("bpx_server_submit: residue code >= 28");
            return BPX_E_ARG;
        }
    // admission: a query that cannot run against the database (diagonal
    // table or LDS) is refused here rather than failing the batch it joins
    uint32_t diag = 0;
    const int fc = bpx_query_footprint(s->ctx, s->db, qlen, &s->cfg.params, &s->cfg.launch, &diag, nullptr, nullptr);
    if (fc != BPX_OK) return fc;
    auto r = std::make_shared<Request>();
    r->diag = diag;
    r->q.assign(query, query + qlen);
    r->t_submit = Clock::now();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stop) return BPX_E_STATE;
        r->id = s->next_id++;
        s->reqs[r->id] = r;
        s->pending.push_back(r);
        *ticket = r->id;
    }
    s->cv_work.notify_one();
    return BPX_OK;
}

int bpx_server_wait(bpx_server* s, uint64_t ticket, int32_t timeout_ms, bpx_request_info* info) {
    if (!s) return BPX_E_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return BPX_E_ARG;
    ReqPtr r = it->second;
    auto pred = [&] { return r->done; };
    if (timeout_ms < 0)
        s->cv_done.wait(lk, pred);
    else if (!s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), pred))
        return BPX_E_STATE;
    if (info) {
        info->num_records = r->recs.size();
        info->stats = r->stats;
        info->latency_ms = ms_between(r->t_submit, r->t_done);
        info->queue_ms = ms_between(r->t_submit, r->t_launch);
        info->batch_size = r->batch_size;
    }
    // a failure of the request's batch is reported on the caller's thread
    if (r->status != BPX_OK) bpx::set_last_error(r->err.c_str());
    return r->status;
}

int bpx_server_records(bpx_server* s, uint64_t ticket, bpx_init_hsp* records, uint32_t* subjects,
                       uint64_t capacity) {
    if (!s) return BPX_E_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return BPX_E_ARG;
    ReqPtr r = it->second;
    if (!r->done) return BPX_E_STATE;
    if (capacity < r->recs.size()) return BPX_E_ARG;
    if (records && !r->recs.empty()) std::memcpy(records, r->recs.data(), r->recs.size() * sizeof(bpx_init_hsp));
    if (subjects && !r->subs.empty()) std::memcpy(subjects, r->subs.data(), r->subs.size() * 4);
    s->reqs.erase(it);
    return BPX_OK;
}

int bpx_server_counters(bpx_server* s, uint64_t* batches, uint64_t* completed) {
    if (!s) return BPX_E_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    if (batches) *batches = s->batches;
    if (completed) *completed = s->completed;
    return BPX_OK;
}

int bpx_server_destroy(bpx_server* s) {
    if (!s) return BPX_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->worker.joinable()) s->worker.join();
    delete s;
    return BPX_OK;
}

}  // extern "C"
