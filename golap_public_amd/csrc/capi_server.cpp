// Serving loop (include/golap_hip.h, golap_server_*). No counterpart in the reference, which
// runs one plan per call on its caller's thread (usecases/ssb/queries/*.hpp).
//
// Threads:
//   - callers submit requests: each is validated and sized, then queued in arrival order;
//   - one scheduler thread admits queued requests while the device bytes of unfinished
//     requests stay within the budget. When a slot is free, it forms a batch of one kind from
//     them and enqueues it as one persistent fused launch on that slot's stream;
//   - one completion thread per slot waits for its batch, copies the results back once and
//     hands every request's result emission to a pool of host threads;
//   - callers wait for their tickets.
// Batches on different slots overlap on the device; requests that arrive while batches run
// form the next batch as soon as a slot frees (continuous batching).
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <functional>
#include <memory>
#include <thread>
#include <unordered_map>
#include <vector>

#include "batch.hpp"

using namespace golap_impl;
using namespace golap_impl::starhost;
using Clock = std::chrono::steady_clock;

namespace {

struct Request {
    uint64_t ticket = 0;
    BatchKind kind = BatchKind::STAR;
    golap_star_query star{};
    golap_scan_plan scan{};
    uint64_t bytes = 0, domain = 1;
    uint32_t n_attrs = 0;
    std::vector<const golap_column *> cols;
    Clock::time_point t_submit, t_done;
    bool done = false;
    bool claimed = false;  // a waiter owns it: a second wait on the ticket is refused
    int status = GOLAP_OK;
    std::string err;
    golap_star_result res{};
    std::unique_ptr<int32_t[]> keys;
    std::unique_ptr<int64_t[]> sums;
    std::unique_ptr<uint64_t[]> counts;
};

struct Slot {
    BatchSlot bs;
    std::thread th;
    std::condition_variable cv;
    golap_batch *batch = nullptr;
    std::vector<Request *> reqs;
    uint64_t bytes = 0;
    bool busy = false;
};

// Fixed pool of host threads for result emission.
class Pool {
   public:
    explicit Pool(uint32_t n, int device) {
        for (uint32_t i = 0; i < n; ++i)
            th_.emplace_back([this, device] {
                (void)hipSetDevice(device);
                for (;;) {
                    std::function<void()> f;
                    {
                        std::unique_lock<std::mutex> lk(m_);
                        cv_.wait(lk, [&] { return stop_ || !q_.empty(); });
                        if (q_.empty()) return;
                        f = std::move(q_.front());
                        q_.pop_front();
                    }
                    f();
                }
            });
    }
    ~Pool() {
        {
            std::lock_guard<std::mutex> lk(m_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto &t : th_) t.join();
    }
    void run(std::function<void()> f) {
        {
            std::lock_guard<std::mutex> lk(m_);
            q_.push_back(std::move(f));
        }
        cv_.notify_one();
    }

   private:
    std::vector<std::thread> th_;
    std::mutex m_;
    std::condition_variable cv_;
    std::deque<std::function<void()>> q_;
    bool stop_ = false;
};

}  // namespace

struct golap_server {
    golap_ctx *ctx = nullptr;
    golap_server_opts opts{};
    std::mutex m;
    std::condition_variable cv_sched, cv_done;
    std::deque<Request *> pending;
    std::unordered_map<uint64_t, Request *> all;
    std::vector<std::unique_ptr<Slot>> slots;
    std::unique_ptr<Pool> pool;
    std::thread sched;
    uint64_t next_ticket = 1, admitted_bytes = 0;
    bool stop = false;
    golap_server_stats stats{};
    uint64_t batched_requests = 0;
};

namespace {

int free_slot(golap_server *s) {
    for (size_t i = 0; i < s->slots.size(); ++i)
        if (!s->slots[i]->busy) return (int)i;
    return -1;
}

bool any_busy(golap_server *s) {
    for (auto &sl : s->slots)
        if (sl->busy) return true;
    return false;
}

// Columns of a forming batch: at most 8 distinct row-aligned columns.
bool fits_columns(std::vector<const golap_column *> &u, const Request *r) {
    std::vector<const golap_column *> v = u;
    for (const golap_column *c : r->cols) {
        if (std::find(v.begin(), v.end(), c) != v.end()) continue;
        if (!v.empty() && (c->n_rows != v[0]->n_rows || c->page_rows != v[0]->page_rows ||
                           c->chunk_rows != v[0]->chunk_rows))
            return false;
        v.push_back(c);
    }
    if (v.size() > STAR_MAX_COLS) return false;
    u.swap(v);
    return true;
}

void finish_request(golap_server *s, Request *r, int status, const std::string &err) {
    std::lock_guard<std::mutex> lk(s->m);
    r->status = status;
    r->err = err;
    r->t_done = Clock::now();
    r->done = true;
    s->cv_done.notify_all();
}

// Creates and enqueues the batch of reqs on slot si; 0 or a status for all of them.
int launch(golap_server *s, Slot &slot, const std::vector<Request *> &reqs, golap_batch **out, std::string &err) {
    golap_ctx *ctx = s->ctx;
    std::lock_guard<std::recursive_mutex> l(ctx->mu);
    if (int r = set_device(ctx)) {
        err = last_error(ctx);
        return r;
    }
    golap_batch *b = nullptr;
    int r;
    if (reqs[0]->kind == BatchKind::STAR) {
        std::vector<golap_star_query> q;
        for (Request *x : reqs) q.push_back(x->star);
        r = batch_create_star(ctx, q.data(), (uint32_t)q.size(), &slot.bs, false, &b);
    } else {
        std::vector<golap_scan_plan> q;
        for (Request *x : reqs) q.push_back(x->scan);
        r = batch_create_scan(ctx, q.data(), (uint32_t)q.size(), &slot.bs, false, &b);
    }
    if (r) {
        err = last_error(ctx);
        return r;
    }
    if ((r = batch_enqueue(b, false))) {
        err = b->err;
        batch_free(b);
        return r;
    }
    *out = b;
    return GOLAP_OK;
}

void release_slot(golap_server *s, Slot &slot, size_t n_done) {
    std::lock_guard<std::mutex> lk(s->m);
    s->admitted_bytes -= slot.bytes;
    slot.bytes = 0;
    slot.busy = false;
    slot.batch = nullptr;
    slot.reqs.clear();
    s->stats.completed += n_done;
    s->cv_sched.notify_all();
}

void completion_loop(golap_server *s, Slot *slot) {
    (void)hipSetDevice(s->ctx->device);
    for (;;) {
        golap_batch *b;
        std::vector<Request *> reqs;
        {
            std::unique_lock<std::mutex> lk(s->m);
            slot->cv.wait(lk, [&] { return s->stop || slot->batch; });
            if (!slot->batch) return;  // stopping with nothing in flight
            b = slot->batch;
            reqs = slot->reqs;
        }
        golap_batch_run_info info{};
        int r = batch_finish(b, &info);
        std::string err = r ? b->err : std::string();
        if (!r) {
            std::lock_guard<std::mutex> lk(s->m);
            s->stats.device_ms += info.ms;
            s->stats.comp_bytes += info.comp_bytes;
            s->stats.shared_comp_bytes += info.shared_comp_bytes;
        }
        if (!r && !reqs.empty()) {  // the whole arena back in one copy, before the parallel emission
            golap_star_result tmp{};
            r = batch_result(b, 0, &tmp, nullptr, nullptr, nullptr, 0);
            if (r == GOLAP_ECAPACITY) r = GOLAP_OK;  // max_groups 0: only the copy was wanted
            if (r) err = b->err;
        }
        if (r) {
            for (Request *x : reqs) finish_request(s, x, r, err);
        } else {
            std::mutex lm;
            std::condition_variable lcv;
            size_t left = reqs.size();
            for (uint32_t i = 0; i < reqs.size(); ++i)
                s->pool->run([&, i] {
                    Request *x = reqs[i];
                    x->keys.reset(new int32_t[std::max<uint64_t>(1, x->domain * std::max(1u, x->n_attrs))]);
                    x->sums.reset(new int64_t[x->domain]);
                    x->counts.reset(new uint64_t[x->domain]);
                    const int rr = batch_result(b, i, &x->res, x->keys.get(), x->sums.get(), x->counts.get(), x->domain);
                    finish_request(s, x, rr, rr ? std::string("emit failed") : std::string());
                    std::lock_guard<std::mutex> g(lm);
                    if (--left == 0) lcv.notify_one();
                });
            std::unique_lock<std::mutex> g(lm);
            lcv.wait(g, [&] { return left == 0; });
        }
        batch_free(b);
        release_slot(s, *slot, reqs.size());
    }
}

void scheduler_loop(golap_server *s) {
    (void)hipSetDevice(s->ctx->device);
    const uint64_t budget = s->opts.hbm_budget;
    std::unique_lock<std::mutex> lk(s->m);
    for (;;) {
        s->cv_sched.wait(lk, [&] {
            if (s->stop && s->pending.empty()) return true;
            return !s->pending.empty() && free_slot(s) >= 0 && s->admitted_bytes + s->pending.front()->bytes <= budget;
        });
        if (s->pending.empty()) {
            if (s->stop) break;
            continue;
        }
        if (s->opts.batch_window_us && s->pending.size() < s->opts.max_batch && !s->stop && any_busy(s))
            s->cv_sched.wait_for(lk, std::chrono::microseconds(s->opts.batch_window_us),
                                 [&] { return s->stop || s->pending.size() >= s->opts.max_batch; });
        const int si = free_slot(s);
        if (si < 0) continue;
        Slot &slot = *s->slots[si];
        // form a batch: requests of the front's kind in arrival order, admitted while the budget
        // holds (the first that does not fit stops admission: no overtaking by size)
        const BatchKind kind = s->pending.front()->kind;
        std::vector<Request *> take;
        std::vector<const golap_column *> cols;
        uint64_t bytes = 0;
        for (auto it = s->pending.begin(); it != s->pending.end() && take.size() < s->opts.max_batch;) {
            Request *r = *it;
            if (s->admitted_bytes + bytes + r->bytes > budget) break;
            if (r->kind != kind || !fits_columns(cols, r)) {
                ++it;
                continue;
            }
            take.push_back(r);
            bytes += r->bytes;
            it = s->pending.erase(it);
        }
        if (take.empty()) continue;
        s->admitted_bytes += bytes;
        s->stats.peak_admitted_bytes = std::max(s->stats.peak_admitted_bytes, s->admitted_bytes);
        s->stats.batches++;
        s->batched_requests += take.size();
        slot.busy = true;
        slot.bytes = bytes;
        lk.unlock();
        golap_batch *b = nullptr;
        std::string err;
        int r = launch(s, slot, take, &b, err);
        if (r && take.size() > 1) {
            // a batch that cannot be formed (e.g. its LDS windows do not fit): run each alone
            for (Request *x : take) {
                golap_batch *one = nullptr;
                std::string e1;
                int r1 = launch(s, slot, std::vector<Request *>{x}, &one, e1);
                golap_batch_run_info info{};
                if (!r1) r1 = batch_finish(one, &info);
                if (!r1) {
                    x->keys.reset(new int32_t[std::max<uint64_t>(1, x->domain * std::max(1u, x->n_attrs))]);
                    x->sums.reset(new int64_t[x->domain]);
                    x->counts.reset(new uint64_t[x->domain]);
                    r1 = batch_result(one, 0, &x->res, x->keys.get(), x->sums.get(), x->counts.get(), x->domain);
                }
                if (r1 && one) e1 = one->err;
                if (one) batch_free(one);
                finish_request(s, x, r1, e1);
            }
            release_slot(s, slot, take.size());
        } else if (r) {
            finish_request(s, take[0], r, err);
            release_slot(s, slot, 1);
        } else {
            std::lock_guard<std::mutex> g(s->m);
            slot.batch = b;
            slot.reqs = take;
            slot.cv.notify_one();
        }
        lk.lock();
    }
}

// Validates a request with the single-query preparation and sizes it.
int size_request(golap_ctx *ctx, Request *r) {
    std::lock_guard<std::recursive_mutex> l(ctx->mu);
    if (int e = set_device(ctx)) return e;
    uint64_t chunks, ppc;
    if (r->kind == BatchKind::STAR) {
        ColumnSet cs;
        StarQueryDev qd;
        QueryHost qh;
        if (r->star.plan.value_minus) return fail(ctx, GOLAP_EINVAL, "server: value_minus is a golap_star_run option");
        if (int e = prepare_query(ctx, &r->star.plan, true, cs, qd, qh)) return e;
        const golap_column *zc = r->star.zone_col;
        if (zc && (zc->ctx != ctx || zc->n_rows != r->star.plan.value->n_rows ||
                   zc->chunk_rows != r->star.plan.value->chunk_rows))
            return fail(ctx, GOLAP_EINVAL, "server: zone column not row-aligned with the fact table");
        r->domain = qh.domain;
        r->n_attrs = (uint32_t)qh.card.size();
        r->cols = cs.cols;
        chunks = r->star.plan.chunk_end - r->star.plan.chunk_begin;
        ppc = r->star.plan.value->chunk_rows / r->star.plan.value->page_rows;
    } else {
        ScanQueryDev qd;
        ScanHost h;
        if (int e = prepare_scan(ctx, &r->scan, qd, h)) return e;
        r->domain = h.domain;
        r->n_attrs = (uint32_t)h.card.size();
        r->cols.assign(r->scan.cols, r->scan.cols + r->scan.n_cols);
        std::sort(r->cols.begin(), r->cols.end());
        r->cols.erase(std::unique(r->cols.begin(), r->cols.end()), r->cols.end());
        chunks = r->scan.chunk_end - r->scan.chunk_begin;
        ppc = r->scan.cols[0]->chunk_rows / r->scan.cols[0]->page_rows;
    }
    r->bytes = batch_query_bytes(r->domain, chunks, (uint32_t)ppc, r->kind == BatchKind::STAR);
    return GOLAP_OK;
}

int submit(golap_server *s, Request *r, uint64_t *ticket) {
    r->t_submit = Clock::now();
    if (int e = size_request(s->ctx, r)) {
        delete r;
        std::lock_guard<std::mutex> lk(s->m);
        s->stats.refused++;
        return e;
    }
    std::lock_guard<std::mutex> lk(s->m);
    if (r->bytes > s->opts.hbm_budget) {
        delete r;
        s->stats.refused++;
        return fail(s->ctx, GOLAP_EBUDGET, "server: request exceeds the HBM budget");
    }
    if (s->stop) {
        delete r;
        return fail(s->ctx, GOLAP_EINVAL, "server: stopping");
    }
    r->ticket = s->next_ticket++;
    *ticket = r->ticket;
    s->all[r->ticket] = r;
    s->pending.push_back(r);
    s->stats.submitted++;
    s->cv_sched.notify_all();
    return GOLAP_OK;
}

}  // namespace

extern "C" {

int golap_server_create(golap_ctx *ctx, const golap_server_opts *opts, golap_server **out) {
    if (!ctx || !out) return GOLAP_EINVAL;
    *out = nullptr;
    auto s = std::make_unique<golap_server>();
    s->ctx = ctx;
    if (opts) s->opts = *opts;
    if (!s->opts.hbm_budget) s->opts.hbm_budget = 4ull << 30;
    if (!s->opts.max_batch) s->opts.max_batch = 64;
    if (!s->opts.slots) s->opts.slots = 2;
    if (!s->opts.emit_threads) s->opts.emit_threads = 8;
    if (s->opts.slots > 16 || s->opts.emit_threads > 256 || s->opts.max_batch > (1u << 16))
        return fail(ctx, GOLAP_EINVAL, "server: too many slots, threads or batch entries");
    {
        std::lock_guard<std::recursive_mutex> l(ctx->mu);
        if (int r = set_device(ctx)) return r;
        for (uint32_t i = 0; i < s->opts.slots; ++i) {
            s->slots.emplace_back(new Slot());
            if (int r = s->slots.back()->bs.init(ctx, true)) {
                for (auto &sl : s->slots) sl->bs.release();
                return fail(ctx, r, s->slots.back()->bs.err);
            }
        }
    }
    s->pool.reset(new Pool(s->opts.emit_threads, ctx->device));
    golap_server *sp = s.get();
    for (auto &sl : s->slots) sl->th = std::thread(completion_loop, sp, sl.get());
    s->sched = std::thread(scheduler_loop, sp);
    *out = s.release();
    return GOLAP_OK;
}

int golap_server_submit_star(golap_server *s, const golap_star_query *q, uint64_t *ticket) {
    if (!s || !q || !ticket) return GOLAP_EINVAL;
    auto *r = new Request();
    r->kind = BatchKind::STAR;
    r->star = *q;
    return submit(s, r, ticket);
}

int golap_server_submit_scan(golap_server *s, const golap_scan_plan *p, uint64_t *ticket) {
    if (!s || !p || !ticket) return GOLAP_EINVAL;
    auto *r = new Request();
    r->kind = BatchKind::SCAN;
    r->scan = *p;
    return submit(s, r, ticket);
}

int golap_server_wait(golap_server *s, uint64_t ticket, golap_star_result *res, int32_t *group_keys, int64_t *sums,
                      uint64_t *counts, uint64_t max_groups, double *latency_ms) {
    if (!s || !res) return GOLAP_EINVAL;
    std::unique_ptr<Request> r;
    {
        std::unique_lock<std::mutex> lk(s->m);
        auto it = s->all.find(ticket);
        if (it == s->all.end() || it->second->claimed)
            return fail(s->ctx, GOLAP_EINVAL, "server: unknown ticket or one already waited on");
        Request *x = it->second;
        x->claimed = true;
        s->cv_done.wait(lk, [&] { return x->done; });
        s->all.erase(ticket);  // looked up again: other tickets may have been erased meanwhile
        r.reset(x);
    }
    if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
    if (r->status) return fail(s->ctx, r->status, r->err);
    *res = r->res;
    const uint64_t n = std::min<uint64_t>(r->res.n_groups, max_groups);
    if (group_keys && r->n_attrs) memcpy(group_keys, r->keys.get(), n * r->n_attrs * sizeof(int32_t));
    if (sums) memcpy(sums, r->sums.get(), n * sizeof(int64_t));
    if (counts) memcpy(counts, r->counts.get(), n * sizeof(uint64_t));
    if (r->res.n_groups > max_groups) return fail(s->ctx, GOLAP_ECAPACITY, "more groups than max_groups");
    return GOLAP_OK;
}

int golap_server_get_stats(golap_server *s, golap_server_stats *out) {
    if (!s || !out) return GOLAP_EINVAL;
    std::lock_guard<std::mutex> lk(s->m);
    *out = s->stats;
    out->mean_batch = s->stats.batches ? (double)s->batched_requests / (double)s->stats.batches : 0.0;
    return GOLAP_OK;
}

int golap_server_destroy(golap_server *s) {
    if (!s) return GOLAP_EINVAL;
    {
        std::lock_guard<std::mutex> lk(s->m);
        s->stop = true;
        s->cv_sched.notify_all();
    }
    s->sched.join();  // leaves once nothing is pending
    {
        std::unique_lock<std::mutex> lk(s->m);
        s->cv_sched.wait(lk, [&] { return !any_busy(s); });
        for (auto &sl : s->slots) sl->cv.notify_all();
    }
    for (auto &sl : s->slots) sl->th.join();
    s->pool.reset();
    {
        std::lock_guard<std::recursive_mutex> l(s->ctx->mu);
        (void)set_device(s->ctx);
        for (auto &sl : s->slots) sl->bs.release();
    }
    for (auto &kv : s->all) delete kv.second;
    delete s;
    return GOLAP_OK;
}

}  // extern "C"
