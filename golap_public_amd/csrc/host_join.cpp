// CPU/GPU co-execution of star plans: one dimension joined on the host.
//
// The reference's query 3 variants keep the customer hash table on the CPU (usecases/ssb/
// queries/03a.hpp, 03c.hpp, 03d.hpp; 03c.hpp:189-488): per chunk the GPU probes the supplier
// and date tables and appends the joined lineorder tuples to a buffer, the buffer is copied to
// the host, and a CPU thread probes the customer HostHashMap (core/host_structs.hpp:9-78) and
// adds into a HostHashAggregate (:82-152).
//
// Here the host-side dimension is a golap_host_hashmap (open addressing over 8-byte slots,
// built by several threads with compare-and-swap), and a run goes in rounds of chunks:
//   GPU   the star kernel in emit mode over round k+1: every dimension but the host one probed
//         exactly, passing rows appended as {host key, group index, value};
//   PCIe  round k's rows copied to pinned memory;
//   CPU   a pool of threads probes round k's rows into the host table and adds into
//         thread-local dense group arrays (shared atomic arrays for large group domains),
// so the host work of a round overlaps the device work of the next. The groups come out dense
// and canonical, exactly as golap_star_run returns them.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "common.hpp"
#include "host_column.hpp"
#include "star_host.hpp"

struct golap_host_hashmap {
    uint64_t ht_size = 0;
    const int32_t *keys = nullptr;  // the build table's key column (caller-owned, as HostHashMap's table)
    uint32_t log2cap = 0;
    std::unique_ptr<std::atomic<uint64_t>[]> slots;  // (uint64(payload) << 32) | (uint32(key) ^ 0x80000000), 0 = empty
    uint64_t filled = 0;
    bool built = false;
};

using namespace golap_impl;
using namespace golap_impl::starhost;

namespace {

using Lock = std::lock_guard<std::recursive_mutex>;

constexpr uint32_t EMPTY_BIAS = 0x80000000u;

inline uint32_t host_hash(int32_t k) { return (uint32_t)k * 0x9E3779B1u; }

inline uint64_t pack_slot(int32_t key, uint32_t payload) {
    return ((uint64_t)payload << 32) | (uint64_t)((uint32_t)key ^ EMPTY_BIAS);
}

// Slot index of key, or -1 when absent.
inline int64_t host_find(const golap_host_hashmap *m, int32_t key) {
    const uint64_t mask = (1ull << m->log2cap) - 1;
    const uint32_t want = (uint32_t)key ^ EMPTY_BIAS;
    for (uint64_t h = host_hash(key) >> (32 - m->log2cap), n = 0; n <= mask; h = (h + 1) & mask, ++n) {
        const uint64_t s = m->slots[h].load(std::memory_order_relaxed);
        if (!s) return -1;
        if ((uint32_t)s == want) return (int64_t)h;
    }
    return -1;
}

unsigned host_threads(uint32_t requested) {
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    return requested ? std::min<unsigned>(requested, 256) : std::min(hw, 16u);
}

bool host_pred_pass(const golap_host_pred &p, uint64_t r) {
    const int64_t v = golap_host::load_value(static_cast<const uint8_t *>(p.col) + r * golap_host::type_size(p.type), p.type);
    return v >= p.lo && v <= p.hi;
}

}  // namespace

extern "C" {

int golap_host_hashmap_create(uint64_t ht_size, const int32_t *keys, golap_host_hashmap **out) {
    if (!out || !keys || !ht_size || ht_size > (1ull << 31)) return GOLAP_EINVAL;
    std::unique_ptr<golap_host_hashmap> m(new golap_host_hashmap());
    m->ht_size = ht_size;
    m->keys = keys;
    m->log2cap = 1;
    while ((1ull << m->log2cap) < 2 * ht_size) ++m->log2cap;
    m->slots.reset(new std::atomic<uint64_t>[1ull << m->log2cap]);
    *out = m.release();
    return GOLAP_OK;
}

int golap_host_hashmap_build(golap_host_hashmap *m, uint64_t n_rows, const golap_host_pred *preds, uint32_t n_preds,
                             const int32_t *payload, uint32_t threads) {
    if (!m || (n_preds && !preds) || n_rows > (1ull << 32)) return GOLAP_EINVAL;
    for (uint32_t i = 0; i < n_preds; ++i)
        if (!preds[i].col || !valid_type(preds[i].type)) return GOLAP_EINVAL;
    const uint64_t cap = 1ull << m->log2cap, mask = cap - 1;
    for (uint64_t i = 0; i < cap; ++i) m->slots[i].store(0, std::memory_order_relaxed);
    std::atomic<uint64_t> filled{0};
    std::atomic<bool> full{false};
    const unsigned nt = (unsigned)std::min<uint64_t>(host_threads(threads), std::max<uint64_t>(1, n_rows / 65536));
    auto work = [&](unsigned t) {
        uint64_t mine = 0;
        for (uint64_t r = n_rows * t / nt; r < n_rows * (t + 1) / nt; ++r) {
            bool pass = true;
            for (uint32_t i = 0; i < n_preds && pass; ++i) pass = host_pred_pass(preds[i], r);
            if (!pass) continue;
            const int32_t key = m->keys[r];
            const uint64_t slot = pack_slot(key, payload ? (uint32_t)payload[r] : (uint32_t)r);
            // insert if absent (one of the rows is kept for duplicate keys)
            for (uint64_t h = host_hash(key) >> (32 - m->log2cap), n = 0;; h = (h + 1) & mask, ++n) {
                if (n > mask) {
                    full = true;
                    break;
                }
                uint64_t cur = m->slots[h].load(std::memory_order_relaxed);
                if (!cur && m->slots[h].compare_exchange_strong(cur, slot, std::memory_order_relaxed)) {
                    ++mine;
                    break;
                }
                if ((uint32_t)cur == (uint32_t)slot) break;
            }
        }
        filled += mine;
    };
    std::vector<std::thread> pool;
    for (unsigned t = 1; t < nt; ++t) pool.emplace_back(work, t);
    work(0);
    for (auto &th : pool) th.join();
    m->filled = filled;
    m->built = !full && m->filled <= m->ht_size;
    return m->built ? GOLAP_OK : GOLAP_ECAPACITY;
}

int golap_host_hashmap_probe(const golap_host_hashmap *m, const int32_t *keys, uint64_t n, uint64_t *out,
                             uint64_t *hits) {
    if (!m || !m->built || (n && (!keys || !out))) return GOLAP_EINVAL;
    uint64_t h = 0;
    for (uint64_t i = 0; i < n; ++i) {
        const int64_t s = host_find(m, keys[i]);
        out[i] = s < 0 ? UINT64_MAX : (m->slots[s].load(std::memory_order_relaxed) >> 32);
        h += s >= 0;
    }
    if (hits) *hits = h;
    return GOLAP_OK;
}

int golap_host_hashmap_filled(const golap_host_hashmap *m, uint64_t *filled) {
    if (!m || !filled) return GOLAP_EINVAL;
    *filled = m->filled;
    return GOLAP_OK;
}

int golap_host_hashmap_free(golap_host_hashmap *m) {
    if (!m) return GOLAP_EINVAL;
    delete m;
    return GOLAP_OK;
}

int golap_star_run_hostdim(golap_ctx *ctx, const golap_star_plan *plan, uint32_t host_dim,
                           const golap_host_hashmap *hmap, uint32_t threads, golap_star_result *res,
                           golap_coexec_stats *stats, int32_t *group_keys, int64_t *sums, uint64_t *counts,
                           uint64_t max_groups) {
    if (!ctx || !plan || !res || !hmap) return GOLAP_EINVAL;
    if (plan->n_dims < 2 || host_dim >= plan->n_dims) return fail(ctx, GOLAP_EINVAL, "star_hostdim: need a device dimension besides the host one");
    if (!hmap->built) return fail(ctx, GOLAP_EINVAL, "star_hostdim: host map not built");
    const auto t_start = std::chrono::steady_clock::now();
    Lock l(ctx->mu);
    if (int r = set_device(ctx)) return r;
    memset(res, 0, sizeof(*res));
    golap_coexec_stats st{};
    ColumnSet cs;
    StarQueryDev qd;
    QueryHost qh;
    if (plan && plan->value_minus) return fail(ctx, GOLAP_EINVAL, "co-execution: value_minus is a golap_star_run option");
    if (int r = prepare_query(ctx, plan, true, cs, qd, qh, (int)host_dim)) return r;
    const golap_column *v = plan->value;
    StarLaunch L{};
    L.n_cols = (uint32_t)cs.cols.size();
    for (uint32_t c = 0; c < L.n_cols; ++c) L.col[c] = view(cs.cols[c]);
    L.n_windows = assign_windows(cs, qd.s.page_begin, qd.s.page_end, L.col_window);
    L.page_rows = v->page_rows;
    L.pages_per_chunk = (uint32_t)(v->chunk_rows / v->page_rows);
    L.n_rows = v->n_rows;
    L.n_pages = v->n_pages;
    L.n_queries = 1;
    L.filter_words = qd.filter_words;
    L.lds_gslots = 0;
    if (star_lds_bytes(L.filter_words, L.page_rows, L.n_windows, 0) > (size_t)ctx->lds_max)
        return fail(ctx, GOLAP_EINVAL, "star_hostdim: LDS budget exceeded (page_rows or coded columns)");
    if (star_blocks_per_cu(L.page_rows, star_lds_bytes(0, L.page_rows, L.n_windows, 0)) >
        star_blocks_per_cu(L.page_rows, star_lds_bytes(L.filter_words, L.page_rows, L.n_windows, 0)))
        L.filter_words = 0;
    qd.use_lds = 0;
    L.group_bits = std::max<uint32_t>(1, bits_for(qh.domain));
    L.err = reinterpret_cast<uint32_t *>(ctx->d_scratch);
    L.emit_dim = host_dim;

    // the host dimension's share of the group index
    const golap_star_dim &hd = plan->dims[host_dim];
    uint64_t h_stride = 0;
    {
        size_t a = 0;
        for (uint32_t d = 0; d < plan->n_dims; ++d)
            if (plan->dims[d].group_card) {
                if (d == host_dim) h_stride = qh.stride[a];
                ++a;
            }
    }
    const uint64_t dom = qh.domain;

    // rounds of whole chunks, about ROUND_ROWS fact rows each, handing over at most `cap` rows:
    // a round that passes more is halved and run again (one chunk always fits: cap >= its rows)
    constexpr uint64_t ROUND_ROWS = 64ull << 20, EMIT_CAP = 4ull << 20;
    const uint64_t cb = plan->chunk_begin, ce = plan->chunk_end;
    const uint64_t rpc = v->chunk_rows, cpr = std::max<uint64_t>(1, ROUND_ROWS / rpc);
    const uint64_t rounds = (ce - cb + cpr - 1) / cpr;
    const uint64_t cap = std::min<uint64_t>(std::max<uint64_t>(EMIT_CAP, rpc), std::max<uint64_t>(1, (ce - cb) * rpc));
    const golap_hashmap *dm = plan->dims[qd.driver].map;
    void *dbuf[2] = {nullptr, nullptr};
    int4 *hbuf[2] = {nullptr, nullptr};
    void *dcount = nullptr;
    uint32_t *hcount = nullptr;
    hipEvent_t copied = nullptr;
    auto release = [&] {
        (void)hipStreamSynchronize(ctx->stream);
        for (int i = 0; i < 2; ++i) {
            if (dbuf[i]) dev_release(ctx, dbuf[i]);
            if (hbuf[i]) (void)hipHostFree(hbuf[i]);
        }
        if (dcount) dev_release(ctx, dcount);
        if (hcount) (void)hipHostFree(hcount);
        if (copied) (void)hipEventDestroy(copied);
    };
    auto fail_release = [&](int code, const std::string &msg) {
        release();
        return fail(ctx, code, "star_hostdim: " + msg);
    };
    const bool any = rounds && dm->filled && hmap->filled;
    if (any) {
        for (int i = 0; i < 2; ++i) {
            if (int r = dev_alloc(ctx, cap * sizeof(int4), &dbuf[i])) {
                release();
                return r;
            }
            if (hipHostMalloc(reinterpret_cast<void **>(&hbuf[i]), cap * sizeof(int4), hipHostMallocDefault) != hipSuccess) {
                hbuf[i] = nullptr;
                return fail_release(GOLAP_ENOMEM, "pinned buffers");
            }
        }
        if (int r = dev_alloc(ctx, 2 * sizeof(uint32_t), &dcount)) {
            release();
            return r;
        }
        if (hipHostMalloc(reinterpret_cast<void **>(&hcount), 2 * sizeof(uint32_t), hipHostMallocDefault) != hipSuccess) {
            hcount = nullptr;
            return fail_release(GOLAP_ENOMEM, "pinned counters");
        }
        if (hipEventCreateWithFlags(&copied, hipEventDisableTiming) != hipSuccess) return fail_release(GOLAP_EHIP, "event");
    }

    // host aggregation: thread-local dense arrays for small domains, shared atomics otherwise
    const unsigned nt = host_threads(threads);
    const bool local = dom <= (1ull << 20);
    std::vector<std::vector<int64_t>> lsum(local ? nt : 0);
    std::vector<std::vector<uint64_t>> lcnt(local ? nt : 0);
    std::unique_ptr<std::atomic<int64_t>[]> gsum;
    std::unique_ptr<std::atomic<uint64_t>[]> gcnt;
    if (local) {
        for (unsigned t = 0; t < nt; ++t) {
            lsum[t].assign(dom, 0);
            lcnt[t].assign(dom, 0);
        }
    } else {
        gsum.reset(new std::atomic<int64_t>[dom]);
        gcnt.reset(new std::atomic<uint64_t>[dom]);
        for (uint64_t i = 0; i < dom; ++i) {
            gsum[i].store(0, std::memory_order_relaxed);
            gcnt[i].store(0, std::memory_order_relaxed);
        }
    }
    std::atomic<bool> keyrange{false};
    double host_ms = 0;
    auto aggregate = [&](const int4 *rows, uint64_t n) {
        const auto h0 = std::chrono::steady_clock::now();
        // part t of `parts` equal shares of the round, into thread t's arrays
        auto work = [&](unsigned t, unsigned parts) {
            int64_t *ls = local ? lsum[t].data() : nullptr;
            uint64_t *lc = local ? lcnt[t].data() : nullptr;
            for (uint64_t i = n * t / parts; i < n * (t + 1) / parts; ++i) {
                const int4 e = rows[i];
                const int64_t s = host_find(hmap, e.x);
                if (s < 0) continue;
                uint64_t g = (uint32_t)e.y;
                if (hd.group_card) {
                    const uint32_t pay = (uint32_t)(hmap->slots[s].load(std::memory_order_relaxed) >> 32);
                    const uint32_t gv = pay - (uint32_t)hd.group_lo;
                    if (gv >= hd.group_card) {
                        keyrange = true;
                        continue;
                    }
                    g += (uint64_t)gv * h_stride;
                }
                const int64_t val = (int64_t)(((uint64_t)(uint32_t)e.w << 32) | (uint32_t)e.z);
                if (local) {
                    ls[g] += val;
                    lc[g] += 1;
                } else {
                    gsum[g].fetch_add(val, std::memory_order_relaxed);
                    gcnt[g].fetch_add(1, std::memory_order_relaxed);
                }
            }
        };
        if (n < 65536 || nt == 1) {
            work(0, 1);
        } else {
            std::vector<std::thread> pool;
            for (unsigned t = 1; t < nt; ++t) pool.emplace_back(work, t, nt);
            work(0, nt);
            for (auto &th : pool) th.join();
        }
        host_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - h0).count();
    };

    // the round over chunks [c0, c1) into buffer `slot`
    auto launch_range = [&](uint64_t c0, uint64_t c1, int slot) -> hipError_t {
        const uint64_t ppc = L.pages_per_chunk;
        StarQueryDev q = qd;
        q.s.chunk_begin = c0;
        q.s.chunk_end = c1;
        q.s.page_begin = c0 * ppc;
        q.s.page_end = std::min<uint64_t>(v->n_pages, c1 * ppc);
        L.q0 = q;
        L.n_items = q.s.page_end - q.s.page_begin;
        L.emit = static_cast<int4 *>(dbuf[slot]);
        L.emit_n = static_cast<uint32_t *>(dcount) + slot;
        L.emit_cap = cap;
        hipError_t e = hipMemsetAsync(L.emit_n, 0, 4, ctx->stream);
        if (e == hipSuccess) e = launch_star(ctx, L, ctx->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(hcount + slot, L.emit_n, 4, hipMemcpyDeviceToHost, ctx->stream);
        return e;
    };

    hipError_t e = hipMemsetAsync(ctx->d_scratch, 0, 4, ctx->stream);
    if (e == hipSuccess) e = hipEventRecord(ctx->ev0, ctx->stream);
    if (any) {
        std::deque<std::pair<uint64_t, uint64_t>> todo;
        for (uint64_t k = 0; k < rounds; ++k) todo.emplace_back(cb + k * cpr, std::min(ce, cb + (k + 1) * cpr));
        int slot = 0;
        std::pair<uint64_t, uint64_t> cur = todo.front();
        todo.pop_front();
        if (e == hipSuccess) e = launch_range(cur.first, cur.second, slot);
        while (e == hipSuccess) {
            e = hipStreamSynchronize(ctx->stream);  // the current round's kernel and count
            if (e != hipSuccess) break;
            const uint64_t n = hcount[slot];
            if (n > cap) {  // more rows than the buffer holds: run the halves instead
                if (cur.second - cur.first < 2) {  // one chunk always fits (cap >= its rows)
                    e = hipErrorInvalidValue;
                    break;
                }
                const uint64_t mid = cur.first + (cur.second - cur.first) / 2;
                todo.emplace_front(mid, cur.second);
                cur.second = mid;
                st.splits++;
                e = launch_range(cur.first, cur.second, slot);
                continue;
            }
            st.emitted += n;
            if (n) e = hipMemcpyAsync(hbuf[slot], dbuf[slot], n * sizeof(int4), hipMemcpyDeviceToHost, ctx->stream);
            if (e == hipSuccess) e = hipEventRecord(copied, ctx->stream);
            // the next round on the device while the host aggregates this one
            const bool more = !todo.empty();
            std::pair<uint64_t, uint64_t> next{0, 0};
            if (more) {
                next = todo.front();
                todo.pop_front();
                if (e == hipSuccess) e = launch_range(next.first, next.second, slot ^ 1);
            }
            if (e == hipSuccess) e = hipEventSynchronize(copied);
            if (e == hipSuccess) aggregate(hbuf[slot], n);
            st.rounds++;
            if (!more) break;
            cur = next;
            slot ^= 1;
        }
    }
    if (e == hipSuccess) e = hipEventRecord(ctx->ev1, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    float span = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&span, ctx->ev0, ctx->ev1);
    uint32_t flags = 0;
    if (e == hipSuccess) e = hipMemcpy(&flags, ctx->d_scratch, 4, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return fail_release(GOLAP_EHIP, hipGetErrorString(e));
    release();
    if (flags) return error_flags_status(ctx, flags, "star_hostdim");
    if (keyrange) return fail(ctx, GOLAP_EINVAL, "star_hostdim: host payload outside its declared group range");

    std::vector<int64_t> hs(dom, 0);
    std::vector<uint64_t> hc(dom, 0);
    if (local) {
        for (unsigned t = 0; t < nt; ++t)
            for (uint64_t g = 0; g < dom; ++g) {
                hs[g] += lsum[t][g];
                hc[g] += lcnt[t][g];
            }
    } else {
        for (uint64_t g = 0; g < dom; ++g) {
            hs[g] = gsum[g].load(std::memory_order_relaxed);
            hc[g] = gcnt[g].load(std::memory_order_relaxed);
        }
    }
    std::vector<uint8_t> run;
    if (plan->dev_chunk_mask && ce > cb) {
        std::vector<uint16_t> mask(ce - cb);
        GOLAP_HIP_CHECK(ctx, hipMemcpy(mask.data(), plan->dev_chunk_mask + cb, mask.size() * 2, hipMemcpyDeviceToHost));
        run.resize(mask.size());
        for (size_t i = 0; i < mask.size(); ++i) run[i] = mask[i] != 0;
    }
    res->driver_dim = qd.driver;
    res->kernel_ms = span;
    slice_stats(*plan, run, res);
    st.host_ms = host_ms;
    st.device_ms = span;
    st.total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    if (stats) *stats = st;
    return emit_groups_host(ctx, dom, hs.data(), hc.data(), qh.lo, qh.card, qh.stride, &res->n_groups, &res->total_sum,
                            group_keys, sums, counts, max_groups, &res->rows_matched);
}

}  // extern "C"
