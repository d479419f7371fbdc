// C ABI: batched star plans (include/golap_hip.h). Many star-join + group-by queries over one
// fact table run as one persistent fused launch:
//   1. a fill kernel zeroes the result arena (work-list counters, error word, every query's
//      dense SUM/COUNT arrays);
//   2. k_batch_items evaluates each query's chunk mask and zone-map predicate per chunk of its
//      slice (check_mmmeta semantics, core/metadata.cuh:539-547) and writes the surviving
//      (chunk, query, part) items;
//   3. k_star (queue schedule) lets workgroups take items with one atomic each.
// The three kernels are captured once into a hipGraph and replayed per run. The launch
// configuration is chosen by a coordinate-descent autotuner that times candidates on the batch
// itself and caches the winner in the context by batch shape.
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "common.hpp"
#include "star_host.hpp"

using namespace golap_impl;
using namespace golap_impl::starhost;
using Lock = std::lock_guard<std::recursive_mutex>;

struct golap_star_batch {
    golap_ctx *ctx = nullptr;
    std::vector<golap_batch_query> q;
    std::vector<QueryHost> qh;
    std::vector<StarQueryDev> qd;  // descriptors as last uploaded
    ColumnSet cs;
    int windows[STAR_MAX_COLS] = {};
    uint32_t n_windows = 0;
    uint32_t page_rows = 0, ppc = 0;
    uint64_t n_rows = 0, n_pages = 0;
    std::vector<uint64_t> chunk_off;  // prefix of slice lengths in chunks, n + 1 entries
    // host view of which chunks survive (chunk masks read at creation, zone maps), per query
    std::vector<std::vector<uint8_t>> survive;
    std::vector<uint64_t> surv_pages;
    uint64_t total_pages = 0, total_bytes = 0;
    uint32_t max_filter_words = 0;
    uint64_t max_domain = 1;
    // device
    StarQueryDev *d_q = nullptr;
    uint64_t *d_chunk_off = nullptr;
    uint2 *d_items = nullptr;
    uint8_t *d_arena = nullptr;  // [0, 256): queue head, item count, error word; then results
    uint64_t arena_bytes = 0;
    uint8_t *h_arena = nullptr;  // pinned host mirror, filled once per run on the first result call
    bool h_valid = false;
    std::vector<uint64_t> res_off;  // per query: byte offset of its sums (counts follow)
    golap_batch_config cfg{};
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::string shape;
    bool ran = false;
    double last_ms = 0;
};

namespace {

constexpr uint64_t ARENA_HEADER = 256;

uint32_t parts_of(const golap_star_batch *b, uint32_t ppi) { return (b->ppc + ppi - 1) / ppi; }

void drop_graph(golap_star_batch *b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

// Batches of the same shape share tuned configurations.
std::string shape_of(const golap_star_batch *b) {
    auto bucket = [](uint64_t v) { return bits_for(std::max<uint64_t>(v, 1)); };
    return std::to_string(b->page_rows) + "/" + std::to_string(b->ppc) + "/" + std::to_string(b->n_windows) + "/" +
           std::to_string(b->cs.cols.size()) + "/" + std::to_string(bucket(b->q.size())) + "/" +
           std::to_string(bucket(b->total_pages)) + "/" + std::to_string(bucket(b->max_domain));
}

uint32_t group_bits_of(const golap_star_batch *b) { return std::max<uint32_t>(1, bits_for(b->max_domain)); }

// LDS group keys are (query << group_bits | group) and must stay below the empty marker.
bool lds_keys_fit(const golap_star_batch *b) { return bits_for(b->q.size()) + group_bits_of(b) <= 31; }

int check_config(const golap_star_batch *b, const golap_batch_config &c) {
    golap_ctx *ctx = b->ctx;
    if (c.pages_per_item < 1 || c.pages_per_item > b->ppc || parts_of(b, c.pages_per_item) > 256)
        return fail(ctx, GOLAP_EINVAL, "batch: pages_per_item out of range");
    if (c.lds_group_slots && (c.lds_group_slots & (c.lds_group_slots - 1) || c.lds_group_slots > LDS_GROUP_SLOTS_MAX))
        return fail(ctx, GOLAP_EINVAL, "batch: lds_group_slots must be 0 or a power of two <= 4096");
    if (c.lds_group_slots && !lds_keys_fit(b))
        return fail(ctx, GOLAP_EINVAL, "batch: too many queries x groups for LDS group keys");
    if (c.filter_lds > 1 || c.compact_tables > 1) return fail(ctx, GOLAP_EINVAL, "batch: bad flag");
    const uint32_t fw = c.filter_lds ? b->max_filter_words : 0;
    if (star_lds_bytes(fw, b->page_rows, b->n_windows, c.lds_group_slots) > (size_t)ctx->lds_max)
        return fail(ctx, GOLAP_EINVAL, "batch: LDS budget exceeded");
    return GOLAP_OK;
}

// Descriptors for a configuration: probe tables (compact or full) and the per-query choice of
// LDS or global aggregation; uploaded to the device.
int upload_descriptors(golap_star_batch *b, const golap_batch_config &c) {
    golap_ctx *ctx = b->ctx;
    const uint32_t nq = (uint32_t)b->q.size();
    for (uint32_t i = 0; i < nq; ++i) {
        StarQueryDev &d = b->qd[i];
        const golap_star_plan &p = b->q[i].plan;
        for (uint32_t k = 0; k < p.n_dims; ++k) {
            const golap_hashmap *m = p.dims[k].map;
            const bool small = c.compact_tables && m->small_log2;
            d.dim[k].slots = small ? m->d_small : m->d_slots;
            d.dim[k].log2cap = small ? m->small_log2 : m->log2cap;
        }
        const double expect = (double)b->surv_pages[i] * b->page_rows * b->qh[i].sel;
        bool lds = p.agg_mode == GOLAP_AGG_LDS ||
                   (p.agg_mode == GOLAP_AGG_AUTO && expect >= (double)c.lds_cutover * (double)b->qh[i].domain);
        d.use_lds = lds && c.lds_group_slots;
        d.g_sum = reinterpret_cast<int64_t *>(b->d_arena + b->res_off[i]);
        d.g_cnt = reinterpret_cast<uint64_t *>(b->d_arena + b->res_off[i] + b->qh[i].domain * 8);
    }
    GOLAP_HIP_CHECK(ctx, hipMemcpyAsync(b->d_q, b->qd.data(), nq * sizeof(StarQueryDev), hipMemcpyHostToDevice,
                                        ctx->stream));
    GOLAP_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return GOLAP_OK;
}

StarLaunch launch_of(const golap_star_batch *b, const golap_batch_config &c) {
    StarLaunch L{};
    L.n_cols = (uint32_t)b->cs.cols.size();
    for (uint32_t i = 0; i < L.n_cols; ++i) {
        L.col[i] = view(b->cs.cols[i]);
        L.col_window[i] = b->windows[i];
    }
    L.n_windows = b->n_windows;
    L.page_rows = b->page_rows;
    L.pages_per_chunk = b->ppc;
    L.n_rows = b->n_rows;
    L.n_pages = b->n_pages;
    L.queries = b->d_q;
    L.n_queries = (uint32_t)b->q.size();
    L.items = b->d_items;
    L.queue = reinterpret_cast<uint32_t *>(b->d_arena);
    L.pages_per_item = c.pages_per_item;
    L.filter_words = c.filter_lds ? b->max_filter_words : 0;
    L.lds_gslots = c.lds_group_slots;
    L.group_bits = group_bits_of(b);
    L.max_blocks_per_cu = c.max_blocks_per_cu;
    L.err = reinterpret_cast<uint32_t *>(b->d_arena + 8);
    return L;
}

// The run's three kernels on one stream.
int enqueue(golap_star_batch *b, hipStream_t s) {
    golap_ctx *ctx = b->ctx;
    GOLAP_HIP_CHECK(ctx, launch_fill_zero(b->d_arena, b->arena_bytes, s));
    GOLAP_HIP_CHECK(ctx, launch_batch_items(b->d_q, (uint32_t)sizeof(StarQueryDev), (uint32_t)b->q.size(),
                                            b->d_chunk_off, b->chunk_off.back(),
                                            parts_of(b, b->cfg.pages_per_item), b->d_items,
                                            reinterpret_cast<uint32_t *>(b->d_arena), s));
    GOLAP_HIP_CHECK(ctx, launch_star(ctx, launch_of(b, b->cfg), s));
    return GOLAP_OK;
}

int capture(golap_star_batch *b) {
    golap_ctx *ctx = b->ctx;
    drop_graph(b);
    GOLAP_HIP_CHECK(ctx, launch_star(ctx, launch_of(b, b->cfg), ctx->stream, true));
    GOLAP_HIP_CHECK(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    const int r = enqueue(b, ctx->stream);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
    if (r) {
        if (g) (void)hipGraphDestroy(g);
        return r;
    }
    GOLAP_HIP_CHECK(ctx, e);
    b->graph = g;
    GOLAP_HIP_CHECK(ctx, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    return GOLAP_OK;
}

int apply_config(golap_star_batch *b, const golap_batch_config &c) {
    if (int r = check_config(b, c)) return r;
    if (int r = upload_descriptors(b, c)) return r;
    b->cfg = c;
    b->ran = false;
    return capture(b);
}

golap_batch_config default_config(const golap_star_batch *b) {
    golap_batch_config c{};
    c.pages_per_item = b->ppc;
    while (parts_of(b, c.pages_per_item) > 256) ++c.pages_per_item;
    c.lds_cutover = 16;
    c.filter_lds = 1;
    c.compact_tables = 1;
    c.max_blocks_per_cu = 0;
    c.lds_group_slots = lds_keys_fit(b) ? LDS_GROUP_SLOTS_MAX : 0;
    while (c.lds_group_slots &&
           star_lds_bytes(b->max_filter_words, b->page_rows, b->n_windows, c.lds_group_slots) > (size_t)b->ctx->lds_max)
        c.lds_group_slots >>= 1;
    return c;
}

int check_builds(golap_star_batch *b) {
    for (const QueryHost &h : b->qh)
        for (const auto &mb : h.builds)
            if (mb.first->build_id != mb.second)
                return fail(b->ctx, GOLAP_EINVAL, "batch: a dimension map was rebuilt after the batch was created");
    return GOLAP_OK;
}

void free_batch(golap_star_batch *b) {
    drop_graph(b);
    dev_release(b->ctx, b->d_q);
    dev_release(b->ctx, b->d_chunk_off);
    dev_release(b->ctx, b->d_items);
    dev_release(b->ctx, b->d_arena);
    if (b->h_arena) (void)hipHostFree(b->h_arena);
    delete b;
}

}  // namespace

extern "C" {

int golap_star_batch_create(golap_ctx *ctx, const golap_batch_query *queries, uint32_t n_queries,
                            golap_star_batch **out) {
    if (!ctx || !queries || !n_queries || !out || n_queries > (1u << 24)) return GOLAP_EINVAL;
    *out = nullptr;
    Lock l(ctx->mu);
    if (int r = set_device(ctx)) return r;
    auto *b = new golap_star_batch();
    b->ctx = ctx;
    b->q.assign(queries, queries + n_queries);
    b->qh.resize(n_queries);
    b->qd.resize(n_queries);
    auto bail = [&](int r) {
        free_batch(b);
        return r;
    };
    for (uint32_t i = 0; i < n_queries; ++i) {
        if (int r = prepare_query(ctx, &b->q[i].plan, true, b->cs, b->qd[i], b->qh[i])) return bail(r);
        const golap_column *zc = b->q[i].zone_col;
        if (zc) {
            if (zc->ctx != ctx || zc->n_rows != b->q[i].plan.value->n_rows ||
                zc->chunk_rows != b->q[i].plan.value->chunk_rows || !zc->d_chunk_min)
                return bail(fail(ctx, GOLAP_EINVAL, "batch: zone column not row-aligned with the fact table"));
            b->qd[i].s.zone[0] = ZonePred{zc->d_chunk_min, zc->d_chunk_max, b->q[i].zone_lo, b->q[i].zone_hi};
        }
        b->max_filter_words = std::max(b->max_filter_words, b->qd[i].filter_words);
        b->max_domain = std::max(b->max_domain, b->qh[i].domain);
    }
    const golap_column *f = b->cs.cols[0];
    b->page_rows = f->page_rows;
    b->ppc = (uint32_t)(f->chunk_rows / f->page_rows);
    b->n_rows = f->n_rows;
    b->n_pages = f->n_pages;
    uint64_t pb = ~0ull, pe = 0;
    b->chunk_off.assign(1, 0);
    for (uint32_t i = 0; i < n_queries; ++i) {
        pb = std::min(pb, b->qd[i].s.page_begin);
        pe = std::max(pe, b->qd[i].s.page_end);
        b->chunk_off.push_back(b->chunk_off.back() + (b->qd[i].s.chunk_end - b->qd[i].s.chunk_begin));
    }
    b->n_windows = assign_windows(b->cs, std::min(pb, pe), pe, b->windows);
    // host view of the surviving chunks: the same test as k_batch_items
    b->survive.resize(n_queries);
    b->surv_pages.assign(n_queries, 0);
    for (uint32_t i = 0; i < n_queries; ++i) {
        const SliceDev &d = b->qd[i].s;
        const uint64_t nch = d.chunk_end - d.chunk_begin;
        std::vector<uint16_t> mask;
        if (d.chunk_mask && nch) {
            mask.resize(nch);
            if (hipMemcpyAsync(mask.data(), d.chunk_mask + d.chunk_begin, nch * 2, hipMemcpyDeviceToHost,
                               ctx->stream) != hipSuccess ||
                hipStreamSynchronize(ctx->stream) != hipSuccess)
                return bail(fail(ctx, GOLAP_EHIP, "batch: reading a chunk mask failed"));
        }
        const golap_column *zc = b->q[i].zone_col;
        auto &sv = b->survive[i];
        sv.resize(nch);
        for (uint64_t k = 0; k < nch; ++k) {
            const uint64_t c = d.chunk_begin + k;
            bool pass = mask.empty() || mask[k];
            if (pass && zc) pass = b->q[i].zone_lo <= zc->chunk_max[c] && b->q[i].zone_hi >= zc->chunk_min[c];
            sv[k] = pass;
            if (pass) b->surv_pages[i] += std::min<uint64_t>(b->n_pages, (c + 1) * b->ppc) - c * b->ppc;
        }
        golap_star_result tmp{};
        slice_stats(b->q[i].plan, sv, &tmp);
        b->total_pages += b->surv_pages[i];
        b->total_bytes += tmp.comp_bytes;
    }
    // device buffers
    uint64_t off = ARENA_HEADER;
    b->res_off.resize(n_queries);
    for (uint32_t i = 0; i < n_queries; ++i) {
        b->res_off[i] = off;
        off += b->qh[i].domain * 16;
    }
    b->arena_bytes = off;
    void *p = nullptr;
    if (int r = dev_alloc(ctx, off, &p)) return bail(r);
    b->d_arena = static_cast<uint8_t *>(p);
    if (hipHostMalloc(reinterpret_cast<void **>(&b->h_arena), off, hipHostMallocDefault) != hipSuccess) {
        b->h_arena = nullptr;
        return bail(fail(ctx, GOLAP_ENOMEM, "batch: pinned result mirror"));
    }
    if (int r = dev_alloc(ctx, n_queries * sizeof(StarQueryDev), &p)) return bail(r);
    b->d_q = static_cast<StarQueryDev *>(p);
    if (int r = dev_alloc(ctx, b->chunk_off.size() * 8, &p)) return bail(r);
    b->d_chunk_off = static_cast<uint64_t *>(p);
    const uint64_t items_cap = std::max<uint64_t>(1, b->chunk_off.back() * b->ppc);
    if (items_cap >= (1ull << 32)) return bail(fail(ctx, GOLAP_EINVAL, "batch: too many work items"));
    if (int r = dev_alloc(ctx, items_cap * sizeof(uint2), &p)) return bail(r);
    b->d_items = static_cast<uint2 *>(p);
    if (hipMemcpyAsync(b->d_chunk_off, b->chunk_off.data(), b->chunk_off.size() * 8, hipMemcpyHostToDevice,
                       ctx->stream) != hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess)
        return bail(fail(ctx, GOLAP_EHIP, "batch: upload failed"));
    b->shape = shape_of(b);
    golap_batch_config c = default_config(b);
    auto it = ctx->tune_cache.find(b->shape);
    if (it != ctx->tune_cache.end() && check_config(b, it->second) == GOLAP_OK) c = it->second;
    if (int r = apply_config(b, c)) return bail(r);
    *out = b;
    return GOLAP_OK;
}

int golap_star_batch_config_get(golap_star_batch *b, golap_batch_config *cfg) {
    if (!b || !cfg) return GOLAP_EINVAL;
    *cfg = b->cfg;
    return GOLAP_OK;
}

int golap_star_batch_config_set(golap_star_batch *b, const golap_batch_config *cfg) {
    if (!b || !cfg) return GOLAP_EINVAL;
    Lock l(b->ctx->mu);
    if (int r = set_device(b->ctx)) return r;
    const golap_batch_config old = b->cfg;
    if (int r = apply_config(b, *cfg)) {
        (void)apply_config(b, old);
        return r;
    }
    return GOLAP_OK;
}

int golap_star_batch_run(golap_star_batch *b, uint32_t use_graph, golap_batch_run_info *info) {
    if (!b) return GOLAP_EINVAL;
    golap_ctx *ctx = b->ctx;
    Lock l(ctx->mu);
    if (int r = set_device(ctx)) return r;
    if (int r = check_builds(b)) return r;
    GOLAP_HIP_CHECK(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    if (use_graph) {
        GOLAP_HIP_CHECK(ctx, hipGraphLaunch(b->exec, ctx->stream));
    } else if (int r = enqueue(b, ctx->stream)) {
        return r;
    }
    GOLAP_HIP_CHECK(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    uint32_t hdr[4] = {0, 0, 0, 0};
    GOLAP_HIP_CHECK(ctx, hipMemcpyAsync(hdr, b->d_arena, 16, hipMemcpyDeviceToHost, ctx->stream));
    GOLAP_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    float ms = 0;
    GOLAP_HIP_CHECK(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    b->last_ms = ms;
    b->ran = false;
    b->h_valid = false;
    if (int r = status_of_flags(ctx, hdr[2], "batch")) return r;
    b->ran = true;
    if (info) {
        info->ms = ms;
        info->items = hdr[1];
        info->pages = b->total_pages;
        info->comp_bytes = b->total_bytes;
    }
    return GOLAP_OK;
}

int golap_star_batch_tune(golap_star_batch *b, uint32_t reps, golap_batch_config *best, double *best_ms) {
    if (!b) return GOLAP_EINVAL;
    golap_ctx *ctx = b->ctx;
    Lock l(ctx->mu);
    if (int r = set_device(ctx)) return r;
    reps = std::max<uint32_t>(reps, 1);
    auto time_cfg = [&](const golap_batch_config &c, double *ms) -> int {
        if (check_config(b, c)) return GOLAP_EINVAL;
        if (int r = apply_config(b, c)) return r;
        if (int r = golap_star_batch_run(b, 1, nullptr)) return r;  // warm-up
        std::vector<double> t;
        for (uint32_t i = 0; i < reps; ++i) {
            golap_batch_run_info info{};
            if (int r = golap_star_batch_run(b, 1, &info)) return r;
            t.push_back(info.ms);
        }
        std::sort(t.begin(), t.end());
        *ms = t[t.size() / 2];
        return GOLAP_OK;
    };
    golap_batch_config cur = b->cfg;
    double cur_ms = 0;
    if (int r = time_cfg(cur, &cur_ms)) return r;
    // coordinate descent over the knobs, in order of expected effect
    std::vector<std::vector<golap_batch_config>> axes;
    {
        std::vector<golap_batch_config> ax;
        for (uint32_t slots : {0u, 1024u, 2048u, 4096u})
            for (uint32_t cut : {0u, 4u, 16u, 256u}) {
                if (!slots && cut) continue;
                golap_batch_config c{};
                c.lds_group_slots = slots;
                c.lds_cutover = cut;
                ax.push_back(c);
            }
        axes.push_back(ax);
    }
    {
        std::vector<golap_batch_config> ax;
        for (uint32_t ppi = 1; ppi <= b->ppc; ppi *= 2) {
            golap_batch_config c{};
            c.pages_per_item = ppi;
            ax.push_back(c);
        }
        axes.push_back(ax);
    }
    axes.push_back({golap_batch_config{0, 0, 0, 1, 0, 0}, golap_batch_config{0, 0, 0, 0, 0, 0}});  // filter
    axes.push_back({golap_batch_config{0, 0, 0, 0, 1, 0}, golap_batch_config{0, 0, 0, 0, 0, 0}});  // tables
    {
        std::vector<golap_batch_config> ax;
        for (uint32_t mb : {0u, 1u, 2u, 3u}) {
            golap_batch_config c{};
            c.max_blocks_per_cu = mb;
            ax.push_back(c);
        }
        axes.push_back(ax);
    }
    for (size_t a = 0; a < axes.size(); ++a)
        for (const golap_batch_config &v : axes[a]) {
            golap_batch_config c = cur;
            switch (a) {
            case 0: c.lds_group_slots = v.lds_group_slots; c.lds_cutover = v.lds_cutover; break;
            case 1: c.pages_per_item = v.pages_per_item; break;
            case 2: c.filter_lds = v.filter_lds; break;
            case 3: c.compact_tables = v.compact_tables; break;
            default: c.max_blocks_per_cu = v.max_blocks_per_cu; break;
            }
            if (!memcmp(&c, &cur, sizeof(c))) continue;
            double ms = 0;
            const int r = time_cfg(c, &ms);
            if (r == GOLAP_EINVAL) continue;  // not valid for this batch (LDS budget, item limits)
            if (r) return r;
            if (ms < cur_ms * 0.98) {  // keep the incumbent on ties (timing noise)
                cur = c;
                cur_ms = ms;
            }
        }
    if (int r = apply_config(b, cur)) return r;
    ctx->tune_cache[b->shape] = cur;
    if (best) *best = cur;
    if (best_ms) *best_ms = cur_ms;
    return GOLAP_OK;
}

int golap_star_batch_result(golap_star_batch *b, uint32_t q, golap_star_result *res, int32_t *group_keys,
                            int64_t *sums, uint64_t *counts, uint64_t max_groups) {
    if (!b || !res || q >= b->q.size()) return GOLAP_EINVAL;
    golap_ctx *ctx = b->ctx;
    Lock l(ctx->mu);
    if (!b->ran) return fail(ctx, GOLAP_EINVAL, "batch: no completed run since creation or reconfiguration");
    if (int r = set_device(ctx)) return r;
    memset(res, 0, sizeof(*res));
    const QueryHost &h = b->qh[q];
    res->kernel_ms = b->last_ms;
    res->driver_dim = b->qd[q].driver;
    res->lds_group_slots = b->qd[q].use_lds ? b->cfg.lds_group_slots : 0;
    slice_stats(b->q[q].plan, b->survive[q], res);
    if (!b->h_valid) {  // every query's results in one copy
        GOLAP_HIP_CHECK(ctx, hipMemcpyAsync(b->h_arena + ARENA_HEADER, b->d_arena + ARENA_HEADER,
                                            b->arena_bytes - ARENA_HEADER, hipMemcpyDeviceToHost, ctx->stream));
        GOLAP_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
        b->h_valid = true;
    }
    const int64_t *hs = reinterpret_cast<const int64_t *>(b->h_arena + b->res_off[q]);
    const uint64_t *hc = reinterpret_cast<const uint64_t *>(b->h_arena + b->res_off[q] + h.domain * 8);
    return emit_groups_host(ctx, h.domain, hs, hc, h.lo, h.card, h.stride, &res->n_groups, &res->total_sum,
                            group_keys, sums, counts, max_groups, &res->rows_matched);
}

int golap_star_batch_free(golap_star_batch *b) {
    if (!b) return GOLAP_EINVAL;
    golap_ctx *ctx = b->ctx;
    Lock l(ctx->mu);
    if (int r = set_device(ctx)) return r;
    free_batch(b);
    return GOLAP_OK;
}

}  // extern "C"
