// Batched plans (batch.hpp, include/golap_hip.h). Many star or scan queries over one fact table
// run as one persistent fused launch:
//   1. a fill kernel zeroes the result arena (work-list counters, error word, every query's
//      dense SUM/COUNT arrays);
//   2. k_batch_items evaluates each query's chunk mask and zone-map predicates per chunk of its
//      slice (check_mmmeta semantics, core/metadata.cuh:539-547) and writes the surviving
//      (chunk, query, part) items;
//   3. k_star or k_scan (queue schedule) lets workgroups take items with one atomic each.
// Public batches capture the three kernels into a hipGraph once and replay it per run. The
// launch configuration is chosen by a coordinate-descent autotuner that times candidates on the
// batch itself and caches the winner in the context by batch shape.
#include "batch.hpp"

#include <algorithm>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

using namespace golap_impl;
using namespace golap_impl::starhost;
using Lock = std::lock_guard<std::recursive_mutex>;

#define BATCH_HIP_CHECK(b, expr)                                                                                 \
    do {                                                                                                         \
        hipError_t e__ = (expr);                                                                                 \
        if (e__ != hipSuccess) return bfail((b), GOLAP_EHIP, std::string(#expr ": ") + hipGetErrorString(e__)); \
    } while (0)

namespace {

constexpr uint64_t ARENA_HEADER = 256;  // queue head, item count, error word

int bfail(golap_batch *b, int code, const std::string &msg) {
    b->err = msg;
    return code;
}

uint32_t parts_of(const golap_batch *b, uint32_t ppi) { return (b->ppc + ppi - 1) / ppi; }
uint32_t desc_size(const golap_batch *b) {
    return b->kind == BatchKind::STAR ? (uint32_t)sizeof(StarQueryDev) : (uint32_t)sizeof(ScanQueryDev);
}
const SliceDev &slice_of(const golap_batch *b, uint32_t i) { return b->kind == BatchKind::STAR ? b->qd[i].s : b->sd[i].s; }
uint64_t domain_of(const golap_batch *b, uint32_t i) {
    return b->kind == BatchKind::STAR ? b->qh[i].domain : b->sh[i].domain;
}

void drop_graph(golap_batch *b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

// Batches of the same shape share tuned configurations.
std::string shape_of(const golap_batch *b) {
    auto bucket = [](uint64_t v) { return bits_for(std::max<uint64_t>(v, 1)); };
    return std::string(b->kind == BatchKind::STAR ? "star/" : "scan/") + std::to_string(b->page_rows) + "/" +
           std::to_string(b->ppc) + "/" + std::to_string(b->n_windows) + "/" + std::to_string(b->win_bytes) + "/" +
           std::to_string(b->cs.cols.size()) + "/" + std::to_string(bucket(b->n_queries())) + "/" +
           std::to_string(bucket(b->total_pages)) + "/" + std::to_string(bucket(b->max_domain));
}

// Columns (of the batch's column set) query i reads.
uint32_t colmask_of(const golap_batch *b, uint32_t i) {
    if (b->kind == BatchKind::SCAN) return b->sd[i].colmask;
    const StarQueryDev &d = b->qd[i];
    uint32_t m = 1u << d.val_col;
    for (uint32_t k = 0; k < d.n_dims; ++k) m |= 1u << d.fk_col[k];
    return m;
}

// The shared star filter: a fact column every query probes (the most common driver among them)
// and each query's membership filter of its dimension on that column. Left out when no column
// qualifies.
int build_shared_filter(golap_ctx *ctx, golap_batch *b) {
    const uint32_t nq = b->n_queries(), ncols = (uint32_t)b->cs.cols.size();
    std::vector<uint32_t> has(ncols, 0), drives(ncols, 0);
    for (uint32_t i = 0; i < nq; ++i) {
        const StarQueryDev &d = b->qd[i];
        uint32_t seen = 0;
        for (uint32_t k = 0; k < d.n_dims; ++k) seen |= 1u << d.fk_col[k];
        for (uint32_t c = 0; c < ncols; ++c) has[c] += (seen >> c) & 1u;
        drives[d.fk_col[d.driver]]++;
    }
    int mcol = -1;
    for (uint32_t c = 0; c < ncols; ++c)
        if (has[c] == nq && (mcol < 0 || drives[c] > drives[mcol])) mcol = (int)c;
    if (mcol < 0) return GOLAP_OK;
    std::vector<DriverFilterView> f(((nq + 63) / 64) * 64, DriverFilterView{});
    for (uint32_t i = 0; i < nq; ++i) {
        const StarQueryDev &d = b->qd[i];
        for (uint32_t k = 0; k < d.n_dims; ++k)
            if (d.fk_col[k] == (uint32_t)mcol) {
                uint32_t words = 0;
                if (int r = filter_view(ctx, b->sq[i].plan.dims[k].map, &f[i], &words)) return r;
                break;
            }
    }
    void *p = nullptr;
    if (int r = dev_alloc(ctx, f.size() * sizeof(DriverFilterView), &p)) return r;
    hipError_t e = hipMemcpyAsync(p, f.data(), f.size() * sizeof(DriverFilterView), hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);  // with the filters built on it
    if (e != hipSuccess) {
        dev_release(ctx, p);
        return fail(ctx, GOLAP_EHIP, std::string("batch shared filter: ") + hipGetErrorString(e));
    }
    b->d_mfilt = static_cast<DriverFilterView *>(p);
    b->mcol = mcol;
    return GOLAP_OK;
}

uint32_t group_bits_of(const golap_batch *b) { return std::max<uint32_t>(1, bits_for(b->max_domain)); }

// LDS group keys are (query << group_bits | group) and must stay below the empty marker.
bool lds_keys_fit(const golap_batch *b) { return bits_for(b->n_queries()) + group_bits_of(b) <= 31; }

size_t lds_bytes(const golap_batch *b, const golap_batch_config &c) {
    if (b->kind == BatchKind::STAR && c.shared_filter) return star_lds_bytes(0, b->page_rows, b->n_windows, 0);
    if (b->kind == BatchKind::STAR)
        return star_lds_bytes(c.filter_lds ? b->max_filter_words : 0, b->page_rows, b->pt_n_windows, c.lds_group_slots,
                              b->pt_n_anchors);
    return scan_lds_bytes(b->win_bytes, c.lds_group_slots, scan_page_waves(b->ctx, b->win_bytes, c.lds_group_slots));
}

int check_config(golap_batch *b, const golap_batch_config &c) {
    if (c.pages_per_item < 1 || c.pages_per_item > b->ppc || parts_of(b, c.pages_per_item) > 256)
        return bfail(b, GOLAP_EINVAL, "batch: pages_per_item out of range");
    const uint32_t max_slots = b->kind == BatchKind::STAR ? LDS_GROUP_SLOTS_MAX : SCAN_AGG_MAX;
    if (c.lds_group_slots && (c.lds_group_slots & (c.lds_group_slots - 1) || c.lds_group_slots > max_slots))
        return bfail(b, GOLAP_EINVAL, "batch: lds_group_slots must be 0 or a power of two <= 4096 (star) / 8192 (scan)");
    if (c.lds_group_slots && b->kind == BatchKind::STAR && !lds_keys_fit(b))
        return bfail(b, GOLAP_EINVAL, "batch: too many queries x groups for LDS group keys");
    if (c.filter_lds > 1 || c.compact_tables > 1 || c.shared_filter > 1) return bfail(b, GOLAP_EINVAL, "batch: bad flag");
    if (c.shared_filter && (b->kind != BatchKind::STAR || !b->d_mfilt))
        return bfail(b, GOLAP_EINVAL, "batch: no shared filter for this batch");
    if (c.rows_per_lane && (b->kind != BatchKind::SCAN || (c.rows_per_lane != 4 && c.rows_per_lane != 8)))
        return bfail(b, GOLAP_EINVAL, "batch: rows_per_lane must be 0, 4 or 8 (scan batches)");
    const size_t budget = b->kind == BatchKind::SCAN ? scan_lds_budget(b->ctx) : (size_t)b->ctx->lds_max;
    if (lds_bytes(b, c) > budget) return bfail(b, GOLAP_EINVAL, "batch: LDS budget exceeded");
    return GOLAP_OK;
}

// Descriptors for a configuration (star: compact or full probe tables; every query: LDS or
// global aggregation and its result arrays), uploaded on the slot stream.
int upload_descriptors(golap_batch *b, const golap_batch_config &c) {
    const uint32_t nq = b->n_queries();
    for (uint32_t i = 0; i < nq; ++i) {
        const uint64_t dom = domain_of(b, i);
        int64_t *gs = reinterpret_cast<int64_t *>(b->slot->d_arena + b->res_off[i]);
        uint64_t *gc = reinterpret_cast<uint64_t *>(b->slot->d_arena + b->res_off[i] + dom * 8);
        if (b->kind == BatchKind::STAR) {
            StarQueryDev &d = b->qd[i];
            const golap_star_plan &p = b->sq[i].plan;
            for (uint32_t k = 0; k < p.n_dims; ++k) {
                const golap_hashmap *m = p.dims[k].map;
                const bool small = c.compact_tables && m->small_log2;
                d.dim[k].slots = small ? m->d_small : m->d_slots;
                d.dim[k].log2cap = small ? m->small_log2 : m->log2cap;
            }
            const double expect = (double)b->surv_pages[i] * b->page_rows * b->qh[i].sel;
            const bool lds = p.agg_mode == GOLAP_AGG_LDS ||
                             (p.agg_mode == GOLAP_AGG_AUTO && expect >= (double)c.lds_cutover * (double)dom);
            d.use_lds = lds && c.lds_group_slots;
            d.g_sum = gs;
            d.g_cnt = gc;
        } else {
            // dense LDS partials per (plan, item) whenever the plan's local domain fits
            ScanQueryDev &d = b->sd[i];
            d.use_lds = b->sh[i].plan.agg_mode != GOLAP_AGG_GLOBAL && c.lds_group_slots;
            d.g_sum = gs;
            d.g_cnt = gc;
        }
    }
    const void *src = b->kind == BatchKind::STAR ? (const void *)b->qd.data() : (const void *)b->sd.data();
    BATCH_HIP_CHECK(b, hipMemcpyAsync(b->slot->d_q, src, (size_t)nq * desc_size(b), hipMemcpyHostToDevice,
                                      b->slot->stream));
    BATCH_HIP_CHECK(b, hipStreamSynchronize(b->slot->stream));
    return GOLAP_OK;
}

hipError_t launch_kernel(golap_batch *b, hipStream_t s, bool prepare_only) {
    const golap_batch_config &c = b->cfg;
    BatchSlot *sl = b->slot;
    if (b->kind == BatchKind::STAR) {
        StarLaunch L{};
        L.n_cols = (uint32_t)b->cs.cols.size();
        for (uint32_t i = 0; i < L.n_cols; ++i) {
            L.col[i] = view(b->cs.cols[i]);
            L.col_window[i] = c.shared_filter ? b->windows[i] : b->pt_windows[i];
            L.col_anchor[i] = c.shared_filter ? 0 : b->pt_anchors[i];
        }
        L.n_windows = c.shared_filter ? b->n_windows : b->pt_n_windows;
        L.n_anchors = c.shared_filter ? 0 : b->pt_n_anchors;
        L.page_rows = b->page_rows;
        L.pages_per_chunk = b->ppc;
        L.n_rows = b->n_rows;
        L.n_pages = b->n_pages;
        L.queries = reinterpret_cast<const StarQueryDev *>(sl->d_q);
        L.n_queries = b->n_queries();
        L.items = sl->d_items;
        L.queue = reinterpret_cast<uint32_t *>(sl->d_arena);
        L.pages_per_item = c.pages_per_item;
        L.filter_words = c.filter_lds ? b->max_filter_words : 0;
        L.lds_gslots = c.lds_group_slots;
        L.group_bits = group_bits_of(b);
        L.max_blocks_per_cu = c.max_blocks_per_cu;
        L.err = reinterpret_cast<uint32_t *>(sl->d_arena + 8);
        if (c.shared_filter) {
            L.items4 = sl->d_items4;
            L.mfilt = b->d_mfilt;
            L.mcol = (uint32_t)b->mcol;
        }
        return launch_star(b->ctx, L, s, prepare_only);
    }
    ScanLaunch L{};
    L.n_cols = (uint32_t)b->cs.cols.size();
    for (uint32_t i = 0; i < L.n_cols; ++i)
        L.col[i] = ScanColView{view(b->cs.cols[i]), b->cs.cols[i]->type, b->windows[i], b->dslots[i], b->wcaps[i]};
    if (b->dec_mask)
        scan_decode_views(b->cs.cols.data(), L.n_cols, b->dec_mask, b->union_begin, b->union_end - b->union_begin,
                          sl->d_dec, L.col);
    L.tok_off = b->tok_off;
    L.tok_cap = b->tok_cap;
    L.page_rows = b->page_rows;
    L.pages_per_chunk = b->ppc;
    L.n_rows = b->n_rows;
    L.n_pages = b->n_pages;
    L.win_bytes = b->win_bytes;
    L.agg_entries = c.lds_group_slots;
    L.page_waves = scan_page_waves(b->ctx, b->win_bytes, c.lds_group_slots);
    L.queries = reinterpret_cast<const ScanQueryDev *>(sl->d_q);
    L.n_queries = b->n_queries();
    L.items = sl->d_items4;
    L.queue = reinterpret_cast<uint32_t *>(sl->d_arena);
    L.pages_per_item = c.pages_per_item;
    L.max_blocks_per_cu = c.max_blocks_per_cu;
    L.rows_per_lane = c.rows_per_lane ? c.rows_per_lane
                                      : scan_rows_per_lane(b->ctx, b->union_pairs * b->ppc,
                                                           scan_lds_bytes(L.win_bytes, L.agg_entries, L.page_waves));
    L.err = reinterpret_cast<uint32_t *>(sl->d_arena + 8);
    return launch_scan(b->ctx, L, s, prepare_only);
}

// The run's three kernels on one stream.
int enqueue_kernels(golap_batch *b, hipStream_t s) {
    BatchSlot *sl = b->slot;
    BATCH_HIP_CHECK(b, launch_fill_zero(sl->d_arena, b->arena_bytes, s));
    if (b->kind == BatchKind::STAR && b->cfg.shared_filter)
        BATCH_HIP_CHECK(b, launch_mask_items(sl->d_q, desc_size(b), b->n_queries(), b->union_begin,
                                             b->union_end - b->union_begin, parts_of(b, b->cfg.pages_per_item),
                                             sl->d_items4, reinterpret_cast<uint32_t *>(sl->d_arena), s));
    else if (b->kind == BatchKind::STAR)
        BATCH_HIP_CHECK(b, launch_batch_items(sl->d_q, desc_size(b), b->n_queries(), sl->d_chunk_off,
                                              b->chunk_off.back(), parts_of(b, b->cfg.pages_per_item), sl->d_items,
                                              reinterpret_cast<uint32_t *>(sl->d_arena), s));
    else
        BATCH_HIP_CHECK(b, launch_scan_items(reinterpret_cast<const ScanQueryDev *>(sl->d_q), b->n_queries(),
                                             b->union_begin, b->union_end - b->union_begin,
                                             parts_of(b, b->cfg.pages_per_item), sl->d_items4,
                                             reinterpret_cast<uint32_t *>(sl->d_arena), s));
    BATCH_HIP_CHECK(b, launch_kernel(b, s, false));
    return GOLAP_OK;
}

int capture(golap_batch *b) {
    drop_graph(b);
    hipStream_t s = b->slot->stream;
    BATCH_HIP_CHECK(b, launch_kernel(b, s, true));  // kernel attributes outside the capture
    BATCH_HIP_CHECK(b, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    const int r = enqueue_kernels(b, s);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &g);
    if (r) {
        if (g) (void)hipGraphDestroy(g);
        return r;
    }
    BATCH_HIP_CHECK(b, e);
    b->graph = g;
    BATCH_HIP_CHECK(b, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    return GOLAP_OK;
}

int apply_config(golap_batch *b, const golap_batch_config &c, bool with_graph) {
    if (int r = check_config(b, c)) return r;
    if (int r = upload_descriptors(b, c)) return r;
    b->cfg = c;
    b->ran = false;
    return with_graph ? capture(b) : GOLAP_OK;
}

golap_batch_config default_config(golap_batch *b) {
    golap_batch_config c{};
    c.pages_per_item = b->ppc;
    while (parts_of(b, c.pages_per_item) > 256) ++c.pages_per_item;
    c.lds_cutover = b->kind == BatchKind::STAR ? 16 : 0;
    c.filter_lds = b->kind == BatchKind::STAR;
    c.compact_tables = 1;
    c.max_blocks_per_cu = 0;
    if (b->kind == BatchKind::STAR) {
        c.lds_group_slots = lds_keys_fit(b) ? LDS_GROUP_SLOTS_MAX : 0;
        while (c.lds_group_slots && lds_bytes(b, c) > (size_t)b->ctx->lds_max) c.lds_group_slots >>= 1;
        c.shared_filter = 0;  // the tuner tries it
    } else {
        // dense partials for several plans per item (4096 entries: about four request
        // templates' zone x hour partials per pass); items small enough to fill the device
        c.lds_group_slots = scan_agg_entries(b->ctx, GOLAP_AGG_AUTO, 0, b->win_bytes,
                                             scan_page_waves(b->ctx, b->win_bytes, 1024),
                                             b->sh.size() > 1 ? std::max(4096u, scan_default_entries(b->max_local))
                                                              : scan_default_entries(b->max_local));
        c.pages_per_item = scan_pages_per_item(b->ctx, b->union_pairs, b->ppc);
    }
    return c;
}

int check_builds(golap_batch *b) {
    for (const QueryHost &h : b->qh)
        for (const auto &mb : h.builds)
            if (mb.first->build_id != mb.second)
                return bfail(b, GOLAP_EINVAL, "batch: a dimension map was rebuilt after the batch was created");
    return GOLAP_OK;
}

// Slices, survivors (host view), arena layout, buffers and the first configuration.
int finish_create(golap_ctx *ctx, golap_batch *b, bool with_graph) {
    const uint32_t nq = b->n_queries();
    const golap_column *f = b->cs.cols[0];
    b->page_rows = f->page_rows;
    b->ppc = (uint32_t)(f->chunk_rows / f->page_rows);
    b->n_rows = f->n_rows;
    b->n_pages = f->n_pages;
    b->chunk_off.assign(1, 0);
    for (uint32_t i = 0; i < nq; ++i) {
        const SliceDev &s = slice_of(b, i);
        b->chunk_off.push_back(b->chunk_off.back() + (s.chunk_end - s.chunk_begin));
        b->max_domain = std::max(b->max_domain, domain_of(b, i));
    }
    b->survive.resize(nq);
    b->surv_pages.assign(nq, 0);
    for (uint32_t i = 0; i < nq; ++i) {
        const SliceDev &d = slice_of(b, i);
        const uint64_t nch = d.chunk_end - d.chunk_begin;
        std::vector<uint16_t> mask;
        if (d.chunk_mask && nch) {
            mask.resize(nch);
            GOLAP_HIP_CHECK(ctx, hipMemcpyAsync(mask.data(), d.chunk_mask + d.chunk_begin, nch * 2,
                                                hipMemcpyDeviceToHost, ctx->stream));
            GOLAP_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
        }
        golap_star_result tmp{};
        if (b->kind == BatchKind::STAR) {
            const golap_star_query &q = b->sq[i];
            auto &sv = b->survive[i];
            sv.resize(nch);
            for (uint64_t k = 0; k < nch; ++k) {
                const uint64_t c = d.chunk_begin + k;
                bool pass = mask.empty() || mask[k];
                if (pass && q.zone_col) pass = q.zone_lo <= q.zone_col->chunk_max[c] && q.zone_hi >= q.zone_col->chunk_min[c];
                sv[k] = pass;
            }
            slice_stats(q.plan, sv, &tmp);
        } else {
            b->survive[i] = scan_survivors(b->sh[i].plan, mask);
            scan_stats(b->sh[i].plan, b->sh[i].colmask, b->survive[i], &tmp);
            b->max_local = std::max(b->max_local, scan_local_domain(b->sh[i].plan, b->sd[i], b->survive[i]));
        }
        for (uint64_t k = 0; k < nch; ++k)
            if (b->survive[i][k]) {
                const uint64_t c = d.chunk_begin + k;
                b->surv_pages[i] += std::min<uint64_t>(b->n_pages, (c + 1) * b->ppc) - c * b->ppc;
            }
        b->total_pages += b->surv_pages[i];
        b->total_bytes += tmp.comp_bytes;
    }
    if (b->kind == BatchKind::SCAN || b->d_mfilt) {
        // shared scan / shared star filter: the union of the slices, the surviving (chunk, plan
        // word) pairs and the bytes read once per chunk for all plans that keep it
        b->union_begin = ~0ull;
        b->union_end = 0;
        for (uint32_t i = 0; i < nq; ++i) {
            const SliceDev &d = slice_of(b, i);
            if (d.chunk_end <= d.chunk_begin) continue;
            b->union_begin = std::min(b->union_begin, d.chunk_begin);
            b->union_end = std::max(b->union_end, d.chunk_end);
        }
        if (b->union_begin > b->union_end) b->union_begin = b->union_end = 0;
        const uint64_t un = b->union_end - b->union_begin;
        const uint32_t nw = (nq + 63) / 64;
        std::vector<uint32_t> cmask(un * nw, 0);  // columns read per (chunk, word); bit 31: kept
        for (uint32_t i = 0; i < nq; ++i) {
            const SliceDev &d = slice_of(b, i);
            for (uint64_t k = 0; k < d.chunk_end - d.chunk_begin; ++k)
                if (b->survive[i][k]) cmask[(d.chunk_begin + k - b->union_begin) * nw + i / 64] |= colmask_of(b, i) | (1u << 31);
        }
        b->dec_runs.clear();
        for (uint64_t k = 0; b->dec_mask && k < un;) {  // chunks where a kept plan reads a decoded column
            auto needs = [&](uint64_t j) {
                for (uint32_t w = 0; w < nw; ++w)
                    if (cmask[j * nw + w] & b->dec_mask) return true;
                return false;
            };
            if (!needs(k)) {
                ++k;
                continue;
            }
            uint64_t e = k + 1;
            while (e < un && needs(e)) ++e;
            b->dec_runs.emplace_back(b->union_begin + k, b->union_begin + e);
            k = e;
        }
        b->union_pairs = 0;
        b->shared_bytes = 0;
        for (uint64_t k = 0; k < un; ++k)
            for (uint32_t w = 0; w < nw; ++w) {
                const uint32_t m = cmask[k * nw + w];
                if (!m) continue;
                ++b->union_pairs;
                for (size_t c = 0; c < b->cs.cols.size(); ++c)
                    if ((m >> c) & 1u) {
                        const auto &ex = b->cs.cols[c]->extra;
                        const uint64_t ch = b->union_begin + k;
                        b->shared_bytes += ex.chunk_payload[ch + 1] - ex.chunk_payload[ch];
                    }
            }
    }
    uint64_t off = ARENA_HEADER;
    b->res_off.resize(nq);
    for (uint32_t i = 0; i < nq; ++i) {
        b->res_off[i] = off;
        off += domain_of(b, i) * 16;
    }
    b->arena_bytes = off;
    const uint64_t items = std::max<uint64_t>(1, b->chunk_off.back() * b->ppc);
    if (items >= (1ull << 32)) return fail(ctx, GOLAP_EINVAL, "batch: too many work items");
    const uint64_t items4 = b->kind == BatchKind::SCAN || b->d_mfilt ? std::max<uint64_t>(1, b->union_pairs * b->ppc) : 0;
    if (items4 >= (1ull << 32)) return fail(ctx, GOLAP_EINVAL, "batch: too many work items");
    if (int r = b->slot->reserve(off, (uint64_t)nq * desc_size(b), b->chunk_off.size(), b->kind == BatchKind::STAR ? items : 1,
                                 items4))
        return fail(ctx, r, b->slot->err);
    if (b->dec_mask) {
        const uint64_t need = scan_decode_bytes(b->cs.cols.data(), (uint32_t)b->cs.cols.size(), b->dec_mask,
                                                b->union_end - b->union_begin);
        if (int r = b->slot->reserve_dec(need)) return fail(ctx, r, b->slot->err);
    }
    GOLAP_HIP_CHECK(ctx, hipMemcpyAsync(b->slot->d_chunk_off, b->chunk_off.data(), b->chunk_off.size() * 8,
                                        hipMemcpyHostToDevice, b->slot->stream));
    b->shape = shape_of(b);
    golap_batch_config c = default_config(b);
    auto it = ctx->tune_cache.find(b->shape);
    if (it != ctx->tune_cache.end() && check_config(b, it->second) == GOLAP_OK) c = it->second;
    if (int r = apply_config(b, c, with_graph)) return fail(ctx, r, b->err);
    return GOLAP_OK;
}

golap_batch *new_batch(golap_ctx *ctx, BatchKind kind, BatchSlot *slot, int *st) {
    auto *b = new golap_batch();
    b->ctx = ctx;
    b->kind = kind;
    *st = GOLAP_OK;
    if (slot) {
        b->slot = slot;
    } else {
        b->slot = new BatchSlot();
        b->own_slot = true;
        if (int r = b->slot->init(ctx, false)) *st = fail(ctx, r, b->slot->err);
    }
    return b;
}

}  // namespace

namespace golap_impl {

int BatchSlot::init(golap_ctx *c, bool own) {
    ctx = c;
    own_stream = own;
    if (own) {
        if (hipStreamCreateWithFlags(&stream, hipStreamNonBlocking) != hipSuccess) {
            err = "batch slot: stream";
            return GOLAP_EHIP;
        }
    } else {
        stream = c->stream;
    }
    if (hipEventCreate(&ev0) != hipSuccess || hipEventCreate(&ev1) != hipSuccess ||
        hipEventCreateWithFlags(&dep, hipEventDisableTiming) != hipSuccess) {
        err = "batch slot: events";
        return GOLAP_EHIP;
    }
    return GOLAP_OK;
}

int BatchSlot::reserve(uint64_t arena, uint64_t qbytes, uint64_t offs, uint64_t items, uint64_t items4) {
    auto grow = [&](auto *&p, uint64_t &cap, uint64_t need, uint64_t unit) -> int {
        if (need <= cap) return GOLAP_OK;
        const uint64_t want = std::max(need, cap * 2);
        dev_release(ctx, p);
        p = nullptr;
        cap = 0;
        void *q = nullptr;
        if (int r = dev_alloc(ctx, want * unit, &q)) {
            err = last_error(ctx);
            return r;
        }
        p = static_cast<std::remove_reference_t<decltype(p)>>(q);
        cap = want;
        return GOLAP_OK;
    };
    int r4 = GOLAP_OK;
    const bool arena_grows = arena > arena_cap;
    if (int r = grow(d_arena, arena_cap, arena, 1)) return r;
    if (arena_grows && h_arena) {
        (void)hipHostFree(h_arena);
        h_arena = nullptr;
    }
    if (!h_arena && hipHostMalloc(reinterpret_cast<void **>(&h_arena), arena_cap, hipHostMallocDefault) != hipSuccess) {
        h_arena = nullptr;
        err = "batch slot: pinned result mirror";
        return GOLAP_ENOMEM;
    }
    if (int r = grow(d_q, q_cap, qbytes, 1)) return r;
    if (int r = grow(d_chunk_off, off_cap, offs, 8)) return r;
    if (int r = grow(d_items, items_cap, items, sizeof(uint2))) return r;
    if (items4 && (r4 = grow(d_items4, items4_cap, items4, sizeof(uint4)))) return r4;
    return GOLAP_OK;
}

int BatchSlot::reserve_dec(uint64_t bytes) {
    if (!d_dec_err) {
        void *q = nullptr;
        if (int r = dev_alloc(ctx, 256, &q)) {
            err = last_error(ctx);
            return r;
        }
        d_dec_err = static_cast<uint32_t *>(q);
    }
    if (bytes <= dec_cap) return GOLAP_OK;
    dev_release(ctx, d_dec);
    d_dec = nullptr;
    dec_cap = 0;
    void *q = nullptr;
    if (int r = dev_alloc(ctx, bytes, &q)) {
        err = last_error(ctx);
        return r;
    }
    d_dec = static_cast<uint8_t *>(q);
    dec_cap = bytes;
    return GOLAP_OK;
}

void BatchSlot::release() {
    dev_release(ctx, d_dec);
    dev_release(ctx, d_dec_err);
    d_dec = nullptr;
    d_dec_err = nullptr;
    dec_cap = 0;
    dev_release(ctx, d_arena);
    dev_release(ctx, d_q);
    dev_release(ctx, d_chunk_off);
    dev_release(ctx, d_items);
    dev_release(ctx, d_items4);
    if (h_arena) (void)hipHostFree(h_arena);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (dep) (void)hipEventDestroy(dep);
    if (own_stream && stream) (void)hipStreamDestroy(stream);
    d_arena = h_arena = d_q = nullptr;
    d_chunk_off = nullptr;
    d_items = nullptr;
    d_items4 = nullptr;
    ev0 = ev1 = dep = nullptr;
    stream = nullptr;
    arena_cap = q_cap = off_cap = items_cap = items4_cap = 0;
}

uint64_t batch_query_bytes(uint64_t domain, uint64_t slice_chunks, uint32_t pages_per_chunk, bool star) {
    return domain * 16 + slice_chunks * pages_per_chunk * sizeof(uint2) + 8 +
           (star ? sizeof(StarQueryDev) : sizeof(ScanQueryDev));
}

int batch_create_star(golap_ctx *ctx, const golap_star_query *q, uint32_t n, BatchSlot *slot, bool with_graph,
                      golap_batch **out) {
    *out = nullptr;
    int st;
    golap_batch *b = new_batch(ctx, BatchKind::STAR, slot, &st);
    auto bail = [&](int r) {
        batch_free(b);
        return r;
    };
    if (st) return bail(st);
    b->sq.assign(q, q + n);
    b->qh.resize(n);
    b->qd.resize(n);
    for (uint32_t i = 0; i < n; ++i) {
        if (b->sq[i].plan.value_minus) return bail(fail(ctx, GOLAP_EINVAL, "batch: value_minus is a golap_star_run option"));
        if (int r = prepare_query(ctx, &b->sq[i].plan, true, b->cs, b->qd[i], b->qh[i])) return bail(r);
        const golap_column *zc = b->sq[i].zone_col;
        if (zc) {
            if (zc->ctx != ctx || zc->n_rows != b->sq[i].plan.value->n_rows ||
                zc->chunk_rows != b->sq[i].plan.value->chunk_rows || !zc->d_chunk_min)
                return bail(fail(ctx, GOLAP_EINVAL, "batch: zone column not row-aligned with the fact table"));
            b->qd[i].s.zone[0] = ZonePred{zc->d_chunk_min, zc->d_chunk_max, b->sq[i].zone_lo, b->sq[i].zone_hi};
        }
        b->max_filter_words = std::max(b->max_filter_words, b->qd[i].filter_words);
    }
    uint64_t pb = ~0ull, pe = 0;
    for (uint32_t i = 0; i < n; ++i) {
        pb = std::min(pb, b->qd[i].s.page_begin);
        pe = std::max(pe, b->qd[i].s.page_end);
    }
    b->n_windows = assign_windows(b->cs, std::min(pb, pe), pe, b->windows);
    uint32_t drivers = 0;
    for (uint32_t i = 0; i < n; ++i) drivers |= 1u << b->qd[i].fk_col[b->qd[i].driver];
    b->pt_n_windows = assign_access(b->cs, std::min(pb, pe), pe, drivers, b->pt_windows, b->pt_anchors, &b->pt_n_anchors);
    if (n > 1)
        if (int r = build_shared_filter(ctx, b)) return bail(r);
    if (int r = finish_create(ctx, b, with_graph)) return bail(r);
    *out = b;
    return GOLAP_OK;
}

int batch_create_scan(golap_ctx *ctx, const golap_scan_plan *q, uint32_t n, BatchSlot *slot, bool with_graph,
                      golap_batch **out) {
    *out = nullptr;
    int st;
    golap_batch *b = new_batch(ctx, BatchKind::SCAN, slot, &st);
    auto bail = [&](int r) {
        batch_free(b);
        return r;
    };
    if (st) return bail(st);
    b->sh.resize(n);
    b->sd.resize(n);
    uint64_t pb = ~0ull, pe = 0;
    uint32_t used = 0;  // launch columns read by some query
    for (uint32_t i = 0; i < n; ++i) {
        if (int r = prepare_scan(ctx, &q[i], b->sd[i], b->sh[i])) return bail(r);
        // plan-local column indexes -> indexes into the launch's columns
        uint8_t map[SCAN_MAX_COLS];
        for (uint32_t c = 0; c < q[i].n_cols; ++c)
            if (int r = b->cs.index(ctx, q[i].cols[c], &map[c])) return bail(r);
        ScanQueryDev &d = b->sd[i];
        for (uint32_t k = 0; k < d.n_preds; ++k) d.pred[k].col = map[d.pred[k].col];
        for (uint32_t k = 0; k < d.n_keys; ++k) d.key[k].col = map[d.key[k].col];
        if (d.val_col >= 0) d.val_col = map[d.val_col];
        uint32_t m = 0;
        for (uint32_t c = 0; c < q[i].n_cols; ++c)
            if ((d.colmask >> c) & 1u) m |= 1u << map[c];
        d.colmask = m;
        used |= m;
        pb = std::min(pb, d.s.page_begin);
        pe = std::max(pe, d.s.page_end);
    }
    // register slots and LDS windows of the used columns over the union of the slices
    {
        int32_t woff[SCAN_MAX_COLS], dslot[SCAN_MAX_COLS];
        uint32_t wcap[SCAN_MAX_COLS];
        b->dec_mask = n >= 2 ? scan_decode_mask(b->cs.cols.data(), (uint32_t)b->cs.cols.size(), used, std::min(pb, pe), pe)
                             : 0u;
        b->win_bytes = scan_layout(b->cs.cols.data(), (uint32_t)b->cs.cols.size(), used & ~b->dec_mask, std::min(pb, pe),
                                   pe, woff, dslot, wcap, &b->tok_off, &b->tok_cap);
        for (size_t c = 0; c < b->cs.cols.size(); ++c) {
            b->windows[c] = woff[c];
            b->dslots[c] = dslot[c];
            b->wcaps[c] = wcap[c];
            b->n_windows += woff[c] >= 0;
        }
    }
    if (int r = finish_create(ctx, b, with_graph)) return bail(r);
    *out = b;
    return GOLAP_OK;
}

int batch_enqueue(golap_batch *b, bool use_graph) {
    if (int r = check_builds(b)) return r;
    BatchSlot *sl = b->slot;
    if (sl->stream != b->ctx->stream) {  // after the context stream's pending work (masks, builds)
        BATCH_HIP_CHECK(b, hipEventRecord(sl->dep, b->ctx->stream));
        BATCH_HIP_CHECK(b, hipStreamWaitEvent(sl->stream, sl->dep, 0));
    }
    BATCH_HIP_CHECK(b, hipEventRecord(sl->ev0, sl->stream));
    if (b->dec_mask) {  // outside the graph: its fill kernel clears the arena's error word
        BATCH_HIP_CHECK(b, hipMemsetAsync(sl->d_dec_err, 0, 4, sl->stream));
        for (const auto &r : b->dec_runs)
            BATCH_HIP_CHECK(b, scan_decode_enqueue(b->ctx, b->cs.cols.data(), (uint32_t)b->cs.cols.size(), b->dec_mask,
                                                   b->union_begin, b->union_end - b->union_begin, sl->d_dec, r.first,
                                                   r.second, sl->stream, sl->d_dec_err));
    }
    if (use_graph && b->exec) {
        BATCH_HIP_CHECK(b, hipGraphLaunch(b->exec, sl->stream));
    } else if (int r = enqueue_kernels(b, sl->stream)) {
        return r;
    }
    BATCH_HIP_CHECK(b, hipEventRecord(sl->ev1, sl->stream));
    b->ran = false;
    b->h_valid = false;
    return GOLAP_OK;
}

int batch_finish(golap_batch *b, golap_batch_run_info *info) {
    BatchSlot *sl = b->slot;
    uint32_t *hdr = reinterpret_cast<uint32_t *>(sl->h_arena);
    BATCH_HIP_CHECK(b, hipMemcpyAsync(hdr, sl->d_arena, 16, hipMemcpyDeviceToHost, sl->stream));
    if (b->dec_mask) BATCH_HIP_CHECK(b, hipMemcpyAsync(hdr + 3, sl->d_dec_err, 4, hipMemcpyDeviceToHost, sl->stream));
    BATCH_HIP_CHECK(b, hipStreamSynchronize(sl->stream));
    if (b->dec_mask && hdr[3]) return bfail(b, GOLAP_ECORRUPT, "batch: corrupt page in a decoded column");
    float ms = 0;
    BATCH_HIP_CHECK(b, hipEventElapsedTime(&ms, sl->ev0, sl->ev1));
    b->last_ms = ms;
    const uint32_t flags = hdr[2];
    if (flags & DEVERR_CORRUPT) return bfail(b, GOLAP_ECORRUPT, "batch: corrupt page");
    if (flags & DEVERR_CAPACITY) return bfail(b, GOLAP_ECAPACITY, "batch: table full");
    if (flags) return bfail(b, GOLAP_EINVAL, "batch: value outside its declared range");
    b->ran = true;
    if (info) {
        info->ms = ms;
        info->items = hdr[1];
        info->pages = b->total_pages;
        info->comp_bytes = b->total_bytes;
        info->shared_comp_bytes = b->kind == BatchKind::SCAN ? b->shared_bytes : b->total_bytes;
    }
    return GOLAP_OK;
}

// Result fields of query q besides its groups; its group attributes' ranges.
void result_meta(const golap_batch *b, uint32_t q, golap_star_result *res, const std::vector<int32_t> *&lo,
                 const std::vector<uint32_t> *&card, const std::vector<uint64_t> *&stride) {
    memset(res, 0, sizeof(*res));
    res->kernel_ms = b->last_ms;
    if (b->kind == BatchKind::STAR) {
        res->driver_dim = b->qd[q].driver;
        res->lds_group_slots = b->qd[q].use_lds ? b->cfg.lds_group_slots : 0;
        slice_stats(b->sq[q].plan, b->survive[q], res);
        lo = &b->qh[q].lo;
        card = &b->qh[q].card;
        stride = &b->qh[q].stride;
    } else {
        res->lds_group_slots = b->sd[q].use_lds ? b->cfg.lds_group_slots : 0;
        scan_stats(b->sh[q].plan, b->sh[q].colmask, b->survive[q], res);
        lo = &b->sh[q].lo;
        card = &b->sh[q].card;
        stride = &b->sh[q].stride;
    }
}

int batch_result(golap_batch *b, uint32_t q, golap_star_result *res, int32_t *keys, int64_t *sums, uint64_t *counts,
                 uint64_t max_groups) {
    if (!b->ran) return bfail(b, GOLAP_EINVAL, "batch: no completed run since creation or reconfiguration");
    BatchSlot *sl = b->slot;
    if (!b->h_valid) {  // every query's results in one copy
        BATCH_HIP_CHECK(b, hipMemcpyAsync(sl->h_arena + ARENA_HEADER, sl->d_arena + ARENA_HEADER,
                                          b->arena_bytes - ARENA_HEADER, hipMemcpyDeviceToHost, sl->stream));
        BATCH_HIP_CHECK(b, hipStreamSynchronize(sl->stream));
        b->h_valid = true;
    }
    const std::vector<int32_t> *lo;
    const std::vector<uint32_t> *card;
    const std::vector<uint64_t> *stride;
    result_meta(b, q, res, lo, card, stride);
    const uint64_t dom = domain_of(b, q);
    const int64_t *hs = reinterpret_cast<const int64_t *>(sl->h_arena + b->res_off[q]);
    const uint64_t *hc = reinterpret_cast<const uint64_t *>(sl->h_arena + b->res_off[q] + dom * 8);
    uint64_t g = 0;
    int64_t total = 0;
    const int r = emit_groups_host(nullptr, dom, hs, hc, *lo, *card, *stride, &g, &total, keys, sums, counts,
                                   max_groups, &res->rows_matched);
    res->n_groups = g;
    res->total_sum = total;
    if (r) return bfail(b, r, "more groups than max_groups");
    return GOLAP_OK;
}

// Every query's non-empty groups compacted on the device and copied back alone, when they are
// few (at most a quarter of the concatenated domains): *done = true and the results written as
// golap_batch_results does. Otherwise *done = false (the dense copy serves).
int batch_results_sparse(golap_batch *b, golap_star_result *res, uint64_t *group_off, int32_t *keys, int64_t *sums,
                         uint64_t *counts, uint64_t max_total, bool *done) {
    *done = false;
    golap_ctx *ctx = b->ctx;
    BatchSlot *sl = b->slot;
    const uint32_t nq = b->n_queries();
    if (!b->d_cmeta) {
        uint64_t total = 0;
        std::vector<uint64_t> meta(3ull * nq);  // {arena offset, domain} per query, then first indexes
        for (uint32_t q = 0; q < nq; ++q) {
            meta[2 * q] = b->res_off[q];
            meta[2 * q + 1] = domain_of(b, q);
            meta[2ull * nq + q] = total;
            total += domain_of(b, q);
        }
        if (total < 4096) return GOLAP_OK;  // small enough to copy densely
        void *m = nullptr, *e = nullptr;
        if (int r = dev_alloc(ctx, meta.size() * 8, &m)) return r;
        const uint64_t bytes = total * 24 + 8;
        if (int r = dev_alloc(ctx, bytes, &e)) {
            dev_release(ctx, m);
            return r;
        }
        uint8_t *h = nullptr;
        if (hipHostMalloc(reinterpret_cast<void **>(&h), bytes, hipHostMallocDefault) != hipSuccess) {
            dev_release(ctx, m);
            dev_release(ctx, e);
            return fail(ctx, GOLAP_ENOMEM, "batch results: pinned buffer");
        }
        GOLAP_HIP_CHECK(ctx, hipMemcpyAsync(m, meta.data(), meta.size() * 8, hipMemcpyHostToDevice, sl->stream));
        GOLAP_HIP_CHECK(ctx, hipStreamSynchronize(sl->stream));
        b->d_cmeta = m;
        b->d_centries = static_cast<uint8_t *>(e);
        b->h_centries = h;
        b->c_total = total;
        b->c_cap = total;
    }
    const uint64_t cap = b->c_cap;
    uint2 *d_ql = reinterpret_cast<uint2 *>(b->d_centries);
    int64_t *d_sum = reinterpret_cast<int64_t *>(b->d_centries + cap * 8);
    uint64_t *d_cnt = reinterpret_cast<uint64_t *>(b->d_centries + cap * 16);
    unsigned long long *d_n = reinterpret_cast<unsigned long long *>(b->d_centries + cap * 24);
    unsigned long long *h_n = reinterpret_cast<unsigned long long *>(b->h_centries + cap * 24);
    GOLAP_HIP_CHECK(ctx, hipMemsetAsync(d_n, 0, 8, sl->stream));
    GOLAP_HIP_CHECK(ctx, launch_compact_groups(sl->d_arena, b->d_cmeta, static_cast<const uint64_t *>(b->d_cmeta) + 2ull * nq,
                                             nq, b->c_total, d_ql, d_sum, d_cnt, cap, d_n, (uint32_t)ctx->cus, sl->stream));
    GOLAP_HIP_CHECK(ctx, hipMemcpyAsync(h_n, d_n, 8, hipMemcpyDeviceToHost, sl->stream));
    GOLAP_HIP_CHECK(ctx, hipStreamSynchronize(sl->stream));
    const uint64_t n = *h_n;
    if (n > cap || n > b->c_total / 4) return GOLAP_OK;
    uint2 *h_ql = reinterpret_cast<uint2 *>(b->h_centries);
    int64_t *h_sum = reinterpret_cast<int64_t *>(b->h_centries + cap * 8);
    uint64_t *h_cnt = reinterpret_cast<uint64_t *>(b->h_centries + cap * 16);
    if (n) {
        GOLAP_HIP_CHECK(ctx, hipMemcpyAsync(h_ql, d_ql, n * 8, hipMemcpyDeviceToHost, sl->stream));
        GOLAP_HIP_CHECK(ctx, hipMemcpyAsync(h_sum, d_sum, n * 8, hipMemcpyDeviceToHost, sl->stream));
        GOLAP_HIP_CHECK(ctx, hipMemcpyAsync(h_cnt, d_cnt, n * 8, hipMemcpyDeviceToHost, sl->stream));
        GOLAP_HIP_CHECK(ctx, hipStreamSynchronize(sl->stream));
    }
    // entries by query, then by group index (the key tuple order)
    std::vector<uint64_t> order(n);
    for (uint64_t i = 0; i < n; ++i) order[i] = i;
    std::sort(order.begin(), order.end(), [&](uint64_t x, uint64_t y) {
        return h_ql[x].x != h_ql[y].x ? h_ql[x].x < h_ql[y].x : h_ql[x].y < h_ql[y].y;
    });
    uint64_t off = 0, koff = 0, e = 0;
    group_off[0] = 0;
    bool over = false;
    for (uint32_t q = 0; q < nq; ++q) {
        const std::vector<int32_t> *lo;
        const std::vector<uint32_t> *card;
        const std::vector<uint64_t> *stride;
        result_meta(b, q, &res[q], lo, card, stride);
        const size_t na = card->size();
        uint64_t g = 0, rows = 0;
        int64_t total = 0;
        for (; e < n && h_ql[order[e]].x == q; ++e) {
            const uint64_t k = order[e];
            total += h_sum[k];
            rows += h_cnt[k];
            if (off + g < max_total && !over) {
                if (keys) {
                    uint64_t rem = h_ql[k].y;
                    for (size_t a = na; a-- > 0;) {
                        keys[koff + g * na + a] = (*lo)[a] + (int32_t)(rem % (*card)[a]);
                        rem /= (*card)[a];
                    }
                }
                if (sums) sums[off + g] = h_sum[k];
                if (counts) counts[off + g] = h_cnt[k];
            }
            ++g;
        }
        (void)stride;
        res[q].n_groups = g;
        res[q].total_sum = total;
        res[q].rows_matched = rows;
        off += g;
        koff += g * na;
        group_off[q + 1] = off;
        if (off > max_total) over = true;
    }
    *done = true;
    if (off > max_total) return fail(ctx, GOLAP_ECAPACITY, "batch_results: more groups than max_total_groups");
    return GOLAP_OK;
}

void batch_free(golap_batch *b) {
    if (!b) return;
    drop_graph(b);
    if (b->d_mfilt) dev_release(b->ctx, b->d_mfilt);
    if (b->d_cmeta) dev_release(b->ctx, b->d_cmeta);
    if (b->d_centries) dev_release(b->ctx, b->d_centries);
    if (b->h_centries) (void)hipHostFree(b->h_centries);
    if (b->own_slot && b->slot) {
        b->slot->release();
        delete b->slot;
    }
    delete b;
}

}  // namespace golap_impl

extern "C" {

int golap_batch_create_star(golap_ctx *ctx, const golap_star_query *queries, uint32_t n_queries, golap_batch **out) {
    if (!ctx || !queries || !n_queries || !out || n_queries > (1u << 24)) return GOLAP_EINVAL;
    Lock l(ctx->mu);
    if (int r = set_device(ctx)) return r;
    return batch_create_star(ctx, queries, n_queries, nullptr, true, out);
}

int golap_batch_create_scan(golap_ctx *ctx, const golap_scan_plan *plans, uint32_t n_plans, golap_batch **out) {
    if (!ctx || !plans || !n_plans || !out || n_plans > (1u << 24)) return GOLAP_EINVAL;
    Lock l(ctx->mu);
    if (int r = set_device(ctx)) return r;
    return batch_create_scan(ctx, plans, n_plans, nullptr, true, out);
}

int golap_batch_config_get(golap_batch *b, golap_batch_config *cfg) {
    if (!b || !cfg) return GOLAP_EINVAL;
    *cfg = b->cfg;
    return GOLAP_OK;
}

int golap_batch_config_set(golap_batch *b, const golap_batch_config *cfg) {
    if (!b || !cfg) return GOLAP_EINVAL;
    golap_ctx *ctx = b->ctx;
    Lock l(ctx->mu);
    if (int r = set_device(ctx)) return r;
    const golap_batch_config old = b->cfg;
    if (int r = apply_config(b, *cfg, true)) {
        const std::string msg = b->err;
        (void)apply_config(b, old, true);
        return fail(ctx, r, msg);
    }
    return GOLAP_OK;
}

int golap_batch_run(golap_batch *b, uint32_t use_graph, golap_batch_run_info *info) {
    if (!b) return GOLAP_EINVAL;
    golap_ctx *ctx = b->ctx;
    Lock l(ctx->mu);
    if (int r = set_device(ctx)) return r;
    if (int r = batch_enqueue(b, use_graph != 0)) return fail(ctx, r, b->err);
    if (int r = batch_finish(b, info)) return fail(ctx, r, b->err);
    return GOLAP_OK;
}

int golap_batch_tune(golap_batch *b, uint32_t reps, golap_batch_config *best, double *best_ms) {
    if (!b) return GOLAP_EINVAL;
    golap_ctx *ctx = b->ctx;
    Lock l(ctx->mu);
    if (int r = set_device(ctx)) return r;
    reps = std::max<uint32_t>(reps, 1);
    auto time_cfg = [&](const golap_batch_config &c, double *ms) -> int {
        if (check_config(b, c)) return GOLAP_EINVAL;
        if (int r = apply_config(b, c, true)) return r;
        if (int r = golap_batch_run(b, 1, nullptr)) return r;  // warm-up
        std::vector<double> t;
        for (uint32_t i = 0; i < reps; ++i) {
            golap_batch_run_info info{};
            if (int r = golap_batch_run(b, 1, &info)) return r;
            t.push_back(info.ms);
        }
        std::sort(t.begin(), t.end());
        *ms = t[t.size() / 2];
        return GOLAP_OK;
    };
    golap_batch_config cur = b->cfg;
    double cur_ms = 0;
    if (int r = time_cfg(cur, &cur_ms)) return fail(ctx, r, b->err);
    const bool star = b->kind == BatchKind::STAR;
    // coordinate descent over the knobs, in order of expected effect
    enum Axis { SHARED, LDS, GRAIN, FILTER, TABLES, OCC, ROWS };
    std::vector<std::pair<Axis, std::vector<golap_batch_config>>> axes;
    {
        std::vector<golap_batch_config> ax;
        for (uint32_t slots : {0u, 1024u, 2048u, 4096u, 8192u})
            for (uint32_t cut : {0u, 4u, 16u, 256u}) {
                if ((!slots || !star) && cut) continue;  // scan plans have no cutover
                if (star && slots > LDS_GROUP_SLOTS_MAX) continue;
                golap_batch_config c{};
                c.lds_group_slots = slots;
                c.lds_cutover = cut;
                ax.push_back(c);
            }
        axes.emplace_back(LDS, ax);
    }
    {
        std::vector<golap_batch_config> ax;
        for (uint32_t ppi = 1; ppi <= b->ppc; ppi *= 2) {
            golap_batch_config c{};
            c.pages_per_item = ppi;
            ax.push_back(c);
        }
        axes.emplace_back(GRAIN, ax);
    }
    if (star) {
        golap_batch_config on{}, off{};
        on.filter_lds = 1;
        axes.emplace_back(FILTER, std::vector<golap_batch_config>{on, off});
        on = off = golap_batch_config{};
        on.compact_tables = 1;
        axes.emplace_back(TABLES, std::vector<golap_batch_config>{on, off});
    }
    if (!star) {  // rows per wave of the scan's sub-tiles
        std::vector<golap_batch_config> ax;
        for (uint32_t rpl : {4u, 8u}) {
            golap_batch_config c{};
            c.rows_per_lane = rpl;
            ax.push_back(c);
        }
        axes.emplace_back(ROWS, ax);
    }
    {
        std::vector<golap_batch_config> ax;
        for (uint32_t mb : {0u, 1u, 2u, 3u}) {
            golap_batch_config c{};
            c.max_blocks_per_cu = mb;
            ax.push_back(c);
        }
        axes.emplace_back(OCC, ax);
    }
    if (star && b->d_mfilt) {
        // the shared filter last, against the tuned per-query schedule; then grain and
        // occupancy once more for whichever schedule won
        golap_batch_config on{}, off{};
        on.shared_filter = 1;
        const auto grain = axes[1], occ = axes.back();
        axes.emplace_back(SHARED, std::vector<golap_batch_config>{on, off});
        axes.push_back(grain);
        axes.push_back(occ);
    }
    for (const auto &axis : axes)
        for (const golap_batch_config &v : axis.second) {
            golap_batch_config c = cur;
            switch (axis.first) {
            case SHARED: c.shared_filter = v.shared_filter; break;
            case LDS: c.lds_group_slots = v.lds_group_slots; c.lds_cutover = v.lds_cutover; break;
            case GRAIN: c.pages_per_item = v.pages_per_item; break;
            case FILTER: c.filter_lds = v.filter_lds; break;
            case TABLES: c.compact_tables = v.compact_tables; break;
            case OCC: c.max_blocks_per_cu = v.max_blocks_per_cu; break;
            case ROWS: c.rows_per_lane = v.rows_per_lane; break;
            }
            if (!memcmp(&c, &cur, sizeof(c))) continue;
            double ms = 0;
            const int r = time_cfg(c, &ms);
            if (r == GOLAP_EINVAL) continue;  // not valid for this batch (LDS budget, item limits)
            if (r) return fail(ctx, r, b->err);
            if (ms < cur_ms * 0.98) {  // keep the incumbent on ties (timing noise)
                cur = c;
                cur_ms = ms;
            }
        }
    if (int r = apply_config(b, cur, true)) return fail(ctx, r, b->err);
    ctx->tune_cache[b->shape] = cur;
    if (best) *best = cur;
    if (best_ms) *best_ms = cur_ms;
    return GOLAP_OK;
}

int golap_batch_result(golap_batch *b, uint32_t q, golap_star_result *res, int32_t *group_keys, int64_t *sums,
                       uint64_t *counts, uint64_t max_groups) {
    if (!b || !res || q >= b->n_queries()) return GOLAP_EINVAL;
    golap_ctx *ctx = b->ctx;
    Lock l(ctx->mu);
    if (int r = set_device(ctx)) return r;
    if (int r = batch_result(b, q, res, group_keys, sums, counts, max_groups)) return fail(ctx, r, b->err);
    return GOLAP_OK;
}

int golap_batch_results(golap_batch *b, golap_star_result *res, uint64_t *group_off, int32_t *group_keys,
                        int64_t *sums, uint64_t *counts, uint64_t max_total_groups) {
    if (!b || !res || !group_off) return GOLAP_EINVAL;
    golap_ctx *ctx = b->ctx;
    Lock l(ctx->mu);
    if (int r = set_device(ctx)) return r;
    if (!b->ran) return fail(ctx, GOLAP_EINVAL, "batch: no completed run since creation or reconfiguration");
    if (!b->h_valid) {
        bool done = false;
        if (int r = batch_results_sparse(b, res, group_off, group_keys, sums, counts, max_total_groups, &done)) return r;
        if (done) return GOLAP_OK;
    }
    const uint32_t nq = b->n_queries();
    uint64_t off = 0, koff = 0;
    group_off[0] = 0;
    for (uint32_t q = 0; q < nq; ++q) {
        const uint64_t room = off < max_total_groups ? max_total_groups - off : 0;
        const size_t na = b->kind == BatchKind::STAR ? b->qh[q].card.size() : b->sh[q].card.size();
        const int r = batch_result(b, q, &res[q], group_keys ? group_keys + koff : nullptr, sums ? sums + off : nullptr,
                                   counts ? counts + off : nullptr, room);
        if (r && r != GOLAP_ECAPACITY) return fail(ctx, r, b->err);
        off += res[q].n_groups;
        koff += res[q].n_groups * na;
        group_off[q + 1] = off;
        if (off > max_total_groups) {  // offsets of the rest still reported
            group_keys = nullptr;
            sums = nullptr;
            counts = nullptr;
        }
    }
    if (off > max_total_groups) return fail(ctx, GOLAP_ECAPACITY, "batch_results: more groups than max_total_groups");
    return GOLAP_OK;
}

int golap_batch_free(golap_batch *b) {
    if (!b) return GOLAP_EINVAL;
    golap_ctx *ctx = b->ctx;
    Lock l(ctx->mu);
    if (int r = set_device(ctx)) return r;
    batch_free(b);
    return GOLAP_OK;
}

}  // extern "C"

namespace golap_impl {

int solo_worklist(golap_ctx *ctx, const void *desc, uint32_t desc_size, uint32_t parts, const void **d_desc,
                  uint32_t **queue, uint2 **items) {
    if (desc_size > 1024) return fail(ctx, GOLAP_EINVAL, "solo work list: descriptor too large");
    const SliceDev &s = *static_cast<const SliceDev *>(desc);
    if (!parts || parts > 256) return fail(ctx, GOLAP_EINVAL, "solo work list: parts per chunk outside 1..256");
    const uint64_t nch = s.chunk_end - s.chunk_begin, n_items = nch * parts;
    if (n_items > ctx->solo_items_cap) {
        dev_release(ctx, ctx->d_solo_items);
        ctx->d_solo_items = nullptr;
        ctx->solo_items_cap = 0;
        if (int r = dev_alloc(ctx, std::max<uint64_t>(n_items, 1024) * sizeof(uint2), &ctx->d_solo_items)) return r;
        ctx->solo_items_cap = std::max<uint64_t>(n_items, 1024);
    }
    uint8_t *dd = ctx->d_scratch + 1024;
    uint64_t *off = reinterpret_cast<uint64_t *>(ctx->d_scratch + 2048);
    uint32_t *q = reinterpret_cast<uint32_t *>(ctx->d_scratch + 2304);
    if (!ctx->h_stage && hipHostMalloc(reinterpret_cast<void **>(&ctx->h_stage), 4096, hipHostMallocDefault) != hipSuccess) {
        ctx->h_stage = nullptr;
        return fail(ctx, GOLAP_ENOMEM, "solo work list: pinned staging");
    }
    // the staging page is reused by the next call only after this call's final synchronisation
    GOLAP_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    memcpy(ctx->h_stage, desc, desc_size);
    const uint64_t host_off[2] = {0, nch};
    memcpy(ctx->h_stage + 1024, host_off, sizeof(host_off));
    GOLAP_HIP_CHECK(ctx, hipMemcpyAsync(dd, ctx->h_stage, desc_size, hipMemcpyHostToDevice, ctx->stream));
    GOLAP_HIP_CHECK(ctx, hipMemcpyAsync(off, ctx->h_stage + 1024, sizeof(host_off), hipMemcpyHostToDevice, ctx->stream));
    GOLAP_HIP_CHECK(ctx, launch_fill_zero(q, 8, ctx->stream));
    uint2 *it = static_cast<uint2 *>(ctx->d_solo_items);
    GOLAP_HIP_CHECK(ctx, launch_batch_items(dd, desc_size, 1, off, nch, parts, it, q, ctx->stream));
    *d_desc = dd;
    *queue = q;
    *items = it;
    return GOLAP_OK;
}

}  // namespace golap_impl
