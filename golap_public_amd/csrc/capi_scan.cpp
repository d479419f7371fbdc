// C ABI: fused scan + filter + group-by plans over typed columns (include/golap_hip.h). Host side
// of the trip-table plans: validate the plan, compute the key strides and division constants,
// place LDS windows for coded columns, and launch the persistent scan kernel.
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "common.hpp"
#include "format.hpp"
#include "star_host.hpp"

using namespace golap_impl;
using namespace golap_impl::starhost;
using Lock = std::lock_guard<std::recursive_mutex>;

namespace golap_impl {
namespace starhost {

int prepare_scan(golap_ctx *ctx, const golap_scan_plan *p, ScanQueryDev &qd, ScanHost &h) {
    if (!p || p->n_cols < 1 || p->n_cols > SCAN_MAX_COLS || p->n_preds > SCAN_MAX_PREDS ||
        p->n_keys > SCAN_MAX_KEYS || p->agg_mode > GOLAP_AGG_GLOBAL || p->zone_prune > 1)
        return fail(ctx, GOLAP_EINVAL, "scan: bad plan");
    const golap_column *f = p->cols[0];
    for (uint32_t c = 0; c < p->n_cols; ++c) {
        const golap_column *col = p->cols[c];
        if (!col || col->ctx != ctx || !is_int_type(col->type))
            return fail(ctx, GOLAP_EINVAL, "scan: missing or non-integer column");
        if (col->n_rows != f->n_rows || col->page_rows != f->page_rows || col->chunk_rows != f->chunk_rows)
            return fail(ctx, GOLAP_EINVAL, "scan: columns are not row-aligned");
    }
    if (f->page_rows > 8192) return fail(ctx, GOLAP_EINVAL, "scan: page_rows above 8192");
    if (p->chunk_begin > p->chunk_end || p->chunk_end > f->n_chunks)
        return fail(ctx, GOLAP_EINVAL, "scan: bad chunk range");
    if (p->value_col >= (int32_t)p->n_cols || p->value_col < -1) return fail(ctx, GOLAP_EINVAL, "scan: bad value column");
    h = ScanHost{};
    h.plan = *p;
    memset(&qd, 0, sizeof(qd));
    qd.n_preds = p->n_preds;
    for (uint32_t i = 0; i < p->n_preds; ++i) {
        if (p->preds[i].col >= p->n_cols) return fail(ctx, GOLAP_EINVAL, "scan: bad predicate column");
        const golap_column *pc = p->cols[p->preds[i].col];
        qd.pred[i] = ScanPredDev{p->preds[i].col, 0, p->preds[i].lo, p->preds[i].hi, pc->d_chunk_min, pc->d_chunk_max};
        qd.colmask |= 1u << p->preds[i].col;
    }
    qd.n_keys = p->n_keys;
    for (uint32_t i = 0; i < p->n_keys; ++i) {
        const golap_scan_key &k = p->keys[i];
        if (k.col >= p->n_cols || !k.div || !k.card) return fail(ctx, GOLAP_EINVAL, "scan: bad key");
        if ((uint64_t)k.card * k.div >= (1ull << 32)) return fail(ctx, GOLAP_EINVAL, "scan: key card * div >= 2^32");
        if (k.base > INT64_MAX - (int64_t)((uint64_t)k.card * k.div - 1))
            return fail(ctx, GOLAP_EINVAL, "scan: key range beyond int64");
        h.domain *= k.card;
        if (h.domain >= MAX_GROUPS) return fail(ctx, GOLAP_EINVAL, "scan: group domain too large");
        h.lo.push_back(k.out_lo);
        h.card.push_back(k.card);
        qd.key[i].col = k.col;
        qd.key[i].card = k.card;
        qd.key[i].base = k.base;
        qd.key[i].span = (uint64_t)k.card * k.div;
        qd.key[i].div = golap_fmt::udiv32_magic(k.div);
        qd.key[i].zmin = p->cols[k.col]->d_chunk_min;
        qd.key[i].zmax = p->cols[k.col]->d_chunk_max;
        qd.colmask |= 1u << k.col;
    }
    h.stride.resize(p->n_keys);
    for (int i = (int)p->n_keys - 1, s = 1; i >= 0; --i) {
        h.stride[i] = (uint64_t)s;
        qd.key[i].stride = (uint64_t)s;
        s *= (int)p->keys[i].card;
    }
    qd.val_col = p->value_col;
    if (p->value_col >= 0) qd.colmask |= 1u << p->value_col;
    const uint64_t ppc = f->chunk_rows / f->page_rows;
    qd.s.chunk_begin = p->chunk_begin;
    qd.s.chunk_end = p->chunk_end;
    qd.s.page_begin = p->chunk_begin * ppc;
    qd.s.page_end = std::min<uint64_t>(f->n_pages, p->chunk_end * ppc);
    qd.s.chunk_mask = p->dev_chunk_mask;
    if (p->zone_prune)
        for (uint32_t z = 0; z < std::min<uint32_t>(p->n_preds, SLICE_MAX_ZONES); ++z) {
            const golap_column *zc = p->cols[p->preds[z].col];
            qd.s.zone[z] = ZonePred{zc->d_chunk_min, zc->d_chunk_max, p->preds[z].lo, p->preds[z].hi};
        }
    qd.n_groups = h.domain;
    h.colmask = qd.colmask;
    return GOLAP_OK;
}

// Host evaluation of which chunks of the slice survive (the device work-list test).
std::vector<uint8_t> scan_survivors(const golap_scan_plan &p, const std::vector<uint16_t> &mask) {
    std::vector<uint8_t> run(p.chunk_end - p.chunk_begin);
    for (uint64_t k = 0; k < run.size(); ++k) {
        const uint64_t c = p.chunk_begin + k;
        bool pass = mask.empty() || mask[k];
        if (p.zone_prune)
            for (uint32_t z = 0; z < std::min<uint32_t>(p.n_preds, SLICE_MAX_ZONES) && pass; ++z) {
                const golap_column *zc = p.cols[p.preds[z].col];
                pass = p.preds[z].lo <= zc->chunk_max[c] && p.preds[z].hi >= zc->chunk_min[c];
            }
        run[k] = pass;
    }
    return run;
}

void scan_stats(const golap_scan_plan &p, uint32_t colmask, const std::vector<uint8_t> &run, golap_star_result *res) {
    // per-chunk payload prefixes of the columns read, looked up once
    std::vector<const uint64_t *> pf;
    for (uint32_t c = 0; c < p.n_cols; ++c)
        if ((colmask >> c) & 1u) pf.push_back(p.cols[c]->extra.chunk_payload.data());
    for (uint64_t k = 0; k < run.size(); ++k) {
        const uint64_t ch = p.chunk_begin + k;
        if (!run[k]) {
            res->chunks_pruned++;
            continue;
        }
        res->chunks_run++;
        for (const uint64_t *f : pf) res->comp_bytes += f[ch + 1] - f[ch];
    }
}

// Register slots and LDS windows of the columns (cols[c], read when bit c of colmask is set)
// over pages [pb, pe): the first SCAN_DBP_SLOTS columns with delta-bitpack pages decode them in
// registers; a column gets a window (16-byte aligned, page_rows * elem bytes) when it has other
// coded pages, or delta pages without a slot. Returns the window bytes.
uint32_t scan_layout(const golap_column *const *cols, uint32_t n_cols, uint32_t colmask, uint64_t pb, uint64_t pe,
                     int32_t *woff, int32_t *dslot, uint32_t *wcap, int32_t *tok_off, uint32_t *tok_cap) {
    uint32_t bytes = 0, slots = 0, tcap = 0;
    for (uint32_t c = 0; c < n_cols; ++c) {
        woff[c] = -1;
        dslot[c] = -1;
        wcap[c] = 0;
        if (!((colmask >> c) & 1u)) continue;
        auto &ex = cols[c]->extra;
        const uint32_t dbp = ex.dbp_prefix[pe] - ex.dbp_prefix[pb];
        uint32_t other = ex.window_prefix[pe] - ex.window_prefix[pb] - dbp;
        if (dbp) {
            if (slots < SCAN_DBP_SLOTS) dslot[c] = (int32_t)slots++;
            else other += dbp;
        }
        if (!other) continue;
        woff[c] = (int32_t)bytes;
        const uint32_t pbytes = cols[c]->page_rows * cols[c]->elem;
        if (ex.has_lz4i && pbytes <= 65536) {  // room to stage the page and decode it in parallel
            uint32_t tc = 0;
            lz4i_window(pbytes, ex.lz4i_max_nseq, wcap[c], tc);
            tcap = std::max(tcap, tc);
            bytes += wcap[c];
        } else {
            bytes += (pbytes + 15) & ~15u;
        }
    }
    *tok_off = tcap ? (int32_t)bytes : -1;
    *tok_cap = tcap;
    return bytes + (tcap ? lz4i_scratch_host(tcap) : 0);
}

uint32_t scan_decode_mask(const golap_column *const *cols, uint32_t n_cols, uint32_t colmask, uint64_t pb, uint64_t pe) {
    uint32_t m = 0;
    for (uint32_t c = 0; c < n_cols; ++c) {
        if (!((colmask >> c) & 1u)) continue;
        const auto &ex = cols[c]->extra;
        const uint64_t pbytes = (uint64_t)cols[c]->page_rows * cols[c]->elem;
        const uint32_t coded = ex.window_prefix[pe] - ex.window_prefix[pb] - (ex.dbp_prefix[pe] - ex.dbp_prefix[pb]);
        if (ex.has_lz4i && pbytes < 65536 && coded) m |= 1u << c;
    }
    return m;
}

namespace {
uint64_t region_bytes(const golap_column *c, uint64_t un) { return (un * c->chunk_rows * c->elem + 255) & ~255ull; }
}  // namespace

uint64_t scan_decode_bytes(const golap_column *const *cols, uint32_t n_cols, uint32_t mask, uint64_t un) {
    uint64_t b = 0;
    for (uint32_t c = 0; c < n_cols; ++c)
        if ((mask >> c) & 1u) b += region_bytes(cols[c], un);
    return b;
}

void scan_decode_views(const golap_column *const *cols, uint32_t n_cols, uint32_t mask, uint64_t c0, uint64_t un,
                       uint8_t *buf, ScanColView *views) {
    uint64_t off = 0;
    for (uint32_t c = 0; c < n_cols; ++c) {
        if (!((mask >> c) & 1u)) continue;
        views[c].decoded = 1;
        views[c].c.comp = buf + off;
        views[c].dec_row0 = c0 * cols[c]->chunk_rows;
        views[c].woff = -1;
        views[c].dslot = -1;
        views[c].wcap = 0;
        off += region_bytes(cols[c], un);
    }
}

hipError_t scan_decode_enqueue(golap_ctx *ctx, const golap_column *const *cols, uint32_t n_cols, uint32_t mask,
                               uint64_t c0, uint64_t un, uint8_t *buf, uint64_t cb, uint64_t ce, hipStream_t st,
                               uint32_t *err) {
    if (cb < c0 || ce > c0 + un || cb >= ce) return cb == ce ? hipSuccess : hipErrorInvalidValue;
    uint64_t off = 0;
    for (uint32_t c = 0; c < n_cols; ++c) {
        if (!((mask >> c) & 1u)) continue;
        const golap_column *col = cols[c];
        const uint64_t ppc = col->chunk_rows / col->page_rows;
        const uint64_t pe = std::min<uint64_t>(col->n_pages, ce * ppc);
        if (cb * ppc < pe) {
            const hipError_t e = launch_decode_pages(ctx, col, cb * ppc, pe, buf + off, c0 * col->chunk_rows, st, err);
            if (e != hipSuccess) return e;
        }
        off += region_bytes(col, un);
    }
    return hipSuccess;
}

// Waves per workgroup that take pages (each needs its own windows): all of them when their
// windows fit beside agg_entries dense partial entries, else fewer.
uint32_t scan_page_waves(golap_ctx *ctx, uint32_t win_bytes, uint32_t agg_entries) {
    uint32_t w = 4;
    while (w > 1 && scan_lds_bytes(win_bytes, agg_entries, w) > scan_lds_budget(ctx)) w >>= 1;
    return w;
}

// Dense LDS aggregation entries of a scan launch: 0 for global aggregation, else the requested
// count or the default `dflt` (powers of two), halved until the launch fits the LDS budget.
uint32_t scan_agg_entries(golap_ctx *ctx, uint32_t agg_mode, uint32_t requested, uint32_t win_bytes, uint32_t page_waves,
                          uint32_t dflt) {
    if (agg_mode == GOLAP_AGG_GLOBAL) return 0;
    uint32_t s = std::min(pow2_at_least(requested ? requested : dflt), SCAN_AGG_MAX);
    while (s && scan_lds_bytes(win_bytes, s, page_waves) > scan_lds_budget(ctx)) s >>= 1;
    return s;
}

// Largest local key domain of plan p over its surviving chunks (run): the dense partial
// entries the kernel gives the plan on a work item, from the key columns' zone maps as the
// kernel computes them (0 when every chunk aggregates in global memory).
uint64_t scan_local_domain(const golap_scan_plan &p, const ScanQueryDev &qd, const std::vector<uint8_t> &run) {
    uint64_t best = 0;
    for (uint64_t k = 0; k < run.size(); ++k) {
        if (!run[k]) continue;
        const uint64_t c = p.chunk_begin + k;
        uint64_t n = 1;
        for (uint32_t i = 0; i < qd.n_keys && n; ++i) {
            const ScanKeyDev &K = qd.key[i];  // key i of p (its column index may be a batch's)
            const golap_column *col = p.cols[p.keys[i].col];
            const int64_t zmin = col->chunk_min[c], zmax = col->chunk_max[c];
            const uint64_t ulo = zmin <= K.base ? 0 : (uint64_t)zmin - (uint64_t)K.base;
            const uint64_t uhi = zmax < K.base ? 0 : (uint64_t)zmax - (uint64_t)K.base;
            if (zmax < K.base || ulo >= K.span) {
                n = 0;
                break;
            }
            const uint64_t div = K.span / K.card;
            const uint64_t klo = ulo / div, khi = std::min<uint64_t>(uhi, K.span - 1) / div;
            const bool inside = zmin >= K.base && uhi < K.span;
            if (!(inside && klo == khi)) n *= khi - klo + 1;
        }
        best = std::max(best, n);
    }
    return best;
}

// Default dense entries for plans whose largest local domain is maxl: at least 1024 (four
// workgroups per CU fit beside small windows), else enough for one such plan.
uint32_t scan_default_entries(uint64_t maxl) {
    return std::min<uint32_t>(SCAN_AGG_MAX, pow2_at_least(std::max<uint64_t>(maxl, 1024)));
}

// Pages per work item: whole chunks when there are enough surviving chunks to fill the device,
// else chunk parts, so a short slice still spreads over every CU.
uint32_t scan_pages_per_item(golap_ctx *ctx, uint64_t surviving_chunks, uint32_t ppc) {
    const uint64_t want = (uint64_t)ctx->cus * 4;
    if (!surviving_chunks || surviving_chunks >= want) return ppc;
    const uint64_t parts = std::min<uint64_t>(ppc, (want + surviving_chunks - 1) / surviving_chunks);
    uint32_t ppi = (uint32_t)((ppc + parts - 1) / parts);
    ppi = std::min(ppc, std::max(ppi, 4u));  // a page for each wave of the workgroup
    while ((ppc + ppi - 1) / ppi > 256) ++ppi;
    return std::max<uint32_t>(ppi, 1);
}

}  // namespace starhost

// One scan call. With fixed == nullptr the device memory beyond inputs and outputs comes from
// the context (its scratch word block, its work-item and group buffers, grown on demand, and a
// pinned staging page); with fixed set it all comes from fixed->base (fixed->bytes bytes, 256-byte
// aligned), laid out as: [0, 64) error word and queue counters, [256, 1280) the plan descriptor,
// then the work items (16 bytes each) and the dense group result (16 bytes per group), and
// nothing is allocated. Plans needing more than fixed->bytes fail with GOLAP_ECAPACITY.
int scan_run_impl(golap_ctx *ctx, const golap_scan_plan *plan, golap_star_result *res, int32_t *group_keys,
                  int64_t *sums, uint64_t *counts, uint64_t max_groups, const ScanFixedScratch *fixed) {
    if (!ctx || !plan || !res) return GOLAP_EINVAL;
    Lock l(ctx->mu);
    if (int r = set_device(ctx)) return r;
    memset(res, 0, sizeof(*res));
    ScanQueryDev qd;
    ScanHost h;
    if (int r = prepare_scan(ctx, plan, qd, h)) return r;
    static_assert(sizeof(ScanQueryDev) <= 1024, "scan descriptor staged in 1 KiB");
    const golap_column *f = plan->cols[0];
    const uint64_t nch = plan->chunk_end - plan->chunk_begin;
    // surviving chunks on the host (the device work list applies the same test): item grain
    std::vector<uint16_t> mask;
    if (plan->dev_chunk_mask && nch) {
        mask.resize(nch);
        GOLAP_HIP_CHECK(ctx, hipMemcpyAsync(mask.data(), plan->dev_chunk_mask + plan->chunk_begin, nch * 2,
                                            hipMemcpyDeviceToHost, ctx->stream));
        GOLAP_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    }
    const std::vector<uint8_t> run = scan_survivors(*plan, mask);
    uint64_t surv = 0;
    for (uint8_t v : run) surv += v;
    ScanLaunch L{};
    L.n_cols = plan->n_cols;
    int32_t woff[SCAN_MAX_COLS], dslot[SCAN_MAX_COLS];
    uint32_t wcap[SCAN_MAX_COLS];
    // one plan: LZ4I pages decode inside the scan, each wave into its window, overlapped with
    // the other waves' work (decoding ahead measured slower here; batches decode ahead)
    L.win_bytes = scan_layout(plan->cols, plan->n_cols, qd.colmask, qd.s.page_begin, qd.s.page_end, woff, dslot, wcap,
                              &L.tok_off, &L.tok_cap);
    for (uint32_t c = 0; c < plan->n_cols; ++c)
        L.col[c] = ScanColView{view(plan->cols[c]), plan->cols[c]->type, woff[c], dslot[c], wcap[c]};
    L.page_rows = f->page_rows;
    L.pages_per_chunk = (uint32_t)(f->chunk_rows / f->page_rows);
    L.n_rows = f->n_rows;
    L.n_pages = f->n_pages;
    L.n_queries = 1;
    L.page_waves = scan_page_waves(ctx, L.win_bytes, 1024);
    L.agg_entries = scan_agg_entries(ctx, plan->agg_mode, plan->lds_group_slots, L.win_bytes, L.page_waves,
                                     scan_default_entries(scan_local_domain(*plan, qd, run)));
    if (scan_lds_bytes(L.win_bytes, L.agg_entries, L.page_waves) > scan_lds_budget(ctx))
        return fail(ctx, GOLAP_EINVAL, "scan: LDS budget exceeded (page_rows or coded columns)");
    qd.use_lds = L.agg_entries != 0;
    L.pages_per_item = scan_pages_per_item(ctx, surv, L.pages_per_chunk);
    L.rows_per_lane = scan_rows_per_lane(ctx, surv * L.pages_per_chunk,
                                         scan_lds_bytes(L.win_bytes, L.agg_entries, L.page_waves));
    const uint32_t parts = (L.pages_per_chunk + L.pages_per_item - 1) / L.pages_per_item;
    const uint64_t n_items = std::max<uint64_t>(nch * parts, 1);
    int64_t *d_sum;
    uint64_t *d_cnt;
    uint32_t *err, *queue;
    ScanQueryDev *dq;
    uint4 *items;
    if (fixed) {
        const uint64_t need = SCAN_FIXED_ITEMS + n_items * sizeof(uint4) + h.domain * 16;
        if (need > fixed->bytes)
            return fail(ctx, GOLAP_ECAPACITY, "scan_run_noalloc: the plan needs " + std::to_string(need) +
                                                  " bytes of scratch (" + std::to_string(n_items) + " work items, " +
                                                  std::to_string(h.domain) + " groups), the maximum is " +
                                                  std::to_string(fixed->bytes));
        uint8_t *b = fixed->base;
        err = reinterpret_cast<uint32_t *>(b);
        queue = reinterpret_cast<uint32_t *>(b + 8);
        dq = reinterpret_cast<ScanQueryDev *>(b + SCAN_FIXED_DESC);
        items = reinterpret_cast<uint4 *>(b + SCAN_FIXED_ITEMS);
        d_sum = reinterpret_cast<int64_t *>(b + SCAN_FIXED_ITEMS + n_items * sizeof(uint4));
        d_cnt = reinterpret_cast<uint64_t *>(d_sum + h.domain);
        // the array keeps the last call's contents: the error word, the counters and the groups
        // start at zero; the descriptor goes in as a kernel argument (no staging memory)
        GOLAP_HIP_CHECK(ctx, launch_fill_zero(b, 64, ctx->stream));
        GOLAP_HIP_CHECK(ctx, launch_fill_zero(d_sum, h.domain * 16, ctx->stream));
        qd.g_sum = d_sum;
        qd.g_cnt = d_cnt;
        GOLAP_HIP_CHECK(ctx, launch_put_words(dq, &qd, sizeof(qd), ctx->stream));
    } else {
        if (int r = group_buffers(ctx, h.domain, &d_sum, &d_cnt)) return r;
        qd.g_sum = d_sum;
        qd.g_cnt = d_cnt;
        err = reinterpret_cast<uint32_t *>(ctx->d_scratch);
        // work list: the descriptor and the queue counters in the context scratch, the items in a
        // context buffer
        if (n_items > ctx->scan_items_cap) {
            dev_release(ctx, ctx->d_scan_items);
            ctx->d_scan_items = nullptr;
            ctx->scan_items_cap = 0;
            const uint64_t cap = std::max<uint64_t>(n_items, 4096);
            if (int r = dev_alloc(ctx, cap * sizeof(uint4), &ctx->d_scan_items)) return r;
            ctx->scan_items_cap = cap;
        }
        if (!ctx->h_stage &&
            hipHostMalloc(reinterpret_cast<void **>(&ctx->h_stage), 4096, hipHostMallocDefault) != hipSuccess) {
            ctx->h_stage = nullptr;
            return fail(ctx, GOLAP_ENOMEM, "scan: pinned staging");
        }
        dq = reinterpret_cast<ScanQueryDev *>(ctx->d_scratch + 1024);
        queue = reinterpret_cast<uint32_t *>(ctx->d_scratch + 2304);
        items = static_cast<uint4 *>(ctx->d_scan_items);
        // the staging page is reused by the next call only after this call's final synchronisation
        GOLAP_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
        memcpy(ctx->h_stage, &qd, sizeof(qd));
        GOLAP_HIP_CHECK(ctx, hipMemcpyAsync(dq, ctx->h_stage, sizeof(qd), hipMemcpyHostToDevice, ctx->stream));
        GOLAP_HIP_CHECK(ctx, launch_fill_zero(queue, 8, ctx->stream));
    }
    L.err = err;
    L.queries = dq;
    L.queue = queue;
    L.items = items;
    GOLAP_HIP_CHECK(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    if (nch) {
        GOLAP_HIP_CHECK(ctx, launch_scan_items(dq, 1, plan->chunk_begin, nch, parts, items, queue, ctx->stream));
        GOLAP_HIP_CHECK(ctx, launch_scan(ctx, L, ctx->stream));
    }
    GOLAP_HIP_CHECK(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    uint32_t flags = 0;
    GOLAP_HIP_CHECK(ctx, hipMemcpyAsync(&flags, err, 4, hipMemcpyDeviceToHost, ctx->stream));
    GOLAP_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    if (flags) return error_flags_status(ctx, flags, "scan");
    float ms = 0;
    GOLAP_HIP_CHECK(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    res->kernel_ms = ms;
    res->lds_group_slots = L.agg_entries;
    scan_stats(*plan, qd.colmask, run, res);
    return emit_groups(ctx, h.domain, d_sum, d_cnt, h.lo, h.card, h.stride, &res->n_groups, &res->total_sum,
                       group_keys, sums, counts, max_groups, &res->rows_matched);
}

}  // namespace golap_impl

extern "C" {

int golap_scan_run(golap_ctx *ctx, const golap_scan_plan *plan, golap_star_result *res, int32_t *group_keys,
                   int64_t *sums, uint64_t *counts, uint64_t max_groups) {
    return scan_run_impl(ctx, plan, res, group_keys, sums, counts, max_groups, nullptr);
}

}  // extern "C"
