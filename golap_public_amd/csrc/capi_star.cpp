// C ABI: group-by, the fused star join + group-by plan and its batched form (include/golap_hip.h).
// Host side of query2 / query3 (usecases/ssb/queries/02.hpp:54-307, 03.hpp:79-677): choose the
// driving dimension and its LDS filter, the LDS windows for coded columns and the aggregation
// table, launch one persistent fused kernel, and return the canonical (sorted) group list.
#include <algorithm>
#include <cmath>
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

uint32_t pow2_at_least(uint64_t v) {
    uint32_t p = 1;
    while (p < v && p < (1u << 31)) p <<= 1;
    return p;
}

// Dense result buffer of the context (sum int64 then count uint64), zeroed for n groups.
int group_buffers(golap_ctx *ctx, uint64_t n, int64_t **sum, uint64_t **cnt) {
    if (n > ctx->groups_cap) {
        dev_release(ctx, ctx->d_groups);
        ctx->d_groups = nullptr;
        ctx->groups_cap = 0;
        if (int r = dev_alloc(ctx, n * 16, &ctx->d_groups)) return r;
        ctx->groups_cap = n;
    }
    *sum = static_cast<int64_t *>(ctx->d_groups);
    *cnt = reinterpret_cast<uint64_t *>(static_cast<uint8_t *>(ctx->d_groups) + n * 8);
    GOLAP_HIP_CHECK(ctx, hipMemsetAsync(ctx->d_groups, 0, n * 16, ctx->stream));
    return GOLAP_OK;
}

uint32_t choose_lds_slots(uint32_t agg_mode, uint32_t requested, uint64_t n_groups) {
    if (agg_mode == GOLAP_AGG_GLOBAL) return 0;
    if (requested) return std::min(pow2_at_least(requested), LDS_GROUP_SLOTS_MAX);
    // twice the domain when it fits (every group finds a slot), else the largest table: rows of
    // a workgroup often share few groups (sorted or selective data), the rest spills
    return (uint32_t)std::min<uint64_t>(pow2_at_least(std::max<uint64_t>(n_groups * 2, 64)), LDS_GROUP_SLOTS_MAX);
}

// Emits the groups with rows of a dense host result, in index (= key tuple) order.
// *rows (optional) = the sum of the counts = the rows that reached the aggregation.
int emit_groups_host(golap_ctx *ctx, uint64_t n, const int64_t *hs, const uint64_t *hc,
                     const std::vector<int32_t> &lo, const std::vector<uint32_t> &card,
                     const std::vector<uint64_t> &stride, uint64_t *n_groups, int64_t *total, int32_t *keys,
                     int64_t *sums, uint64_t *counts, uint64_t max_groups, uint64_t *rows) {
    (void)stride;  // index i = sum of digit[a] * stride[a], the last attribute varying fastest
    uint64_t g = 0, nr = 0;
    int64_t t = 0;
    const size_t na = card.size();
    // digits of index `at`: stepped like an odometer to the next group with rows, or recomputed
    // by division when that group lies far ahead (sparse results skip the empty stretches)
    std::vector<uint32_t> digit(na, 0);
    uint64_t at = 0;
    auto move_to = [&](uint64_t idx) {
        if (idx - at > 64) {
            uint64_t rem = idx;
            for (size_t a = na; a-- > 0;) {
                digit[a] = (uint32_t)(rem % card[a]);
                rem /= card[a];
            }
            at = idx;
            return;
        }
        for (; at < idx; ++at)
            for (size_t a = na; a-- > 0;) {  // increment the last digit, carry left
                if (++digit[a] < card[a]) break;
                digit[a] = 0;
            }
    };
    for (uint64_t i = 0; i < n; ++i) {
        if (!hc[i]) continue;
        t += hs[i];
        nr += hc[i];
        if (g < max_groups) {
            if (keys) {
                move_to(i);
                for (size_t a = 0; a < na; ++a) keys[g * na + a] = lo[a] + (int32_t)digit[a];
            }
            if (sums) sums[g] = hs[i];
            if (counts) counts[g] = hc[i];
        }
        ++g;
    }
    *n_groups = g;
    if (total) *total = t;
    if (rows) *rows = nr;
    if (g > max_groups) return fail(ctx, GOLAP_ECAPACITY, "more groups than max_groups");
    return GOLAP_OK;
}

// Copies a dense device result back and emits it.
int emit_groups(golap_ctx *ctx, uint64_t n, const int64_t *d_sum, const uint64_t *d_cnt,
                const std::vector<int32_t> &lo, const std::vector<uint32_t> &card,
                const std::vector<uint64_t> &stride, uint64_t *n_groups, int64_t *total,
                int32_t *keys, int64_t *sums, uint64_t *counts, uint64_t max_groups, uint64_t *rows) {
    std::vector<int64_t> hs(n);
    std::vector<uint64_t> hc(n);
    GOLAP_HIP_CHECK(ctx, hipMemcpyAsync(hs.data(), d_sum, n * 8, hipMemcpyDeviceToHost, ctx->stream));
    GOLAP_HIP_CHECK(ctx, hipMemcpyAsync(hc.data(), d_cnt, n * 8, hipMemcpyDeviceToHost, ctx->stream));
    GOLAP_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return emit_groups_host(ctx, n, hs.data(), hc.data(), lo, card, stride, n_groups, total, keys, sums, counts,
                            max_groups, rows);
}

bool is_int_type(uint32_t t) { return t >= GOLAP_I8 && t <= GOLAP_I64; }

// Smallest b with 2^b >= n.
uint32_t bits_for(uint64_t n) {
    uint32_t b = 0;
    while (b < 63 && (1ull << b) < n) ++b;
    return b;
}



// Blocked Bloom filter of a map's keys, cached until the map is rebuilt.
int ensure_hfilter(golap_ctx *ctx, const golap_hashmap *dm) {
    const uint32_t words = 1u << HASH_FILTER_WORDS_LOG2;
    if (!dm->d_hfilter) {
        void *p = nullptr;
        if (int r = dev_alloc(ctx, words * 4, &p)) return r;
        dm->d_hfilter = static_cast<uint32_t *>(p);
        dm->hfilter_build = 0;
    }
    if (dm->hfilter_build != dm->build_id) {
        GOLAP_HIP_CHECK(ctx, hipMemsetAsync(dm->d_hfilter, 0, words * 4, ctx->stream));
        GOLAP_HIP_CHECK(ctx, launch_hash_filter(ctx, dm->small_log2 ? dm->d_small : dm->d_slots,
                                                dm->small_log2 ? dm->small_log2 : dm->log2cap, dm->d_hfilter,
                                                HASH_FILTER_WORDS_LOG2));
        dm->hfilter_build = dm->build_id;
    }
    return GOLAP_OK;
}

int filter_view(golap_ctx *ctx, const golap_hashmap *dm, DriverFilterView *out, uint32_t *words) {
    if (dm->use_bitmap && (dm->krange + 31) / 32 <= (1u << HASH_FILTER_WORDS_LOG2)) {
        *out = DriverFilterView{dm->d_bitmap, dm->kmin, dm->krange, 0};
        *words = (dm->krange + 31) / 32;
        return GOLAP_OK;
    }
    if (int r = ensure_hfilter(ctx, dm)) return r;
    *out = DriverFilterView{dm->d_hfilter, 0, 0, HASH_FILTER_WORDS_LOG2};
    *words = 1u << HASH_FILTER_WORDS_LOG2;
    return GOLAP_OK;
}

// Validates a plan and fills its device descriptor (all but use_lds, item_begin and the result
// pointers): fact column indexes, driver, probe tables, driver filter, slice and group strides.
int prepare_query(golap_ctx *ctx, const golap_star_plan *plan, bool compact, ColumnSet &cs, StarQueryDev &qd,
                  QueryHost &qh, int host_dim) {
    if (!plan || plan->n_dims < 1 || plan->n_dims > STAR_MAX_DIMS || !plan->value || plan->agg_mode > GOLAP_AGG_GLOBAL)
        return fail(ctx, GOLAP_EINVAL, "star: bad plan");
    const uint32_t nd = plan->n_dims;
    const golap_column *v = plan->value;
    auto is32 = [](const golap_column *c) { return c->type == GOLAP_I32 || c->type == GOLAP_U32; };
    if (v->ctx != ctx || !is32(v)) return fail(ctx, GOLAP_EINVAL, "star: value column must be 32-bit");
    for (uint32_t d = 0; d < nd; ++d) {
        const golap_star_dim &sd = plan->dims[d];
        const bool on_host = (int)d == host_dim;  // joined on the host: no device map
        if (!sd.fk || sd.fk->ctx != ctx || (!on_host && (!sd.map || sd.map->ctx != ctx || !sd.map->built)))
            return fail(ctx, GOLAP_EINVAL, "star: missing column or unbuilt map");
        if (!is32(sd.fk)) return fail(ctx, GOLAP_EINVAL, "star: foreign keys must be 32-bit");
        if (sd.fk->n_rows != v->n_rows || sd.fk->page_rows != v->page_rows || sd.fk->chunk_rows != v->chunk_rows)
            return fail(ctx, GOLAP_EINVAL, "star: columns are not row-aligned");
        if (sd.fk->type == GOLAP_U32)
            for (int64_t mx : sd.fk->chunk_max)
                if (mx > INT32_MAX) return fail(ctx, GOLAP_EINVAL, "star: uint32 keys above INT32_MAX");
    }
    if (v->page_rows != 1024 && v->page_rows != 2048 && v->page_rows != 4096 && v->page_rows != 8192)
        return fail(ctx, GOLAP_EINVAL, "star: page_rows must be 1024, 2048, 4096 or 8192");
    if (plan->chunk_begin > plan->chunk_end || plan->chunk_end > v->n_chunks)
        return fail(ctx, GOLAP_EINVAL, "star: bad chunk range");
    qh = QueryHost{};
    qh.plan = *plan;
    memset(&qd, 0, sizeof(qd));
    // group domain, attributes in dimension order (first = most significant)
    for (uint32_t d = 0; d < nd; ++d)
        if (plan->dims[d].group_card) {
            qh.lo.push_back(plan->dims[d].group_lo);
            qh.card.push_back(plan->dims[d].group_card);
            qh.domain *= plan->dims[d].group_card;
            if (qh.domain >= MAX_GROUPS) return fail(ctx, GOLAP_EINVAL, "star: group domain too large");
        }
    qh.stride.resize(qh.card.size());
    for (int a = (int)qh.card.size() - 1, s = 1; a >= 0; --a) {
        qh.stride[a] = (uint64_t)s;
        s *= (int)qh.card[a];
    }
    // driver: the dimension whose filter passes the smallest share of its key domain
    double best = 2.0;
    for (uint32_t d = 0; d < nd; ++d) {
        if ((int)d == host_dim) continue;
        const golap_hashmap *m = plan->dims[d].map;
        const double sel = m->ht_size ? (double)m->filled / (double)m->ht_size : 1.0;
        qh.sel *= sel;
        qh.builds.emplace_back(m, m->build_id);
        if (sel < best) {
            best = sel;
            qd.driver = d;
        }
    }
    size_t a_idx = 0;
    for (uint32_t d = 0; d < nd; ++d) {
        const golap_star_dim &sd = plan->dims[d];
        const golap_hashmap *m = sd.map;
        if (int r = cs.index(ctx, sd.fk, &qd.fk_col[d])) return r;
        if ((int)d != host_dim) {
            const bool small = compact && m->small_log2;
            qd.dim[d].slots = small ? m->d_small : m->d_slots;
            qd.dim[d].log2cap = small ? m->small_log2 : m->log2cap;
        }
        qd.dim[d].group_lo = sd.group_lo;
        qd.dim[d].group_card = sd.group_card;
        qd.dim[d].group_stride = sd.group_card ? qh.stride[a_idx++] : 0;
    }
    if (int r = cs.index(ctx, v, &qd.val_col)) return r;
    qd.n_dims = (uint8_t)nd;
    qd.val_signed = v->type == GOLAP_I32;
    if (const golap_column *vm = plan->value_minus) {
        if (vm->ctx != ctx || !is32(vm)) return fail(ctx, GOLAP_EINVAL, "star: value_minus column must be 32-bit");
        if (vm->n_rows != v->n_rows || vm->page_rows != v->page_rows || vm->chunk_rows != v->chunk_rows)
            return fail(ctx, GOLAP_EINVAL, "star: columns are not row-aligned");
        if (int r = cs.index(ctx, vm, &qd.val2_col)) return r;
        qd.has_val2 = 1;
        qd.val2_signed = vm->type == GOLAP_I32;
    }
    // driver filter: dense exact bitmap, or the blocked Bloom filter of the map's keys
    if (int r = filter_view(ctx, plan->dims[qd.driver].map, &qd.filter, &qd.filter_words)) return r;
    const uint64_t ppc = v->chunk_rows / v->page_rows;
    qd.s.chunk_begin = plan->chunk_begin;
    qd.s.chunk_end = plan->chunk_end;
    qd.s.page_begin = plan->chunk_begin * ppc;
    qd.s.page_end = std::min<uint64_t>(v->n_pages, plan->chunk_end * ppc);
    qd.s.chunk_mask = plan->dev_chunk_mask;
    qd.n_groups = qh.domain;
    return GOLAP_OK;
}

// LDS windows: a column gets one when some page of [pb, pe) is neither read in place nor
// point-readable.
uint32_t assign_windows(const ColumnSet &cs, uint64_t pb, uint64_t pe, int *win) {
    uint32_t n = 0;
    for (size_t c = 0; c < cs.cols.size(); ++c) {
        auto &ex = cs.cols[c]->extra;
        win[c] = ex.window_prefix[pe] != ex.window_prefix[pb] ? (int)n++ : -1;
    }
    return n;
}

// Windows and prefix-sum regions of a launch's columns over pages [pb, pe). A column gets a
// window when some page there is neither read in place, point-readable nor delta-bitpacked, or
// when it has delta-bitpack pages and is a driver's column (bit c of driver_cols: every row of it
// is tested, so a full decode is cheaper); other delta-bitpack pages are point-read through
// per-tile prefix sums (anchor[c] = 1 + region). Returns the windows.
uint32_t assign_access(const ColumnSet &cs, uint64_t pb, uint64_t pe, uint32_t driver_cols, int *win, int *anchor,
                       uint32_t *n_anchors) {
    uint32_t n = 0, na = 0;
    for (size_t c = 0; c < cs.cols.size(); ++c) {
        auto &ex = cs.cols[c]->extra;
        const uint32_t coded = ex.window_prefix[pe] - ex.window_prefix[pb];
        const uint32_t dbp = ex.dbp_prefix[pe] - ex.dbp_prefix[pb];
        const bool needs_window = coded > dbp || (dbp && ((driver_cols >> c) & 1u));
        win[c] = needs_window ? (int)n++ : -1;
        anchor[c] = !needs_window && dbp ? (int)++na : 0;
    }
    *n_anchors = na;
    return n;
}

// chunks_run / chunks_pruned / comp_bytes of a plan over its slice, given the chunk flags
// (empty = all run).
void slice_stats(const golap_star_plan &plan, const std::vector<uint8_t> &run, golap_star_result *res) {
    // per-chunk payload prefixes of the plan's columns, looked up once
    const uint64_t *pf[STAR_MAX_DIMS + 2];
    uint32_t n = 0;
    for (uint32_t d = 0; d < plan.n_dims; ++d) pf[n++] = plan.dims[d].fk->extra.chunk_payload.data();
    pf[n++] = plan.value->extra.chunk_payload.data();
    if (plan.value_minus) pf[n++] = plan.value_minus->extra.chunk_payload.data();
    for (uint64_t ch = plan.chunk_begin; ch < plan.chunk_end; ++ch) {
        if (!run.empty() && !run[ch - plan.chunk_begin]) {
            res->chunks_pruned++;
            continue;
        }
        res->chunks_run++;
        for (uint32_t c = 0; c < n; ++c) res->comp_bytes += pf[c][ch + 1] - pf[c][ch];
    }
}

int status_of_flags(golap_ctx *ctx, uint32_t flags, const char *what) {
    if (flags & DEVERR_CORRUPT) return fail(ctx, GOLAP_ECORRUPT, std::string(what) + ": corrupt page");
    if (flags & DEVERR_CAPACITY) return fail(ctx, GOLAP_ECAPACITY, std::string(what) + ": table full");
    if (flags) return fail(ctx, GOLAP_EINVAL, std::string(what) + ": value outside its declared range");
    return GOLAP_OK;
}

int ColumnSet::index(golap_ctx *ctx, const golap_column *c, uint8_t *idx) {
    for (size_t i = 0; i < cols.size(); ++i)
        if (cols[i] == c) {
            *idx = (uint8_t)i;
            return GOLAP_OK;
        }
    if (!cols.empty()) {
        const golap_column *f = cols[0];
        if (c->n_rows != f->n_rows || c->page_rows != f->page_rows || c->chunk_rows != f->chunk_rows)
            return fail(ctx, GOLAP_EINVAL, "star: columns are not row-aligned");
    }
    if (cols.size() >= STAR_MAX_COLS) return fail(ctx, GOLAP_EINVAL, "star: more than 8 distinct fact columns");
    *idx = (uint8_t)cols.size();
    cols.push_back(c);
    return GOLAP_OK;
}

}  // namespace starhost
}  // namespace golap_impl

extern "C" {

int golap_group_by(golap_ctx *ctx, const golap_group_key *keys, uint32_t n_keys, const void *dev_val,
                   uint32_t val_type, uint64_t n_rows, uint32_t agg_mode, uint64_t *n_groups,
                   int32_t *group_keys, int64_t *sums, uint64_t *counts, uint64_t max_groups) {
    if (!ctx || !keys || n_keys < 1 || n_keys > 4 || !n_groups || (n_rows && !dev_val) ||
        !is_int_type(val_type) || agg_mode > GOLAP_AGG_GLOBAL)
        return GOLAP_EINVAL;
    Lock l(ctx->mu);
    if (int r = set_device(ctx)) return r;
    std::vector<int32_t> lo(n_keys);
    std::vector<uint32_t> card(n_keys);
    std::vector<uint64_t> stride(n_keys);
    uint64_t domain = 1;
    for (uint32_t k = 0; k < n_keys; ++k) {
        if ((n_rows && !keys[k].dev_col) || !is_int_type(keys[k].type) || !keys[k].card)
            return fail(ctx, GOLAP_EINVAL, "group_by: bad key");
        lo[k] = keys[k].lo;
        card[k] = keys[k].card;
        domain *= keys[k].card;
        if (domain >= MAX_GROUPS) return fail(ctx, GOLAP_EINVAL, "group_by: group domain too large");
    }
    for (int k = (int)n_keys - 1, s = 1; k >= 0; --k) {
        stride[k] = (uint64_t)s;
        s *= (int)card[k];
    }
    int64_t *d_sum;
    uint64_t *d_cnt;
    if (int r = group_buffers(ctx, domain, &d_sum, &d_cnt)) return r;
    GroupByLaunch g{};
    for (uint32_t k = 0; k < n_keys; ++k) {
        g.key[k] = keys[k].dev_col;
        g.key_type[k] = keys[k].type;
        g.key_lo[k] = lo[k];
        g.key_card[k] = card[k];
        g.key_stride[k] = stride[k];
    }
    g.n_keys = n_keys;
    g.val = dev_val;
    g.val_type = val_type;
    g.n = n_rows;
    g.lds_gslots = choose_lds_slots(agg_mode, 0, domain);
    g.g_sum = d_sum;
    g.g_cnt = d_cnt;
    g.err = reinterpret_cast<uint32_t *>(ctx->d_scratch);
    if (n_rows) GOLAP_HIP_CHECK(ctx, launch_group_by(ctx, g));
    if (int r = check_device_errors(ctx, "group_by")) return r;
    return emit_groups(ctx, domain, d_sum, d_cnt, lo, card, stride, n_groups, nullptr, group_keys, sums,
                       counts, max_groups);
}

int golap_star_run(golap_ctx *ctx, const golap_star_plan *plan, golap_star_result *res,
                   int32_t *group_keys, int64_t *sums, uint64_t *counts, uint64_t max_groups) {
    if (!ctx || !plan || !res) return GOLAP_EINVAL;
    Lock l(ctx->mu);
    if (int r = set_device(ctx)) return r;
    memset(res, 0, sizeof(*res));
    ColumnSet cs;
    StarQueryDev qd;
    QueryHost qh;
    if (int r = prepare_query(ctx, plan, true, cs, qd, qh)) return r;
    const golap_column *v = plan->value;
    StarLaunch L{};
    L.n_cols = (uint32_t)cs.cols.size();
    for (uint32_t c = 0; c < L.n_cols; ++c) L.col[c] = view(cs.cols[c]);
    L.n_windows = assign_access(cs, qd.s.page_begin, qd.s.page_end, 1u << qd.fk_col[qd.driver], L.col_window,
                                L.col_anchor, &L.n_anchors);
    L.page_rows = v->page_rows;
    L.pages_per_chunk = (uint32_t)(v->chunk_rows / v->page_rows);
    L.n_rows = v->n_rows;
    L.n_pages = v->n_pages;
    L.n_queries = 1;
    L.n_items = qd.s.page_end - qd.s.page_begin;
    L.filter_words = qd.filter_words;
    // AUTO: LDS partial tables pay off when many rows land in each group
    const double expect = (double)L.n_items * L.page_rows * qh.sel;
    uint32_t agg = plan->agg_mode;
    if (agg == GOLAP_AGG_AUTO)
        agg = expect >= 16.0 * (double)qh.domain && expect >= 65536.0 ? GOLAP_AGG_LDS : GOLAP_AGG_GLOBAL;
    L.lds_gslots = choose_lds_slots(agg, plan->lds_group_slots, qh.domain);
    const uint32_t na = L.n_anchors;
    while (L.lds_gslots && star_lds_bytes(L.filter_words, L.page_rows, L.n_windows, L.lds_gslots, na) > (size_t)ctx->lds_max)
        L.lds_gslots >>= 1;
    if (star_lds_bytes(L.filter_words, L.page_rows, L.n_windows, L.lds_gslots, na) > (size_t)ctx->lds_max)
        return fail(ctx, GOLAP_EINVAL, "star: LDS budget exceeded (page_rows or coded columns)");
    // the driver filter stays in global memory (L2) when pinning it in LDS would cost
    // workgroups per CU: the kernel is latency-bound, and tiles in flight matter more
    if (star_blocks_per_cu(L.page_rows, star_lds_bytes(0, L.page_rows, L.n_windows, L.lds_gslots, na)) >
        star_blocks_per_cu(L.page_rows, star_lds_bytes(L.filter_words, L.page_rows, L.n_windows, L.lds_gslots, na)))
        L.filter_words = 0;
    qd.use_lds = L.lds_gslots != 0;
    L.group_bits = std::max<uint32_t>(1, bits_for(qh.domain));
    int64_t *d_sum;
    uint64_t *d_cnt;
    if (int r = group_buffers(ctx, qh.domain, &d_sum, &d_cnt)) return r;
    qd.g_sum = d_sum;
    qd.g_cnt = d_cnt;
    L.err = reinterpret_cast<uint32_t *>(ctx->d_scratch);
    const golap_hashmap *dm = plan->dims[qd.driver].map;
    // the chunk flags (for the slice statistics, and to size the work items by the chunks that
    // survive)
    std::vector<uint16_t> mask;
    uint64_t surviving = 0;
    if (plan->dev_chunk_mask && plan->chunk_end > plan->chunk_begin) {
        mask.resize(plan->chunk_end - plan->chunk_begin);
        GOLAP_HIP_CHECK(ctx, hipMemcpyAsync(mask.data(), plan->dev_chunk_mask + plan->chunk_begin, mask.size() * 2,
                                            hipMemcpyDeviceToHost, ctx->stream));
        GOLAP_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
        for (uint16_t f : mask) surviving += f != 0;
    }
    GOLAP_HIP_CHECK(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    // with a chunk mask, the surviving chunks may sit anywhere in the slice: a device-built work
    // list keeps every workgroup busy (the static page ranges would leave most of them idle)
    if (plan->dev_chunk_mask && L.n_items && dm->filled && surviving) {
        // few surviving chunks are cut into page ranges, so that their pages spread over about
        // four workgroups per CU instead of queueing chunk after chunk on a few of them
        const uint32_t parts = (uint32_t)std::min<uint64_t>(
            {(uint64_t)L.pages_per_chunk, 256, std::max<uint64_t>(1, ((uint64_t)ctx->cus * 4 + surviving - 1) / surviving)});
        const void *dq;
        if (int r = solo_worklist(ctx, &qd, sizeof(qd), parts, &dq, &L.queue, const_cast<uint2 **>(&L.items)))
            return r;
        L.queries = static_cast<const StarQueryDev *>(dq);
        L.pages_per_item = (L.pages_per_chunk + parts - 1) / parts;
    } else {
        L.q0 = qd;
    }
    if (L.n_items && dm->filled && (!plan->dev_chunk_mask || surviving))
        GOLAP_HIP_CHECK(ctx, launch_star(ctx, L, ctx->stream));
    GOLAP_HIP_CHECK(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    uint32_t flags = 0;
    GOLAP_HIP_CHECK(ctx, hipMemcpyAsync(&flags, ctx->d_scratch, 4, hipMemcpyDeviceToHost, ctx->stream));
    GOLAP_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    if (flags) return error_flags_status(ctx, flags, "star");
    float ms = 0;
    GOLAP_HIP_CHECK(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    res->kernel_ms = ms;
    res->driver_dim = qd.driver;
    res->lds_group_slots = L.lds_gslots;
    std::vector<uint8_t> run(mask.size());
    for (size_t i = 0; i < mask.size(); ++i) run[i] = mask[i] != 0;
    slice_stats(*plan, run, res);
    return emit_groups(ctx, qh.domain, d_sum, d_cnt, qh.lo, qh.card, qh.stride, &res->n_groups, &res->total_sum,
                       group_keys, sums, counts, max_groups, &res->rows_matched);
}

}  // extern "C"
