// C ABI: group-by and the fused star join + group-by plan (include/golap_hip.h). Host side of
// query2 / query3 (usecases/ssb/queries/02.hpp:54-307, 03.hpp:79-677): choose the driving
// dimension and its LDS filter, the LDS windows for coded columns and the aggregation table,
// launch one fused kernel, and return the canonical (sorted) group list.
#include <algorithm>
#include <cstring>
#include <vector>

#include "common.hpp"
#include "format.hpp"

using namespace golap_impl;
using Lock = std::lock_guard<std::recursive_mutex>;

namespace {

constexpr uint64_t MAX_GROUPS = 1ull << 31;
constexpr uint32_t LDS_GROUP_SLOTS_MAX = 4096;  // 64 KiB of LDS (key, count, int64 sum)

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

// Copies the dense result back and emits the groups with rows, in index (= key tuple) order.
int emit_groups(golap_ctx *ctx, uint64_t n, const int64_t *d_sum, const uint64_t *d_cnt,
                const std::vector<int32_t> &lo, const std::vector<uint32_t> &card,
                const std::vector<uint64_t> &stride, uint64_t *n_groups, int64_t *total,
                int32_t *keys, int64_t *sums, uint64_t *counts, uint64_t max_groups) {
    std::vector<int64_t> hs(n);
    std::vector<uint64_t> hc(n);
    GOLAP_HIP_CHECK(ctx, hipMemcpyAsync(hs.data(), d_sum, n * 8, hipMemcpyDeviceToHost, ctx->stream));
    GOLAP_HIP_CHECK(ctx, hipMemcpyAsync(hc.data(), d_cnt, n * 8, hipMemcpyDeviceToHost, ctx->stream));
    GOLAP_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    uint64_t g = 0;
    int64_t t = 0;
    const size_t na = card.size();
    for (uint64_t i = 0; i < n; ++i) {
        if (!hc[i]) continue;
        t += hs[i];
        if (g < max_groups) {
            for (size_t a = 0; a < na && keys; ++a)
                keys[g * na + a] = lo[a] + (int32_t)((i / stride[a]) % card[a]);
            if (sums) sums[g] = hs[i];
            if (counts) counts[g] = hc[i];
        }
        ++g;
    }
    *n_groups = g;
    if (total) *total = t;
    if (g > max_groups) return fail(ctx, GOLAP_ECAPACITY, "more groups than max_groups");
    return GOLAP_OK;
}

bool is_int_type(uint32_t t) { return t >= GOLAP_I8 && t <= GOLAP_I64; }

}  // namespace

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
    if (!ctx || !plan || !res || plan->n_dims < 1 || plan->n_dims > STAR_MAX_DIMS || !plan->value ||
        plan->agg_mode > GOLAP_AGG_GLOBAL)
        return GOLAP_EINVAL;
    const uint32_t nd = plan->n_dims;
    const golap_column *v = plan->value;
    auto is32 = [](const golap_column *c) { return c->type == GOLAP_I32 || c->type == GOLAP_U32; };
    if (v->ctx != ctx || !is32(v)) return fail(ctx, GOLAP_EINVAL, "star: value column must be 32-bit");
    for (uint32_t d = 0; d < nd; ++d) {
        const golap_star_dim &sd = plan->dims[d];
        if (!sd.fk || !sd.map || sd.fk->ctx != ctx || sd.map->ctx != ctx || !sd.map->built)
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
    Lock l(ctx->mu);
    if (int r = set_device(ctx)) return r;
    memset(res, 0, sizeof(*res));

    // group domain, attributes in dimension order (first = most significant)
    std::vector<int32_t> lo;
    std::vector<uint32_t> card;
    std::vector<uint32_t> gdim;
    uint64_t domain = 1;
    for (uint32_t d = 0; d < nd; ++d)
        if (plan->dims[d].group_card) {
            lo.push_back(plan->dims[d].group_lo);
            card.push_back(plan->dims[d].group_card);
            gdim.push_back(d);
            domain *= plan->dims[d].group_card;
            if (domain >= MAX_GROUPS) return fail(ctx, GOLAP_EINVAL, "star: group domain too large");
        }
    std::vector<uint64_t> stride(card.size());
    for (int a = (int)card.size() - 1, s = 1; a >= 0; --a) {
        stride[a] = (uint64_t)s;
        s *= (int)card[a];
    }

    StarLaunch L{};
    L.n_dims = nd;
    // driver: the dimension whose filter passes the smallest share of its key domain
    double best = 2.0;
    for (uint32_t d = 0; d < nd; ++d) {
        const golap_hashmap *m = plan->dims[d].map;
        const double sel = m->ht_size ? (double)m->filled / (double)m->ht_size : 1.0;
        if (sel < best) {
            best = sel;
            L.driver = d;
        }
    }
    const uint64_t ppc = v->chunk_rows / v->page_rows;
    L.page_rows = v->page_rows;
    L.n_rows = v->n_rows;
    L.page_begin = plan->chunk_begin * ppc;
    L.page_end = std::min<uint64_t>(v->n_pages, plan->chunk_end * ppc);
    L.chunk_mask = plan->dev_chunk_mask;
    L.pages_per_chunk = (uint32_t)ppc;
    uint32_t windows = 0;
    auto window_of = [&](const golap_column *c) {
        std::lock_guard<std::mutex> g(g_extra_mu);
        auto &ex = g_extra[c];
        return ex.window_prefix[L.page_end] != ex.window_prefix[L.page_begin] ? (int)windows++ : -1;
    };
    size_t a_idx = 0;
    for (uint32_t d = 0; d < nd; ++d) {
        const golap_star_dim &sd = plan->dims[d];
        const golap_hashmap *m = sd.map;
        L.fk[d] = view(sd.fk);
        L.fk_window[d] = window_of(sd.fk);
        L.dim[d].slots = m->small_log2 ? m->d_small : m->d_slots;
        L.dim[d].log2cap = m->small_log2 ? m->small_log2 : m->log2cap;
        L.dim[d].group_lo = sd.group_lo;
        L.dim[d].group_card = sd.group_card;
        L.dim[d].group_stride = sd.group_card ? stride[a_idx++] : 0;
    }
    L.val = view(v);
    L.val_window = window_of(v);
    L.val_signed = v->type == GOLAP_I32;
    L.n_windows = windows;

    // driver filter: dense exact bitmap, or the hashed one-bit filter of the map's keys
    const golap_hashmap *dm = plan->dims[L.driver].map;
    if (dm->use_bitmap && (dm->krange + 31) / 32 <= (1u << HASH_FILTER_WORDS_LOG2)) {
        L.filter = dm->d_bitmap;
        L.filter_words = (dm->krange + 31) / 32;
        L.fkmin = dm->kmin;
        L.fkrange = dm->krange;
        L.fhash_log2 = 0;
    } else {
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
                                                    dm->small_log2 ? dm->small_log2 : dm->log2cap,
                                                    dm->d_hfilter, HASH_FILTER_WORDS_LOG2));
            dm->hfilter_build = dm->build_id;
        }
        L.filter = dm->d_hfilter;
        L.filter_words = words;
        L.fhash_log2 = HASH_FILTER_WORDS_LOG2;
    }
    L.n_groups = domain;
    // AUTO: LDS partial tables pay off when many rows land in each group; estimate the
    // qualifying rows from the maps' fill ratios (uniform foreign keys)
    double expect = (double)(L.page_end - L.page_begin) * L.page_rows;
    for (uint32_t d = 0; d < nd; ++d) {
        const golap_hashmap *m = plan->dims[d].map;
        expect *= m->ht_size ? (double)m->filled / (double)m->ht_size : 1.0;
    }
    uint32_t agg = plan->agg_mode;
    if (agg == GOLAP_AGG_AUTO) agg = expect >= 16.0 * (double)domain && expect >= 65536.0 ? GOLAP_AGG_LDS : GOLAP_AGG_GLOBAL;
    L.lds_gslots = choose_lds_slots(agg, plan->lds_group_slots, domain);
    while (L.lds_gslots && star_lds_bytes(L.filter_words, L.page_rows, L.n_windows, L.lds_gslots) >
                               (size_t)ctx->lds_max)
        L.lds_gslots >>= 1;
    if (star_lds_bytes(L.filter_words, L.page_rows, L.n_windows, L.lds_gslots) > (size_t)ctx->lds_max)
        return fail(ctx, GOLAP_EINVAL, "star: LDS budget exceeded (page_rows or coded columns)");
    int64_t *d_sum;
    uint64_t *d_cnt;
    if (int r = group_buffers(ctx, domain, &d_sum, &d_cnt)) return r;
    L.g_sum = d_sum;
    L.g_cnt = d_cnt;
    L.rows_matched = reinterpret_cast<uint64_t *>(ctx->d_scratch + 16);
    L.err = reinterpret_cast<uint32_t *>(ctx->d_scratch);
    GOLAP_HIP_CHECK(ctx, hipMemsetAsync(ctx->d_scratch + 16, 0, 8, ctx->stream));
    GOLAP_HIP_CHECK(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    if (L.page_end > L.page_begin && dm->filled) GOLAP_HIP_CHECK(ctx, launch_star(ctx, L));
    GOLAP_HIP_CHECK(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    uint64_t host[3] = {0, 0, 0};
    GOLAP_HIP_CHECK(ctx, hipMemcpyAsync(host, ctx->d_scratch, 24, hipMemcpyDeviceToHost, ctx->stream));
    std::vector<uint16_t> mask;
    if (plan->dev_chunk_mask && plan->chunk_end > plan->chunk_begin) {
        mask.resize(plan->chunk_end);
        GOLAP_HIP_CHECK(ctx, hipMemcpyAsync(mask.data() + plan->chunk_begin, plan->dev_chunk_mask + plan->chunk_begin,
                                            (plan->chunk_end - plan->chunk_begin) * 2, hipMemcpyDeviceToHost,
                                            ctx->stream));
    }
    GOLAP_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    if ((uint32_t)host[0]) return error_flags_status(ctx, (uint32_t)host[0], "star");
    float ms = 0;
    GOLAP_HIP_CHECK(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    res->kernel_ms = ms;
    res->rows_matched = host[2];
    res->driver_dim = L.driver;
    res->lds_group_slots = L.lds_gslots;
    {
        std::lock_guard<std::mutex> g(g_extra_mu);
        for (uint64_t ch = plan->chunk_begin; ch < plan->chunk_end; ++ch) {
            if (!mask.empty() && !mask[ch]) {
                res->chunks_pruned++;
                continue;
            }
            res->chunks_run++;
            auto add = [&](const golap_column *c) {
                auto &ex = g_extra[c];
                res->comp_bytes += ex.chunk_payload[ch + 1] - ex.chunk_payload[ch];
            };
            for (uint32_t d = 0; d < nd; ++d) add(plan->dims[d].fk);
            add(v);
        }
    }
    return emit_groups(ctx, domain, d_sum, d_cnt, lo, card, stride, &res->n_groups, &res->total_sum, group_keys,
                       sums, counts, max_groups);
}

}  // extern "C"
