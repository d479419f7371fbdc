// Host-side helpers of the star plans shared by the single-query entry point (capi_star.cpp)
// and the batched form (capi_batch.cpp).
#pragma once
#include <cstdint>
#include <utility>
#include <vector>

#include "common.hpp"

namespace golap_impl {
namespace starhost {

constexpr uint64_t MAX_GROUPS = 1ull << 31;
constexpr uint32_t LDS_GROUP_SLOTS_MAX = 4096;  // 64 KiB of LDS (key, count, int64 sum)

// Distinct fact columns of one launch; all row-aligned with the first.
struct ColumnSet {
    std::vector<const golap_column *> cols;
    int index(golap_ctx *ctx, const golap_column *c, uint8_t *idx);
};

// Host side of one prepared query.
struct QueryHost {
    golap_star_plan plan{};
    std::vector<int32_t> lo;
    std::vector<uint32_t> card;
    std::vector<uint64_t> stride;
    uint64_t domain = 1;
    double sel = 1.0;  // expected share of fact rows that match (uniform foreign keys)
    std::vector<std::pair<const golap_hashmap *, uint64_t>> builds;  // maps and their build ids
};

// Host side of one prepared scan plan.
struct ScanHost {
    golap_scan_plan plan{};
    uint32_t colmask = 0;  // columns read, as indexes of plan.cols
    std::vector<int32_t> lo;
    std::vector<uint32_t> card;
    std::vector<uint64_t> stride;
    uint64_t domain = 1;
};

uint32_t pow2_at_least(uint64_t v);
uint32_t bits_for(uint64_t n);
uint32_t choose_lds_slots(uint32_t agg_mode, uint32_t requested, uint64_t n_groups);
int group_buffers(golap_ctx *ctx, uint64_t n, int64_t **sum, uint64_t **cnt);
int emit_groups(golap_ctx *ctx, uint64_t n, const int64_t *d_sum, const uint64_t *d_cnt,
                const std::vector<int32_t> &lo, const std::vector<uint32_t> &card,
                const std::vector<uint64_t> &stride, uint64_t *n_groups, int64_t *total, int32_t *keys,
                int64_t *sums, uint64_t *counts, uint64_t max_groups, uint64_t *rows = nullptr);
int emit_groups_host(golap_ctx *ctx, uint64_t n, const int64_t *hs, const uint64_t *hc,
                     const std::vector<int32_t> &lo, const std::vector<uint32_t> &card,
                     const std::vector<uint64_t> &stride, uint64_t *n_groups, int64_t *total, int32_t *keys,
                     int64_t *sums, uint64_t *counts, uint64_t max_groups, uint64_t *rows = nullptr);
// host_dim >= 0: that dimension is joined on the host (emit mode): its map may be null, it is
// never the driver and the kernel does not probe it.
int prepare_query(golap_ctx *ctx, const golap_star_plan *plan, bool compact, ColumnSet &cs, StarQueryDev &qd,
                  QueryHost &qh, int host_dim = -1);
uint32_t assign_windows(const ColumnSet &cs, uint64_t pb, uint64_t pe, int *win);
uint32_t assign_access(const ColumnSet &cs, uint64_t pb, uint64_t pe, uint32_t driver_cols, int *win, int *anchor,
                       uint32_t *n_anchors);
int ensure_hfilter(golap_ctx *ctx, const golap_hashmap *dm);
// Membership filter of a map's keys: its dense bitmap when that is small, else its Bloom filter.
int filter_view(golap_ctx *ctx, const golap_hashmap *dm, DriverFilterView *out, uint32_t *words);
void slice_stats(const golap_star_plan &plan, const std::vector<uint8_t> &run, golap_star_result *res);
int status_of_flags(golap_ctx *ctx, uint32_t flags, const char *what);
bool is_int_type(uint32_t t);

int prepare_scan(golap_ctx *ctx, const golap_scan_plan *p, ScanQueryDev &qd, ScanHost &h);
std::vector<uint8_t> scan_survivors(const golap_scan_plan &p, const std::vector<uint16_t> &mask);
void scan_stats(const golap_scan_plan &p, uint32_t colmask, const std::vector<uint8_t> &run, golap_star_result *res);
// LDS windows of one wave over columns colmask and pages [pb, pe): woff (-1 = none), delta-bitpack
// register slots dslot (-1 = none), window bytes wcap of columns whose LZ4I pages decode in
// parallel (0 = the page bytes), and the offset (-1 = none) and capacity of the LZ4I token
// offsets behind the windows. Returns the bytes of one wave's region.
uint32_t scan_layout(const golap_column *const *cols, uint32_t n_cols, uint32_t colmask, uint64_t pb, uint64_t pe,
                     int32_t *woff, int32_t *dslot, uint32_t *wcap, int32_t *tok_off, uint32_t *tok_cap);
// Columns of colmask decoded ahead of a scan over pages [pb, pe): those with LZ4I pages there that
// a wave would otherwise decode into its LDS window page by page. The chunk decompressor (a
// workgroup per page, pointer jumping for long copy chains) writes them to an HBM buffer and the
// scan reads them in place.
uint32_t scan_decode_mask(const golap_column *const *cols, uint32_t n_cols, uint32_t colmask, uint64_t pb, uint64_t pe);
// Buffer bytes for the decoded columns of mask over un chunks (each column's region 256-byte aligned).
uint64_t scan_decode_bytes(const golap_column *const *cols, uint32_t n_cols, uint32_t mask, uint64_t un);
// Points the views of the decoded columns at their regions of buf (chunks from chunk c0 on).
void scan_decode_views(const golap_column *const *cols, uint32_t n_cols, uint32_t mask, uint64_t c0, uint64_t un,
                       uint8_t *buf, ScanColView *views);
// Enqueues the decodes of the chunks [cb, ce) of the decoded columns into their regions (chunks
// from c0 on); corrupt pages set err.
hipError_t scan_decode_enqueue(golap_ctx *ctx, const golap_column *const *cols, uint32_t n_cols, uint32_t mask,
                               uint64_t c0, uint64_t un, uint8_t *buf, uint64_t cb, uint64_t ce, hipStream_t st,
                               uint32_t *err);
uint32_t scan_page_waves(golap_ctx *ctx, uint32_t win_bytes, uint32_t agg_entries);
uint32_t scan_agg_entries(golap_ctx *ctx, uint32_t agg_mode, uint32_t requested, uint32_t win_bytes, uint32_t page_waves,
                          uint32_t dflt);
uint64_t scan_local_domain(const golap_scan_plan &p, const ScanQueryDev &qd, const std::vector<uint8_t> &run);
uint32_t scan_default_entries(uint64_t maxl);
uint32_t scan_pages_per_item(golap_ctx *ctx, uint64_t surviving_chunks, uint32_t ppc);

}  // namespace starhost
}  // namespace golap_impl
