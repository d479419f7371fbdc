// Host-side objects behind the C ABI handles.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/golap_hip.h"
#include "host_column.hpp"
#include "udiv.hpp"

// Device-side error flags raised by kernels (bit-or into ctx scratch word 0).
enum : uint32_t { DEVERR_CORRUPT = 1u, DEVERR_CAPACITY = 2u, DEVERR_KEYRANGE = 4u };

struct golap_ctx {
    int device = 0;
    int cus = 0;
    int lds_max = 65536;  // dynamic LDS bytes a workgroup may request
    char name[128] = {0};
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // Serialises the calls of one context (its stream and scratch). The allocation registry and
    // the error string have locks of their own: the serving loop's completion threads free batch
    // buffers and report errors while the scheduler holds mu to launch the next batch.
    std::recursive_mutex mu;
    std::mutex err_mu;
    std::string err;
    std::mutex alloc_mu;
    // Device scratch: word 0 error flags; [8,16) int64 sum; [16,24) u64 count; [64,80) hash
    // build; [1024,2048) a single query's descriptor, [2048,2064) its chunk offsets, [2304,2312)
    // its work-list counters (pruned single-query plans).
    uint8_t *d_scratch = nullptr;
    static constexpr uint64_t SCRATCH_BYTES = 4096;
    uint64_t allocated = 0;
    std::unordered_map<void *, uint64_t> allocs;
    // dense group-by result buffers (sum int64, count uint64), grown on demand
    void *d_groups = nullptr;
    uint64_t groups_cap = 0;
    // autotuned batch launch configurations by batch shape
    std::unordered_map<std::string, golap_batch_config> tune_cache;
    // work list of pruned single-query plans, grown on demand
    void *d_solo_items = nullptr;
    uint64_t solo_items_cap = 0;
    // work list of single scan plans (uint4 items), grown on demand
    void *d_scan_items = nullptr;
    uint64_t scan_items_cap = 0;
    uint8_t *h_stage = nullptr;  // pinned host staging for small uploads (4 KiB)
    // the allocation-free scan's device array on this context's device (its address there),
    // or the error met looking it up
    uint8_t *d_fixed = nullptr;
    uint64_t fixed_bytes = 0;
    hipError_t fixed_err = hipSuccess;
};

// A caller's stream (golap_stream_create) with its own device error word, so work enqueued on it
// reports its errors when the stream is synchronised, not through the context's scratch.
struct golap_stream {
    golap_ctx *ctx = nullptr;
    hipStream_t stream = nullptr;
    uint32_t *d_err = nullptr;
};
struct golap_event {
    golap_ctx *ctx = nullptr;
    hipEvent_t event = nullptr;
};
// A group-by accumulator kept on the device across calls (golap_hashagg_create).
struct golap_hashagg {
    golap_ctx *ctx = nullptr;
    uint32_t n_keys = 0, agg_mode = 0;
    std::vector<int32_t> lo;
    std::vector<uint32_t> card, type;
    std::vector<uint64_t> stride;
    uint64_t domain = 0;
    int64_t *d_sum = nullptr;  // [domain] then counts [domain]
    uint64_t *d_cnt = nullptr;
};

// Host planning tables of a device column, from its host catalog: launches are planned from
// them without a device round trip.
struct ColumnExtra {
    std::vector<uint32_t> nondirect_prefix;  // pages that cannot be read in place, prefix count
    std::vector<uint32_t> window_prefix;     // pages neither in place nor point-readable
    std::vector<uint32_t> dbp_prefix;        // delta-bitpack pages that are not point-readable
    std::vector<uint64_t> chunk_payload;     // payload bytes per chunk, prefix
    // most sequences of any LZ4I page (sizes the token arrays of parallel decodes), 0 when the
    // page data was not seen on the host
    uint32_t lz4i_max_nseq = 0;
    uint32_t lz4i_max_csize = 0;  // largest LZ4I page (bytes), from the catalog
    // largest LZ4I page of more than LZ4I_SERIAL_MAX sequences (the pages a cooperative decode
    // stages in LDS); known when the page headers were read on the host
    uint32_t lz4i_coop_csize = 0;
    bool lz4i_coop_known = false;
    bool has_lz4i = false;
};

struct golap_column {
    golap_ctx *ctx = nullptr;
    uint32_t type = 0, elem = 0, page_rows = 0, chunk_rows = 0;
    uint64_t n_rows = 0, n_pages = 0, n_chunks = 0, comp_bytes = 0, payload_bytes = 0;
    // host copies of the catalog, used to plan launches without device round trips
    std::vector<uint64_t> page_off;
    std::vector<uint32_t> page_csize;
    std::vector<uint8_t> page_codec;
    std::vector<int64_t> chunk_min, chunk_max;
    // device copies
    uint8_t *d_comp = nullptr;
    uint64_t *d_page_off = nullptr;
    uint32_t *d_page_csize = nullptr;
    uint8_t *d_page_codec = nullptr;
    int64_t *d_chunk_min = nullptr, *d_chunk_max = nullptr;
    // planning tables, rebuilt whenever the catalog changes (column creation, stream windows)
    ColumnExtra extra;
};

// Open addressing over 8-byte slots: slot = (uint64(payload) << 32) | (uint32(key) ^ 0x80000000).
// The all-zero slot is empty, so INT32_MIN cannot be a build key.
struct golap_hashmap {
    golap_ctx *ctx = nullptr;
    uint64_t ht_size = 0;
    const void *d_keys = nullptr;
    uint32_t key_type = 0;
    uint32_t log2cap = 0;          // main table: 2^log2cap slots
    uint64_t *d_slots = nullptr;
    uint64_t filled = 0;           // rows inserted by the last build
    // compact copy sized to twice `filled` (2^small_log2 slots, 0 = none): probed by the star
    // plans, pinned in LDS by the Q1 plan when it has at most 4096 slots
    uint32_t small_log2 = 0;
    uint32_t small_cap_log2 = 0;  // allocated size of d_small
    uint64_t *d_small = nullptr;
    // blocked Bloom filter (2^HASH_FILTER_WORDS_LOG2 words), built on demand when the key
    // range is too wide for the dense bitmap; a cache, valid for the build it was made from
    mutable uint32_t *d_hfilter = nullptr;
    mutable uint64_t hfilter_build = 0;
    uint64_t build_id = 0;
    // dense form when the inserted keys span a small range: membership bits over
    // [kmin, kmin + krange)
    int32_t kmin = 0, kmax = -1;
    uint32_t krange = 0;
    bool use_bitmap = false;
    uint32_t *d_bitmap = nullptr;
    uint64_t bitmap_words = 0;  // allocated size of d_bitmap
    bool built = false;
};

namespace golap_impl {

inline int fail(golap_ctx *ctx, int code, const std::string &msg) {
    if (ctx) {
        std::lock_guard<std::mutex> g(ctx->err_mu);
        ctx->err = msg;
    }
    return code;
}
// A copy of the context's last error message, taken under its lock.
inline std::string last_error(golap_ctx *ctx) {
    std::lock_guard<std::mutex> g(ctx->err_mu);
    return ctx->err;
}

#define GOLAP_HIP_CHECK(ctx, expr)                                                          \
    do {                                                                                    \
        hipError_t e__ = (expr);                                                            \
        if (e__ != hipSuccess)                                                              \
            return golap_impl::fail((ctx), GOLAP_EHIP,                                      \
                                    std::string(#expr ": ") + hipGetErrorString(e__));      \
    } while (0)

// Runtime helpers shared by the C ABI translation units (capi.cpp).
bool page_direct_host(uint8_t codec);
int dev_alloc(golap_ctx *ctx, uint64_t bytes, void **p);
// A device column from a host column's catalog and zone maps, with comp_bytes of page data
// copied from comp, or only allocated (and left for the caller to fill) when comp is null.
// Call with the context lock held.
int create_device_column(golap_ctx *ctx, const golap_host_column *h, uint64_t comp_bytes, const uint8_t *comp,
                         golap_column **out);
// The host planning tables of a column, from its host catalog.
void register_extra(golap_column *c);
uint32_t lz4i_max_nseq(const golap_host_column &h, const uint8_t *comp, uint32_t *coop_csize = nullptr);
// Bytes an LZ4I stage region needs for a column's pages (a multiple of 16).
uint32_t lz4i_stage_bytes(const ColumnExtra &ex);
// Points device column c (created with catalog capacity for at least w's pages and chunks) at a
// window w of a larger column: catalog and zone maps copied on stream without a wait (from the
// column's own host copy, kept until its next window), planning tables rebuilt; the caller fills
// d_comp with comp_bytes of page data and orders every reader of the column after the copies.
int set_column_window(golap_ctx *ctx, golap_column *c, const golap_host_column &w, uint64_t comp_bytes,
                      hipStream_t stream);
void dev_release(golap_ctx *ctx, void *p);
int set_device(golap_ctx *ctx);
int check_device_errors(golap_ctx *ctx, const char *what);
int error_flags_status(golap_ctx *ctx, uint32_t flags, const char *what);
bool valid_type(uint32_t t);

// Device views passed to kernels by value.
struct ColView {
    const uint8_t *comp;
    const uint64_t *page_off;
    const uint32_t *page_csize;
    const uint8_t *page_codec;
    uint32_t elem;
};
inline ColView view(const golap_column *c) {
    return ColView{c->d_comp, c->d_page_off, c->d_page_csize, c->d_page_codec, c->elem};
}

constexpr uint32_t STAR_MAX_DIMS = 4;
constexpr uint32_t HASH_FILTER_WORDS_LOG2 = 12;  // blocked Bloom filter: 4096 words = 16 KiB of LDS

// One dimension of a star plan as the kernel sees it: exact table and group-key contribution.
struct StarDimProbe {
    const uint64_t *slots;
    uint32_t log2cap;
    int32_t group_lo;
    uint32_t group_card;    // 0: the dimension only filters
    uint64_t group_stride;
};

// Device view of a driver filter (see stardev::filter_test).
struct DriverFilterView {
    const uint32_t *f;
    int32_t fkmin;
    uint32_t fkrange;
    uint32_t fhash_log2;
};

constexpr uint32_t STAR_MAX_COLS = 8;  // fact columns one star launch can read
constexpr uint32_t SLICE_MAX_ZONES = 2;

// Zone-map predicate of check_mmmeta (core/metadata.cuh:539-547): chunk c may hold matching
// rows iff lo <= zmax[c] && hi >= zmin[c].
struct ZonePred {
    const int64_t *zmin, *zmax;  // zmin null = unused
    int64_t lo, hi;
};

// A query's fact slice; the first member of every query descriptor, so one work-list builder
// serves every query kind.
struct SliceDev {
    uint64_t chunk_begin, chunk_end;
    uint64_t page_begin, page_end;  // the same slice in pages (static schedule)
    const uint16_t *chunk_mask;     // zone-map flags per chunk, may be null
    ZonePred zone[SLICE_MAX_ZONES];
    uint64_t item_begin;            // static schedule: pages of the queries before this one
};

// One query of the persistent star kernel. Its columns are indexes into the launch's fact
// columns, so queries over different foreign keys of one fact table share a launch.
struct StarQueryDev {
    SliceDev s;
    StarDimProbe dim[STAR_MAX_DIMS];
    uint8_t fk_col[STAR_MAX_DIMS];
    uint8_t val_col;
    uint8_t n_dims;
    uint8_t val_signed;
    uint8_t use_lds;                // aggregate in the LDS table (else straight to global atomics)
    uint32_t driver;                // dimension whose filter is tested first
    DriverFilterView filter;
    uint32_t filter_words;
    uint8_t has_val2;               // SUM(value - value2): the second value column val2_col
    uint8_t val2_col;
    uint8_t val2_signed;
    uint8_t pad0;
    uint64_t n_groups;
    int64_t *g_sum;
    uint64_t *g_cnt;
};

struct StarLaunch {
    ColView col[STAR_MAX_COLS];
    int col_window[STAR_MAX_COLS];
    // 1 + the LDS prefix-sum region of a column read without a window whose delta-bitpack
    // pages the tiles point-read; 0 = none (the k_star schedules only; the shared-filter kernel
    // decodes such pages into windows)
    int col_anchor[STAR_MAX_COLS];
    uint32_t n_anchors;
    uint32_t n_cols;
    uint32_t page_rows;
    uint64_t n_rows, n_pages;
    uint32_t pages_per_chunk;
    uint32_t n_windows;
    // static schedule: q0 alone (queries == nullptr) or a device array; n_items pages in all
    StarQueryDev q0;
    const StarQueryDev *queries;
    uint32_t n_queries;
    uint64_t n_items;
    // queue schedule (items != nullptr): device-built (chunk, query << 8 | part) items,
    // queue[0] = next item, queue[1] = item count; an item covers pages_per_item pages
    const uint2 *items;
    uint32_t *queue;
    uint32_t pages_per_item;
    uint32_t filter_words;        // LDS words for the driver filter, 0 = read it from global memory
    uint32_t lds_gslots;
    uint32_t group_bits;
    uint32_t max_blocks_per_cu;   // 0 = occupancy limit
    uint32_t *err;
    // shared-filter schedule (items4 != nullptr): items {chunk, part | word << 8, mask lo, mask
    // hi} over the queries of one word of 64; a page's keys of fact column mcol are read once
    // and tested against mfilt[q], the membership filter of query q's dimension on that column
    const uint4 *items4;
    const DriverFilterView *mfilt;
    uint32_t mcol;
    // emit mode (emit != nullptr; one query, no LDS group table): rows passing every dimension
    // but emit_dim are appended to emit as {emit_dim's key, group index, value lo, value hi};
    // emit_n counts them all, appends beyond emit_cap are dropped
    int4 *emit;
    uint32_t *emit_n;
    uint64_t emit_cap;
    uint32_t emit_dim;
};
// prepare_only: set the kernel's LDS attribute and check the launch without launching (graph
// capture then records the launch alone).
hipError_t launch_star(golap_ctx *ctx, const StarLaunch &l, hipStream_t stream, bool prepare_only = false);

// Batch work list: one thread per (query, chunk of its slice); surviving chunks (chunk mask and
// zone predicates) emit parts items each into items, counted in queue[1]. Query q's slice is
// the SliceDev at qbase + q * qstride.
hipError_t launch_batch_items(const void *qbase, uint32_t qstride, uint32_t n_queries, const uint64_t *chunk_off,
                              uint64_t total_chunks, uint32_t parts, uint2 *items, uint32_t *queue,
                              hipStream_t stream);
hipError_t launch_fill_zero(void *p, uint64_t bytes, hipStream_t stream);
// Copies bytes (a multiple of 4, at most PUT_WORDS_MAX) from host memory to dst (4-byte aligned)
// as a kernel argument: no staging memory, a kernel node under capture.
constexpr uint32_t PUT_WORDS_MAX = 2048;
hipError_t launch_put_words(void *dst, const void *src, uint32_t bytes, hipStream_t stream);
// Non-empty groups of a batch's dense results (query q's sums at arena + qoff[q].x, its counts
// qoff[q].y entries later; qfirst[q] = its first index of the concatenated domains, total in
// all): {q, group} + sum + count appended in any order, *n_out counted (also beyond cap).
hipError_t launch_compact_groups(const uint8_t *arena, const void *qoff, const uint64_t *qfirst, uint32_t nq,
                                 uint64_t total, uint2 *out_ql, int64_t *out_sum, uint64_t *out_cnt, uint64_t cap,
                                 unsigned long long *n_out, uint32_t cus, hipStream_t stream);
// The keys map m holds, ascending, into keys (capacity cap); *n = their count (also when it
// exceeds cap: then GOLAP_ECAPACITY). Synchronises the context stream.
int collect_sorted_keys(golap_ctx *ctx, const golap_hashmap *m, int32_t *keys, uint64_t cap, uint64_t *n);
// flags[c] = some key of the ascending array keys[0, n_keys) lies in [min[c], max[c]].
hipError_t launch_check_points(golap_ctx *ctx, const golap_column *col, const int32_t *keys, uint64_t n_keys,
                               uint16_t *flags);
hipError_t launch_read_probe(golap_ctx *ctx, const void *p, uint64_t bytes, uint32_t blocks_per_cu, uint32_t *sink,
                             hipStream_t stream);

// Work list of one pruned query (a batch of one): its descriptor (desc_size <= 1024 bytes,
// beginning with its SliceDev) goes to the context scratch, the surviving chunks' items to a
// context buffer, `parts` items per chunk (1..256; pages_per_item = pages per chunk / parts,
// rounded up). On return *d_desc, *queue and *items are the device pointers for the queue schedule.
int solo_worklist(golap_ctx *ctx, const void *desc, uint32_t desc_size, uint32_t parts, const void **d_desc,
                  uint32_t **queue, uint2 **items);
size_t star_lds_bytes(uint32_t filter_words, uint32_t page_rows, uint32_t n_windows, uint32_t lds_gslots,
                      uint32_t n_anchors = 0);
// Workgroups per CU of the star kernel for pages of page_rows rows and `lds` bytes of dynamic LDS.
uint32_t star_blocks_per_cu(uint32_t page_rows, size_t lds);
hipError_t launch_hash_filter(golap_ctx *ctx, const uint64_t *slots, uint32_t log2cap, uint32_t *filter,
                              uint32_t words_log2);

// ---- scan + group-by over typed columns (trip-table plans)
constexpr uint32_t SCAN_MAX_COLS = 8, SCAN_MAX_PREDS = 4, SCAN_MAX_KEYS = 3;
constexpr uint64_t COMP_TAIL = 16;  // zero bytes after a device column's page data

// Delta-bitpack pages of up to SCAN_DBP_SLOTS columns of a launch are decoded in registers (a
// running sum per thread); other coded pages, and delta pages of further columns, go through an
// LDS window.
constexpr uint32_t SCAN_DBP_SLOTS = 1;
struct ScanColView {
    ColView c;
    uint32_t type;
    int32_t woff;   // byte offset of the column's LDS window, -1 = none
    int32_t dslot;  // register slot for its delta-bitpack pages, -1 = none
    uint32_t wcap;  // window bytes when it decodes LZ4I pages in parallel (output + staging), else 0
    uint32_t decoded;   // 1: decoded before the launch; row r is read in place at c.comp + (r - dec_row0) * elem
    uint32_t pad;
    uint64_t dec_row0;
};
struct ScanPredDev {
    uint32_t col, pad;
    int64_t lo, hi;
    const int64_t *zmin, *zmax;  // the column's zone maps: a chunk inside [lo, hi] skips the test
};
// Computed group key k = (v - base) / div over [0, card): span = card * div < 2^32, so the
// offset fits 32 bits and divides by the multiply-high method.
struct ScanKeyDev {
    uint32_t col, card;
    int64_t base;
    uint64_t span;
    golap_fmt::UDiv32 div;
    uint32_t pad;
    uint64_t stride;
    const int64_t *zmin, *zmax;  // the column's zone maps: the key's range within a chunk
};
struct ScanQueryDev {
    SliceDev s;
    ScanPredDev pred[SCAN_MAX_PREDS];
    ScanKeyDev key[SCAN_MAX_KEYS];
    uint32_t n_preds, n_keys;
    int32_t val_col;     // -1: COUNT only
    uint32_t use_lds;
    uint32_t colmask;    // fact columns the query reads
    uint32_t pad;
    uint64_t n_groups;
    int64_t *g_sum;
    uint64_t *g_cnt;
};
// Shared-scan launch: work items are (chunk, part, query word) with a 64-bit mask of the
// queries of that word that need the chunk; a workgroup reads the item's pages once for all of
// them. queue[0] = next item, queue[1] = item count (device-built by launch_scan_items).
struct ScanLaunch {
    ScanColView col[SCAN_MAX_COLS];
    uint32_t n_cols;
    uint32_t page_rows, pages_per_chunk;
    uint64_t n_rows, n_pages;
    uint32_t win_bytes;    // LDS bytes of one wave's windows
    uint32_t agg_entries;  // dense LDS aggregation entries (SUM int64 + COUNT u32), 0 = global atomics
    uint32_t page_waves;   // waves of a workgroup that take pages, each with its own windows (1..4)
    const ScanQueryDev *queries;
    uint32_t n_queries;
    const uint4 *items;
    uint32_t *queue;
    uint32_t pages_per_item;
    uint32_t max_blocks_per_cu;
    uint32_t rows_per_lane;  // 8: short launches (fewer steps per page); 4: long ones (more workgroups per CU)
    uint32_t *err;
    // LZ4I token offsets and sink of the wave's parallel decodes: byte offset in the wave's LDS
    // region (-1 = none) and the most sequences they hold
    int32_t tok_off = -1;
    uint32_t tok_cap = 0;
};
constexpr uint32_t SCAN_AGG_MAX = 8192;  // 96 KiB of LDS
hipError_t launch_scan(golap_ctx *ctx, const ScanLaunch &l, hipStream_t stream, bool prepare_only = false);
// Work list of a scan launch: one thread per (chunk in [c0, c0 + nch), query word) sets the
// bits of the word's queries whose slice keeps the chunk and, if any, appends `parts` items
// {chunk, part | word << 8, mask lo, mask hi}.
hipError_t launch_scan_items(const ScanQueryDev *queries, uint32_t n_queries, uint64_t c0, uint64_t nch,
                             uint32_t parts, uint4 *items, uint32_t *queue, hipStream_t stream);
// The same for any descriptors that start with their SliceDev (query q's at qbase + q * qstride).
hipError_t launch_mask_items(const void *qbase, uint32_t qstride, uint32_t n_queries, uint64_t c0, uint64_t nch,
                             uint32_t parts, uint4 *items, uint32_t *queue, hipStream_t stream);
size_t scan_lds_bytes(uint32_t win_bytes, uint32_t agg_entries, uint32_t page_waves);
// Dynamic LDS a scan launch may use (the device limit less the kernel's own LDS).
size_t scan_lds_budget(golap_ctx *ctx);
// Rows per lane of a scan launch over `pages` surviving pages with `lds` bytes of dynamic LDS.
uint32_t scan_rows_per_lane(golap_ctx *ctx, uint64_t pages, size_t lds);


struct GroupByLaunch {
    const void *key[4];
    uint32_t key_type[4];
    int32_t key_lo[4];
    uint32_t key_card[4];
    uint64_t key_stride[4];
    uint32_t n_keys;
    const void *val;
    uint32_t val_type;
    uint64_t n;
    uint32_t lds_gslots;
    int64_t *g_sum;
    uint64_t *g_cnt;
    uint32_t *err;
    hipStream_t stream = nullptr;  // nullptr: the context's stream
};
hipError_t launch_group_by(golap_ctx *ctx, const GroupByLaunch &l);

struct PredView {
    const void *col;
    uint32_t type;
    int64_t lo, hi;
};
constexpr int MAX_PREDS = 4;

// kernel launchers (kernels.hip)
// stream / err: where to launch and raise device errors (nullptr: the context's stream and scratch)
hipError_t launch_decode_pages(golap_ctx *ctx, const golap_column *col, uint64_t page_begin,
                               uint64_t page_end, uint8_t *dev_out, uint64_t out_row0,
                               hipStream_t stream = nullptr, uint32_t *err = nullptr);
hipError_t launch_hash_clear(golap_ctx *ctx, uint64_t *slots, uint64_t n);
hipError_t launch_hash_build(golap_ctx *ctx, golap_hashmap *m, uint64_t n_rows,
                             const PredView *preds, uint32_t n_preds, const int32_t *payload,
                             uint64_t *d_filled, int32_t *d_kminmax);
hipError_t launch_hash_bitmap(golap_ctx *ctx, golap_hashmap *m);
hipError_t launch_hash_compact(golap_ctx *ctx, golap_hashmap *m);
hipError_t launch_hash_probe(golap_ctx *ctx, const golap_hashmap *m, const void *keys,
                             uint32_t key_type, uint64_t n, uint64_t *out_rid, uint64_t *hits);
hipError_t launch_minmax_check(golap_ctx *ctx, const golap_column *col, int64_t lo, int64_t hi,
                               uint16_t *flags);
hipError_t launch_combine_and(golap_ctx *ctx, uint16_t *a, const uint16_t *b, uint64_t n);

struct Q1Launch {
    ColView date, disc, qty, price;
    uint32_t page_rows;
    uint64_t n_rows;
    uint64_t page_begin, page_end;  // row tiles (page indices) to process
    const uint16_t *chunk_mask;     // chunk flags, 0 = pruned; may be null
    uint32_t pages_per_chunk;
    uint32_t disc_lo, disc_span, qty_lo, qty_span, price_lo, price_span;  // lo <= v <= lo+span
    int probe;                      // 0 hash in HBM, 1 hash in LDS, 2 bitmap in LDS
    const uint64_t *slots;
    uint32_t log2cap;
    const uint32_t *bitmap;
    int32_t kmin;
    uint32_t krange;
    int64_t *sum;
    uint64_t *count;
    uint32_t *err;
    // tiles with coded pages: LDS decode window (a multiple of 16 bytes, 0 = the four-wave
    // kernel) and the LZ4I token offsets it holds
    uint32_t win_cap = 0, tok_cap = 0;
    // the workgroup kernel: the largest LZ4I page of the four columns (bytes, a multiple of 16)
    // and the most sequences of one (0: unknown, bounded by the page size)
    uint32_t stage_cap = 0, max_nseq = 0;
};
hipError_t launch_q1(golap_ctx *ctx, const Q1Launch &a, bool direct);
// Tiles with coded pages by one workgroup each, LZ4I pages decoded by the whole workgroup
// (q1wg.hip); hipErrorNotSupported when the page shape or LDS does not allow it.
hipError_t launch_q1_wg(golap_ctx *ctx, const Q1Launch &a);
// Decode window and token capacity for LZ4I pages of at most page_bytes output bytes: the
// window holds the output and, at its tail, the staged page (LZ4I pages shrink, so a slack of an
// eighth of the page, at least 1 KiB, is enough for the in-place order); max_nseq 0 = unknown.
inline void lz4i_window(uint32_t page_bytes, uint32_t max_nseq, uint32_t &win_cap, uint32_t &tok_cap) {
    const uint32_t slack = std::max<uint32_t>(1024, page_bytes / 8);  // in-place room, see lz4i::decode_par
    win_cap = (page_bytes + slack + 15) & ~15u;
    tok_cap = max_nseq ? max_nseq : std::min<uint32_t>(page_bytes / 4 + 1, std::max<uint32_t>(64, page_bytes / 8));
}
// LDS beside the windows for LZ4I pages of up to tok_cap sequences: 2-byte token offsets
// (rounded up to 16 bytes), then a 256-byte sink for the copies of idle lanes
// (lz4i::scratch_bytes on the device side).
inline uint32_t lz4i_scratch_host(uint32_t tok_cap) { return ((2 * tok_cap + 15) & ~15u) + 256; }

// Device memory handed to one scan call in place of the context's buffers (the allocation-free
// entry point): base is 256-byte aligned; the layout is fixed by scan_run_impl.
struct ScanFixedScratch {
    uint8_t *base;
    uint64_t bytes;
};
constexpr uint64_t SCAN_FIXED_DESC = 256;                                  // the plan descriptor
constexpr uint64_t SCAN_FIXED_ITEMS = SCAN_FIXED_DESC + 1024;              // then the work items
int scan_run_impl(golap_ctx *ctx, const golap_scan_plan *plan, golap_star_result *res, int32_t *group_keys,
                  int64_t *sums, uint64_t *counts, uint64_t max_groups, const ScanFixedScratch *fixed);
// Looks up the allocation-free scan's device array on ctx's device (which must be current) and
// keeps its address in ctx (ctx->fixed_err on failure). Called once by golap_ctx_create.
void fixed_scratch_init(golap_ctx *ctx);

}  // namespace golap_impl
