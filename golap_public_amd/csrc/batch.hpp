// Batched plans (internal interface shared by the batch C ABI and the serving loop).
//
// A batch holds many queries of one kind (star join + group-by, or scan + group-by) over one
// fact table and runs them as three kernels on one stream:
//   1. zero the result arena;
//   2. build the work list;
//   3. the persistent fused kernel.
// A BatchSlot owns the stream, the events and the device and pinned buffers a batch runs in.
// The public batch API gives each batch its own slot on the context stream. The serving loop
// keeps a few slots with their own streams and reuses their buffers across batches, so batches
// on different slots overlap on the device.
#pragma once
#include <string>
#include <vector>

#include "common.hpp"
#include "star_host.hpp"

namespace golap_impl {

struct BatchSlot {
    golap_ctx *ctx = nullptr;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, dep = nullptr;
    uint8_t *d_arena = nullptr, *h_arena = nullptr, *d_q = nullptr;
    uint64_t *d_chunk_off = nullptr;
    uint2 *d_items = nullptr;   // star work list
    uint4 *d_items4 = nullptr;  // scan work list
    uint64_t arena_cap = 0, q_cap = 0, off_cap = 0, items_cap = 0, items4_cap = 0;
    uint8_t *d_dec = nullptr;        // scan batches: columns decoded ahead of the launch
    uint32_t *d_dec_err = nullptr;   // their decode's error word
    uint64_t dec_cap = 0;
    std::string err;

    // own_stream: a new stream (serving slots); otherwise the context stream.
    int init(golap_ctx *c, bool own);
    // Grows the buffers to at least these sizes (bytes, bytes, entries, entries, entries).
    int reserve(uint64_t arena, uint64_t qbytes, uint64_t offs, uint64_t items, uint64_t items4);
    // Grows the decode buffer to at least bytes (and allocates its error word).
    int reserve_dec(uint64_t bytes);
    void release();
};

enum class BatchKind { STAR, SCAN };

}  // namespace golap_impl

struct golap_batch {
    golap_ctx *ctx = nullptr;
    golap_impl::BatchKind kind = golap_impl::BatchKind::STAR;
    golap_impl::BatchSlot *slot = nullptr;
    bool own_slot = false;
    // star queries
    std::vector<golap_star_query> sq;
    std::vector<golap_impl::starhost::QueryHost> qh;
    std::vector<golap_impl::StarQueryDev> qd;
    uint32_t max_filter_words = 0;
    // scan queries
    std::vector<golap_impl::starhost::ScanHost> sh;
    std::vector<golap_impl::ScanQueryDev> sd;
    uint32_t win_bytes = 0;
    // fact columns of the launch and their LDS windows
    golap_impl::starhost::ColumnSet cs;
    int windows[golap_impl::STAR_MAX_COLS] = {};
    int dslots[golap_impl::STAR_MAX_COLS] = {};  // scan: register slots of delta-bitpack columns
    uint32_t wcaps[golap_impl::STAR_MAX_COLS] = {};  // scan: window bytes of parallel LZ4I decodes
    int32_t tok_off = -1;                        // scan: LZ4I token offsets in the wave's region
    uint32_t tok_cap = 0;
    uint32_t n_windows = 0;
    // star, k_star schedules: windows only for the drivers' delta-bitpack columns and other coded
    // pages, prefix-sum regions (1 + region, 0 = none) for the rest; the shared-filter schedule
    // uses `windows` above
    int pt_windows[golap_impl::STAR_MAX_COLS] = {};
    int pt_anchors[golap_impl::STAR_MAX_COLS] = {};
    uint32_t pt_n_windows = 0, pt_n_anchors = 0;
    // slices and work list
    uint32_t page_rows = 0, ppc = 0;
    uint64_t n_rows = 0, n_pages = 0;
    std::vector<uint64_t> chunk_off;
    std::vector<std::vector<uint8_t>> survive;  // host view of surviving chunks per query
    std::vector<uint64_t> surv_pages;
    uint64_t total_pages = 0, total_bytes = 0;
    // scan batches: union of the slices, surviving (chunk, plan word) pairs, bytes read once
    uint64_t union_begin = 0, union_end = 0, union_pairs = 0, shared_bytes = 0;
    uint64_t max_local = 0;  // largest local key domain of a plan on a surviving chunk
    // scan batches of several plans: LZ4I columns decoded ahead of the launch (the scan reads them
    // in place; its passes over an item do not decode them again), over these chunk runs
    uint32_t dec_mask = 0;
    std::vector<std::pair<uint64_t, uint64_t>> dec_runs;
    // star batches: fact column mcol that every query probes, and each query's membership filter
    // of its dimension on that column (64 entries per word of queries), for the shared filter
    golap_impl::DriverFilterView *d_mfilt = nullptr;
    int mcol = -1;
    uint64_t max_domain = 1;
    std::vector<uint64_t> res_off;
    // compaction of the non-empty groups (golap_batch_results), allocated on first use: per-query
    // {arena offset, domain} and first flattened index, the entries on the device and pinned
    void *d_cmeta = nullptr;
    uint8_t *d_centries = nullptr, *h_centries = nullptr;
    uint64_t c_total = 0, c_cap = 0;
    uint64_t arena_bytes = 0;
    golap_batch_config cfg{};
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::string shape;
    bool ran = false, h_valid = false;
    double last_ms = 0;
    std::string err;

    uint32_t n_queries() const { return (uint32_t)(kind == golap_impl::BatchKind::STAR ? sq.size() : sh.size()); }
};

namespace golap_impl {

// Creation validates and prepares every query (call with the context lock held). slot null:
// the batch gets its own slot on the context stream. capture: record the hipGraph (public
// batches; one-shot serving batches launch plainly).
int batch_create_star(golap_ctx *ctx, const golap_star_query *q, uint32_t n, BatchSlot *slot, bool capture,
                      golap_batch **out);
int batch_create_scan(golap_ctx *ctx, const golap_scan_plan *q, uint32_t n, BatchSlot *slot, bool capture,
                      golap_batch **out);
// Enqueues one run on the slot stream after the context stream's pending work (lock held).
int batch_enqueue(golap_batch *b, bool use_graph);
// Waits for the run and checks its error word; no context lock needed.
int batch_finish(golap_batch *b, golap_batch_run_info *info);
// Query q's groups of the last finished run; no context lock needed.
int batch_result(golap_batch *b, uint32_t q, golap_star_result *res, int32_t *keys, int64_t *sums, uint64_t *counts,
                 uint64_t max_groups);
// Device bytes a query adds to a batch (results, work items, descriptor): the serving loop's
// admission unit.
uint64_t batch_query_bytes(uint64_t domain, uint64_t slice_chunks, uint32_t pages_per_chunk, bool star);
void batch_free(golap_batch *b);

}  // namespace golap_impl
