/*
 * golap_hip.h - C ABI of the MI355X-native in-data-path query pipeline.
 *
 * The reference (nboeschen/golap_public) exposes this path only as header-only CUDA/C++
 * templates (core/apps.cuh, core/comp.cuh, core/dev_structs.cuh, core/join.cuh, core/group.cuh,
 * core/metadata.cuh, usecases/ssb/queries/). This header is the binary boundary beneath a C++
 * facade (include/golap/golap.hpp) that keeps the reference's names. Each entry point names the
 * reference interface it replaces.
 *
 * Conventions: every function returns an int status (GOLAP_OK or a negative GOLAP_E* code; the
 * reference aborts via checkCudaErrors instead). Handles are opaque, owned by the caller and
 * freed explicitly. Pointers named dev_* are device (HBM) addresses obtained from
 * golap_dev_alloc; all others are host addresses. A context is bound to one device and may be
 * used from several threads; calls on one context are serialised internally.
 */
#ifndef GOLAP_HIP_H
#define GOLAP_HIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GOLAP_API_VERSION 1

enum golap_status {
    GOLAP_OK = 0,
    GOLAP_EINVAL = -1,    /* bad argument or handle */
    GOLAP_ENOMEM = -2,    /* host or device allocation failed */
    GOLAP_EHIP = -3,      /* HIP runtime error (message in golap_last_error) */
    GOLAP_ECORRUPT = -4,  /* a compressed page failed its bounds checks while decoding */
    GOLAP_ECAPACITY = -5, /* a fixed-capacity structure (hash table, group table) overflowed */
    GOLAP_EBUDGET = -6,   /* admission refused: request exceeds the HBM budget */
    GOLAP_ENODEV = -7     /* no HIP device */
};

/* Value types of columns and predicates. */
enum golap_type {
    GOLAP_I8 = 1, GOLAP_U8 = 2, GOLAP_I16 = 3, GOLAP_U16 = 4,
    GOLAP_I32 = 5, GOLAP_U32 = 6, GOLAP_I64 = 7
};

/* Page codecs (CompInfo::comp_algo, core/comp.cuh:35-78). */
enum golap_codec {
    GOLAP_CODEC_RAW = 0, GOLAP_CODEC_LZ4 = 1, GOLAP_CODEC_DBP = 2, GOLAP_CODEC_RLE = 3,
    GOLAP_CODEC_FOR = 4,   /* bit-packed frame of reference (point-readable) */
    GOLAP_CODEC_AUTO = 255 /* smallest of LZ4/DBP/RLE/FOR per page */
};

typedef struct golap_ctx golap_ctx;
typedef struct golap_host_column golap_host_column;
typedef struct golap_column golap_column;
typedef struct golap_hashmap golap_hashmap;
typedef struct golap_stream golap_stream;
typedef struct golap_event golap_event;
typedef struct golap_hashagg golap_hashagg;

/* ------------------------------------------------------------------ library / context */
int golap_api_version(void);
const char *golap_status_string(int status);

/* Replaces cudaSetDevice + the implicit per-process CUDA context (01.hpp:61). */
int golap_ctx_create(int device, golap_ctx **out);
int golap_ctx_destroy(golap_ctx *ctx);
/* Replaces util::Log::error_fmt (core/util.hpp:67-144): the last error of this context. */
const char *golap_last_error(golap_ctx *ctx);
/* cudaStreamSynchronize of the context's stream (core/core.hpp:110 finish_step); also reports the
 * device errors of work enqueued there without a wait (GOLAP_ECORRUPT / GOLAP_EINVAL). */
int golap_sync(golap_ctx *ctx);
/* Number of compute units and the device name (for reports). */
int golap_device_info(golap_ctx *ctx, int *cus, char *name, int name_len);

/* ------------------------------------------------------------------ streams and events
 * CStream / CEvent (core/core.hpp:192-235): a worker's stream and the events the Op protocol
 * passes between streams (Op::step(stream, parent_event), Op::event, core/core.hpp:94-124;
 * Collector::step's cudaStreamWaitEvent fan-in, core.hpp:252-307). A stream has its own device
 * error word: work enqueued on it (golap_decompress_chunks_async, golap_hashagg_add) reports
 * corrupt pages or out-of-range keys at golap_stream_sync. In every call taking a stream, NULL
 * means the context's own stream (where the plan entry points run), so plans can wait on
 * worker streams through events. Streams have the default priority. */
int golap_stream_create(golap_ctx *ctx, golap_stream **out);
int golap_stream_destroy(golap_stream *s);
/* cudaStreamSynchronize; then GOLAP_ECORRUPT / GOLAP_EINVAL for errors raised on the stream
 * (the error word is cleared). */
int golap_stream_sync(golap_stream *s);
int golap_event_create(golap_ctx *ctx, golap_event **out);
int golap_event_destroy(golap_event *e);
/* cudaEventRecord(e, s) / cudaStreamWaitEvent(s, e) / cudaEventSynchronize(e). */
int golap_event_record(golap_event *e, golap_stream *s);
int golap_stream_wait_event(golap_ctx *ctx, golap_stream *s, golap_event *e);
int golap_event_sync(golap_event *e);
/* cudaMemcpyAsync on a stream (any direction; pinned host memory keeps it asynchronous). */
int golap_stream_copy(golap_ctx *ctx, golap_stream *s, void *dst, const void *src, uint64_t bytes);

/* ------------------------------------------------------------------ device memory
 * DeviceMem / MirrorMem RAII (core/mem.hpp:136-371) and the global DEVICE_ALLOCATED counter
 * (core/mem.hpp:14-15). */
int golap_dev_alloc(golap_ctx *ctx, uint64_t bytes, void **dev_ptr);
int golap_dev_free(golap_ctx *ctx, void *dev_ptr);
int golap_h2d(golap_ctx *ctx, void *dev_dst, const void *src, uint64_t bytes);
int golap_d2h(golap_ctx *ctx, void *dst, const void *dev_src, uint64_t bytes);
int golap_memset(golap_ctx *ctx, void *dev_dst, int value, uint64_t bytes);
/* DataCopy (core/access.hpp:16-67): device-to-device copy on the context's stream. */
int golap_d2d(golap_ctx *ctx, void *dev_dst, const void *dev_src, uint64_t bytes);
/* No counterpart in the reference: a streaming read of `bytes` (16-byte aligned) of device
   memory by blocks_per_cu workgroups per CU; *ms = its device time. The HBM read bandwidth a
   plain read kernel reaches, reported beside the fused scans' roofline fraction. */
int golap_read_probe(golap_ctx *ctx, const void *dev_buf, uint64_t bytes, uint32_t blocks_per_cu, double *ms);
uint64_t golap_device_allocated(golap_ctx *ctx);

/* ------------------------------------------------------------------ write path
 * prepare_compressed (core/apps.cuh:148-255): the column is cut into row-aligned chunks and
 * pages, each page coded independently, the catalog (BlockInfo, core/comp.cuh:23-32) and the
 * per-chunk min/max zone map (MinMaxMeta::init_chunk, core/metadata.cuh:157-160) recorded. */
typedef struct {
    uint32_t codec;          /* golap_codec */
    uint32_t page_rows;      /* rows per independently coded page */
    uint32_t chunk_rows;     /* rows per chunk; a multiple of page_rows */
    uint32_t lz4_min_match;  /* shortest match the LZ4 encoder emits (>= 4) */
    uint32_t threads;        /* host encoder threads, 0 = all hardware threads */
    uint32_t lz4_unit;       /* LZ4 sequences in whole units of this many bytes: 1 (the default,
                              * byte-granular as LZ4_compress_default, core/comp_cpu.hpp:30), 2, 4;
                              * 0 = the element width, 8-byte elements in 4-byte units, like the
                              * reference's typed nvcomp LZ4 (core/apps.cuh:262-270,
                              * core/comp.cuh:230-236) */
} golap_encode_opts;

/* Default options: LZ4, 4096-row pages, 262,144-row chunks (1 MiB of an int32 column, the
 * reference's tuples_per_chunk = chunk_bytes / max sizeof, usecases/ssb/queries/01.hpp:62-63). */
void golap_encode_opts_default(golap_encode_opts *opts);

int golap_encode(const void *host_values, uint64_t n_rows, uint32_t type,
                 const golap_encode_opts *opts, golap_host_column **out);
/* Assemble a column from pages coded elsewhere (e.g. by the reference's CPU codec liblz4);
 * pages[i] holds csize[i] bytes coded with codec[i]. No statistics are known, so every chunk's
 * zone map is [INT64_MIN, INT64_MAX] and no chunk is ever pruned. */
int golap_host_column_assemble(uint32_t type, uint64_t n_rows, uint32_t page_rows,
                               uint32_t chunk_rows, const uint8_t *const *pages,
                               const uint32_t *csize, const uint8_t *codec,
                               golap_host_column **out);
int golap_host_column_free(golap_host_column *col);

/* Catalog of an encoded column. For a host column the arrays are host memory; for a device
 * column (golap_column_info) they are device memory. */
typedef struct {
    uint64_t n_rows, n_pages, n_chunks;
    uint64_t comp_bytes;    /* bytes of the column buffer (pages + alignment gaps) */
    uint64_t payload_bytes; /* sum of page sizes */
    uint64_t uncomp_bytes;  /* n_rows * elem */
    uint32_t type, elem, page_rows, chunk_rows;
    const uint8_t *comp;
    const uint64_t *page_off;
    const uint32_t *page_csize;
    const uint8_t *page_codec; /* codec | PAGE_LITERAL_ONLY (0x80) */
    const int64_t *chunk_min;
    const int64_t *chunk_max;
} golap_column_desc;
int golap_host_column_info(const golap_host_column *col, golap_column_desc *out);

/* Upload an encoded column to HBM (prepare_compressed_device's device-resident result,
 * core/apps.cuh:257-275; the configs keep compressed chunks resident instead of on SSD). */
int golap_column_upload(golap_ctx *ctx, const golap_host_column *col, golap_column **out);

/* ------------------------------------------------------------------ column files, SSD -> HBM
 * DenseWriter (core/access.hpp:116-157) / .dat serialization (core/hl/serialization.hpp:60-126)
 * analogue: a column file holds a header, the catalog and zone maps, then the page data as one
 * 4 KiB-aligned region. golap_host_column_load reads it back to the host.
 * golap_column_load_file: the SSD -> CPU -> GPU path (SSD2CPU2GPU, core/apps.cuh:315-320;
 * VarLoader, core/access.hpp:189-435): O_DIRECT preads of segment_bytes (rounded up to 4 KiB;
 * 0 = 64 MiB) into a ring of n_buffers (2..16; 0 = 4) pinned buffers on a reader thread,
 * overlapped with hipMemcpyAsync of the filled buffers into the new device column. File
 * systems without O_DIRECT are read with buffered preads (stats->direct_io = 0). */
typedef struct {
    uint64_t bytes;       /* page data moved to HBM */
    uint64_t segments;
    uint32_t direct_io;   /* 1: O_DIRECT reads */
    double read_ms;       /* time inside preads (reader thread) */
    double total_ms;      /* open to device column ready */
} golap_load_stats;
int golap_host_column_save(const golap_host_column *col, const char *path);
int golap_host_column_load(const char *path, golap_host_column **out);
int golap_column_load_file(golap_ctx *ctx, const char *path, uint64_t segment_bytes, uint32_t n_buffers,
                           golap_column **out, golap_load_stats *stats);
int golap_column_free(golap_ctx *ctx, golap_column *col);
int golap_column_info(const golap_column *col, golap_column_desc *out);
/* Encode + upload; *comp_ms receives the host encode time like prepare_compressed_device's
 * return value. */
int golap_prepare_compressed_device(golap_ctx *ctx, const void *host_values, uint64_t n_rows,
                                    uint32_t type, const golap_encode_opts *opts,
                                    golap_column **out, double *comp_ms);

/* prepare_compressed's compression on the device (ChunkCompressor, core/comp.cuh:81-149, in
 * the chain of core/apps.cuh:148-255; nvcomp there): n_rows values of `type` already in HBM
 * (dev_values, 16-byte aligned) coded into a new device column, zone maps included. The column
 * is byte-identical to golap_encode + golap_column_upload of the same values with the same
 * options: same codec per page, same LZ4 parse, same layout. Pages up to 64 KiB (1 MiB for RAW
 * and RLE). *kernel_ms = device time of the coding, zone-map and placement kernels. */
int golap_encode_device(golap_ctx *ctx, const void *dev_values, uint64_t n_rows, uint32_t type,
                        const golap_encode_opts *opts, golap_column **out, double *kernel_ms);

/* ------------------------------------------------------------------ decompression
 * ChunkDecompressor::step (core/comp.cuh:171-195), nvcomp decompress (core/comp.cuh:186):
 * decode chunks [chunk_begin, chunk_end) into dev_out, which holds
 * (chunk_end-chunk_begin)*chunk_rows*elem bytes; rows past n_rows are zero, like the
 * reference's zero-padded last chunk (core/apps.cuh:185-187). GOLAP_ECORRUPT if any page fails
 * its bounds checks. *kernel_ms (may be NULL) gets the device time. */
int golap_decompress_chunks(golap_ctx *ctx, const golap_column *col, uint64_t chunk_begin,
                            uint64_t chunk_end, void *dev_out, double *kernel_ms);
/* The same decode enqueued on stream s without waiting for it (ChunkDecompressor::step on a
 * worker's stream, core/comp.cuh:171-195, core/apps.cuh:347): corrupt pages are reported by
 * golap_stream_sync(s). Concurrent calls on different streams of one context may overlap on the
 * device. */
int golap_decompress_chunks_async(golap_ctx *ctx, const golap_column *col, uint64_t chunk_begin,
                                  uint64_t chunk_end, void *dev_out, golap_stream *s);

/* ------------------------------------------------------------------ predicates
 * PredInfo / ConstPred (core/dev_structs.cuh:21-33) generalised to an inclusive range:
 * a row passes iff lo <= col[row] <= hi (ConstPred<C> is lo == hi == C). */
typedef struct {
    const void *dev_col;
    uint32_t type;
    int64_t lo, hi;
} golap_pred;

/* ------------------------------------------------------------------ hash join
 * HashMap (core/dev_structs.cuh:36-132). The device table is open addressing over 8-byte
 * slots {int32 key, int32 payload}; probe semantics match HashMap::probe for unique build keys:
 * the rid of the build row with that key, or UINT64_MAX. ht_size is the reference's
 * constructor argument (number of build rows the map must hold). */
int golap_hashmap_create(golap_ctx *ctx, uint64_t ht_size, const void *dev_keys,
                         uint32_t key_type, golap_hashmap **out);
/* hash_map_build_pred_other (core/join.cuh:107-135) over rows [0, n_rows) of the key column:
 * a row is inserted iff every predicate passes. dev_payload (int32 per build row, or NULL)
 * is stored beside the key and returned by the fused plans' probes; NULL stores the rid.
 * GOLAP_ECAPACITY if more than ht_size rows qualify. */
int golap_hashmap_build(golap_ctx *ctx, golap_hashmap *map, uint64_t n_rows,
                        const golap_pred *preds, uint32_t n_preds, const int32_t *dev_payload);
/* HashMap::probe for n keys (hash_join_rids, core/join.cuh:81-97, without compaction):
 * dev_out_rid[i] = rid or UINT64_MAX. */
int golap_hashmap_probe(golap_ctx *ctx, const golap_hashmap *map, const void *dev_probe_keys,
                        uint32_t key_type, uint64_t n, uint64_t *dev_out_rid);
/* hash_join_count (core/join.cuh:45-56). */
int golap_hash_join_count(golap_ctx *ctx, const golap_hashmap *map, const void *dev_probe_keys,
                          uint32_t key_type, uint64_t n, uint64_t *hits);
/* HashMap::filled (core/dev_structs.cuh:52,98). */
int golap_hashmap_filled(golap_ctx *ctx, const golap_hashmap *map, uint64_t *filled);
int golap_hashmap_free(golap_ctx *ctx, golap_hashmap *map);

/* ------------------------------------------------------------------ zone maps
 * check_mmmeta (core/metadata.cuh:539-547): dev_flags[c] = (lo <= max[c] && hi >= min[c]).
 * combine_and (core/metadata.cuh:632-636): dev_a[c] = dev_a[c] && dev_b[c]. */
int golap_minmax_check(golap_ctx *ctx, const golap_column *col, int64_t lo, int64_t hi,
                       uint16_t *dev_flags);
int golap_combine_and(golap_ctx *ctx, uint16_t *dev_a, const uint16_t *dev_b, uint64_t n);
/* Chunk metadata beyond min/max, built on the device from a column's values (dev_values: n_rows
 * values of an integer golap_type) in chunks of chunk_rows rows, as prepare_compressed fills it
 * per chunk (core/apps.cuh:212-224).
 * EqHistogram (core/metadata.cuh:225-382): per chunk its min, max and bucket_num (1..8192)
 * equal-width bucket counts, bucket = (v - min) * bucket_num / (max - min) in double, the
 * maximum clamped to the last bucket. check_hist (:573-585): dev_flags[c] = 1 if a bucket from
 * (lo - min) * inv to (hi - min) * inv (int64 truncation as in the reference; from the last
 * bucket when lo == max) is non-empty; a constant chunk when lo <= v <= hi.
 * BloomMeta (:386-537): a (p, m) Bloom filter of m_bytes (a multiple of 4, <= 65536) per chunk
 * with n = ceil(8m ln(0.5^ln 2) / ln p) and k = ceil(8m ln 2 / n) hashes of value + i (the
 * 64-bit MurmurHash3 finalizer modulo 8m). check_bloom (:593-600): dev_flags[c] = 1 if all k
 * bits of value (taken as the column type) are set. Flags are written, not only set. */
typedef struct golap_chunkmeta golap_chunkmeta;
typedef struct {
    uint64_t n_chunks;
    uint32_t buckets;           /* histogram: buckets per chunk */
    uint64_t bloom_bytes, bloom_k, bloom_n;
    const int64_t *mins, *maxs; /* device: histogram chunk bounds */
    const uint64_t *hist;       /* device: n_chunks x buckets counts */
    const uint32_t *bloom_state;/* device: n_chunks x bloom_bytes / 4 words */
} golap_chunkmeta_desc;
int golap_eqhist_build(golap_ctx *ctx, const void *dev_values, uint32_t type, uint64_t n_rows, uint64_t chunk_rows,
                       uint32_t bucket_num, golap_chunkmeta **out);
int golap_eqhist_check(golap_ctx *ctx, const golap_chunkmeta *h, int64_t lo, int64_t hi, uint16_t *dev_flags);
int golap_bloommeta_build(golap_ctx *ctx, const void *dev_values, uint32_t type, uint64_t n_rows,
                          uint64_t chunk_rows, double p, uint64_t m_bytes, golap_chunkmeta **out);
int golap_bloommeta_check(golap_ctx *ctx, const golap_chunkmeta *b, int64_t value, uint16_t *dev_flags);
/* Prejoined metadata: SortHelper::prejoin_tables (usecases/ssb/helper.hpp:115-152) with
 * MinMaxMeta over the prejoined column (MetaFlags::META, usecases/ssb/queries/01.hpp:70-78,146-150).
 * Per chunk of the 32-bit foreign-key column fk, the min and max of the dimension attribute its
 * rows join to: the payload of map (built with the attribute as dev_payload) for every row whose
 * key the map holds. A chunk with no joining row has min > max and never qualifies. The reference
 * materialises the prejoined column on the host; only its per-chunk range is kept here.
 * golap_chunkmeta_minmax_check: dev_flags[c] = (lo <= max[c] && hi >= min[c]) (check_mmmeta,
 * core/metadata.cuh:539-547) for a min/max meta (prejoined or histogram). */
int golap_prejoin_minmax(golap_ctx *ctx, const golap_column *fk, const golap_hashmap *map, golap_chunkmeta **out);
int golap_chunkmeta_minmax_check(golap_ctx *ctx, const golap_chunkmeta *m, int64_t lo, int64_t hi,
                                 uint16_t *dev_flags);
int golap_chunkmeta_info(const golap_chunkmeta *m, golap_chunkmeta_desc *out);
int golap_chunkmeta_free(golap_ctx *ctx, golap_chunkmeta *m);

/* Join-key pruning (usecases/ssb/queries/01.hpp:246-270).
 * collect_hashmap_keys (core/dev_structs.cuh:136-150): the keys map m holds, written to
 * dev_keys in ascending order (the reference appends them in arbitrary order); *n = their
 * count. GOLAP_ECAPACITY (with *n set) when more than cap keys are held.
 * check_mmmeta_multiple_pts (core/metadata.cuh:550-559): dev_flags[c] = 1 if some key of the
 * ascending dev_keys[0, n) lies in [min[c], max[c]] of column col's zone map, else 0 (the
 * reference only sets flags, on an array its caller zeroed). */
int golap_hashmap_keys(golap_ctx *ctx, const golap_hashmap *m, int32_t *dev_keys, uint64_t cap, uint64_t *n);
int golap_minmax_check_points(golap_ctx *ctx, const golap_column *col, const int32_t *dev_keys, uint64_t n,
                              uint16_t *dev_flags);

/* ------------------------------------------------------------------ Q1 plan
 * query1 (usecases/ssb/queries/01.hpp:53-449) with pipeline_q1_lineorder
 * (usecases/ssb/queries/01.cuh:26-54) fused with decompression: over chunks
 * [chunk_begin, chunk_end) of four row-aligned columns, rows with disc, qty and price in their
 * ranges whose date key probes into date_map contribute price*disc to an int64 sum. If
 * dev_chunk_mask is set, chunks whose flag is 0 are skipped (skip_step, 01.hpp:368-378). */
typedef struct {
    const golap_column *date, *disc, *qty, *price;
    const golap_hashmap *date_map;
    int64_t disc_lo, disc_hi, qty_lo, qty_hi, price_lo, price_hi;
    uint64_t chunk_begin, chunk_end;
    const uint16_t *dev_chunk_mask;
} golap_q1_plan;
typedef struct {
    int64_t sum;
    uint64_t matches;
    uint64_t chunks_run, chunks_pruned;
    uint64_t comp_bytes;   /* compressed bytes of the non-pruned pages read */
    double kernel_ms;      /* device time of the fused kernels */
} golap_q1_result;
int golap_q1_run(golap_ctx *ctx, const golap_q1_plan *plan, golap_q1_result *res);

/* Q1 streamed from column files (query1's SSD -> GPU pipeline, usecases/ssb/queries/01.hpp:
 * 350-424, with VarLoader reads, core/access.hpp:189-260, and skip_step for pruned blocks,
 * 01.hpp:368-378): the four files (paths: date, disc, qty, price; row-aligned) are scanned in
 * windows of window_chunks chunks (0: about 64 MiB of pages per window). Reader threads
 * pread the next window into pinned buffers (O_DIRECT where the file system allows it) while
 * the current window is copied to HBM and scanned; chunks that are pruned are never read: those
 * whose host_chunk_mask flag (optional, one per chunk) is 0 and, with zone_prune set, those whose
 * disc, qty or price zone map misses the predicate. The result equals golap_q1_run over the same
 * columns resident in HBM; kernel_ms sums the windows' kernels. */
typedef struct {
    const char *paths[4];
    const golap_hashmap *date_map;
    int64_t disc_lo, disc_hi, qty_lo, qty_hi, price_lo, price_hi;
    const uint16_t *host_chunk_mask;
    uint64_t host_chunk_mask_len;  /* entries in host_chunk_mask, at least the files' chunk count */
    uint32_t window_chunks;
    uint32_t zone_prune;
} golap_q1_file_plan;
typedef struct {
    uint64_t windows;          /* windows scanned */
    uint64_t windows_skipped;  /* windows with every chunk pruned (nothing read) */
    uint64_t bytes_read;       /* bytes read from the files, 4 KiB-aligned reads included */
    uint64_t bytes_copied;     /* page bytes copied to HBM */
    uint32_t direct_io;        /* 1: every read was O_DIRECT */
    uint32_t window_chunks;    /* window size used */
    double read_ms;            /* time inside preads (reader thread) */
    double total_ms;           /* open to result */
} golap_stream_stats;
int golap_q1_run_files(golap_ctx *ctx, const golap_q1_file_plan *plan, golap_q1_result *res,
                       golap_stream_stats *stats);

/* ------------------------------------------------------------------ group-by
 * Results of the grouping entry points are dense over a declared group domain: each grouping
 * attribute a has a range [lo, lo + card); the group index is the mixed-radix number of the
 * attributes in the order given (first = most significant), so groups come out sorted by their
 * key tuple. Groups with at least one row are returned: keys [n_groups][n_attrs] int32, SUM as
 * int64 and COUNT. The product of the cards must stay below 2^31.
 * Aggregation: GOLAP_AGG_LDS keeps workgroup-private partial tables in LDS (hashed, probing a
 * few slots) and spills rows that find no slot to global int64 atomics; GOLAP_AGG_GLOBAL adds
 * every row with global atomics; GOLAP_AGG_AUTO picks. */
enum golap_agg_mode { GOLAP_AGG_AUTO = 0, GOLAP_AGG_LDS = 1, GOLAP_AGG_GLOBAL = 2 };

typedef struct {
    const void *dev_col;
    uint32_t type;
    int32_t lo;
    uint32_t card;
} golap_group_key;

/* group_by_agg_2col / group_by_agg_4col (core/group.cuh:14-36) with HashAggregate::add
 * (core/dev_structs.cuh:209-244) and an atomic SUM functor: n_keys (1..4) key columns, value
 * column (any integer type), n_rows. GOLAP_ECAPACITY if more than max_groups groups exist
 * (*n_groups is still set); GOLAP_EINVAL if a key lies outside its declared range. */
int golap_group_by(golap_ctx *ctx, const golap_group_key *keys, uint32_t n_keys, const void *dev_val,
                   uint32_t val_type, uint64_t n_rows, uint32_t agg_mode, uint64_t *n_groups,
                   int32_t *group_keys, int64_t *sums, uint64_t *counts, uint64_t max_groups);

/* HashAggregate (core/dev_structs.cuh:157-253): a group-by accumulator that stays on the device
 * across calls, as the reference's HashAggregate object outlives the per-chunk pipeline kernels.
 * keys[k] give each key's type, lo and card (dev_col is ignored here); the groups are the dense
 * mixed-radix domain of the cards (product below 2^31). golap_hashagg_add adds n_rows rows of
 * key columns dev_keys[0..n_keys) and a value column (HashAggregate::add with an atomic SUM
 * functor and a count, dev_structs.cuh:209-244, group_by_agg_*col core/group.cuh:14-36) on
 * stream s (NULL: the context's stream); a key outside its range is reported by the stream's
 * sync (golap_sync for NULL). golap_hashagg_results waits for the context's stream and copies
 * the groups with at least one row out as golap_group_by does; golap_hashagg_clear zeroes them. */
int golap_hashagg_create(golap_ctx *ctx, const golap_group_key *keys, uint32_t n_keys, uint32_t agg_mode,
                         golap_hashagg **out);
int golap_hashagg_add(golap_hashagg *h, const void *const *dev_keys, const void *dev_val, uint32_t val_type,
                      uint64_t n_rows, golap_stream *s);
int golap_hashagg_results(golap_hashagg *h, uint64_t *n_groups, int32_t *group_keys, int64_t *sums,
                          uint64_t *counts, uint64_t max_groups);
int golap_hashagg_clear(golap_hashagg *h);
int golap_hashagg_destroy(golap_hashagg *h);

/* ------------------------------------------------------------------ star join + group-by
 * query2 / query3 (usecases/ssb/queries/02.hpp:54-307, 03.hpp:79-677) with pipeline_q2 /
 * pipeline_q3 (02.cuh:102-193, 03.cuh:118-170) fused with decompression: over chunks
 * [chunk_begin, chunk_end) of row-aligned fact columns, a row qualifies if every dimension's
 * foreign key probes into its built map; the group key is made of the payload attributes of the
 * dimensions with group_card > 0 (their maps must be built with that attribute as payload), in
 * dimension order; SUM(value) as int64 and COUNT per group. */
#define GOLAP_STAR_MAX_DIMS 4
typedef struct {
    const golap_column *fk;
    const golap_hashmap *map;
    int32_t group_lo;
    uint32_t group_card;  /* 0: this dimension only filters */
} golap_star_dim;
typedef struct {
    golap_star_dim dims[GOLAP_STAR_MAX_DIMS];
    uint32_t n_dims;
    const golap_column *value;
    uint64_t chunk_begin, chunk_end;
    const uint16_t *dev_chunk_mask;
    uint32_t agg_mode;         /* golap_agg_mode */
    uint32_t lds_group_slots;  /* LDS partial-table slots (power of two); 0 = library choice */
    /* optional: SUM(value - value_minus), the two row-aligned 32-bit columns each widened by
     * its own signedness (pipeline_q4's lo_revenue - lo_supplycost, 04.cuh:99-153); NULL = SUM(value).
     * golap_star_run only (batches, co-execution and streamed plans refuse it). */
    const golap_column *value_minus;
} golap_star_plan;
typedef struct {
    uint64_t n_groups;
    uint64_t rows_matched;
    int64_t total_sum;
    uint64_t chunks_run, chunks_pruned;
    uint64_t comp_bytes;      /* compressed bytes of the non-pruned pages of the fact columns */
    double kernel_ms;
    uint32_t driver_dim;      /* dimension whose filter is pinned in LDS */
    uint32_t lds_group_slots; /* LDS partial-table slots used (0 = global atomics) */
} golap_star_result;
int golap_star_run(golap_ctx *ctx, const golap_star_plan *plan, golap_star_result *res,
                   int32_t *group_keys, int64_t *sums, uint64_t *counts, uint64_t max_groups);

/* A star plan (golap_star_run) streamed from column files in the windows of golap_q1_run_files: the fact columns
 * are files (fk_paths[d] per dimension, value_path; row-aligned), the dimension maps are device
 * maps. Each window runs golap_star_run on the window's columns; the groups of all windows are
 * added into one dense result and returned as golap_star_run returns them (kernel_ms sums the
 * windows'). Chunks whose host_chunk_mask flag (optional) is 0 are never read. */
typedef struct {
    const char *fk_paths[GOLAP_STAR_MAX_DIMS];
    const golap_hashmap *maps[GOLAP_STAR_MAX_DIMS];
    int32_t group_lo[GOLAP_STAR_MAX_DIMS];
    uint32_t group_card[GOLAP_STAR_MAX_DIMS];  /* 0: the dimension only filters */
    uint32_t n_dims;
    const char *value_path;
    const uint16_t *host_chunk_mask;
    uint64_t host_chunk_mask_len;
    uint32_t window_chunks;
    uint32_t agg_mode;
} golap_star_file_plan;
int golap_star_run_files(golap_ctx *ctx, const golap_star_file_plan *plan, golap_star_result *res,
                         golap_stream_stats *stats, int32_t *group_keys, int64_t *sums, uint64_t *counts,
                         uint64_t max_groups);

/* ------------------------------------------------------------------ CPU/GPU co-execution
 * HostHashMap (core/host_structs.hpp:9-78): a host-side table over a caller-owned int32 key
 * column (which must stay valid until the build returns). build inserts the rows of
 * [0, n_rows) that pass every predicate (inclusive bounds on host columns), storing payload[r]
 * (or the rid r when payload is NULL), with up to `threads` threads (0 = up to 16);
 * GOLAP_ECAPACITY if more than ht_size rows qualify. probe writes the stored payload (its 32
 * bits, zero-extended) or rid per key, UINT64_MAX for a miss. */
typedef struct golap_host_hashmap golap_host_hashmap;
typedef struct {
    const void *col;  /* host column */
    uint32_t type;
    int64_t lo, hi;
} golap_host_pred;
int golap_host_hashmap_create(uint64_t ht_size, const int32_t *keys, golap_host_hashmap **out);
int golap_host_hashmap_build(golap_host_hashmap *map, uint64_t n_rows, const golap_host_pred *preds,
                             uint32_t n_preds, const int32_t *payload, uint32_t threads);
int golap_host_hashmap_probe(const golap_host_hashmap *map, const int32_t *keys, uint64_t n, uint64_t *out,
                             uint64_t *hits);
int golap_host_hashmap_filled(const golap_host_hashmap *map, uint64_t *filled);
int golap_host_hashmap_free(golap_host_hashmap *map);

/* A star plan with dimension host_dim joined on the host: the query 3 variants that keep the
 * customer table on the CPU (usecases/ssb/queries/03a.hpp, 03c.hpp, 03d.hpp; 03c.hpp:189-488,
 * HostHashAggregate, core/host_structs.hpp:82-152). plan->dims[host_dim].map is ignored (may
 * be NULL); its fk column, group_lo and group_card apply, with hmap's payload as the group
 * attribute. The GPU probes every other dimension exactly and hands the passing rows
 * {host key, group index, value} to the host in rounds of chunks; host threads (`threads`,
 * 0 = up to 16) probe hmap and aggregate while the GPU runs the next round. Results, groups
 * and result fields as golap_star_run (kernel_ms = device span of the rounds); at least one
 * dimension must stay on the device. */
typedef struct {
    uint64_t rounds;
    uint64_t splits;    /* rounds halved because they passed more rows than one hand-over holds */
    uint64_t emitted;   /* fact rows handed to the host */
    double host_ms;     /* host probe + aggregation time */
    double device_ms;   /* first round's launch to the last round's end */
    double total_ms;
} golap_coexec_stats;
int golap_star_run_hostdim(golap_ctx *ctx, const golap_star_plan *plan, uint32_t host_dim,
                           const golap_host_hashmap *hmap, uint32_t threads, golap_star_result *res,
                           golap_coexec_stats *stats, int32_t *group_keys, int64_t *sums, uint64_t *counts,
                           uint64_t max_groups);

/* ------------------------------------------------------------------ scan + group-by plans
 * Filter + group-by over row-aligned typed fact columns, fused with decompression:
 *   - the range predicates of pipeline_q1_lineorder (usecases/ssb/queries/01.cuh:26-54,
 *     inclusive bounds) on any columns;
 *   - the grouping of group_by_agg_2col/4col with HashAggregate::add (core/group.cuh:14-36,
 *     core/dev_structs.cuh:209-244), here over keys computed per row:
 *     k = (value - base) / div in [0, card). For example the hour of a timestamp, or a zone id
 *     with div 1. A qualifying row whose key is outside its range is GOLAP_EINVAL;
 *   - SUM(value column) as int64 and COUNT per group, returned like golap_group_by, with each
 *     key reported as out_lo + k.
 * The product of the cards must stay below 2^31 and card * div below 2^32.
 * With zone_prune, chunks whose zone map fails a predicate on the first two predicate columns
 * are skipped (check_mmmeta semantics, core/metadata.cuh:539-547). */
#define GOLAP_SCAN_MAX_COLS 8
typedef struct {
    uint32_t col;          /* index into cols */
    int64_t lo, hi;
} golap_scan_pred;
typedef struct {
    uint32_t col;
    int64_t base;
    uint32_t div;          /* >= 1 */
    uint32_t card;
    int32_t out_lo;
} golap_scan_key;
typedef struct {
    const golap_column *cols[GOLAP_SCAN_MAX_COLS];
    uint32_t n_cols;
    golap_scan_pred preds[4];
    uint32_t n_preds;
    golap_scan_key keys[3];
    uint32_t n_keys;
    int32_t value_col;     /* index into cols, -1 = COUNT only */
    uint64_t chunk_begin, chunk_end;
    const uint16_t *dev_chunk_mask;
    uint32_t agg_mode;     /* golap_agg_mode (AUTO = LDS tables, flushed as they fill) */
    uint32_t zone_prune;
    uint32_t lds_group_slots; /* 0 = library choice */
} golap_scan_plan;
/* golap_star_result fields as for star plans; driver_dim is 0. */
int golap_scan_run(golap_ctx *ctx, const golap_scan_plan *plan, golap_star_result *res, int32_t *group_keys,
                   int64_t *sums, uint64_t *counts, uint64_t max_groups);

/* golap_scan_run without device allocation. No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). Arguments, results and errors are golap_scan_run's. The
 * device memory the call needs beyond its inputs and outputs (error word, work-list counters,
 * plan descriptor, work items, dense group result) is one library array of
 * GOLAP_NOALLOC_SCRATCH_BYTES per device, held until the process ends; a plan needing more
 * (1,280 bytes + 16 per work item + 16 per group of the key domain) fails with GOLAP_ECAPACITY
 * and runs nothing. The array keeps the last call's contents and is cleared where needed by
 * each call. All calls on one GPU share it: calls that overlap on one GPU would race, so the
 * library serialises them within a process (a call returns with its work finished). */
#define GOLAP_NOALLOC_SCRATCH_BYTES (8u << 20)
int golap_scan_run_noalloc(golap_ctx *ctx, const golap_scan_plan *plan, golap_star_result *res,
                           int32_t *group_keys, int64_t *sums, uint64_t *counts, uint64_t max_groups);

/* ------------------------------------------------------------------ batched plans
 * Many plans of one kind over one fact table in one persistent fused launch. The reference
 * runs one plan per call, with a worker loop over chunks (query2/query3, 02.hpp:54-307,
 * 03.hpp:79-677), and prunes with MetaChecker before the loop (01.hpp:305-332).
 * Each query brings:
 *   - a star query: its own dimension maps, chunk range, optional chunk mask and optional
 *     zone-map range predicate on zone_col;
 *   - a scan plan: its own predicates and keys, and zone pruning on its first two predicates.
 * The zone-map test is check_mmmeta (metadata.cuh:539-547): a chunk survives iff
 * lo <= max && hi >= min.
 * One run is three kernels:
 *   1. zero the result arena;
 *   2. build the work list: one entry per surviving (query, chunk part);
 *   3. the persistent kernel, whose workgroups take work-list items with one atomic each.
 * The three are captured once into a hipGraph that every run replays.
 * Every query's results equal golap_star_run / golap_scan_run of its plan alone. All queries'
 * fact columns (at most 8 distinct columns) must be row-aligned.
 * golap_batch_tune times candidate launch configurations and keeps the fastest (autotuner):
 *   - pages per work item (tile grain);
 *   - LDS group-table slots and the LDS-versus-global aggregation cutover;
 *   - star plans: the driver filter in LDS or in L2, and compact or full probe tables;
 *   - workgroups per CU.
 * The chosen configuration is cached in the context for batches of the same shape.
 * The maps and columns a batch reads must outlive it. A star batch fails to run once one of
 * its maps has been rebuilt. */
typedef struct golap_batch golap_batch;
typedef struct {
    golap_star_plan plan;           /* plan.lds_group_slots is ignored: see golap_batch_config */
    const golap_column *zone_col;   /* NULL: no zone-map predicate */
    int64_t zone_lo, zone_hi;
} golap_star_query;
typedef struct {
    uint32_t pages_per_item;     /* pages of one chunk per work item, 1..pages per chunk */
    uint32_t lds_group_slots;    /* LDS group table (power of two <= 4096), 0 = global atomics */
    uint32_t lds_cutover;        /* AUTO queries aggregate in LDS when their expected matching
                                    rows reach lds_cutover x their group count */
    uint32_t filter_lds;         /* star: 1 = copy each query's driver filter into LDS */
    uint32_t compact_tables;     /* star: 1 = probe the compact tables (2 x filled slots) */
    uint32_t max_blocks_per_cu;  /* 0 = occupancy limit */
    uint32_t shared_filter;      /* star: 1 = test each page's rows once for all queries against a
                                    membership table of one dimension column they all probe
                                    (available when every query has that dimension and its
                                    keys span at most 2^27 values); candidates then probe
                                    every dimension per query, aggregating in global memory */
    uint32_t rows_per_lane;      /* scan: rows of a lane per sub-tile of a page (rows per wave = 64 x
                                    this): 4 or 8, 0 = chosen from the launch size */
} golap_batch_config;
typedef struct {
    double ms;                   /* device time of the run (graph or plain launches) */
    uint64_t items;              /* work items built on the device */
    uint64_t pages;              /* page tiles of the surviving chunks, all queries */
    uint64_t comp_bytes;         /* compressed bytes of the surviving pages' columns, all queries */
    uint64_t shared_comp_bytes;  /* scan batches: the bytes read once per chunk for all queries that
                                    keep it (union of their columns); star batches: comp_bytes */
} golap_batch_run_info;
int golap_batch_create_star(golap_ctx *ctx, const golap_star_query *queries, uint32_t n_queries,
                            golap_batch **out);
int golap_batch_create_scan(golap_ctx *ctx, const golap_scan_plan *plans, uint32_t n_plans,
                            golap_batch **out);
int golap_batch_config_get(golap_batch *b, golap_batch_config *cfg);
int golap_batch_config_set(golap_batch *b, const golap_batch_config *cfg);
/* use_graph 1: replay the captured graph; 0: the same three kernels as plain launches. */
int golap_batch_run(golap_batch *b, uint32_t use_graph, golap_batch_run_info *info);
/* Times each candidate reps times (after one warm-up), keeps and caches the fastest. */
int golap_batch_tune(golap_batch *b, uint32_t reps, golap_batch_config *best, double *best_ms);
/* Results of query q of the last run, as the single-query entry points return them. */
int golap_batch_result(golap_batch *b, uint32_t q, golap_star_result *res, int32_t *group_keys,
                       int64_t *sums, uint64_t *counts, uint64_t max_groups);
/* Results of every query of the last run in one call: query q's groups are entries
 * [group_off[q], group_off[q+1]) of sums / counts (group_off has n_queries + 1 entries), its keys
 * the next n_groups x n_attrs int32 of group_keys after the previous queries' keys.
 * GOLAP_ECAPACITY (with group_off and res complete) if more than max_total_groups in all. */
int golap_batch_results(golap_batch *b, golap_star_result *res, uint64_t *group_off, int32_t *group_keys,
                        int64_t *sums, uint64_t *counts, uint64_t max_total_groups);
int golap_batch_free(golap_batch *b);

/* ------------------------------------------------------------------ serving loop
 * No counterpart in the reference, which runs one plan per call on its caller's thread.
 * A request loop over one context:
 *   - Submitted star queries and scan plans wait in arrival order.
 *   - They are admitted while the device bytes of admitted, unfinished requests (result arrays,
 *     work lists, descriptors) stay within hbm_budget.
 *   - As soon as a slot is free, the admitted requests of one kind form a batch, run as one
 *     persistent fused launch on that slot's own stream (golap_batch_*).
 *   - Up to `slots` batches are on the device at once; requests arriving meanwhile form the next
 *     batch (continuous batching).
 *   - The dimension maps and compressed columns a request names stay resident across requests
 *     (the caller builds them once).
 *   - Results are the single-query entry points' results, emitted by a pool of host threads.
 *   - golap_server_wait blocks until the request is done and reports its latency from submit
 *     to results ready.
 * A request whose own device bytes exceed hbm_budget is refused with GOLAP_EBUDGET; an invalid
 * plan is refused at submit with the single-query entry point's status. */
typedef struct golap_server golap_server;
typedef struct {
    uint64_t hbm_budget;       /* bytes; 0 = 4 GiB */
    uint32_t max_batch;        /* requests per launch; 0 = 64 */
    uint32_t slots;            /* launches in flight, one stream each; 0 = 2 */
    uint32_t emit_threads;     /* host threads that emit results; 0 = 8 */
    uint32_t batch_window_us;  /* with a free slot and fewer than max_batch requests waiting,
                                  wait this long for more arrivals before launching; 0 = none */
} golap_server_opts;
typedef struct {
    uint64_t submitted, completed, refused, batches;
    uint64_t peak_admitted_bytes;
    double mean_batch;         /* requests per launch */
    double device_ms;          /* device time of all launches (events around each batch) */
    uint64_t comp_bytes;       /* compressed bytes of the surviving pages, summed over requests */
    uint64_t shared_comp_bytes; /* the bytes the shared launches read (per batch, once per chunk) */
} golap_server_stats;
int golap_server_create(golap_ctx *ctx, const golap_server_opts *opts, golap_server **out);
int golap_server_submit_star(golap_server *s, const golap_star_query *q, uint64_t *ticket);
int golap_server_submit_scan(golap_server *s, const golap_scan_plan *p, uint64_t *ticket);
/* Blocks until the request is done; output as golap_star_run / golap_scan_run. The ticket is
 * released when this returns. latency_ms may be NULL. */
int golap_server_wait(golap_server *s, uint64_t ticket, golap_star_result *res, int32_t *group_keys,
                      int64_t *sums, uint64_t *counts, uint64_t max_groups, double *latency_ms);
int golap_server_get_stats(golap_server *s, golap_server_stats *out);
/* Finishes the submitted requests (unwaited results are dropped) and stops the loop. */
int golap_server_destroy(golap_server *s);

/* ------------------------------------------------------------------ synthetic data
 * Seeded column generators of the benchmark workloads (the reference seeds its generator with
 * util::uniform_int, core/util.hpp:177-181; these use a SplitMix64 counter hash so row ranges
 * are generated independently). Date dimension: n_days consecutive days from 1992-01-01, key
 * yyyymmdd. Q1 fact rows [row0, row0+n): date_key uniform over date_keys, discount U[0,10],
 * quantity U[1,50], price U[90000,10500000]; any output may be NULL. threads 0 = all. */
int golap_gen_date_dim(uint32_t n_days, int32_t *key, int16_t *year, uint8_t *month);
int golap_gen_q1_fact(uint64_t seed, uint64_t row0, uint64_t n, const int32_t *date_keys,
                      uint32_t n_dates, int32_t *date_key, uint8_t *discount, uint8_t *quantity,
                      int32_t *price, uint32_t threads);
/* Star-schema fact rows [row0, row0+n) of n_total: date_key sorted (row r on day
 * floor(r*n_dates/n_total)), customer/supplier/part keys uniform over [0, n_*), revenue
 * U[100, 5000000]. */
int golap_gen_star_fact(uint64_t seed, uint64_t row0, uint64_t n, uint64_t n_total,
                        const int32_t *date_keys, uint32_t n_dates, uint32_t n_cust,
                        uint32_t n_supp, uint32_t n_part, int32_t *date_key, int32_t *cust_key,
                        int32_t *supp_key, int32_t *part_key, int32_t *revenue, uint32_t threads);
/* Trip-table rows [row0, row0+n) of n_total: pickup_ts (unix seconds) sorted over the 366 days
 * of 2024, row r at GOLAP_TRIP_T0 + floor(r * GOLAP_TRIP_SPAN / n_total); passenger_count
 * Zipf(s = 1.2) over 1..6 (inverse CDF on 53 random bits); zone uniform over 1..265; fare_cents
 * lognormal(mu 7.2, sigma 0.6) and distance_m lognormal(mu 8, sigma 0.8), normal deviates by
 * Box-Muller, rounded half up, clipped to [250, 50000] and [1, INT32_MAX]. */
#define GOLAP_TRIP_T0 1704067200LL /* 2024-01-01 00:00:00 UTC */
#define GOLAP_TRIP_SPAN 31622400LL /* 366 days in seconds */
int golap_gen_trip(uint64_t seed, uint64_t row0, uint64_t n, uint64_t n_total, int64_t *pickup_ts,
                   uint8_t *passenger, uint16_t *zone, int32_t *fare, int32_t *distance, uint32_t threads);
/* out[i] uniform over [lo, hi], rows [0, n), generator column id col. */
int golap_gen_uniform_i32(uint64_t seed, uint32_t col, uint64_t n, int32_t lo, int32_t hi,
                          int32_t *out, uint32_t threads);

#ifdef __cplusplus
}
#endif
#endif
