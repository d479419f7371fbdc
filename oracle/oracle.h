/*
 * CPU oracle for the compressed-chunk -> filter -> join-probe -> int64 aggregate path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker: only tests/, the smoke() entry
 * point and bench.py's cpu_baseline leg may load it. The product library
 * (golap_public_amd/) never links, loads or calls it.
 *
 * Every function restates the reference algorithm it names (file:line in
 * nboeschen/golap_public). Parity pins per function are listed in oracle/README.md.
 */
#ifndef GOLAP_ORACLE_H
#define GOLAP_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ SSB generator
 * Restates SSB_Tables_col::populate_* (usecases/ssb/data/SSB_def.hpp:179-399) with the LCG
 * util::uniform_int (core/util.hpp:177-181), seed 0xBEE5BEE5 (SSB_def.hpp:416), and
 * partkey_to_price (usecases/ssb/data/SSB_helper.hpp:67-73). */
typedef struct {
    uint64_t lineorder_rows;
    uint64_t q1[3];          /* Q1.1, Q1.2, Q1.3: SUM(lo_extendedprice*lo_discount) */
    uint64_t q2_sum[3];      /* Q2.1..Q2.3: sum over groups of SUM(lo_revenue) */
    uint64_t q2_groups[3];   /* Q2.1..Q2.3: number of (d_year, p_brand1) groups */
    uint64_t q34_sum[7];     /* Q3.1..Q3.4 SUM(lo_revenue), Q4.1..Q4.3 SUM(lo_revenue - lo_supplycost) */
    uint64_t q34_groups[7];  /* their numbers of groups */
} oracle_ssb_golden;

/* Streams the whole SSB database at scale factor sf and evaluates Q1.x..Q4.x on the fly. */
int oracle_ssb_golden_run(uint32_t sf, oracle_ssb_golden *out);

/* Table sizes of SSB_Tables_col (SSB_def.hpp:144-153); lineorder rows are only known after
 * generation (returned by oracle_ssb_generate). */
void oracle_ssb_sizes(uint32_t sf, uint64_t *orders, uint64_t *parts, uint64_t *suppliers,
                      uint64_t *customers, uint64_t *dates);

/* Columns of the generated database; any pointer may be NULL. Fact arrays hold orders*7 rows
 * (an upper bound; the row count is returned). Dimension arrays: d_* [2555]; p_* [parts]:
 * p_category = num1*10+num2, p_brand1 = num1*1000+num2*100+num3, p_mfgr = num1; s_* [suppliers]
 * and c_* [customers]: region = index into AFRICA, AMERICA, ASIA, EUROPE, MIDDLE EAST, nation =
 * index into the sorted nation names, city = nation*10 + city code. */
typedef struct {
    int32_t *lo_orderdate;
    uint8_t *lo_discount, *lo_quantity;
    uint32_t *lo_extendedprice, *lo_partkey, *lo_suppkey, *lo_custkey, *lo_revenue, *lo_supplycost;
    int32_t *d_key;
    int16_t *d_year;
    int32_t *d_yearmonthnum;
    int16_t *d_weeknuminyear;
    int16_t *p_category, *p_brand1;
    uint8_t *p_mfgr;
    uint8_t *s_region, *s_nation;
    int16_t *s_city;
    uint8_t *c_region, *c_nation;
    int16_t *c_city;
} oracle_ssb_cols;
uint64_t oracle_ssb_generate(uint32_t sf, const oracle_ssb_cols *out);

/* ------------------------------------------------------------------ synthetic data
 * Counter-based generator (SplitMix64 of (seed, column, row)) for the star-schema workloads.
 * value = lo + ((splitmix64(seed ^ col*GOLDEN ^ row*MIX) >> 32) * (hi - lo + 1) >> 32). */
uint64_t oracle_splitmix64(uint64_t x);
uint32_t oracle_uniform(uint64_t seed, uint32_t col, uint64_t row, uint32_t lo, uint32_t hi);
/* Q1 fact columns: date_key uniform over date_keys[0..n_dates), discount U[0,10], quantity
 * U[1,50], price_cents U[90000,10500000]. Rows [row0, row0+n). */
void oracle_gen_q1_fact(uint64_t seed, uint64_t row0, uint64_t n, const int32_t *date_keys,
                        uint32_t n_dates, int32_t *date_key, uint8_t *discount,
                        uint8_t *quantity, int32_t *price);
/* Date dimension of n_days consecutive days starting 1992-01-01: key yyyymmdd, year, month. */
void oracle_gen_date_dim(uint32_t n_days, int32_t *key, int16_t *year, uint8_t *month);
/* Star-schema fact: date_key sorted (day floor(r*n_dates/n_total)), cust/supp/part keys uniform
 * over the dimension sizes (columns 10, 11, 12), revenue U[100,5000000] (column 13). */
void oracle_gen_star_fact(uint64_t seed, uint64_t row0, uint64_t n, uint64_t n_total,
                          const int32_t *date_keys, uint32_t n_dates, uint32_t n_cust,
                          uint32_t n_supp, uint32_t n_part, int32_t *date_key, int32_t *cust_key,
                          int32_t *supp_key, int32_t *part_key, int32_t *revenue);
/* Trip table rows [row0, row0+n) of n_total (definition in include/golap_hip.h, golap_gen_trip). */
void oracle_gen_trip(uint64_t seed, uint64_t row0, uint64_t n, uint64_t n_total, int64_t *pickup_ts,
                     uint8_t *passenger, uint16_t *zone, int32_t *fare, int32_t *distance);
/* out[i] uniform over [lo, hi] (column col), rows [0, n). */
void oracle_gen_uniform_i32(uint64_t seed, uint32_t col, uint64_t n, int32_t lo, int32_t hi,
                            int32_t *out);

/* ------------------------------------------------------------------ codecs (CPU decode)
 * LZ4 block decode with the semantics the reference's CPU codec relies on:
 * LZ4_decompress_safe_partial(src, dst, csize, cap, cap) (core/comp_cpu.hpp:22-32).
 * Returns the number of bytes written, or -1 on a malformed/overrunning block. */
int64_t oracle_lz4_decode(const uint8_t *src, uint64_t csize, uint8_t *dst, uint64_t cap);
/* An LZ4I page (parse index + LZ4 block); -1 when the index disagrees with the block. */
int64_t oracle_lz4i_decode(const uint8_t *src, uint64_t csize, uint8_t *dst, uint64_t cap);
/* Bit-packed frame-of-reference delta page and RLE page decoders (format in oracle/README.md).
 * Return rows written or -1. */
int64_t oracle_dbp_decode(const uint8_t *src, uint64_t csize, uint8_t *dst, uint64_t cap);
int64_t oracle_rle_decode(const uint8_t *src, uint64_t csize, uint8_t *dst, uint64_t cap);
int64_t oracle_for_decode(const uint8_t *src, uint64_t csize, uint8_t *dst, uint64_t cap);
/* Decodes one page of any codec id (0 raw, 1 lz4, 2 delta-bitpack, 3 rle, 4 frame of reference). */
int64_t oracle_page_decode(uint32_t codec, const uint8_t *src, uint64_t csize, uint8_t *dst,
                           uint64_t cap);

/* Decodes a whole paged column (chunk catalog layout in oracle/README.md) into dst, which holds
 * chunks*chunk_rows*elem bytes; the tail of the last chunk is zero-filled like the reference's
 * zero-padded last chunk (core/apps.cuh:185-187). Returns 0 or -1. */
int oracle_decode_column(const uint8_t *comp, const uint64_t *page_off, const uint32_t *page_csize,
                         const uint8_t *page_codec, uint64_t n_rows, uint32_t elem,
                         uint32_t page_rows, uint32_t chunk_rows, uint8_t *dst);

/* ------------------------------------------------------------------ hash map (chained)
 * HostHashMap (core/host_structs.hpp:9-78): head[key % ht_size] = rid+1, chain[rid] = prev.
 * Build is serial here (deterministic chain order); probe returns the first rid on the chain
 * whose key equals, else UINT64_MAX. */
typedef struct {
    uint64_t ht_size;
    uint64_t *head;
    uint64_t *chain;
    const int32_t *keys;
} oracle_hashmap;
int oracle_hashmap_init(oracle_hashmap *m, uint64_t ht_size, const int32_t *keys);
void oracle_hashmap_free(oracle_hashmap *m);
void oracle_hashmap_insert(oracle_hashmap *m, uint64_t rid, int64_t key);
uint64_t oracle_hashmap_probe(const oracle_hashmap *m, int64_t key);

/* ------------------------------------------------------------------ predicates
 * A conjunction of inclusive ranges over dimension columns, the PredInfo/ConstPred chain of
 * hash_map_build_pred_other (core/join.cuh:107-135, core/dev_structs.cuh:21-33) generalised to
 * lo <= v <= hi (ConstPred C is lo == hi == C). Column types: 1 i8, 2 u8, 3 i16, 4 u16,
 * 5 i32, 6 u32, 7 i64. */
typedef struct {
    const void *col;
    uint32_t type;
    int64_t lo, hi;
} oracle_pred;
int oracle_pred_eval(const oracle_pred *p, uint64_t row);

/* ------------------------------------------------------------------ plans
 * Q1-style plan of query1inmem (usecases/ssb/queries/01inmem.hpp:51-141) and pipeline_q1_lineorder
 * (usecases/ssb/queries/01.cuh:26-54): build a chained map over the dimension rows passing
 * dim_preds, then for each fact row with disc in [dlo,dhi], qty in [qlo,qhi], price in [plo,phi]
 * and a probe hit, add price*disc (int64). threads>1 uses thread-local partials merged at the
 * end (01inmem.hpp:109-131). Returns the sum; *matches gets the qualifying-row count. */
typedef struct {
    int64_t disc_lo, disc_hi, qty_lo, qty_hi, price_lo, price_hi;
} oracle_q1_consts;
int64_t oracle_q1_plan(const int32_t *dim_key, uint64_t n_dim, const oracle_pred *dim_preds,
                       uint32_t n_dim_preds, const int32_t *fact_date, const uint8_t *disc,
                       const uint8_t *qty, const int32_t *price, uint64_t n_fact,
                       const oracle_q1_consts *c, uint32_t threads, uint64_t *matches);

/* Full CPU path of the Q1 plan over LZ4 paged columns: per chunk decode (comp_cpu.hpp:101-135)
 * of the four columns, then the plan above, chunk-parallel over threads. */
typedef struct {
    const uint8_t *comp;
    const uint64_t *page_off;
    const uint32_t *page_csize;
    const uint8_t *page_codec;
} oracle_colref;
int64_t oracle_q1_compressed(const int32_t *dim_key, uint64_t n_dim, const oracle_pred *dim_preds,
                             uint32_t n_dim_preds, const oracle_colref *cols /* date,disc,qty,price */,
                             uint64_t n_fact, uint32_t page_rows, uint32_t chunk_rows,
                             uint64_t chunk_begin, uint64_t chunk_end,
                             const oracle_q1_consts *c, uint32_t threads, uint64_t *matches);

/* Decode timing (the CPU baseline's decode lines): every page of n_cols row-aligned paged
 * columns (elem[c] bytes per value) decoded on `threads` threads into per-thread buffers, by the
 * oracle's decoders, or with use_liblz4 by liblz4's LZ4_decompress_safe_partial for LZ4 and
 * LZ4I pages (core/comp_cpu.hpp:22-32; liblz4.so.1 loaded at run time). *checksum: a
 * position-weighted sum of the decoded words, equal for both decoders on the same pages;
 * *seconds: wall time of the decode. Returns 0, -1 if a page fails, -2 without liblz4. */
int oracle_decode_timing(const oracle_colref *cols, const uint32_t *elem, uint32_t n_cols, uint64_t n_rows,
                         uint32_t page_rows, uint32_t use_liblz4, uint32_t threads, uint64_t *checksum,
                         double *seconds);

/* Star join + group-by, the plan of 02inmem.hpp:58-206 (query2 CPU) / pipeline_q2 (02.cuh:102-193):
 * n_dims dimensions, each with a key column, predicate list and a group attribute column
 * (int32, or NULL for none); fact FK columns fk[d]; value column val (int32 or uint32 as given).
 * A row qualifies if every dim probe hits. Group key = tuple of the group attributes of the
 * dims with group_attr != NULL, in dim order. Output groups are sorted ascending by key tuple;
 * returns the number of groups, or -1 if it exceeds max_groups. keys_out is
 * [n_groups][n_group_attrs] int32, sums_out int64, counts_out uint64. */
typedef struct {
    const int32_t *key;
    uint64_t n;
    const oracle_pred *preds;
    uint32_t n_preds;
    const int32_t *group_attr;
} oracle_dim;
int64_t oracle_star_groupby(const oracle_dim *dims, uint32_t n_dims, const int32_t *const *fk,
                            const int32_t *val, uint64_t n_fact, uint32_t threads,
                            uint64_t max_groups, int32_t *keys_out, int64_t *sums_out,
                            uint64_t *counts_out);

/* HashAggregate::add semantics (core/dev_structs.cuh:209-244) for a plain group-by over up to
 * 4 int32 key columns with an int64 SUM and a COUNT; canonical (sorted) output as above. */
int64_t oracle_groupby(const int32_t *const *keys, uint32_t n_keys, const int64_t *val,
                       uint64_t n, uint64_t max_groups, int32_t *keys_out, int64_t *sums_out,
                       uint64_t *counts_out);

/* Scan + filter + group-by over typed columns (the trip-table plans). A row qualifies if
 * lo <= col <= hi for every predicate (inclusive bounds like the reference's range predicates,
 * usecases/ssb/queries/01.cuh:26-54). Its group is the mixed-radix index (first key most
 * significant) of k_i = floor((col_i - base_i) / div_i), each required in [0, card_i): the
 * grouping of group_by_agg_2col/4col with HashAggregate::add (core/group.cuh:14-36,
 * core/dev_structs.cuh:209-244) over computed attributes. SUM(value column) as wrapping int64
 * and COUNT per group; val_col < 0 sums nothing. Output: groups with rows in index order, key
 * tuple out_lo_i + k_i. Returns the number of groups, -1 if a qualifying row's key is out of
 * range, -2 if there are more than max_groups groups. types: golap_type ids (1..7). */
typedef struct {
    uint32_t col;
    int64_t lo, hi;
} oracle_scan_pred;
typedef struct {
    uint32_t col;
    int64_t base;
    uint32_t div, card;
    int32_t out_lo;
} oracle_scan_key;
int64_t oracle_scan_groupby(uint64_t n, const void *const *cols, const uint32_t *types,
                            const oracle_scan_pred *preds, uint32_t n_preds, const oracle_scan_key *keys,
                            uint32_t n_keys, int32_t val_col, uint32_t threads, uint64_t max_groups,
                            int32_t *keys_out, int64_t *sums_out, uint64_t *counts_out);

/* ------------------------------------------------------------------ zone maps
 * MinMaxMeta::check_pred (core/metadata.cuh:195-211) / check_mmmeta (:539-547): a chunk qualifies
 * iff lo <= max && hi >= min. out[c] = 1/0. combine_and (:632-636) is out[c] &= flag. */
void oracle_minmax_check_i64(const int64_t *mins, const int64_t *maxs, uint64_t n_chunks,
                             int64_t lo, int64_t hi, uint16_t *out);
/* Per-chunk min/max of an int column (MinMaxMeta::init_chunk, metadata.cuh:157-160). */
void oracle_minmax_build_i32(const int32_t *col, uint64_t n, uint64_t chunk_rows, int64_t *mins,
                             int64_t *maxs);

/* ------------------------------------------------------------------ chunk metadata (meta.c)
 * EqHistogram (core/metadata.cuh:225-382) and BloomMeta (:386-537) with fill_hist / check_hist /
 * fill_bloom / check_bloom (:561-603), over int64 column values in chunks of chunk_rows. */
void oracle_eqhist_build(const int64_t *v, uint64_t n, uint64_t chunk_rows, uint64_t bucket_num, int64_t *mins,
                         int64_t *maxs, uint64_t *hist);
void oracle_eqhist_check(const int64_t *mins, const int64_t *maxs, const uint64_t *hist, uint64_t n_chunks,
                         uint64_t bucket_num, int64_t lo, int64_t hi, uint16_t *out);
void oracle_bloom_params(double p, uint64_t m, uint64_t *n, uint64_t *k);
void oracle_bloom_build(const int64_t *v, uint64_t n, uint64_t chunk_rows, uint64_t m, uint64_t k, uint32_t *state);
void oracle_bloom_check(const uint32_t *state, uint64_t n_chunks, uint64_t m, uint64_t k, int64_t value,
                        uint16_t *out);

#ifdef __cplusplus
}
#endif
#endif
