/*
 * CPU plans (TEST INFRASTRUCTURE: the oracle, never the product).
 *
 * HostHashMap (core/host_structs.hpp:9-78), the Q1 plan of query1inmem
 * (usecases/ssb/queries/01inmem.hpp:51-141), the star-join group-by of query2 CPU
 * (usecases/ssb/queries/02inmem.hpp:58-206), HashAggregate grouping semantics
 * (core/dev_structs.cuh:209-244) and MinMaxMeta pruning (core/metadata.cuh:195-211,539-547).
 */
#include "oracle.h"
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ---------------------------------------------------------------- chained hash map */
int oracle_hashmap_init(oracle_hashmap *m, uint64_t ht_size, const int32_t *keys) {
    m->ht_size = ht_size ? ht_size : 1;
    m->head = (uint64_t *)calloc(m->ht_size, sizeof(uint64_t));
    m->chain = (uint64_t *)calloc(m->ht_size, sizeof(uint64_t));
    m->keys = keys;
    return (m->head && m->chain) ? 0 : -1;
}
void oracle_hashmap_free(oracle_hashmap *m) {
    free(m->head);
    free(m->chain);
    m->head = m->chain = NULL;
}
/* host_structs.hpp:45-55: hash = key % ht_size (key taken as uint64), exchange the head. */
void oracle_hashmap_insert(oracle_hashmap *m, uint64_t rid, int64_t key) {
    uint64_t h = (uint64_t)key % m->ht_size;
    uint64_t prev = m->head[h];
    m->head[h] = rid + 1;
    m->chain[rid] = prev;
}
/* host_structs.hpp:60-69: walk the chain from the head, return the first equal rid. */
uint64_t oracle_hashmap_probe(const oracle_hashmap *m, int64_t key) {
    uint64_t h = (uint64_t)key % m->ht_size;
    for (uint64_t hit = m->head[h]; hit > 0; hit = m->chain[hit - 1])
        if ((int64_t)m->keys[hit - 1] == key) return hit - 1;
    return UINT64_MAX;
}

/* ---------------------------------------------------------------- predicates */
static inline int64_t col_value(const void *col, uint32_t type, uint64_t row) {
    switch (type) {
    case 1: return ((const int8_t *)col)[row];
    case 2: return ((const uint8_t *)col)[row];
    case 3: return ((const int16_t *)col)[row];
    case 4: return ((const uint16_t *)col)[row];
    case 5: return ((const int32_t *)col)[row];
    case 6: return ((const uint32_t *)col)[row];
    case 7: return ((const int64_t *)col)[row];
    default: return 0;
    }
}
int oracle_pred_eval(const oracle_pred *p, uint64_t row) {
    int64_t v = col_value(p->col, p->type, row);
    return v >= p->lo && v <= p->hi;
}
static int preds_pass(const oracle_pred *ps, uint32_t n, uint64_t row) {
    for (uint32_t i = 0; i < n; ++i)
        if (!oracle_pred_eval(&ps[i], row)) return 0;
    return 1;
}

/* hash_map_build_pred_other (core/join.cuh:107-135): rid passes every predicate -> insert. The
 * map size is the dimension size (01.hpp:173 constructs HashMap(d_key.size(), ...)). */
static int build_dim_map(oracle_hashmap *m, const int32_t *key, uint64_t n, const oracle_pred *ps,
                         uint32_t np) {
    if (oracle_hashmap_init(m, n, key)) return -1;
    for (uint64_t r = 0; r < n; ++r)
        if (preds_pass(ps, np, r)) oracle_hashmap_insert(m, r, key[r]);
    return 0;
}

/* ---------------------------------------------------------------- Q1 plan */
typedef struct {
    const oracle_hashmap *map;
    const int32_t *date;
    const uint8_t *disc, *qty;
    const int32_t *price;
    uint64_t begin, end;
    const oracle_q1_consts *c;
    int64_t sum;
    uint64_t matches;
} q1_job;

/* pipeline_q1_lineorder row body (01.cuh:41-51) / query1inmem probe loop (01inmem.hpp:111-127). */
static void q1_rows(q1_job *j) {
    int64_t sum = 0;
    uint64_t matches = 0;
    const oracle_q1_consts *c = j->c;
    for (uint64_t r = j->begin; r < j->end; ++r) {
        int64_t d = j->disc[r], q = j->qty[r], p = j->price[r];
        if (d < c->disc_lo || d > c->disc_hi || q < c->qty_lo || q > c->qty_hi || p < c->price_lo ||
            p > c->price_hi)
            continue;
        if (oracle_hashmap_probe(j->map, j->date[r]) == UINT64_MAX) continue;
        sum += p * d;
        matches++;
    }
    j->sum = sum;
    j->matches = matches;
}
static void *q1_thread(void *arg) {
    q1_rows((q1_job *)arg);
    return NULL;
}

int64_t oracle_q1_plan(const int32_t *dim_key, uint64_t n_dim, const oracle_pred *dim_preds,
                       uint32_t n_dim_preds, const int32_t *fact_date, const uint8_t *disc,
                       const uint8_t *qty, const int32_t *price, uint64_t n_fact,
                       const oracle_q1_consts *c, uint32_t threads, uint64_t *matches) {
    oracle_hashmap m;
    if (build_dim_map(&m, dim_key, n_dim, dim_preds, n_dim_preds)) return 0;
    if (threads < 1) threads = 1;
    q1_job *jobs = (q1_job *)calloc(threads, sizeof(q1_job));
    pthread_t *tids = (pthread_t *)calloc(threads, sizeof(pthread_t));
    /* util::SliceSeq split (core/util.hpp:337-356): first work%workers slices get one extra. */
    uint64_t base = n_fact / threads, rest = n_fact % threads, cur = 0;
    for (uint32_t t = 0; t < threads; ++t) {
        uint64_t len = base + (t < rest ? 1 : 0);
        jobs[t] = (q1_job){&m, fact_date, disc, qty, price, cur, cur + len, c, 0, 0};
        cur += len;
        if (threads > 1) pthread_create(&tids[t], NULL, q1_thread, &jobs[t]);
        else q1_rows(&jobs[t]);
    }
    int64_t sum = 0;
    uint64_t mt = 0;
    for (uint32_t t = 0; t < threads; ++t) {
        if (threads > 1) pthread_join(tids[t], NULL);
        sum += jobs[t].sum; /* fetch_add merge of thread partials (01inmem.hpp:128) */
        mt += jobs[t].matches;
    }
    if (matches) *matches = mt;
    free(jobs);
    free(tids);
    oracle_hashmap_free(&m);
    return sum;
}

/* ---------------------------------------------------------------- Q1 over compressed chunks */
typedef struct {
    const oracle_hashmap *map;
    const oracle_colref *cols;
    uint64_t n_fact;
    uint32_t page_rows, chunk_rows;
    uint64_t chunk_begin, chunk_end;
    const oracle_q1_consts *c;
    int64_t sum;
    uint64_t matches;
    int err;
} q1c_job;

static const uint32_t Q1_ELEM[4] = {4, 1, 1, 4};

static void *q1c_thread(void *arg) {
    q1c_job *j = (q1c_job *)arg;
    uint8_t *buf[4];
    for (int k = 0; k < 4; ++k) buf[k] = (uint8_t *)malloc((uint64_t)j->chunk_rows * Q1_ELEM[k]);
    uint32_t ppc = j->chunk_rows / j->page_rows;
    for (uint64_t ch = j->chunk_begin; ch < j->chunk_end && !j->err; ++ch) {
        uint64_t r0 = ch * j->chunk_rows;
        uint64_t rows = (j->n_fact - r0 < j->chunk_rows) ? j->n_fact - r0 : j->chunk_rows;
        /* CPUChunkDecompressor::step per column (comp_cpu.hpp:107-125): decode every page. */
        for (int k = 0; k < 4; ++k) {
            for (uint32_t pi = 0; pi < ppc; ++pi) {
                uint64_t pr0 = (uint64_t)pi * j->page_rows;
                if (pr0 >= rows) break;
                uint64_t prows = rows - pr0 < j->page_rows ? rows - pr0 : j->page_rows;
                uint64_t p = ch * ppc + pi;
                const oracle_colref *cr = &j->cols[k];
                int64_t got = oracle_page_decode(cr->page_codec[p], cr->comp + cr->page_off[p],
                                                 cr->page_csize[p], buf[k] + pr0 * Q1_ELEM[k],
                                                 prows * Q1_ELEM[k]);
                if (got < 0) j->err = 1;
            }
        }
        q1_job sub = {j->map, (const int32_t *)buf[0], buf[1], buf[2], (const int32_t *)buf[3],
                      0, rows, j->c, 0, 0};
        q1_rows(&sub);
        j->sum += sub.sum;
        j->matches += sub.matches;
    }
    for (int k = 0; k < 4; ++k) free(buf[k]);
    return NULL;
}

int64_t oracle_q1_compressed(const int32_t *dim_key, uint64_t n_dim, const oracle_pred *dim_preds,
                             uint32_t n_dim_preds, const oracle_colref *cols, uint64_t n_fact,
                             uint32_t page_rows, uint32_t chunk_rows, uint64_t chunk_begin,
                             uint64_t chunk_end, const oracle_q1_consts *c, uint32_t threads,
                             uint64_t *matches) {
    if (!page_rows || chunk_rows % page_rows) return 0;
    oracle_hashmap m;
    if (build_dim_map(&m, dim_key, n_dim, dim_preds, n_dim_preds)) return 0;
    if (threads < 1) threads = 1;
    uint64_t work = chunk_end - chunk_begin;
    q1c_job *jobs = (q1c_job *)calloc(threads, sizeof(q1c_job));
    pthread_t *tids = (pthread_t *)calloc(threads, sizeof(pthread_t));
    uint64_t base = work / threads, rest = work % threads, cur = chunk_begin;
    for (uint32_t t = 0; t < threads; ++t) {
        uint64_t len = base + (t < rest ? 1 : 0);
        jobs[t] = (q1c_job){&m, cols, n_fact, page_rows, chunk_rows, cur, cur + len, c, 0, 0, 0};
        cur += len;
        pthread_create(&tids[t], NULL, q1c_thread, &jobs[t]);
    }
    int64_t sum = 0;
    uint64_t mt = 0;
    int err = 0;
    for (uint32_t t = 0; t < threads; ++t) {
        pthread_join(tids[t], NULL);
        sum += jobs[t].sum;
        mt += jobs[t].matches;
        err |= jobs[t].err;
    }
    if (matches) *matches = err ? UINT64_MAX : mt;
    free(jobs);
    free(tids);
    oracle_hashmap_free(&m);
    return sum;
}

/* ---------------------------------------------------------------- grouping */
/* Open-addressing table keyed by up to 4 int32 attributes; grouping semantics of
 * HashAggregate::add (dev_structs.cuh:209-244): equal keys share a slot, SUM and COUNT add. */
typedef struct {
    uint64_t cap, n;
    uint32_t nk;
    int32_t *keys;
    int64_t *sums;
    uint64_t *counts;
    uint8_t *used;
} gtab;

static int gtab_init(gtab *g, uint32_t nk, uint64_t expect) {
    uint64_t cap = 16;
    while (cap < expect * 2) cap <<= 1;
    g->cap = cap;
    g->n = 0;
    g->nk = nk;
    g->keys = (int32_t *)calloc(cap * (nk ? nk : 1), sizeof(int32_t));
    g->sums = (int64_t *)calloc(cap, sizeof(int64_t));
    g->counts = (uint64_t *)calloc(cap, sizeof(uint64_t));
    g->used = (uint8_t *)calloc(cap, 1);
    return (g->keys && g->sums && g->counts && g->used) ? 0 : -1;
}
static void gtab_free(gtab *g) {
    free(g->keys);
    free(g->sums);
    free(g->counts);
    free(g->used);
}
static uint64_t hash_keys(const int32_t *k, uint32_t nk) {
    uint64_t h = 0x12345;
    for (uint32_t i = 0; i < nk; ++i) h = oracle_splitmix64(h ^ (uint32_t)k[i]);
    return h;
}
static int gtab_add(gtab *g, const int32_t *k, int64_t v, uint64_t cnt);
static int gtab_grow(gtab *g) {
    gtab n;
    if (gtab_init(&n, g->nk, g->cap)) return -1;
    for (uint64_t s = 0; s < g->cap; ++s)
        if (g->used[s]) gtab_add(&n, g->keys + s * g->nk, g->sums[s], g->counts[s]);
    gtab_free(g);
    *g = n;
    return 0;
}
static int gtab_add(gtab *g, const int32_t *k, int64_t v, uint64_t cnt) {
    if ((g->n + 1) * 2 > g->cap && gtab_grow(g)) return -1;
    uint64_t s = hash_keys(k, g->nk) & (g->cap - 1);
    for (;;) {
        if (!g->used[s]) {
            g->used[s] = 1;
            memcpy(g->keys + s * g->nk, k, g->nk * sizeof(int32_t));
            g->n++;
            break;
        }
        if (!memcmp(g->keys + s * g->nk, k, g->nk * sizeof(int32_t))) break;
        s = (s + 1) & (g->cap - 1);
    }
    g->sums[s] += v;
    g->counts[s] += cnt;
    return 0;
}

static uint32_t g_sort_nk;
static const int32_t *g_sort_keys;
static int cmp_slot(const void *a, const void *b) {
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    for (uint32_t i = 0; i < g_sort_nk; ++i) {
        int32_t u = g_sort_keys[x * g_sort_nk + i], w = g_sort_keys[y * g_sort_nk + i];
        if (u != w) return u < w ? -1 : 1;
    }
    return 0;
}
/* Canonical output: groups sorted ascending by key tuple. qsort's comparator has no context
 * argument, so this helper is not re-entrant (the oracle is called from one thread). */
static int64_t gtab_emit(gtab *g, uint64_t max_groups, int32_t *keys_out, int64_t *sums_out,
                         uint64_t *counts_out) {
    if (g->n > max_groups) return -1;
    uint64_t *idx = (uint64_t *)malloc((g->n ? g->n : 1) * sizeof(uint64_t)), m = 0;
    for (uint64_t s = 0; s < g->cap; ++s)
        if (g->used[s]) idx[m++] = s;
    g_sort_nk = g->nk;
    g_sort_keys = g->keys;
    qsort(idx, m, sizeof(uint64_t), cmp_slot);
    for (uint64_t i = 0; i < m; ++i) {
        if (keys_out) memcpy(keys_out + i * g->nk, g->keys + idx[i] * g->nk, g->nk * sizeof(int32_t));
        if (sums_out) sums_out[i] = g->sums[idx[i]];
        if (counts_out) counts_out[i] = g->counts[idx[i]];
    }
    free(idx);
    return (int64_t)m;
}

int64_t oracle_groupby(const int32_t *const *keys, uint32_t n_keys, const int64_t *val,
                       uint64_t n, uint64_t max_groups, int32_t *keys_out, int64_t *sums_out,
                       uint64_t *counts_out) {
    if (n_keys > 4) return -1;
    gtab g;
    if (gtab_init(&g, n_keys, 64)) return -1;
    int32_t k[4];
    for (uint64_t r = 0; r < n; ++r) {
        for (uint32_t i = 0; i < n_keys; ++i) k[i] = keys[i][r];
        if (gtab_add(&g, k, val[r], 1)) {
            gtab_free(&g);
            return -1;
        }
    }
    int64_t res = gtab_emit(&g, max_groups, keys_out, sums_out, counts_out);
    gtab_free(&g);
    return res;
}

/* ---------------------------------------------------------------- star join + group-by */
typedef struct {
    const oracle_dim *dims;
    uint32_t n_dims;
    const oracle_hashmap *maps;
    const int32_t *const *fk;
    const int32_t *val;
    uint64_t begin, end;
    uint32_t ng;
    gtab g;
    int err;
} star_job;

/* pipeline_q2 / 02inmem.hpp:150-176: probe each dimension in order; a miss drops the row; the
 * group is made of the matched rows' attributes; thread-local tables merged at the end. */
static void *star_thread(void *arg) {
    star_job *j = (star_job *)arg;
    int32_t k[8];
    for (uint64_t r = j->begin; r < j->end; ++r) {
        uint32_t gi = 0;
        int ok = 1;
        for (uint32_t d = 0; d < j->n_dims; ++d) {
            uint64_t hit = oracle_hashmap_probe(&j->maps[d], j->fk[d][r]);
            if (hit == UINT64_MAX) {
                ok = 0;
                break;
            }
            if (j->dims[d].group_attr) k[gi++] = j->dims[d].group_attr[hit];
        }
        if (!ok) continue;
        if (gtab_add(&j->g, k, (int64_t)j->val[r], 1)) {
            j->err = 1;
            return NULL;
        }
    }
    return NULL;
}

int64_t oracle_star_groupby(const oracle_dim *dims, uint32_t n_dims, const int32_t *const *fk,
                            const int32_t *val, uint64_t n_fact, uint32_t threads,
                            uint64_t max_groups, int32_t *keys_out, int64_t *sums_out,
                            uint64_t *counts_out) {
    if (n_dims > 8) return -1;
    oracle_hashmap maps[8];
    uint32_t ng = 0;
    for (uint32_t d = 0; d < n_dims; ++d) {
        if (build_dim_map(&maps[d], dims[d].key, dims[d].n, dims[d].preds, dims[d].n_preds))
            return -1;
        if (dims[d].group_attr) ng++;
    }
    if (threads < 1) threads = 1;
    star_job *jobs = (star_job *)calloc(threads, sizeof(star_job));
    pthread_t *tids = (pthread_t *)calloc(threads, sizeof(pthread_t));
    uint64_t base = n_fact / threads, rest = n_fact % threads, cur = 0;
    for (uint32_t t = 0; t < threads; ++t) {
        uint64_t len = base + (t < rest ? 1 : 0);
        jobs[t].dims = dims;
        jobs[t].n_dims = n_dims;
        jobs[t].maps = maps;
        jobs[t].fk = fk;
        jobs[t].val = val;
        jobs[t].begin = cur;
        jobs[t].end = cur + len;
        jobs[t].ng = ng;
        gtab_init(&jobs[t].g, ng, 64);
        cur += len;
        pthread_create(&tids[t], NULL, star_thread, &jobs[t]);
    }
    gtab all;
    gtab_init(&all, ng, 64);
    int err = 0;
    for (uint32_t t = 0; t < threads; ++t) {
        pthread_join(tids[t], NULL);
        err |= jobs[t].err;
        /* merge thread-local groups (HostHashAggregate add of the local map, 02inmem.hpp:178) */
        for (uint64_t s = 0; s < jobs[t].g.cap; ++s)
            if (jobs[t].g.used[s])
                err |= gtab_add(&all, jobs[t].g.keys + s * ng, jobs[t].g.sums[s], jobs[t].g.counts[s]);
        gtab_free(&jobs[t].g);
    }
    int64_t res = err ? -1 : gtab_emit(&all, max_groups, keys_out, sums_out, counts_out);
    gtab_free(&all);
    free(jobs);
    free(tids);
    for (uint32_t d = 0; d < n_dims; ++d) oracle_hashmap_free(&maps[d]);
    return res;
}

/* ---------------------------------------------------------------- zone maps */
void oracle_minmax_check_i64(const int64_t *mins, const int64_t *maxs, uint64_t n_chunks,
                             int64_t lo, int64_t hi, uint16_t *out) {
    for (uint64_t c = 0; c < n_chunks; ++c) out[c] = (lo <= maxs[c] && hi >= mins[c]) ? 1 : 0;
}

void oracle_minmax_build_i32(const int32_t *col, uint64_t n, uint64_t chunk_rows, int64_t *mins,
                             int64_t *maxs) {
    uint64_t nc = (n + chunk_rows - 1) / chunk_rows;
    for (uint64_t c = 0; c < nc; ++c) {
        uint64_t r0 = c * chunk_rows, r1 = r0 + chunk_rows < n ? r0 + chunk_rows : n;
        int64_t mn = INT64_MAX, mx = INT64_MIN;
        for (uint64_t r = r0; r < r1; ++r) {
            if (col[r] < mn) mn = col[r];
            if (col[r] > mx) mx = col[r];
        }
        mins[c] = mn;
        maxs[c] = mx;
    }
}

/* ------------------------------------------------------------------ scan + group-by */
static int64_t typed_at(const void *col, uint32_t type, uint64_t i) {
    switch (type) {
    case 1: return ((const int8_t *)col)[i];
    case 2: return ((const uint8_t *)col)[i];
    case 3: return ((const int16_t *)col)[i];
    case 4: return ((const uint16_t *)col)[i];
    case 5: return ((const int32_t *)col)[i];
    case 6: return ((const uint32_t *)col)[i];
    default: return ((const int64_t *)col)[i];
    }
}

typedef struct {
    uint64_t n, r0, r1, domain;
    const void *const *cols;
    const uint32_t *types;
    const oracle_scan_pred *preds;
    uint32_t n_preds;
    const oracle_scan_key *keys;
    uint32_t n_keys;
    int32_t val_col;
    uint64_t *sums, *counts; /* thread-local dense arrays, or shared ones updated atomically */
    int shared;
    int bad;
} scan_job;

static void *scan_worker(void *arg) {
    scan_job *j = (scan_job *)arg;
    for (uint64_t i = j->r0; i < j->r1; ++i) {
        int ok = 1;
        for (uint32_t p = 0; p < j->n_preds && ok; ++p) {
            int64_t v = typed_at(j->cols[j->preds[p].col], j->types[j->preds[p].col], i);
            ok = v >= j->preds[p].lo && v <= j->preds[p].hi;
        }
        if (!ok) continue;
        uint64_t g = 0;
        for (uint32_t k = 0; k < j->n_keys; ++k) {
            const oracle_scan_key *K = &j->keys[k];
            __int128 d = (__int128)typed_at(j->cols[K->col], j->types[K->col], i) - K->base;
            if (d < 0 || d / K->div >= K->card) {
                j->bad = 1;
                return NULL;
            }
            g = g * K->card + (uint64_t)(d / K->div);
        }
        uint64_t v = j->val_col >= 0 ? (uint64_t)typed_at(j->cols[j->val_col], j->types[j->val_col], i) : 0;
        if (j->shared) {
            __atomic_fetch_add(&j->sums[g], v, __ATOMIC_RELAXED);
            __atomic_fetch_add(&j->counts[g], 1, __ATOMIC_RELAXED);
        } else {
            j->sums[g] += v;
            j->counts[g] += 1;
        }
    }
    return NULL;
}

int64_t oracle_scan_groupby(uint64_t n, const void *const *cols, const uint32_t *types,
                            const oracle_scan_pred *preds, uint32_t n_preds, const oracle_scan_key *keys,
                            uint32_t n_keys, int32_t val_col, uint32_t threads, uint64_t max_groups,
                            int32_t *keys_out, int64_t *sums_out, uint64_t *counts_out) {
    uint64_t domain = 1;
    for (uint32_t k = 0; k < n_keys; ++k) {
        if (!keys[k].div || !keys[k].card) return -1;
        domain *= keys[k].card;
    }
    if (threads < 1) threads = 1;
    if (threads > 64) threads = 64;
    const int shared = domain * 16 * threads > (512ull << 20);
    const uint32_t copies = shared ? 1 : threads;
    uint64_t *sums = calloc(domain * copies, 8), *counts = calloc(domain * copies, 8);
    scan_job jobs[64];
    pthread_t th[64];
    for (uint32_t t = 0; t < threads; ++t) {
        scan_job j = {n, n * t / threads, n * (t + 1) / threads, domain, cols, types, preds, n_preds, keys,
                      n_keys, val_col, sums + (shared ? 0 : domain * t), counts + (shared ? 0 : domain * t),
                      shared, 0};
        jobs[t] = j;
        pthread_create(&th[t], NULL, scan_worker, &jobs[t]);
    }
    int bad = 0;
    for (uint32_t t = 0; t < threads; ++t) {
        pthread_join(th[t], NULL);
        bad |= jobs[t].bad;
    }
    int64_t out = 0;
    if (bad) {
        out = -1;
    } else {
        for (uint64_t g = 0; g < domain; ++g) {
            uint64_t s = 0, c = 0;
            for (uint32_t t = 0; t < copies; ++t) {
                s += sums[domain * t + g];
                c += counts[domain * t + g];
            }
            if (!c) continue;
            if ((uint64_t)out >= max_groups) {
                out = -2;
                break;
            }
            uint64_t rest = g;
            for (int k = (int)n_keys - 1; k >= 0; --k) {
                keys_out[(uint64_t)out * n_keys + k] = keys[k].out_lo + (int32_t)(rest % keys[k].card);
                rest /= keys[k].card;
            }
            sums_out[out] = (int64_t)s;
            counts_out[out] = c;
            ++out;
        }
    }
    free(sums);
    free(counts);
    return out;
}
