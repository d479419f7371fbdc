/* Decode timing of paged columns on the CPU (test infrastructure: the decode lines of the CPU
 * baseline). Pages are decoded page-parallel into per-thread buffers, either by the oracle's
 * restated decoders or, for LZ4 and LZ4I pages, by liblz4 itself: the reference's CPU codec,
 * LZ4_decompress_safe_partial (core/comp_cpu.hpp:22-32), loaded at run time with dlopen so the
 * oracle builds without it. Both report the same checksum when they decode the same bytes. */
#define _GNU_SOURCE
#include <dlfcn.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "oracle.h"

typedef int (*lz4_partial_fn)(const char *src, char *dst, int csize, int target, int cap);

typedef struct {
    const oracle_colref *cols;
    const uint32_t *elem;
    uint32_t n_cols;
    uint64_t n_rows, n_pages;
    uint32_t page_rows;
    lz4_partial_fn lz4;
    atomic_uint_fast64_t next;
    atomic_int failed;
} timing_job;

/* Position-weighted sum of the page's 64-bit words (tail zero padded): a page decoded into
 * other bytes or another order changes it. */
static uint64_t page_sum(const uint8_t *p, uint64_t n, uint64_t salt) {
    uint64_t s = 0, i = 0;
    for (; i + 8 <= n; i += 8) {
        uint64_t w;
        memcpy(&w, p + i, 8);
        s += w * (2 * (i / 8 + salt) + 1);
    }
    if (i < n) {
        uint64_t w = 0;
        memcpy(&w, p + i, n - i);
        s += w * (2 * (i / 8 + salt) + 1);
    }
    return s;
}

/* The LZ4 block of an LZ4I page: behind the header of (4 + 2 * (segments - 1)) bytes rounded
 * up to 16; segments = ceil(nseq / 2^seg_log2). */
static uint64_t lz4i_block_offset(const uint8_t *pg, uint64_t csize) {
    if (csize < 4) return csize;
    uint64_t nseq = pg[0] | ((uint64_t)pg[1] << 8), sl = pg[2];
    if (sl > 16) return csize;
    uint64_t nseg = (nseq + (1ull << sl) - 1) >> sl;
    uint64_t hb = (4 + 2 * (nseg ? nseg - 1 : 0) + 15) & ~15ull;
    return hb;
}

static void *timing_worker(void *arg) {
    timing_job *j = (timing_job *)arg;
    uint32_t max_elem = 1;
    for (uint32_t c = 0; c < j->n_cols; ++c)
        if (j->elem[c] > max_elem) max_elem = j->elem[c];
    uint8_t *buf = (uint8_t *)malloc((uint64_t)j->page_rows * max_elem + 64);
    uint64_t sum = 0;
    if (!buf) {
        atomic_store(&j->failed, 1);
        return NULL;
    }
    const uint64_t total = j->n_pages * j->n_cols, grain = 32;
    for (;;) {
        uint64_t t0 = atomic_fetch_add(&j->next, grain);
        if (t0 >= total || atomic_load(&j->failed)) break;
        uint64_t t1 = t0 + grain < total ? t0 + grain : total;
        for (uint64_t t = t0; t < t1; ++t) {
            uint32_t c = (uint32_t)(t / j->n_pages);
            uint64_t p = t % j->n_pages;
            const oracle_colref *col = &j->cols[c];
            uint64_t r0 = p * j->page_rows;
            uint64_t rows = j->n_rows - r0 < j->page_rows ? j->n_rows - r0 : j->page_rows;
            uint64_t nb = rows * j->elem[c];
            const uint8_t *src = col->comp + col->page_off[p];
            uint64_t cs = col->page_csize[p];
            uint32_t cd = col->page_codec[p] & 0x0f;
            int64_t got;
            if (j->lz4 && (cd == 1 || cd == 5)) {
                uint64_t off = cd == 5 ? lz4i_block_offset(src, cs) : 0;
                got = off < cs ? j->lz4((const char *)src + off, (char *)buf, (int)(cs - off), (int)nb, (int)nb) : -1;
            } else {
                got = oracle_page_decode(col->page_codec[p], src, cs, buf, nb);
                if (got >= 0 && cd != 0 && cd != 1 && cd != 5) got = (int64_t)(got * j->elem[c]);
            }
            if (got != (int64_t)nb) {
                atomic_store(&j->failed, 1);
                break;
            }
            sum += page_sum(buf, nb, r0 * j->elem[c] / 8 + ((uint64_t)c << 40));
        }
    }
    free(buf);
    return (void *)(uintptr_t)sum;
}

int oracle_decode_timing(const oracle_colref *cols, const uint32_t *elem, uint32_t n_cols, uint64_t n_rows,
                         uint32_t page_rows, uint32_t use_liblz4, uint32_t threads, uint64_t *checksum,
                         double *seconds) {
    if (!cols || !elem || !n_cols || !page_rows || !checksum || !seconds) return -1;
    timing_job j;
    memset(&j, 0, sizeof j);
    j.cols = cols;
    j.elem = elem;
    j.n_cols = n_cols;
    j.n_rows = n_rows;
    j.n_pages = (n_rows + page_rows - 1) / page_rows;
    j.page_rows = page_rows;
    if (use_liblz4) {
        void *h = dlopen("liblz4.so.1", RTLD_NOW | RTLD_LOCAL);
        if (!h) return -2;
        j.lz4 = (lz4_partial_fn)dlsym(h, "LZ4_decompress_safe_partial");
        if (!j.lz4) return -2;
    }
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    pthread_t th[256];
    struct timespec a, b;
    clock_gettime(CLOCK_MONOTONIC, &a);
    uint32_t started = 0;
    for (; started < threads; ++started)
        if (pthread_create(&th[started], NULL, timing_worker, &j)) break;
    uint64_t sum = 0;
    for (uint32_t t = 0; t < started; ++t) {
        void *r;
        pthread_join(th[t], &r);
        sum += (uint64_t)(uintptr_t)r;
    }
    clock_gettime(CLOCK_MONOTONIC, &b);
    *seconds = (double)(b.tv_sec - a.tv_sec) + 1e-9 * (double)(b.tv_nsec - a.tv_nsec);
    *checksum = sum;
    if (!started) return -1;
    return atomic_load(&j.failed) ? -1 : 0;
}
