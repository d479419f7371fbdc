/* Chunk metadata beyond min/max: the CPU restatement of EqHistogram and BloomMeta
 * (core/metadata.cuh:225-537) with their fill and check kernels (fill_hist / check_hist /
 * fill_bloom / check_bloom, :561-603) and the per-chunk fill of prepare_compressed
 * (core/apps.cuh:212-224). Test infrastructure only.
 *
 * Values are given as int64 (the column's integer values); the reference computes
 * val - min in the column type, which equals the int64 difference whenever that does not
 * overflow the type. */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "oracle.h"

/* EqHistogram::init_chunk (metadata.cuh:281-285) over each chunk's real tuples, then
 * fill_hist -> EqHistogram::add (:333-347): bucket = (val - min) * (B / double(max - min)),
 * clamped to B - 1; a chunk with min == max counts into bucket 0. */
void oracle_eqhist_build(const int64_t *v, uint64_t n, uint64_t chunk_rows, uint64_t bucket_num, int64_t *mins,
                         int64_t *maxs, uint64_t *hist) {
    const uint64_t nch = chunk_rows ? (n + chunk_rows - 1) / chunk_rows : 0;
    memset(hist, 0, nch * bucket_num * sizeof(uint64_t));
    for (uint64_t c = 0; c < nch; ++c) {
        const uint64_t r0 = c * chunk_rows, r1 = r0 + chunk_rows < n ? r0 + chunk_rows : n;
        int64_t mn = v[r0], mx = v[r0];
        for (uint64_t r = r0; r < r1; ++r) {
            if (v[r] < mn) mn = v[r];
            if (v[r] > mx) mx = v[r];
        }
        mins[c] = mn;
        maxs[c] = mx;
        for (uint64_t r = r0; r < r1; ++r) {
            uint64_t b = 0;
            if (mx != mn) {
                const double inv = (double)bucket_num / (double)(mx - mn);
                b = (uint64_t)((double)(v[r] - mn) * inv);
                if (b > bucket_num - 1) b = bucket_num - 1;
            }
            hist[c * bucket_num + b] += 1;
        }
    }
}

/* check_hist (metadata.cuh:573-585) with EqHistogram::check_pred (:349-372): chunk c is kept
 * iff a bucket between first = (lo - min) * inv and last = (hi - min) * inv (first = B - 1
 * when lo == max) is non-empty; a chunk with min == max iff lo <= min, hi >= max and its
 * bucket 0 is non-empty. out[c] = 1 / 0. */
void oracle_eqhist_check(const int64_t *mins, const int64_t *maxs, const uint64_t *hist, uint64_t n_chunks,
                         uint64_t bucket_num, int64_t lo, int64_t hi, uint16_t *out) {
    for (uint64_t c = 0; c < n_chunks; ++c) {
        uint16_t keep = 0;
        if (mins[c] == maxs[c]) {
            keep = lo <= mins[c] && hi >= maxs[c] && hist[c * bucket_num] != 0;
        } else {
            const double inv = (double)bucket_num / (double)(maxs[c] - mins[c]);
            int64_t first = (int64_t)(((double)lo - (double)mins[c]) * inv);
            const int64_t last = (int64_t)(((double)hi - (double)mins[c]) * inv);
            if (lo == maxs[c]) first = (int64_t)bucket_num - 1;
            for (uint64_t b = 0; b < bucket_num && !keep; ++b)
                keep = first <= (int64_t)b && last >= (int64_t)b && hist[c * bucket_num + b] != 0;
        }
        out[c] = keep;
    }
}

/* BloomMeta's (p, m) parameters (metadata.cuh:398-401): n = ceil(mbits * ln(0.5^ln 2) / ln p),
 * k = ceil(ln 2 * mbits / n). */
void oracle_bloom_params(double p, uint64_t m, uint64_t *n, uint64_t *k) {
    const uint64_t mbits = m << 3;
    *n = (uint64_t)ceil((double)mbits * log(pow(0.5, log(2.0))) / log(p));
    *k = (uint64_t)ceil(log(2.0) * (double)mbits / (double)*n);
}

/* BloomMeta::hash (metadata.cuh:457-465): MurmurHash3's 64-bit finalizer, modulo the bits. */
static uint64_t bloom_hash(uint64_t x, uint64_t mbits) {
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdull;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ull;
    x ^= x >> 33;
    return x % mbits;
}

/* fill_bloom -> BloomMeta::add (metadata.cuh:467-481): bit hash(val + i) for i < k in the
 * chunk's m-byte filter (state: m / 4 words per chunk). */
void oracle_bloom_build(const int64_t *v, uint64_t n, uint64_t chunk_rows, uint64_t m, uint64_t k, uint32_t *state) {
    const uint64_t nch = chunk_rows ? (n + chunk_rows - 1) / chunk_rows : 0;
    memset(state, 0, nch * m);
    for (uint64_t r = 0; r < n; ++r) {
        uint32_t *f = state + ((r / chunk_rows) * m >> 2);
        for (uint64_t i = 0; i < k; ++i) {
            const uint64_t bit = bloom_hash((uint64_t)v[r] + i, m << 3);
            f[bit >> 5] |= 1u << (bit & 31);
        }
    }
}

/* check_bloom -> BloomMeta::check_pred (metadata.cuh:483-501): chunk c is kept iff all k bits
 * of value are set. out[c] = 1 / 0. */
void oracle_bloom_check(const uint32_t *state, uint64_t n_chunks, uint64_t m, uint64_t k, int64_t value,
                        uint16_t *out) {
    for (uint64_t c = 0; c < n_chunks; ++c) {
        const uint32_t *f = state + (c * m >> 2);
        uint16_t keep = 1;
        for (uint64_t i = 0; i < k; ++i) {
            const uint64_t bit = bloom_hash((uint64_t)value + i, m << 3);
            if (!((f[bit >> 5] >> (bit & 31)) & 1u)) keep = 0;
        }
        out[c] = keep;
    }
}
