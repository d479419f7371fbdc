// Fused star join + group-by over paged fact columns (gfx950).
//
// Replaces the reference's per-chunk pipeline_q2 / pipeline_q3 kernels
// (usecases/ssb/queries/02.cuh:102-193, 03.cuh:118-170) that probe chained HashMaps and
// aggregate with HashAggregate::add (core/dev_structs.cuh:209-244) after nvcomp decompressed
// each chunk. One workgroup handles one page-row tile:
//   1. the driving dimension's foreign keys are tested against its membership filter pinned in
//      LDS (exact bitmap for a dense key range, one-hash bit filter otherwise);
//   2. passing rows are compacted into an LDS candidate list (wave prefix sum + one LDS atomic
//      per wave);
//   3. only for tiles with candidates, coded pages of the other columns are decoded into LDS;
//      in-place pages are read at the candidate rows only;
//   4. candidates probe every dimension's exact table (open addressing, 8-byte slots), form a
//      dense group index from the dimensions' payload attributes and add SUM/COUNT into an
//      LDS-private hashed group table that spills to global int64 atomics when a slot run is
//      full; the table is merged into the global dense result once per workgroup.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "common.hpp"
#include "device_codecs.hpp"
#include "star_dev.hpp"

namespace golap_impl {
using namespace golap_dev;
using namespace stardev;

// Arguments of the persistent star kernel (see StarLaunch).
struct StarArgs {
    StarCol col[STAR_MAX_COLS];
    uint32_t page_rows;
    uint32_t pages_per_chunk;
    uint64_t n_rows, n_pages;
    uint32_t n_windows;
    uint32_t n_anchors;  // LDS prefix-sum regions (page_rows / 32 words each)
    uint32_t n_queries;
    StarQueryDev q0;
    const StarQueryDev *queries;
    uint64_t n_items;
    const uint2 *items;
    uint32_t *queue;
    uint32_t pages_per_item;
    uint32_t filter_words;
    uint32_t lds_gslots;
    uint32_t group_bits;
    uint32_t *err;
    // shared-filter schedule (see StarLaunch)
    const uint4 *items4;
    const DriverFilterView *mfilt;
    uint32_t mcol;
    uint32_t n_cols;
    // emit mode (see StarLaunch)
    int4 *emit;
    uint32_t *emit_n;
    uint64_t emit_cap;
    uint32_t emit_dim;
};

constexpr uint32_t QDESC_WORDS = sizeof(StarQueryDev) / 4;
static_assert(offsetof(StarQueryDev, s) == 0 && offsetof(ScanQueryDev, s) == 0,
              "work lists read a descriptor's slice at its start");
static_assert(sizeof(StarQueryDev) % 8 == 0, "query descriptor copied as words");

// LDS of one workgroup besides the dynamic arrays.
struct StarShared {
    StarQueryDev q;                        // current query
    ColAccess acc[STAR_MAX_DIMS + 2];      // per tile: foreign keys of q in dimension order, value, value2
    uint32_t cand_n, tile_bad;
    uint32_t item;
};

// Column c of the tile's access list: foreign keys of the current query in dimension order, then
// the value column, then the second value column.
__device__ __forceinline__ uint32_t tile_col(const StarQueryDev &q, uint32_t c) {
    return c < q.n_dims ? q.fk_col[c] : c == q.n_dims ? q.val_col : q.val2_col;
}
// The row's value: value, or value - value2 (each widened by its own signedness).
__device__ __forceinline__ int64_t row_value(const StarShared &S, uint32_t r) {
    const uint32_t nd = S.q.n_dims;
    const int32_t raw = col_value(S.acc[nd], r);
    int64_t v = S.q.val_signed ? (int64_t)raw : (int64_t)(uint32_t)raw;
    if (S.q.has_val2) {
        const int32_t raw2 = col_value(S.acc[nd + 1], r);
        v -= S.q.val2_signed ? (int64_t)raw2 : (int64_t)(uint32_t)raw2;
    }
    return v;
}

// Emit mode: candidates probe every dimension but emit_dim exactly; the rows that pass all of
// them are appended as {emit_dim's key, group index over the probed dimensions, value lo, hi}
// with one atomic per wave (ballot + prefix within the wave). Appends beyond emit_cap are
// dropped; emit_n still counts them, so the host sees the overflow and reruns smaller ranges.
__device__ __forceinline__ void star_emit(const StarArgs &a, StarShared &S, const uint16_t *cand, uint32_t nc) {
    const uint32_t nd = S.q.n_dims, drv = S.q.driver, hd = a.emit_dim;
    const uint32_t lane = threadIdx.x % WAVE;
    for (uint32_t i0 = 0; i0 < nc; i0 += BLOCK) {  // every lane takes every round: ballots stay whole
        const uint32_t i = i0 + threadIdx.x;
        bool ok = i < nc;
        int32_t hkey = 0;
        uint64_t gidx = 0;
        int64_t v = 0;
        if (ok) {
            const uint32_t r = cand[i];
            uint32_t pay = 0;
            const StarDimProbe &dd = S.q.dim[drv];
            const int32_t dk = col_value(S.acc[drv], r);
            const uint32_t h0 = hash32((uint32_t)dk) >> (32 - dd.log2cap);
            ok = probe_exact(dd.slots, dd.log2cap, dk, dd.slots[h0], h0, pay);
            for (uint32_t d = 0; d < nd && ok; ++d) {
                const int32_t key = col_value(S.acc[d], r);
                if (d == hd) {
                    hkey = key;
                    continue;
                }
                const StarDimProbe &pd = S.q.dim[d];
                const uint32_t h = hash32((uint32_t)key) >> (32 - pd.log2cap);
                ok = probe_exact(pd.slots, pd.log2cap, key, pd.slots[h], h, pay);
                if (ok && pd.group_card) {
                    const uint32_t gv = pay - (uint32_t)pd.group_lo;
                    if (gv >= pd.group_card) {
                        raise_err(a.err, DEVERR_KEYRANGE);
                        ok = false;
                    }
                    gidx += (uint64_t)gv * pd.group_stride;
                }
            }
            v = row_value(S, r);
        }
        const uint64_t m = __ballot(ok);
        if (!m) continue;
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(a.emit_n, (uint32_t)__popcll(m));
        base = __shfl(base, 0, WAVE);
        const uint64_t pos = (uint64_t)base + (uint64_t)__popcll(m & ((1ull << lane) - 1));
        if (ok && pos < a.emit_cap)
            a.emit[pos] = make_int4(hkey, (int32_t)(uint32_t)gidx, (int32_t)(uint32_t)(uint64_t)v,
                                    (int32_t)(uint32_t)((uint64_t)v >> 32));
    }
}

// One page tile of one query: driver filter, compaction, decode of the other coded columns
// for tiles with candidates, exact probes, group index, aggregation (or, in emit mode, the
// rows handed to the host).
template <int PER, bool EMIT>
__device__ __forceinline__ void star_tile(const StarArgs &a, StarShared &S, uint32_t q, uint64_t p,
                                          const uint32_t *lfilter, uint16_t *cand, uint8_t *win, uint32_t *ancs,
                                          const GroupTable &T) {
    constexpr int G = PER / 4;
    const uint32_t P = a.page_rows;
    const uint32_t wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    const uint32_t nd = S.q.n_dims, ncols = nd + 1 + S.q.has_val2;
    const uint32_t rows = (uint32_t)min<uint64_t>(P, a.n_rows - p * P);
    if (threadIdx.x == 0) {
        S.cand_n = 0;
        S.tile_bad = 0;
    }
    if (threadIdx.x < ncols) {
        const StarCol &sc = a.col[tile_col(S.q, threadIdx.x)];
        S.acc[threadIdx.x] = col_access(sc, p, rows, sc.window >= 0 ? win + (size_t)sc.window * P * 4 : nullptr,
                                        sc.anchors >= 0 ? ancs + (size_t)sc.anchors * (P / 32) : nullptr);
    }
    __syncthreads();
    bool bad = false;
    for (uint32_t c = 0; c < ncols; ++c) bad |= S.acc[c].mode == ACC_BAD;
    if (bad) {  // uniform: every thread read the same entries
        if (threadIdx.x == 0) raise_err(a.err, DEVERR_CORRUPT);
        __syncthreads();
        return;
    }
    const uint32_t drv = S.q.driver;
    DriverFilter df = S.q.filter;
    if (lfilter) df.f = lfilter;
    const ColAccess dacc = S.acc[drv];
    if (dacc.mode == ACC_WINDOW) {
        if (wave == 0) {
            const ColView &cv = a.col[S.q.fk_col[drv]].c;
            const int64_t got = page_decode_wave(cv.page_codec[p], cv.comp + cv.page_off[p], cv.page_csize[p],
                                                 const_cast<uint8_t *>(dacc.base), rows * 4, 4);
            if (got != (int64_t)rows * 4 && lane == 0) {
                S.tile_bad = 1;
                raise_err(a.err, DEVERR_CORRUPT);
            }
        }
        __syncthreads();
    }
    // filter + compaction: thread t tests rows 4(t + k*BLOCK) + j; a wave prefix sum over the
    // per-thread counts gives every thread its slots in the candidate list
    uint32_t passbits = 0;
    if (!S.tile_bad) {
        if (rows == (uint32_t)(PER * BLOCK) && dacc.mode <= ACC_WINDOW) {
            int4 kv[G];
#pragma unroll
            for (int k = 0; k < G; ++k)
                kv[k] = *reinterpret_cast<const int4 *>(dacc.base + 16 * ((size_t)threadIdx.x + (size_t)k * BLOCK));
#pragma unroll
            for (int k = 0; k < G; ++k) {
                passbits |= (uint32_t)filter_test(df, kv[k].x) << (4 * k);
                passbits |= (uint32_t)filter_test(df, kv[k].y) << (4 * k + 1);
                passbits |= (uint32_t)filter_test(df, kv[k].z) << (4 * k + 2);
                passbits |= (uint32_t)filter_test(df, kv[k].w) << (4 * k + 3);
            }
        } else {
            for (int k = 0; k < G; ++k)
                for (int j = 0; j < 4; ++j) {
                    const uint32_t r = 4 * (threadIdx.x + k * BLOCK) + j;
                    if (r < rows && filter_test(df, col_value(dacc, r))) passbits |= 1u << (4 * k + j);
                }
        }
    }
    const uint32_t mine = __popc(passbits);
    uint32_t incl = mine;
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) {
        const uint32_t t = __shfl_up(incl, d, WAVE);
        if (lane >= (uint32_t)d) incl += t;
    }
    const uint32_t wtotal = __shfl(incl, WAVE - 1, WAVE);
    uint32_t wbase = 0;
    if (lane == WAVE - 1 && wtotal) wbase = atomicAdd(&S.cand_n, wtotal);
    wbase = __shfl(wbase, WAVE - 1, WAVE);
    uint32_t slot = wbase + incl - mine;
    for (uint32_t bits = passbits; bits; bits &= bits - 1) {
        const uint32_t b = __ffs(bits) - 1;
        cand[slot++] = (uint16_t)(4 * (threadIdx.x + (b >> 2) * BLOCK) + (b & 3));
    }
    __syncthreads();
    const uint32_t nc = S.cand_n;
    if (nc == 0 || S.tile_bad) {
        __syncthreads();
        return;
    }
    // other columns that need a decode: wave w decodes columns w, w+4, ... (delta-bitpack
    // pages without a window: their prefix sums only)
    for (uint32_t c = wave; c < ncols; c += BLOCK / WAVE) {
        if (S.acc[c].mode == ACC_DBPA) {
            dbp_anchors_wave(S.acc[c], rows);
            continue;
        }
        if (c == drv || S.acc[c].mode != ACC_WINDOW) continue;
        const ColView &cv = a.col[tile_col(S.q, c)].c;
        const int64_t got = page_decode_wave(cv.page_codec[p], cv.comp + cv.page_off[p], cv.page_csize[p],
                                             const_cast<uint8_t *>(S.acc[c].base), rows * 4, 4);
        if (got != (int64_t)rows * 4 && lane == 0) {
            S.tile_bad = 1;
            raise_err(a.err, DEVERR_CORRUPT);
        }
    }
    __syncthreads();
    if (EMIT) {
        if (!S.tile_bad) star_emit(a, S, cand, nc);
    } else if (!S.tile_bad) {
        const bool use_lds = S.q.use_lds && T.slots;
        const uint32_t qkey = q << a.group_bits;
        for (uint32_t i = threadIdx.x; i < nc; i += BLOCK) {
            const uint32_t r = cand[i];
            // exact probe of the driver first (its small table stays in L2): filter false
            // positives stop here, before any other column is read
            uint32_t pay = 0;
            {
                const StarDimProbe &dd = S.q.dim[drv];
                const int32_t dk = col_value(S.acc[drv], r);
                const uint32_t h = hash32((uint32_t)dk) >> (32 - dd.log2cap);
                if (!probe_exact(dd.slots, dd.log2cap, dk, dd.slots[h], h, pay)) continue;
            }
            // then every column value of the row and the first slot of every table: two
            // memory round trips for the row instead of two per dimension
            int32_t key[STAR_MAX_DIMS];
            for (uint32_t d = 0; d < nd; ++d) key[d] = col_value(S.acc[d], r);
            const int64_t v = row_value(S, r);
            uint32_t h[STAR_MAX_DIMS];
            uint64_t first[STAR_MAX_DIMS];
            for (uint32_t d = 0; d < nd; ++d) {
                h[d] = hash32((uint32_t)key[d]) >> (32 - S.q.dim[d].log2cap);
                first[d] = S.q.dim[d].slots[h[d]];
            }
            uint64_t gidx = 0;
            bool ok = true;
            for (uint32_t d = 0; d < nd && ok; ++d) {
                const StarDimProbe &dd = S.q.dim[d];
                ok = probe_exact(dd.slots, dd.log2cap, key[d], first[d], h[d], pay);
                if (ok && dd.group_card) {
                    const uint32_t gv = pay - (uint32_t)dd.group_lo;
                    if (gv >= dd.group_card) {
                        raise_err(a.err, DEVERR_KEYRANGE);
                        ok = false;
                    }
                    gidx += (uint64_t)gv * dd.group_stride;
                }
            }
            if (!ok) continue;
            if (!use_lds || !group_add_lds(T, qkey | (uint32_t)gidx, v)) group_add_global(S.q.g_sum, S.q.g_cnt, gidx, v);
        }
    }
    __syncthreads();  // candidate list, windows and access entries are reused by the next tile
}

// Makes query q current: descriptor into LDS and, when the filter is pinned, its words too.
__device__ __forceinline__ void star_set_query(const StarArgs &a, StarShared &S, uint32_t q, uint32_t *lfilter) {
    __syncthreads();
    const uint32_t *src = reinterpret_cast<const uint32_t *>(a.queries ? a.queries + q : &a.q0);
    uint32_t *dst = reinterpret_cast<uint32_t *>(&S.q);
    for (uint32_t i = threadIdx.x; i < QDESC_WORDS; i += BLOCK) dst[i] = src[i];
    __syncthreads();
    if (lfilter) {
        const uint32_t n = S.q.filter_words, *f = S.q.filter.f;
        for (uint32_t i = threadIdx.x; i < n; i += BLOCK) lfilter[i] = f[i];
        __syncthreads();
    }
}

// Persistent star kernel. PER = page_rows / BLOCK rows per thread (4, 8, 16 or 32).
// Static schedule (QUEUE false): the queries' pages, flattened in query order; workgroup b takes
// the contiguous range [b*n/G, (b+1)*n/G), so on sorted tables its rows share a narrow key range
// (few distinct groups per workgroup); pages of chunks whose mask flag is 0 are skipped.
// Queue schedule: workgroups take items from a device-built list of surviving (query, chunk
// part) items with one atomic each, which balances slices and pruning of very uneven size.
// The LDS group table is keyed by (query, group); it is merged into every query's dense global
// result once at the end.
// Register bounds for more workgroups per CU (LDS allows six on the usual launches): six at
// 4,096- and 8,192-row pages (80 registers, a 20-byte spill that costs less than the extra
// workgroup gains), five at 1,024-row pages and four at 2,048-row pages, where tighter bounds
// spill more.
template <int PER, bool QUEUE, bool EMIT>
__global__ __launch_bounds__(BLOCK, PER == 8 ? 4 : (PER == 4 ? 5 : 6)) void k_star(StarArgs a) {
    extern __shared__ __align__(16) uint8_t lds[];
    __shared__ StarShared S;
    const uint32_t P = a.page_rows;
    size_t off = 0;
    uint32_t *lfilter = a.filter_words ? reinterpret_cast<uint32_t *>(lds) : nullptr;
    off += ((size_t)a.filter_words * 4 + 15) & ~(size_t)15;
    uint16_t *cand = reinterpret_cast<uint16_t *>(lds + off);
    off += ((size_t)P * 2 + 15) & ~(size_t)15;
    uint8_t *win = lds + off;
    off += (size_t)a.n_windows * P * 4;
    uint32_t *ancs = reinterpret_cast<uint32_t *>(lds + off);
    off += (size_t)a.n_anchors * (P / 32) * 4;
    GroupTable T;
    T.slots = a.lds_gslots;
    T.gkey = reinterpret_cast<uint32_t *>(lds + off);
    off += (size_t)a.lds_gslots * 4;
    T.gcnt = reinterpret_cast<uint32_t *>(lds + off);
    off += (size_t)a.lds_gslots * 4;
    T.gsum = reinterpret_cast<unsigned long long *>(lds + off);
    for (uint32_t i = threadIdx.x; i < a.lds_gslots; i += BLOCK) {
        T.gkey[i] = GROUP_EMPTY;
        T.gcnt[i] = 0;
        T.gsum[i] = 0;
    }
    uint32_t cur = 0xffffffffu;
    if (!QUEUE) {
        const uint64_t i0 = a.n_items * blockIdx.x / gridDim.x, i1 = a.n_items * (blockIdx.x + 1) / gridDim.x;
        uint32_t q = 0;
        for (uint64_t it = i0; it < i1; ++it) {
            if (a.queries)
                while (q + 1 < a.n_queries && it >= a.queries[q + 1].s.item_begin) ++q;
            if (q != cur) {
                star_set_query(a, S, q, lfilter);
                cur = q;
            }
            const uint64_t p = S.q.s.page_begin + (it - S.q.s.item_begin);
            if (!slice_keeps(S.q.s, p / a.pages_per_chunk)) continue;
            star_tile<PER, EMIT>(a, S, q, p, lfilter, cand, win, ancs, T);
        }
    } else {
        const uint32_t n_items = a.queue[1];
        for (;;) {
            __syncthreads();
            if (threadIdx.x == 0) S.item = atomicAdd(&a.queue[0], 1u);
            __syncthreads();
            const uint32_t it = S.item;
            if (it >= n_items) break;
            const uint2 w = a.items[it];
            const uint32_t q = w.y >> 8, part = w.y & 0xff;
            if (q != cur) {
                star_set_query(a, S, q, lfilter);
                cur = q;
            }
            const uint64_t c0 = (uint64_t)w.x * a.pages_per_chunk;
            const uint64_t pb = c0 + (uint64_t)part * a.pages_per_item;
            const uint64_t pe = min(min(pb + a.pages_per_item, c0 + a.pages_per_chunk), a.n_pages);
            for (uint64_t p = pb; p < pe; ++p) star_tile<PER, EMIT>(a, S, q, p, lfilter, cand, win, ancs, T);
        }
    }
    // merge the LDS-private table into every query's global dense result
    __syncthreads();
    const uint32_t gmask = (1u << a.group_bits) - 1;
    for (uint32_t i = threadIdx.x; i < a.lds_gslots; i += BLOCK) {
        const uint32_t k = T.gkey[i];
        if (k == GROUP_EMPTY) continue;
        const StarQueryDev *Q = a.queries ? a.queries + (k >> a.group_bits) : &a.q0;
        atomicAdd(reinterpret_cast<unsigned long long *>(&Q->g_sum[k & gmask]), T.gsum[i]);
        atomicAdd(reinterpret_cast<unsigned long long *>(&Q->g_cnt[k & gmask]), (unsigned long long)T.gcnt[i]);
    }
}

// Shared-filter star kernel (batches whose queries all probe one dimension column): work items
// are (chunk part, word of 64 queries, mask of the queries that keep the chunk). Per page tile:
//   1. the access plans of the columns the item's queries read;
//   2. the shared column's keys are read once into registers and tested against the membership
//      filter of every query of the item (each query's filter of its dimension on that column);
//   3. rows that pass any filter are compacted into the LDS candidate list (as in k_star);
//   4. a candidate row tests the filters again query by query, and for each query that passes
//      probes every dimension of that query exactly (the shared one too, for its payload) and
//      adds SUM/COUNT to the query's dense global result.
// A page's catalog, shared-column reads, compaction and barriers are paid once however many
// queries keep the chunk.
struct SharedLds {
    ColAccess acc[STAR_MAX_COLS];
    uint32_t cand_n, tile_bad, item, colmask;
};

template <int PER>
__global__ __launch_bounds__(BLOCK) void k_star_shared(StarArgs a) {
    constexpr int G = PER / 4;
    extern __shared__ __align__(16) uint8_t lds[];
    __shared__ SharedLds S;
    const uint32_t P = a.page_rows;
    const uint32_t wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    uint16_t *cand = reinterpret_cast<uint16_t *>(lds);
    uint8_t *win = lds + (((size_t)P * 2 + 15) & ~(size_t)15);
    const uint32_t n_items = a.queue[1];
    const uint32_t mcol = a.mcol;
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) S.item = atomicAdd(&a.queue[0], 1u);
        __syncthreads();
        const uint32_t it = S.item;
        if (it >= n_items) break;
        const uint4 w = a.items4[it];
        const uint64_t c = w.x;
        const uint32_t part = w.y & 0xffu, word = w.y >> 8;
        const uint64_t mask = (uint64_t)w.z | ((uint64_t)w.w << 32);
        const DriverFilterView *mf = a.mfilt + (size_t)word * WAVE;
        // the columns the item's queries read (wave 0, one lane per query of the word)
        if (threadIdx.x < WAVE) {
            const uint32_t q = word * WAVE + lane;
            uint32_t cm = 0;
            if (((mask >> lane) & 1u) && q < a.n_queries) {
                const StarQueryDev &Q = a.queries[q];
                cm = 1u << Q.val_col;
                for (uint32_t d = 0; d < Q.n_dims; ++d) cm |= 1u << Q.fk_col[d];
            }
#pragma unroll
            for (int o = WAVE / 2; o > 0; o >>= 1) cm |= __shfl_xor(cm, o, WAVE);
            if (lane == 0) S.colmask = cm | (1u << mcol);
        }
        const uint64_t c0 = c * a.pages_per_chunk;
        const uint64_t pb = c0 + (uint64_t)part * a.pages_per_item;
        const uint64_t pe = min(min(pb + a.pages_per_item, c0 + a.pages_per_chunk), a.n_pages);
        for (uint64_t p = pb; p < pe; ++p) {
            const uint32_t rows = (uint32_t)min<uint64_t>(P, a.n_rows - p * P);
            __syncthreads();  // the previous tile's readers are done; colmask is visible
            const uint32_t colmask = S.colmask;
            if (threadIdx.x == 0) {
                S.cand_n = 0;
                S.tile_bad = 0;
            }
            if (threadIdx.x < a.n_cols && ((colmask >> threadIdx.x) & 1u)) {
                const StarCol &sc = a.col[threadIdx.x];
                S.acc[threadIdx.x] = col_access(sc, p, rows, sc.window >= 0 ? win + (size_t)sc.window * P * 4 : nullptr);
            }
            __syncthreads();
            bool bad = false;
            for (uint32_t k = 0; k < a.n_cols; ++k) bad |= ((colmask >> k) & 1u) && S.acc[k].mode == ACC_BAD;
            if (bad) {  // uniform
                if (threadIdx.x == 0) raise_err(a.err, DEVERR_CORRUPT);
                continue;
            }
            const ColAccess dacc = S.acc[mcol];
            if (dacc.mode == ACC_WINDOW) {
                if (wave == 0) {
                    const ColView &cv = a.col[mcol].c;
                    const int64_t got = page_decode_wave(cv.page_codec[p], cv.comp + cv.page_off[p], cv.page_csize[p],
                                                         const_cast<uint8_t *>(dacc.base), rows * 4, 4);
                    if (got != (int64_t)rows * 4 && lane == 0) {
                        S.tile_bad = 1;
                        raise_err(a.err, DEVERR_CORRUPT);
                    }
                }
                __syncthreads();
            }
            // the shared column's keys once, tested against every query's filter
            uint32_t passbits = 0;
            if (!S.tile_bad) {
                int32_t kv[PER];
                uint32_t valid = 0;
                if (rows == (uint32_t)(PER * BLOCK) && dacc.mode <= ACC_WINDOW) {
#pragma unroll
                    for (int k = 0; k < G; ++k) {
                        const int4 v = *reinterpret_cast<const int4 *>(dacc.base + 16 * ((size_t)threadIdx.x + (size_t)k * BLOCK));
                        kv[4 * k] = v.x;
                        kv[4 * k + 1] = v.y;
                        kv[4 * k + 2] = v.z;
                        kv[4 * k + 3] = v.w;
                    }
                    valid = PER == 32 ? 0xffffffffu : (1u << PER) - 1;
                } else {
#pragma unroll
                    for (int k = 0; k < G; ++k)
#pragma unroll
                        for (int j = 0; j < 4; ++j) {
                            const uint32_t r = 4 * (threadIdx.x + k * BLOCK) + j;
                            kv[4 * k + j] = r < rows ? col_value<false>(dacc, r) : 0;
                            valid |= (uint32_t)(r < rows) << (4 * k + j);
                        }
                }
                for (uint64_t m = mask; m; m &= m - 1) {  // uniform
                    const DriverFilterView f = mf[__ffsll((long long)m) - 1];
#pragma unroll
                    for (int j = 0; j < PER; ++j) passbits |= (uint32_t)filter_test(f, kv[j]) << j;
                }
                passbits &= valid;
            }
            const uint32_t mine = __popc(passbits);
            uint32_t incl = mine;
#pragma unroll
            for (int d = 1; d < WAVE; d <<= 1) {
                const uint32_t t = __shfl_up(incl, d, WAVE);
                if (lane >= (uint32_t)d) incl += t;
            }
            const uint32_t wtotal = __shfl(incl, WAVE - 1, WAVE);
            uint32_t wbase = 0;
            if (lane == WAVE - 1 && wtotal) wbase = atomicAdd(&S.cand_n, wtotal);
            wbase = __shfl(wbase, WAVE - 1, WAVE);
            uint32_t slot = wbase + incl - mine;
            for (uint32_t bits = passbits; bits; bits &= bits - 1) {
                const uint32_t b = __ffs(bits) - 1;
                cand[slot++] = (uint16_t)(4 * (threadIdx.x + (b >> 2) * BLOCK) + (b & 3));
            }
            __syncthreads();
            const uint32_t nc = S.cand_n;
            if (nc == 0 || S.tile_bad) continue;
            // other columns that need a decode: wave w decodes columns w, w+4, ...
            for (uint32_t k = wave; k < a.n_cols; k += BLOCK / WAVE) {
                if (k == mcol || !((colmask >> k) & 1u) || S.acc[k].mode != ACC_WINDOW) continue;
                const ColView &cv = a.col[k].c;
                const int64_t got = page_decode_wave(cv.page_codec[p], cv.comp + cv.page_off[p], cv.page_csize[p],
                                                     const_cast<uint8_t *>(S.acc[k].base), rows * 4, 4);
                if (got != (int64_t)rows * 4 && lane == 0) {
                    S.tile_bad = 1;
                    raise_err(a.err, DEVERR_CORRUPT);
                }
            }
            __syncthreads();
            if (S.tile_bad) continue;
            for (uint32_t i = threadIdx.x; i < nc; i += BLOCK) {
                const uint32_t r = cand[i];
                const int32_t mk = col_value<false>(S.acc[mcol], r);
                for (uint64_t qm = mask; qm; qm &= qm - 1) {
                    const uint32_t b = (uint32_t)__ffsll((long long)qm) - 1;
                    if (!filter_test(mf[b], mk)) continue;
                    const StarQueryDev &Q = a.queries[word * WAVE + b];
                    const uint32_t nd = Q.n_dims;
                    int32_t key[STAR_MAX_DIMS];
                    uint32_t h[STAR_MAX_DIMS];
                    uint64_t first[STAR_MAX_DIMS];
                    for (uint32_t d = 0; d < nd; ++d) {
                        key[d] = col_value<false>(S.acc[Q.fk_col[d]], r);
                        h[d] = hash32((uint32_t)key[d]) >> (32 - Q.dim[d].log2cap);
                        first[d] = Q.dim[d].slots[h[d]];
                    }
                    uint64_t gidx = 0;
                    bool ok = true;
                    for (uint32_t d = 0; d < nd && ok; ++d) {
                        const StarDimProbe &dd = Q.dim[d];
                        uint32_t pay = 0;
                        ok = probe_exact(dd.slots, dd.log2cap, key[d], first[d], h[d], pay);
                        if (ok && dd.group_card) {
                            const uint32_t gv = pay - (uint32_t)dd.group_lo;
                            if (gv >= dd.group_card) {
                                raise_err(a.err, DEVERR_KEYRANGE);
                                ok = false;
                            }
                            gidx += (uint64_t)gv * dd.group_stride;
                        }
                    }
                    if (!ok) continue;
                    const int32_t raw = col_value<false>(S.acc[Q.val_col], r);
                    const int64_t v = Q.val_signed ? (int64_t)raw : (int64_t)(uint32_t)raw;
                    group_add_global(Q.g_sum, Q.g_cnt, gidx, v);
                }
            }
        }
    }
}

size_t star_lds_bytes(uint32_t filter_words, uint32_t page_rows, uint32_t n_windows, uint32_t lds_gslots,
                      uint32_t n_anchors) {
    return (((size_t)filter_words * 4 + 15) & ~(size_t)15) + (((size_t)page_rows * 2 + 15) & ~(size_t)15) +
           (size_t)n_windows * page_rows * 4 + (size_t)n_anchors * (page_rows / 32) * 4 + (size_t)lds_gslots * 16;
}

template <typename K>
static hipError_t launch_star_t(golap_ctx *ctx, K kernel, const StarArgs &a, size_t lds, hipStream_t stream,
                                uint32_t max_blocks_per_cu, uint64_t max_grid, bool prepare_only) {
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    int occ = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kernel, BLOCK, lds);
    if (e != hipSuccess) return e;
    occ = std::max(occ, 1);
    if (max_blocks_per_cu) occ = std::min<int>(occ, (int)max_blocks_per_cu);
    const uint64_t grid = std::max<uint64_t>(1, std::min<uint64_t>(max_grid, (uint64_t)ctx->cus * occ));
    if (prepare_only) return hipSuccess;
    hipLaunchKernelGGL(kernel, dim3((uint32_t)grid), dim3(BLOCK), lds, stream, a);
    return hipGetLastError();
}

uint32_t star_blocks_per_cu(uint32_t page_rows, size_t lds) {
    int occ = 0;
    hipError_t e = hipErrorInvalidValue;
    switch (page_rows) {
    case 1024: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_star<4, false, false>, BLOCK, lds); break;
    case 2048: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_star<8, false, false>, BLOCK, lds); break;
    case 4096: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_star<16, false, false>, BLOCK, lds); break;
    case 8192: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_star<32, false, false>, BLOCK, lds); break;
    default: break;
    }
    return e == hipSuccess ? (uint32_t)std::max(occ, 0) : 0u;
}

hipError_t launch_star(golap_ctx *ctx, const StarLaunch &l, hipStream_t stream, bool prepare_only) {
    if (l.n_cols > STAR_MAX_COLS || (!l.queries && l.n_queries != 1) || (l.items && !l.queue) ||
        (l.items && (!l.pages_per_item || l.pages_per_item > l.pages_per_chunk)))
        return hipErrorInvalidValue;
    if (l.items4 && (!l.queries || !l.queue || !l.mfilt || l.mcol >= l.n_cols || !l.pages_per_item ||
                     l.pages_per_item > l.pages_per_chunk))
        return hipErrorInvalidValue;
    StarArgs a{};
    for (uint32_t c = 0; c < l.n_cols; ++c) a.col[c] = StarCol{l.col[c], l.col_window[c], l.col_anchor[c] - 1};
    a.page_rows = l.page_rows;
    a.pages_per_chunk = l.pages_per_chunk;
    a.n_rows = l.n_rows;
    a.n_pages = l.n_pages;
    a.n_windows = l.n_windows;
    a.n_anchors = l.n_anchors;
    a.n_queries = l.n_queries;
    a.q0 = l.q0;
    a.queries = l.queries;
    a.n_items = l.n_items;
    a.items = l.items;
    a.queue = l.queue;
    a.pages_per_item = l.pages_per_item;
    a.filter_words = l.filter_words;
    a.lds_gslots = l.lds_gslots;
    a.group_bits = l.group_bits;
    a.err = l.err;
    a.items4 = l.items4;
    a.mfilt = l.mfilt;
    a.mcol = l.mcol;
    a.n_cols = l.n_cols;
    a.emit = l.emit;
    a.emit_n = l.emit_n;
    a.emit_cap = l.emit_cap;
    a.emit_dim = l.emit_dim;
    if (l.emit && (l.items4 || l.n_queries != 1 || !l.emit_n || l.lds_gslots)) return hipErrorInvalidValue;
    if (l.items4 && l.n_anchors) return hipErrorInvalidValue;  // the shared-filter kernel decodes into windows
    const uint32_t mb = l.max_blocks_per_cu;
    if (l.items4) {  // shared filter: candidate list and windows only
        a.filter_words = 0;
        a.lds_gslots = 0;
        const size_t lds = star_lds_bytes(0, l.page_rows, l.n_windows, 0);
        if (lds > (size_t)ctx->lds_max) return hipErrorInvalidConfiguration;
        switch (l.page_rows) {
        case 1024: return launch_star_t(ctx, k_star_shared<4>, a, lds, stream, mb, ~0ull, prepare_only);
        case 2048: return launch_star_t(ctx, k_star_shared<8>, a, lds, stream, mb, ~0ull, prepare_only);
        case 4096: return launch_star_t(ctx, k_star_shared<16>, a, lds, stream, mb, ~0ull, prepare_only);
        case 8192: return launch_star_t(ctx, k_star_shared<32>, a, lds, stream, mb, ~0ull, prepare_only);
        default: return hipErrorInvalidValue;
        }
    }
    if (!l.items && !a.n_items) return hipSuccess;
    const size_t lds = star_lds_bytes(a.filter_words, l.page_rows, l.n_windows, l.lds_gslots, l.n_anchors);
    if (lds > (size_t)ctx->lds_max) return hipErrorInvalidConfiguration;
    // the queue schedule's item count is only known on the device: fill the chip
    const uint64_t max_grid = l.items ? ~0ull : a.n_items;
#define GOLAP_STAR_CASE(ROWS, PER)                                                                              \
    case ROWS:                                                                                                  \
        if (l.emit)                                                                                             \
            return l.items ? launch_star_t(ctx, k_star<PER, true, true>, a, lds, stream, mb, max_grid, prepare_only)   \
                           : launch_star_t(ctx, k_star<PER, false, true>, a, lds, stream, mb, max_grid, prepare_only); \
        return l.items ? launch_star_t(ctx, k_star<PER, true, false>, a, lds, stream, mb, max_grid, prepare_only)      \
                       : launch_star_t(ctx, k_star<PER, false, false>, a, lds, stream, mb, max_grid, prepare_only);
    switch (l.page_rows) {
        GOLAP_STAR_CASE(1024, 4)
        GOLAP_STAR_CASE(2048, 8)
        GOLAP_STAR_CASE(4096, 16)
        GOLAP_STAR_CASE(8192, 32)
    default: return hipErrorInvalidValue;
    }
#undef GOLAP_STAR_CASE
}

// Non-empty groups of a batch's dense results: entry i of the queries' concatenated domains
// (qmeta[q] = {result offset in the arena, domain, first flattened index}) with a count is
// appended as {query, group index} + sum + count, one atomic per wave. Appends stop at cap (the
// count still grows, so the host sees it).
__global__ __launch_bounds__(BLOCK) void k_compact_groups(const uint8_t *arena, const ulonglong2 *qoff,
                                                          const uint64_t *qfirst, uint32_t nq, uint64_t total,
                                                          uint2 *out_ql, int64_t *out_sum, uint64_t *out_cnt,
                                                          uint64_t cap, unsigned long long *n_out) {
    const uint32_t lane = threadIdx.x % WAVE;
    const uint64_t stride = (uint64_t)gridDim.x * BLOCK;
    for (uint64_t i0 = (uint64_t)blockIdx.x * BLOCK + (threadIdx.x - lane); i0 < total; i0 += stride) {
        const uint64_t i = i0 + lane;
        uint32_t q = 0;
        uint64_t cnt = 0, g = 0;
        if (i < total) {
            uint32_t lo = 0, hi = nq - 1;  // last query whose first index is <= i
            while (lo < hi) {
                const uint32_t mid = (lo + hi + 1) >> 1;
                if (qfirst[mid] <= i) lo = mid;
                else hi = mid - 1;
            }
            q = lo;
            g = i - qfirst[q];
            cnt = reinterpret_cast<const uint64_t *>(arena + qoff[q].x)[qoff[q].y + g];
        }
        const uint64_t m = __ballot(cnt != 0);
        if (!m) continue;
        unsigned long long base = 0;
        if (lane == 0) base = atomicAdd(n_out, (unsigned long long)__popcll(m));
        base = __shfl(base, 0, WAVE);
        const uint64_t pos = base + __popcll(m & ((1ull << lane) - 1));
        if (cnt && pos < cap) {
            out_ql[pos] = make_uint2(q, (uint32_t)g);
            out_sum[pos] = reinterpret_cast<const int64_t *>(arena + qoff[q].x)[g];
            out_cnt[pos] = cnt;
        }
    }
}

hipError_t launch_compact_groups(const uint8_t *arena, const void *qoff, const uint64_t *qfirst, uint32_t nq,
                                 uint64_t total, uint2 *out_ql, int64_t *out_sum, uint64_t *out_cnt, uint64_t cap,
                                 unsigned long long *n_out, uint32_t cus, hipStream_t stream) {
    if (!nq || !total) return hipSuccess;
    const uint64_t blocks = std::min<uint64_t>((total + BLOCK - 1) / BLOCK, (uint64_t)cus * 8);
    hipLaunchKernelGGL(k_compact_groups, dim3((uint32_t)blocks), dim3(BLOCK), 0, stream, arena,
                       static_cast<const ulonglong2 *>(qoff), qfirst, nq, total, out_ql, out_sum, out_cnt, cap, n_out);
    return hipGetLastError();
}

// Batch work list. Wave w handles the flattened chunk indexes [64w, 64w + 64) of the queries'
// slices (chunk_off: prefix of slice lengths in chunks); survivors are compacted with a wave
// ballot and one global atomic per wave, then write `parts` items each.
__global__ __launch_bounds__(BLOCK) void k_batch_items(const uint8_t *qbase, uint32_t qstride, uint32_t nq,
                                                       const uint64_t *chunk_off, uint64_t total, uint32_t parts,
                                                       uint2 *items, uint32_t *queue) {
    const uint32_t lane = threadIdx.x % WAVE;
    const uint64_t waves = (uint64_t)gridDim.x * (BLOCK / WAVE);
    for (uint64_t w = (uint64_t)blockIdx.x * (BLOCK / WAVE) + threadIdx.x / WAVE; w * WAVE < total; w += waves) {
        const uint64_t i = w * WAVE + lane;
        bool pass = false;
        uint32_t q = 0;
        uint64_t c = 0;
        if (i < total) {
            uint32_t lo = 0, hi = nq;  // last query with chunk_off[q] <= i
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) / 2;
                if (chunk_off[mid] <= i) lo = mid;
                else hi = mid;
            }
            q = lo;
            const SliceDev &Q = *reinterpret_cast<const SliceDev *>(qbase + (size_t)q * qstride);
            c = Q.chunk_begin + (i - chunk_off[q]);
            pass = slice_keeps(Q, c);
        }
        const uint64_t ball = __ballot(pass);
        uint32_t base = 0;
        if (lane == 0 && ball) base = atomicAdd(&queue[1], (uint32_t)__popcll(ball) * parts);
        base = __shfl(base, 0, WAVE);
        if (pass) {
            const uint64_t below = lane ? (ball << (64 - lane)) : 0ull;
            uint32_t slot = base + (uint32_t)__popcll(below) * parts;
            for (uint32_t t = 0; t < parts; ++t) items[slot + t] = make_uint2((uint32_t)c, (q << 8) | t);
        }
    }
}

hipError_t launch_batch_items(const void *qbase, uint32_t qstride, uint32_t n_queries, const uint64_t *chunk_off,
                              uint64_t total_chunks, uint32_t parts, uint2 *items, uint32_t *queue,
                              hipStream_t stream) {
    if (!n_queries || !parts || parts > 256 || n_queries > (1u << 24)) return hipErrorInvalidValue;
    if (!total_chunks) return hipSuccess;
    const uint64_t waves = (total_chunks + WAVE - 1) / WAVE;
    const uint32_t grid = (uint32_t)std::min<uint64_t>((waves + BLOCK / WAVE - 1) / (BLOCK / WAVE), 4096);
    hipLaunchKernelGGL(k_batch_items, dim3(grid), dim3(BLOCK), 0, stream, static_cast<const uint8_t *>(qbase),
                       qstride, n_queries, chunk_off, total_chunks, parts, items, queue);
    return hipGetLastError();
}

// group_by_agg_2col / _4col (core/group.cuh:14-36) over plain device columns: one thread per
// row forms the dense group index of up to four int key columns and adds the value into the
// LDS-private table (spilling to global atomics), merged once per workgroup.
struct GroupByArgs {
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
};

__device__ __forceinline__ int64_t ld_typed(const void *col, uint32_t type, uint64_t row) {
    switch (type) {
    case GOLAP_I8: return static_cast<const int8_t *>(col)[row];
    case GOLAP_U8: return static_cast<const uint8_t *>(col)[row];
    case GOLAP_I16: return static_cast<const int16_t *>(col)[row];
    case GOLAP_U16: return static_cast<const uint16_t *>(col)[row];
    case GOLAP_I32: return static_cast<const int32_t *>(col)[row];
    case GOLAP_U32: return static_cast<const uint32_t *>(col)[row];
    default: return static_cast<const int64_t *>(col)[row];
    }
}

__global__ __launch_bounds__(BLOCK) void k_group_by(GroupByArgs g) {
    extern __shared__ __align__(16) uint8_t lds[];
    GroupTable L;
    L.slots = g.lds_gslots;
    L.gkey = reinterpret_cast<uint32_t *>(lds);
    L.gcnt = L.gkey + g.lds_gslots;
    L.gsum = reinterpret_cast<unsigned long long *>(lds + (size_t)g.lds_gslots * 8);
    for (uint32_t i = threadIdx.x; i < g.lds_gslots; i += BLOCK) {
        L.gkey[i] = GROUP_EMPTY;
        L.gcnt[i] = 0;
        L.gsum[i] = 0;
    }
    __syncthreads();
    for (uint64_t r = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; r < g.n; r += (uint64_t)gridDim.x * BLOCK) {
        uint64_t gidx = 0;
        bool ok = true;
        for (uint32_t k = 0; k < g.n_keys; ++k) {
            const uint64_t v = (uint64_t)(ld_typed(g.key[k], g.key_type[k], r) - (int64_t)g.key_lo[k]);
            if (v >= g.key_card[k]) ok = false;
            gidx += v * g.key_stride[k];
        }
        if (!ok) {
            raise_err(g.err, DEVERR_KEYRANGE);
            continue;
        }
        const int64_t v = ld_typed(g.val, g.val_type, r);
        if (!group_add_lds(L, (uint32_t)gidx, v)) group_add_global(g.g_sum, g.g_cnt, gidx, v);
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < g.lds_gslots; i += BLOCK) {
        const uint32_t k = L.gkey[i];
        if (k == GROUP_EMPTY) continue;
        atomicAdd(reinterpret_cast<unsigned long long *>(&g.g_sum[k]), L.gsum[i]);
        atomicAdd(reinterpret_cast<unsigned long long *>(&g.g_cnt[k]), (unsigned long long)L.gcnt[i]);
    }
}

hipError_t launch_group_by(golap_ctx *ctx, const GroupByLaunch &l) {
    GroupByArgs g{};
    for (uint32_t k = 0; k < l.n_keys; ++k) {
        g.key[k] = l.key[k];
        g.key_type[k] = l.key_type[k];
        g.key_lo[k] = l.key_lo[k];
        g.key_card[k] = l.key_card[k];
        g.key_stride[k] = l.key_stride[k];
    }
    g.n_keys = l.n_keys;
    g.val = l.val;
    g.val_type = l.val_type;
    g.n = l.n;
    g.lds_gslots = l.lds_gslots;
    g.g_sum = l.g_sum;
    g.g_cnt = l.g_cnt;
    g.err = l.err;
    const size_t lds = (size_t)l.lds_gslots * 16;
    int occ = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_group_by, BLOCK, lds);
    if (e != hipSuccess) return e;
    const uint64_t grid = std::max<uint64_t>(1, std::min<uint64_t>((l.n + BLOCK - 1) / BLOCK,
                                                                   (uint64_t)ctx->cus * std::max(occ, 1)));
    hipLaunchKernelGGL(k_group_by, dim3((uint32_t)grid), dim3(BLOCK), lds, l.stream ? l.stream : ctx->stream, g);
    return hipGetLastError();
}

// Blocked Bloom filter (2^wlog2 words) over the keys of a table: the driver filter when the key
// range is too wide for the dense bitmap.
__global__ void k_hash_filter(const uint64_t *slots, uint64_t n, uint32_t *filter, uint32_t wlog2) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t s = slots[i];
        if (s == 0) continue;
        const uint32_t h = fmix32((uint32_t)s ^ 0x80000000u);
        atomicOr(&filter[bloom_word(h, wlog2)], bloom_bits(h));
    }
}

hipError_t launch_hash_filter(golap_ctx *ctx, const uint64_t *slots, uint32_t log2cap, uint32_t *filter,
                              uint32_t words_log2) {
    const uint64_t n = 1ull << log2cap;
    const uint32_t grid = (uint32_t)std::min<uint64_t>((n + BLOCK - 1) / BLOCK, 4096);
    hipLaunchKernelGGL(k_hash_filter, dim3(std::max(grid, 1u)), dim3(BLOCK), 0, ctx->stream, slots, n, filter,
                       words_log2);
    return hipGetLastError();
}

}  // namespace golap_impl
