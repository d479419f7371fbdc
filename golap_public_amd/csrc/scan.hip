// Fused scan + filter + group-by over paged typed columns (gfx950): the trip-table plans.
//
// The reference groups with group_by_agg_2col / 4col over decompressed attribute columns and
// HashAggregate::add (core/group.cuh:14-36, core/dev_structs.cuh:209-244) after its range
// predicates (usecases/ssb/queries/01.cuh:26-54). Here one persistent kernel reads the
// compressed pages of every column a query uses:
//   - in place (raw and literal-only LZ4 pages), by point access (frame-of-reference and RLE
//     pages, delta pages with bw 0), or decoded into an LDS window by one wave (other pages);
//   - range predicates on any column; group keys computed per row as (v - base) / div with a
//     magic-number division (hour of a timestamp, zone id), mixed-radix group index;
//   - SUM/COUNT into an LDS-private hashed table keyed by (query, group) that spills rows to
//     global int64 atomics when a slot run is full, and is flushed to the global dense result
//     when it fills up at a work-item end (sorted time columns give every stretch of rows its
//     own small set of groups).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "common.hpp"
#include "device_codecs.hpp"
#include "star_dev.hpp"

namespace golap_impl {
using namespace golap_dev;
using namespace stardev;

namespace {

struct ScanArgs {
    ScanColView col[SCAN_MAX_COLS];
    uint32_t n_cols;
    uint32_t page_rows, pages_per_chunk;
    uint64_t n_rows, n_pages;
    uint32_t win_bytes;
    uint32_t n_queries;
    ScanQueryDev q0;
    const ScanQueryDev *queries;
    uint64_t n_items;
    const uint2 *items;
    uint32_t *queue;
    uint32_t pages_per_item;
    uint32_t lds_gslots, group_bits, flush_fill;
    uint32_t *err;
};

constexpr uint32_t SQ_WORDS = sizeof(ScanQueryDev) / 4;
static_assert(sizeof(ScanQueryDev) % 8 == 0, "query descriptor copied as words");

// How one typed column is read at a tile row.
enum : uint32_t { TA_DIRECT = 0, TA_WINDOW = 1, TA_AFFINE = 2, TA_RLE = 3, TA_FOR = 4, TA_BAD = 5 };
struct TAcc {
    const uint8_t *base;  // DIRECT: payload; WINDOW: LDS window; RLE: page start; FOR: fields
    int64_t a, b;         // AFFINE: v[r] = a + r*b; FOR: a = frame of reference
    uint32_t mode;
    uint32_t aux;         // RLE: runs; FOR: bits per field
};

struct ScanShared {
    ScanQueryDev q;
    TAcc acc[SCAN_MAX_COLS];
    uint32_t tile_bad, fill, item;
};

__device__ __forceinline__ uint32_t elem_of(uint32_t type) {
    return type <= GOLAP_U8 ? 1u : type <= GOLAP_U16 ? 2u : type <= GOLAP_U32 ? 4u : 8u;
}

// The low elem bytes of raw as a value of type (sign or zero extension).
__device__ __forceinline__ int64_t typed(uint64_t raw, uint32_t type) {
    switch (type) {
    case GOLAP_I8: return (int8_t)raw;
    case GOLAP_U8: return (uint8_t)raw;
    case GOLAP_I16: return (int16_t)raw;
    case GOLAP_U16: return (uint16_t)raw;
    case GOLAP_I32: return (int32_t)raw;
    case GOLAP_U32: return (uint32_t)raw;
    default: return (int64_t)raw;
    }
}

__device__ __forceinline__ uint64_t load_elem(const uint8_t *p, uint32_t elem) {
    switch (elem) {
    case 1: return *p;
    case 2: return *reinterpret_cast<const uint16_t *>(p);
    case 4: return *reinterpret_cast<const uint32_t *>(p);
    default: return *reinterpret_cast<const uint64_t *>(p);
    }
}

__device__ inline TAcc tcol_access(const ScanColView &v, uint64_t p, uint32_t rows, uint8_t *wwin) {
    TAcc x{};
    const uint32_t elem = v.c.elem;
    const uint8_t codec = v.c.page_codec[p];
    const uint8_t *src = v.c.comp + v.c.page_off[p];
    const uint32_t csize = v.c.page_csize[p];
    const uint32_t cd = codec & PAGE_CODEC_MASK;
    if (is_direct(codec)) {
        const uint32_t hdr = direct_hdr(codec, rows * elem);
        x.base = src + hdr;
        x.mode = (csize == hdr + rows * elem && !(reinterpret_cast<uintptr_t>(x.base) % elem)) ? TA_DIRECT : TA_BAD;
    } else if (cd == golap_fmt::DBP && (codec & golap_fmt::PAGE_POINT)) {
        if (csize < 32 || src[20] != 0 || src[21] != elem || ld_u32_unaligned(src + 16) != rows) {
            x.mode = TA_BAD;
        } else {
            x.a = (int64_t)ld_u64_unaligned(src);
            x.b = (int64_t)ld_u64_unaligned(src + 8);
            x.mode = TA_AFFINE;
        }
    } else if (cd == golap_fmt::FOR) {
        if (for_check(src, csize, elem) != (int64_t)rows) {
            x.mode = TA_BAD;
        } else {
            x.base = src + 32;
            x.a = (int64_t)ld_u64_unaligned(src);
            x.aux = src[20];
            x.mode = TA_FOR;
        }
    } else if (cd == golap_fmt::RLE) {
        const uint32_t nruns = csize >= 16 ? ld_u32_unaligned(src) : 0u;
        const uint64_t vbytes = ((uint64_t)nruns * elem + 3) & ~3ull;
        const bool ok = csize >= 16 && src[8] == elem && ld_u32_unaligned(src + 4) == rows && nruns >= 1 &&
                        16ull + vbytes + 4ull * nruns <= csize;
        x.base = src;
        x.aux = nruns;
        x.b = (int64_t)vbytes;
        x.mode = ok ? TA_RLE : TA_BAD;
    } else {
        x.base = wwin;
        x.mode = wwin ? TA_WINDOW : TA_BAD;
    }
    return x;
}

__device__ __forceinline__ int64_t tcol_value(const TAcc &x, uint32_t r, uint32_t type) {
    const uint32_t elem = elem_of(type);
    uint64_t raw;
    switch (x.mode) {
    case TA_AFFINE: raw = (uint64_t)x.a + (uint64_t)r * (uint64_t)x.b; break;
    case TA_FOR: raw = for_value(x.base, (uint64_t)x.a, x.aux, r); break;
    case TA_RLE: {
        const uint8_t *ends = x.base + 16 + x.b;
        uint32_t lo = 0, hi = x.aux - 1;
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (ld_u32_unaligned(ends + 4ull * mid) > r) hi = mid;
            else lo = mid + 1;
        }
        raw = 0;
        for (uint32_t b = 0; b < elem; ++b) raw |= (uint64_t)x.base[16 + (uint64_t)lo * elem + b] << (8 * b);
        break;
    }
    default: raw = load_elem(x.base + (size_t)r * elem, elem); break;
    }
    return typed(raw, type);
}

// group_add_lds that also counts the keys the table holds (fill), for the flush decision.
__device__ __forceinline__ bool group_add_counted(const GroupTable &T, uint32_t key, int64_t v, uint32_t *fill) {
    const uint32_t mask = T.slots - 1;
    uint32_t h = (hash32(key) >> 7) & mask;
    for (int i = 0; i < GROUP_PROBES; ++i) {
        uint32_t k = T.gkey[h];
        if (k == GROUP_EMPTY) {
            const uint32_t old = atomicCAS(&T.gkey[h], GROUP_EMPTY, key);
            if (old == GROUP_EMPTY) atomicAdd(fill, 1u);
            k = old == GROUP_EMPTY ? key : old;
        }
        if (k == key) {
            atomicAdd(&T.gsum[h], (unsigned long long)v);
            atomicAdd(&T.gcnt[h], 1u);
            return true;
        }
        h = (h + 1) & mask;
    }
    return false;
}

// Moves every LDS entry into its query's dense global result and empties the table.
__device__ void scan_flush(const ScanArgs &a, const GroupTable &T, ScanShared &S) {
    __syncthreads();
    const uint32_t gmask = (1u << a.group_bits) - 1;
    for (uint32_t i = threadIdx.x; i < T.slots; i += BLOCK) {
        const uint32_t k = T.gkey[i];
        if (k == GROUP_EMPTY) continue;
        const ScanQueryDev *Q = a.queries ? a.queries + (k >> a.group_bits) : &a.q0;
        atomicAdd(reinterpret_cast<unsigned long long *>(&Q->g_sum[k & gmask]), T.gsum[i]);
        atomicAdd(reinterpret_cast<unsigned long long *>(&Q->g_cnt[k & gmask]), (unsigned long long)T.gcnt[i]);
        T.gkey[i] = GROUP_EMPTY;
        T.gcnt[i] = 0;
        T.gsum[i] = 0;
    }
    if (threadIdx.x == 0) S.fill = 0;
    __syncthreads();
}

__device__ void scan_set_query(const ScanArgs &a, ScanShared &S, uint32_t q) {
    __syncthreads();
    const uint32_t *src = reinterpret_cast<const uint32_t *>(a.queries ? a.queries + q : &a.q0);
    uint32_t *dst = reinterpret_cast<uint32_t *>(&S.q);
    for (uint32_t i = threadIdx.x; i < SQ_WORDS; i += BLOCK) dst[i] = src[i];
    __syncthreads();
}

__device__ void scan_tile(const ScanArgs &a, ScanShared &S, uint32_t q, uint64_t p, uint8_t *win,
                          const GroupTable &T) {
    const uint32_t wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    const uint32_t rows = (uint32_t)min<uint64_t>(a.page_rows, a.n_rows - p * a.page_rows);
    const uint32_t used = S.q.colmask;
    if (threadIdx.x == 0) S.tile_bad = 0;
    if (threadIdx.x < a.n_cols && ((used >> threadIdx.x) & 1u)) {
        const ScanColView &v = a.col[threadIdx.x];
        S.acc[threadIdx.x] = tcol_access(v, p, rows, v.woff >= 0 ? win + v.woff : nullptr);
    }
    __syncthreads();
    bool bad = false;
    for (uint32_t c = 0; c < a.n_cols; ++c) bad |= ((used >> c) & 1u) && S.acc[c].mode == TA_BAD;
    if (bad) {  // uniform
        if (threadIdx.x == 0) raise_err(a.err, DEVERR_CORRUPT);
        __syncthreads();
        return;
    }
    for (uint32_t c = wave; c < a.n_cols; c += BLOCK / WAVE) {
        if (!((used >> c) & 1u) || S.acc[c].mode != TA_WINDOW) continue;
        const ColView &cv = a.col[c].c;
        const uint32_t want = rows * cv.elem;
        const int64_t got = page_decode_wave(cv.page_codec[p], cv.comp + cv.page_off[p], cv.page_csize[p],
                                             const_cast<uint8_t *>(S.acc[c].base), want, cv.elem);
        if (got != (int64_t)want && lane == 0) {
            S.tile_bad = 1;
            raise_err(a.err, DEVERR_CORRUPT);
        }
    }
    __syncthreads();
    if (!S.tile_bad) {
        const uint32_t np = S.q.n_preds, nk = S.q.n_keys;
        const int32_t vc = S.q.val_col;
        const bool use_lds = S.q.use_lds && T.slots;
        const uint32_t qkey = q << a.group_bits;
        for (uint32_t r = threadIdx.x; r < rows; r += BLOCK) {
            bool ok = true;
            for (uint32_t i = 0; i < np; ++i) {
                const ScanPredDev &pr = S.q.pred[i];
                const int64_t v = tcol_value(S.acc[pr.col], r, a.col[pr.col].type);
                ok &= v >= pr.lo && v <= pr.hi;
            }
            if (!ok) continue;
            uint64_t g = 0;
            for (uint32_t i = 0; i < nk && ok; ++i) {
                const ScanKeyDev &k = S.q.key[i];
                const uint64_t u = (uint64_t)(tcol_value(S.acc[k.col], r, a.col[k.col].type) - k.base);
                if (u >= k.span) {
                    raise_err(a.err, DEVERR_KEYRANGE);
                    ok = false;
                } else {
                    g += (uint64_t)golap_fmt::udiv32((uint32_t)u, k.div) * k.stride;
                }
            }
            if (!ok) continue;
            const int64_t v = vc >= 0 ? tcol_value(S.acc[vc], r, a.col[vc].type) : 0;
            if (!use_lds || !group_add_counted(T, qkey | (uint32_t)g, v, &S.fill))
                group_add_global(S.q.g_sum, S.q.g_cnt, g, v);
        }
    }
    __syncthreads();
}

// Persistent scan kernel. Static schedule: the queries' pages flattened in query order, workgroup
// b takes [b*n/G, (b+1)*n/G). Queue schedule: device-built (chunk, query, part) items taken with
// one atomic each.
template <bool QUEUE>
__global__ __launch_bounds__(BLOCK) void k_scan(ScanArgs a) {
    extern __shared__ __align__(16) uint8_t lds[];
    __shared__ ScanShared S;
    uint8_t *win = lds;
    GroupTable T;
    T.slots = a.lds_gslots;
    T.gkey = reinterpret_cast<uint32_t *>(lds + a.win_bytes);
    T.gcnt = T.gkey + a.lds_gslots;
    T.gsum = reinterpret_cast<unsigned long long *>(lds + a.win_bytes + (size_t)a.lds_gslots * 8);
    for (uint32_t i = threadIdx.x; i < a.lds_gslots; i += BLOCK) {
        T.gkey[i] = GROUP_EMPTY;
        T.gcnt[i] = 0;
        T.gsum[i] = 0;
    }
    if (threadIdx.x == 0) S.fill = 0;
    uint32_t cur = 0xffffffffu;
    if (!QUEUE) {
        const uint64_t i0 = a.n_items * blockIdx.x / gridDim.x, i1 = a.n_items * (blockIdx.x + 1) / gridDim.x;
        uint32_t q = 0;
        for (uint64_t it = i0; it < i1; ++it) {
            if (a.queries)
                while (q + 1 < a.n_queries && it >= a.queries[q + 1].s.item_begin) ++q;
            if (q != cur) {
                scan_set_query(a, S, q);
                cur = q;
            }
            const uint64_t p = S.q.s.page_begin + (it - S.q.s.item_begin);
            if (!slice_keeps(S.q.s, p / a.pages_per_chunk)) continue;
            scan_tile(a, S, q, p, win, T);
            if (T.slots && S.fill > a.flush_fill) scan_flush(a, T, S);  // S.fill read after the tile's barrier
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
                scan_set_query(a, S, q);
                cur = q;
            }
            const uint64_t c0 = (uint64_t)w.x * a.pages_per_chunk;
            const uint64_t pb = c0 + (uint64_t)part * a.pages_per_item;
            const uint64_t pe = min(min(pb + a.pages_per_item, c0 + a.pages_per_chunk), a.n_pages);
            for (uint64_t p = pb; p < pe; ++p) scan_tile(a, S, q, p, win, T);
            if (T.slots && S.fill > a.flush_fill) scan_flush(a, T, S);
        }
    }
    if (T.slots) scan_flush(a, T, S);
}

template <typename K>
hipError_t launch_scan_t(golap_ctx *ctx, K kernel, const ScanArgs &a, size_t lds, hipStream_t stream,
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

}  // namespace

size_t scan_lds_bytes(uint32_t win_bytes, uint32_t lds_gslots) { return (size_t)win_bytes + (size_t)lds_gslots * 16; }

hipError_t launch_scan(golap_ctx *ctx, const ScanLaunch &l, hipStream_t stream, bool prepare_only) {
    if (l.n_cols > SCAN_MAX_COLS || (!l.queries && l.n_queries != 1) || (l.items && !l.queue) ||
        (l.items && (!l.pages_per_item || l.pages_per_item > l.pages_per_chunk)) || l.win_bytes % 16 ||
        (l.lds_gslots & (l.lds_gslots - 1)) || l.group_bits > 31)
        return hipErrorInvalidValue;
    ScanArgs a{};
    for (uint32_t c = 0; c < l.n_cols; ++c) a.col[c] = l.col[c];
    a.n_cols = l.n_cols;
    a.page_rows = l.page_rows;
    a.pages_per_chunk = l.pages_per_chunk;
    a.n_rows = l.n_rows;
    a.n_pages = l.n_pages;
    a.win_bytes = l.win_bytes;
    a.n_queries = l.n_queries;
    a.q0 = l.q0;
    a.queries = l.queries;
    a.n_items = l.n_items;
    a.items = l.items;
    a.queue = l.queue;
    a.pages_per_item = l.pages_per_item;
    a.lds_gslots = l.lds_gslots;
    a.group_bits = l.group_bits;
    a.flush_fill = l.flush_fill;
    a.err = l.err;
    if (!l.items && !a.n_items) return hipSuccess;
    const size_t lds = scan_lds_bytes(l.win_bytes, l.lds_gslots);
    if (lds > (size_t)ctx->lds_max) return hipErrorInvalidConfiguration;
    const uint64_t max_grid = l.items ? ~0ull : a.n_items;
    return l.items ? launch_scan_t(ctx, k_scan<true>, a, lds, stream, l.max_blocks_per_cu, max_grid, prepare_only)
                   : launch_scan_t(ctx, k_scan<false>, a, lds, stream, l.max_blocks_per_cu, max_grid, prepare_only);
}

}  // namespace golap_impl
