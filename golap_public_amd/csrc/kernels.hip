// gfx950 kernels of the compressed-chunk -> filter -> join-probe -> int64-sum path.
//
// Launch shape everywhere: 256-thread workgroups (4 wave64s). Scan kernels are persistent
// (grid = CUs x resident workgroups) and walk pages with a grid stride, so per-workgroup set-up
// (copying a dimension table into LDS) is paid once per workgroup, not once per page.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "common.hpp"
#include "device_codecs.hpp"
#include "q1_dev.hpp"

namespace golap_impl {

__global__ void k_fill_zero(uint64_t *p, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

// Zeroing by a kernel rather than hipMemsetAsync, so that captured graphs hold kernel nodes only.
hipError_t launch_fill_zero(void *p, uint64_t bytes, hipStream_t stream) {
    if (bytes % 8 || reinterpret_cast<uintptr_t>(p) % 8) return hipErrorInvalidValue;
    const uint64_t n = bytes / 8;
    const uint32_t grid = (uint32_t)std::min<uint64_t>((n + BLOCK - 1) / BLOCK, 4096);
    hipLaunchKernelGGL(k_fill_zero, dim3(std::max(grid, 1u)), dim3(BLOCK), 0, stream, static_cast<uint64_t *>(p), n);
    return hipGetLastError();
}

struct PutWords {
    uint32_t w[PUT_WORDS_MAX / 4];
};
__global__ void k_put_words(uint32_t *dst, PutWords v, uint32_t n) {
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) dst[i] = v.w[i];
}

hipError_t launch_put_words(void *dst, const void *src, uint32_t bytes, hipStream_t stream) {
    if (bytes % 4 || bytes > PUT_WORDS_MAX || reinterpret_cast<uintptr_t>(dst) % 4) return hipErrorInvalidValue;
    if (!bytes) return hipSuccess;
    PutWords v;
    memcpy(v.w, src, bytes);
    hipLaunchKernelGGL(k_put_words, dim3(1), dim3(BLOCK), 0, stream, static_cast<uint32_t *>(dst), v, bytes / 4);
    return hipGetLastError();
}

// Streaming read of n16 16-byte words, grid-stride with four loads in flight per thread: the
// HBM read bandwidth a plain read kernel reaches, beside which the fused scans are reported.
// The sink is written only when the checksum equals `never` (a value the host picks), so the
// loads stay.
__global__ __launch_bounds__(BLOCK) void k_read_probe(const uint4 *p, uint64_t n16, uint32_t *sink, uint32_t never) {
    uint32_t acc = 0;
    const uint64_t step = (uint64_t)gridDim.x * BLOCK;
    uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
    for (; i + 3 * step < n16; i += 4 * step) {
        const uint4 a = p[i], b = p[i + step], c = p[i + 2 * step], d = p[i + 3 * step];
        acc ^= a.x ^ a.y ^ a.z ^ a.w ^ b.x ^ b.y ^ b.z ^ b.w ^ c.x ^ c.y ^ c.z ^ c.w ^ d.x ^ d.y ^ d.z ^ d.w;
    }
    for (; i < n16; i += step) {
        const uint4 a = p[i];
        acc ^= a.x ^ a.y ^ a.z ^ a.w;
    }
    if (acc == never) sink[0] = acc;
}

hipError_t launch_read_probe(golap_ctx *ctx, const void *p, uint64_t bytes, uint32_t blocks_per_cu, uint32_t *sink,
                             hipStream_t stream) {
    if (bytes % 16 || reinterpret_cast<uintptr_t>(p) % 16 || !blocks_per_cu || blocks_per_cu > 32)
        return hipErrorInvalidValue;
    const uint32_t grid = (uint32_t)ctx->cus * blocks_per_cu;
    hipLaunchKernelGGL(k_read_probe, dim3(grid), dim3(BLOCK), 0, stream, static_cast<const uint4 *>(p), bytes / 16,
                       sink, 0x9e3779b9u);
    return hipGetLastError();
}

hipError_t launch_hash_clear(golap_ctx *ctx, uint64_t *slots, uint64_t n) {
    const uint32_t grid = (uint32_t)std::min<uint64_t>((n + BLOCK - 1) / BLOCK, 4096);
    hipLaunchKernelGGL(k_fill_zero, dim3(std::max(grid, 1u)), dim3(BLOCK), 0, ctx->stream, slots, n);
    return hipGetLastError();
}

struct BuildArgs {
    const void *keys;
    uint32_t key_type;
    uint64_t n;
    PredView preds[MAX_PREDS];
    uint32_t n_preds;
    const int32_t *payload;
    uint64_t *slots;
    uint32_t log2cap;
    uint64_t *filled;
    int32_t *kminmax;  // [0] min, [1] max of the inserted keys
    uint32_t *err;
};

template <typename T>
__device__ __forceinline__ T wave_min(T v) {
#pragma unroll
    for (int d = WAVE / 2; d > 0; d >>= 1) v = min(v, (T)__shfl_xor(v, d, WAVE));
    return v;
}
template <typename T>
__device__ __forceinline__ T wave_max(T v) {
#pragma unroll
    for (int d = WAVE / 2; d > 0; d >>= 1) v = max(v, (T)__shfl_xor(v, d, WAVE));
    return v;
}

// hash_map_build_pred_other (core/join.cuh:107-135): one thread per build row; the row is
// inserted iff every predicate passes. Insert is one 64-bit CAS on an empty slot (the key is
// the occupancy marker), so no thread ever waits on another.
__global__ __launch_bounds__(BLOCK) void k_hash_build(BuildArgs a) {
    uint32_t inserted = 0;
    int32_t kmin = INT32_MAX, kmax = INT32_MIN;
    for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < a.n;
         r += (uint64_t)gridDim.x * blockDim.x) {
        bool ok = true;
        for (uint32_t k = 0; k < a.n_preds; ++k) {
            const int64_t v = load_typed(a.preds[k].col, a.preds[k].type, r);
            ok = ok && v >= a.preds[k].lo && v <= a.preds[k].hi;
        }
        if (!ok) continue;
        const int64_t key = load_typed(a.keys, a.key_type, r);
        if (key <= (int64_t)INT32_MIN || key > (int64_t)INT32_MAX) {
            raise_err(a.err, DEVERR_KEYRANGE);
            continue;
        }
        const uint32_t pay = a.payload ? (uint32_t)a.payload[r] : (uint32_t)r;
        if (ht_insert(a.slots, a.log2cap, (int32_t)key, pay)) {
            ++inserted;
            kmin = min(kmin, (int32_t)key);
            kmax = max(kmax, (int32_t)key);
        } else {
            raise_err(a.err, DEVERR_CAPACITY);
        }
    }
    const uint32_t tot = wave_sum(inserted);
    kmin = wave_min(kmin);
    kmax = wave_max(kmax);
    if (__lane_id() == 0 && tot) {
        atomicAdd(reinterpret_cast<unsigned long long *>(a.filled), (unsigned long long)tot);
        atomicMin(&a.kminmax[0], kmin);
        atomicMax(&a.kminmax[1], kmax);
    }
}

// Exact membership bitmap over [kmin, kmin + krange) of the keys in a built table: the dense
// form of the join that the fused scans pin in LDS when the key range is small.
__global__ void k_bitmap_from_table(const uint64_t *slots, uint64_t n, uint32_t *bitmap,
                                    int32_t kmin, uint32_t krange, uint32_t *err) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t s = slots[i];
        if (s == 0) continue;
        const int32_t key = (int32_t)((uint32_t)s ^ 0x80000000u);
        const uint32_t off = (uint32_t)key - (uint32_t)kmin;
        if (off >= krange) {
            raise_err(err, DEVERR_CAPACITY);
            continue;
        }
        atomicOr(&bitmap[off >> 5], 1u << (off & 31));
    }
}

hipError_t launch_hash_bitmap(golap_ctx *ctx, golap_hashmap *m) {
    const uint64_t n = 1ull << m->log2cap;
    const uint32_t grid = (uint32_t)std::min<uint64_t>((n + BLOCK - 1) / BLOCK, 4096);
    hipLaunchKernelGGL(k_bitmap_from_table, dim3(std::max(grid, 1u)), dim3(BLOCK), 0, ctx->stream,
                       m->d_slots, n, m->d_bitmap, m->kmin, m->krange,
                       reinterpret_cast<uint32_t *>(ctx->d_scratch));
    return hipGetLastError();
}

hipError_t launch_hash_build(golap_ctx *ctx, golap_hashmap *m, uint64_t n_rows,
                             const PredView *preds, uint32_t n_preds, const int32_t *payload,
                             uint64_t *d_filled, int32_t *d_kminmax) {
    BuildArgs a{};
    a.keys = m->d_keys;
    a.key_type = m->key_type;
    a.n = n_rows;
    for (uint32_t i = 0; i < n_preds; ++i) a.preds[i] = preds[i];
    a.n_preds = n_preds;
    a.payload = payload;
    a.slots = m->d_slots;
    a.log2cap = m->log2cap;
    a.filled = d_filled;
    a.kminmax = d_kminmax;
    a.err = reinterpret_cast<uint32_t *>(ctx->d_scratch);
    const uint32_t grid = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((n_rows + BLOCK - 1) / BLOCK, 8192));
    hipLaunchKernelGGL(k_hash_build, dim3(grid), dim3(BLOCK), 0, ctx->stream, a);
    return hipGetLastError();
}

__global__ void k_hash_compact(const uint64_t *big, uint64_t n_big, uint64_t *small,
                               uint32_t small_log2, uint32_t *err) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_big;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t s = big[i];
        if (s == 0) continue;
        const int32_t key = (int32_t)((uint32_t)s ^ 0x80000000u);
        if (!ht_insert(small, small_log2, key, (uint32_t)(s >> 32))) raise_err(err, DEVERR_CAPACITY);
    }
}

hipError_t launch_hash_compact(golap_ctx *ctx, golap_hashmap *m) {
    const uint64_t n_big = 1ull << m->log2cap;
    const uint32_t grid = (uint32_t)std::min<uint64_t>((n_big + BLOCK - 1) / BLOCK, 4096);
    hipLaunchKernelGGL(k_hash_compact, dim3(std::max(grid, 1u)), dim3(BLOCK), 0, ctx->stream,
                       m->d_slots, n_big, m->d_small, m->small_log2,
                       reinterpret_cast<uint32_t *>(ctx->d_scratch));
    return hipGetLastError();
}

struct ProbeArgs {
    const uint64_t *slots;
    uint32_t log2cap;
    const void *keys;
    uint32_t key_type;
    uint64_t n;
    uint64_t *out_rid;
    uint64_t *hits;
};

// HashMap::probe per key (core/dev_structs.cuh:104-113), hits counted like hash_join_count
// (core/join.cuh:45-56) but aggregated per wave before the global atomic.
__global__ __launch_bounds__(BLOCK) void k_hash_probe(ProbeArgs a) {
    uint32_t hits = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < a.n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const int64_t key = load_typed(a.keys, a.key_type, i);
        uint32_t pay = 0;
        bool found = key > (int64_t)INT32_MIN && key <= (int64_t)INT32_MAX &&
                     ht_probe(a.slots, a.log2cap, (int32_t)key, pay);
        if (a.out_rid) a.out_rid[i] = found ? (uint64_t)pay : UINT64_MAX;
        hits += found;
    }
    const uint32_t tot = wave_sum(hits);
    if (a.hits && __lane_id() == 0 && tot) atomicAdd(reinterpret_cast<unsigned long long *>(a.hits), (unsigned long long)tot);
}

hipError_t launch_hash_probe(golap_ctx *ctx, const golap_hashmap *m, const void *keys,
                             uint32_t key_type, uint64_t n, uint64_t *out_rid, uint64_t *hits) {
    if (!n) return hipSuccess;
    ProbeArgs a{m->d_slots, m->log2cap, keys, key_type, n, out_rid, hits};
    const uint32_t grid = (uint32_t)std::min<uint64_t>((n + BLOCK - 1) / BLOCK, (uint64_t)ctx->cus * 8);
    hipLaunchKernelGGL(k_hash_probe, dim3(std::max(grid, 1u)), dim3(BLOCK), 0, ctx->stream, a);
    return hipGetLastError();
}

// ------------------------------------------------------------------------ zone maps
// check_mmmeta (core/metadata.cuh:539-547) with MinMaxMeta::check_pred (:203-211).
__global__ void k_minmax_check(const int64_t *mins, const int64_t *maxs, uint64_t n, int64_t lo,
                               int64_t hi, uint16_t *flags) {
    for (uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n;
         c += (uint64_t)gridDim.x * blockDim.x)
        flags[c] = (lo <= maxs[c] && hi >= mins[c]) ? 1 : 0;
}
// combine_and (core/metadata.cuh:632-636).
__global__ void k_combine_and(uint16_t *a, const uint16_t *b, uint64_t n) {
    for (uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n;
         c += (uint64_t)gridDim.x * blockDim.x)
        a[c] = (a[c] && b[c]) ? 1 : 0;
}

hipError_t launch_minmax_check(golap_ctx *ctx, const golap_column *col, int64_t lo, int64_t hi,
                               uint16_t *flags) {
    if (!col->n_chunks) return hipSuccess;
    const uint32_t grid = (uint32_t)std::min<uint64_t>((col->n_chunks + BLOCK - 1) / BLOCK, 1024);
    hipLaunchKernelGGL(k_minmax_check, dim3(grid), dim3(BLOCK), 0, ctx->stream, col->d_chunk_min,
                       col->d_chunk_max, col->n_chunks, lo, hi, flags);
    return hipGetLastError();
}
hipError_t launch_combine_and(golap_ctx *ctx, uint16_t *a, const uint16_t *b, uint64_t n) {
    if (!n) return hipSuccess;
    const uint32_t grid = (uint32_t)std::min<uint64_t>((n + BLOCK - 1) / BLOCK, 1024);
    hipLaunchKernelGGL(k_combine_and, dim3(grid), dim3(BLOCK), 0, ctx->stream, a, b, n);
    return hipGetLastError();
}


// Per-tile catalog entries of the four columns. Every load is issued unconditionally and the
// decision is taken on the loaded values, so the twelve loads share one memory round trip.
struct Q1Tile {
    const uint8_t *dp, *sp, *qp, *pp;
    uint32_t rows;
    uint32_t state;  // 0 skip (pruned or not in place), 1 run, 2 corrupt catalog
};

__device__ __forceinline__ Q1Tile q1_tile(const Q1Args &a, uint64_t p) {
    const uint32_t P = a.page_rows;
    const uint8_t cd = a.date.page_codec[p], cs = a.disc.page_codec[p], cq = a.qty.page_codec[p],
                  cp = a.price.page_codec[p];
    const uint32_t zd = a.date.page_csize[p], zs = a.disc.page_csize[p], zq = a.qty.page_csize[p],
                   zp = a.price.page_csize[p];
    const uint64_t od = a.date.page_off[p], os = a.disc.page_off[p], oq = a.qty.page_off[p],
                   op = a.price.page_off[p];
    const uint32_t keep = a.chunk_mask ? (uint32_t)a.chunk_mask[p / a.pages_per_chunk] : 1u;
    Q1Tile t;
    t.rows = (uint32_t)min<uint64_t>(P, a.n_rows - p * P);
    const uint32_t hd = direct_header(cd, t.rows * 4), hs = direct_header(cs, t.rows),
                   hq = direct_header(cq, t.rows), hp = direct_header(cp, t.rows * 4);
    t.dp = a.date.comp + od + hd;
    t.sp = a.disc.comp + os + hs;
    t.qp = a.qty.comp + oq + hq;
    t.pp = a.price.comp + op + hp;
    const uint32_t direct = (uint32_t)page_direct(cd) & (uint32_t)page_direct(cs) &
                            (uint32_t)page_direct(cq) & (uint32_t)page_direct(cp);
    const uint32_t sizes_ok = (uint32_t)(zd == hd + t.rows * 4) & (uint32_t)(zs == hs + t.rows) &
                              (uint32_t)(zq == hq + t.rows) & (uint32_t)(zp == hp + t.rows * 4);
    const uint32_t aligned = ((reinterpret_cast<uintptr_t>(t.dp) | reinterpret_cast<uintptr_t>(t.sp) |
                               reinterpret_cast<uintptr_t>(t.qp) | reinterpret_cast<uintptr_t>(t.pp)) & 15) == 0;
    t.state = (keep == 0 || !direct) ? 0u : ((sizes_ok & aligned) ? 1u : 2u);
    return t;
}

// Tiles whose four pages are readable in place. G = page_rows / (4 * BLOCK) row groups of
// four per thread: a full tile issues all 4*G loads of a thread before any row is evaluated,
// and the catalog entries of the workgroup's next tile are loaded before this tile's data.
template <int PROBE, int G>
__global__ __launch_bounds__(BLOCK) void k_q1_direct(Q1Args a) {
    extern __shared__ __align__(16) uint8_t lds[];
    const void *tab = q1_table<PROBE>(a, lds);
    int64_t sum = 0;
    uint32_t cnt = 0;
    const XcdRange xr = xcd_range(a.page_begin, a.page_end);
    uint64_t p = xr.first;
    Q1Tile cur{};
    if (p < xr.end) cur = q1_tile(a, p);
    while (p < xr.end) {
        const uint64_t pn = p + xr.stride;
        Q1Tile nxt{};
        if (pn < xr.end) nxt = q1_tile(a, pn);
        if (cur.state == 2) {
            if (threadIdx.x == 0) raise_err(a.err, DEVERR_CORRUPT);
        } else if (cur.state == 1) {
            if (G > 0 && cur.rows == (uint32_t)(G * 4 * BLOCK)) {
                int4 dv[G > 0 ? G : 1], pv[G > 0 ? G : 1];
                uint32_t sv[G > 0 ? G : 1], qv[G > 0 ? G : 1];
#pragma unroll
                for (int k = 0; k < G; ++k) {
                    const size_t g = threadIdx.x + (size_t)k * BLOCK;
                    dv[k] = *reinterpret_cast<const int4 *>(cur.dp + 16 * g);
                    pv[k] = *reinterpret_cast<const int4 *>(cur.pp + 16 * g);
                    sv[k] = *reinterpret_cast<const uint32_t *>(cur.sp + 4 * g);
                    qv[k] = *reinterpret_cast<const uint32_t *>(cur.qp + 4 * g);
                }
#pragma unroll
                for (int k = 0; k < G; ++k) q1_group<PROBE>(a, tab, dv[k], pv[k], sv[k], qv[k], sum, cnt);
            } else {
                q1_rows_generic<PROBE>(a, tab, cur.dp, cur.sp, cur.qp, cur.pp, cur.rows, sum, cnt);
            }
        }
        cur = nxt;
        p = pn;
    }
    q1_flush(a, sum, cnt);
}

// Tiles with at least one coded page: wave w decodes column w's page into an LDS window, then
// the workgroup evaluates the tile from LDS.
template <int PROBE>
__global__ __launch_bounds__(BLOCK) void k_q1_general(Q1Args a) {
    extern __shared__ __align__(16) uint8_t lds[];
    __shared__ uint32_t tile_bad;
    const void *tab = q1_table<PROBE>(a, lds);
    const uint32_t tab_bytes = PROBE == PROBE_HASH_GLOBAL ? 0u
                               : PROBE == PROBE_BITMAP ? ((a.krange + 31) / 32 * 4 + 15) & ~15u
                                                       : (8u << a.log2cap);
    const uint32_t P = a.page_rows;
    uint8_t *w_date = lds + tab_bytes;
    uint8_t *w_price = w_date + (size_t)P * 4;
    uint8_t *w_disc = w_price + (size_t)P * 4;
    uint8_t *w_qty = w_disc + P;
    const uint32_t wave = threadIdx.x / WAVE;
    int64_t sum = 0;
    uint32_t cnt = 0;
    for (uint64_t p = a.page_begin + blockIdx.x; p < a.page_end; p += gridDim.x) {
        if (a.chunk_mask && a.chunk_mask[p / a.pages_per_chunk] == 0) continue;
        const bool direct = page_direct(a.date.page_codec[p]) && page_direct(a.disc.page_codec[p]) &&
                            page_direct(a.qty.page_codec[p]) && page_direct(a.price.page_codec[p]);
        if (direct) continue;  // k_q1_direct's tile
        const uint32_t rows = (uint32_t)min<uint64_t>(P, a.n_rows - p * P);
        if (threadIdx.x == 0) tile_bad = 0;
        __syncthreads();
        const ColView *cv = wave == 0 ? &a.date : wave == 1 ? &a.disc : wave == 2 ? &a.qty : &a.price;
        uint8_t *win = wave == 0 ? w_date : wave == 1 ? w_disc : wave == 2 ? w_qty : w_price;
        const uint32_t nb = rows * cv->elem;
        // in-place pages decode as RAW/LZ4 too, so this single call covers every codec
        const int64_t got = page_decode_wave(cv->page_codec[p], cv->comp + cv->page_off[p],
                                             cv->page_csize[p], win, nb, cv->elem);
        if (got != (int64_t)nb && __lane_id() == 0) {
            tile_bad = 1;
            raise_err(a.err, DEVERR_CORRUPT);
        }
        __syncthreads();
        if (!tile_bad) q1_rows_generic<PROBE>(a, tab, w_date, w_disc, w_qty, w_price, rows, sum, cnt);
        __syncthreads();  // windows and tile_bad are reused by the next tile
    }
    q1_flush(a, sum, cnt);
}

// ------------------------------------------------------------------ one wave per tile
// Tiles with coded pages (LZ4I, LZ4, DBP, RLE, FOR) at page sizes of 1024..8192 rows. The
// workgroup is one wave with one LDS decode window and token array: the tile's columns are
// decoded into the window one after another (LZ4I pages by the parallel parse), and each lane
// keeps the state of its rows in registers between them: a pass bit per row and the discount
// bytes. Lane l owns rows 256 i + 4 l + {0..3} (i < G4), so every column access is one
// 16-byte (int32) or 4-byte (uint8) load per four rows, consecutive across the lanes. The date
// probe goes first when it is the LDS bitmap (cheap, the most selective test); a hash probe goes
// after the quantity and discount tests, for the rows that are still candidates. Once no row of
// the tile is a candidate, the remaining columns are not decoded.

// Raw page bytes of one column in registers, loaded ahead of their use: slot j of lane l holds
// bytes [16 (l + 64 j), +16) of the page (or of an in-place page's rows). n16 = 0 when the page
// does not fit the slots (then it is read from global memory where it is used).
template <int NU>
struct RawPage {
    u32x4 v[NU];
    uint32_t n16;
    // every slot is loaded (indexes past the page read its last 16 bytes again), so the loads
    // carry no condition and all stay in flight
    __device__ __forceinline__ void load(const uint8_t *src, uint32_t nbytes) {
        const uint32_t n = (nbytes + 15) / 16;
        n16 = n <= (uint32_t)(NU * WAVE) ? n : 0u;
        const u32x4 *s4 = reinterpret_cast<const u32x4 *>(src);
        const uint32_t last = n16 ? n16 - 1 : 0u;
#pragma unroll
        for (int j = 0; j < NU; ++j) v[j] = s4[min(__lane_id() + (uint32_t)j * WAVE, last)];
    }
    __device__ __forceinline__ void store(uint8_t *dst) const {
        u32x4 *d4 = reinterpret_cast<u32x4 *>(dst);
#pragma unroll
        for (int j = 0; j < NU; ++j) {
            const uint32_t i = __lane_id() + (uint32_t)j * WAVE;
            if (i < n16) d4[i] = v[j];
        }
    }
};

// The bytes to load ahead for a page: the rows of an in-place page, an LZ4I page whole; other
// codecs are decoded from global memory (nothing loaded).
__device__ __forceinline__ void q1_ahead(const Q1Page &pg, uint32_t nb, const uint8_t *&from, uint32_t &bytes) {
    if (page_direct(pg.code)) {
        from = pg.src + direct_header(pg.code, nb);
        bytes = nb;
    } else {
        from = pg.src;
        bytes = (pg.code & PAGE_CODEC_MASK) == golap_fmt::LZ4I ? pg.csize : 0u;
    }
}

// The rows of one column of a tile into the window, in two steps so that other loads can be
// issued between them. Stage: an in-place page's rows and an LZ4I page are copied from their
// registers into the window (an LZ4I page at its tail). Decode: the staged LZ4I page by the
// parallel parse; any other coded page from global memory. False for a corrupt page.
enum : int { Q1_IN_PLACE = 0, Q1_STAGED = 1, Q1_GLOBAL = 2, Q1_BAD = 3 };
template <int NU>
__device__ __forceinline__ int q1_stage(const Q1Args &a, const Q1Page &pg, uint32_t nb, const RawPage<NU> &raw,
                                        uint8_t *win) {
    if (page_direct(pg.code)) {
        const uint32_t hd = direct_header(pg.code, nb);
        if (pg.csize != hd + nb || ((reinterpret_cast<uintptr_t>(pg.src + hd) & 15) != 0) || raw.n16 == 0) return Q1_BAD;
        raw.store(win);
        __builtin_amdgcn_wave_barrier();
        return Q1_IN_PLACE;
    }
    if ((pg.code & PAGE_CODEC_MASK) == golap_fmt::LZ4I && raw.n16 && raw.n16 * 16 <= a.win_cap &&
        (reinterpret_cast<uintptr_t>(pg.src) & 15) == 0) {
        raw.store(win + lz4i_stage_offset(a.win_cap, pg.csize));
        __builtin_amdgcn_wave_barrier();
        return Q1_STAGED;
    }
    return Q1_GLOBAL;
}
__device__ __forceinline__ bool q1_decode(const Q1Args &a, int how, const Q1Page &pg, uint32_t nb, uint32_t elem,
                                          uint8_t *win, uint16_t *tok) {
    if (how == Q1_IN_PLACE) return true;
    if (how == Q1_BAD) return false;
    if (how == Q1_STAGED) return lz4i_decode_staged(pg.src, pg.csize, win, a.win_cap, nb, tok, a.tok_cap) == (int64_t)nb;
    return page_decode_wave(pg.code, pg.src, pg.csize, win, nb, elem, a.win_cap, tok, a.tok_cap) == (int64_t)nb;
}

// LDS allows at most two of these waves per SIMD, so each may use up to 256 VGPRs (the pages
// loaded ahead live in registers).
template <int PROBE, int G4>
__global__ __launch_bounds__(WAVE) __attribute__((amdgpu_waves_per_eu(1, 2))) void k_q1_wave(Q1Args a) {
    extern __shared__ __align__(16) uint8_t lds[];
    constexpr bool HASH = PROBE != PROBE_BITMAP;
    const uint32_t *tab;
    if (PROBE == PROBE_HASH_GLOBAL) {
        tab = reinterpret_cast<const uint32_t *>(a.slots);
    } else {
        uint32_t *t = reinterpret_cast<uint32_t *>(lds);
        const uint32_t *src = PROBE == PROBE_BITMAP ? a.bitmap : reinterpret_cast<const uint32_t *>(a.slots);
        const uint32_t words = PROBE == PROBE_BITMAP ? (a.krange + 31) / 32 : (2u << a.log2cap);
        for (uint32_t i = threadIdx.x; i < words; i += WAVE) t[i] = src[i];
        __syncthreads();
        tab = t;
    }
    uint8_t *win = lds + a.tab_bytes;
    uint16_t *tok = reinterpret_cast<uint16_t *>(win + a.win_cap);
    const uint32_t P = a.page_rows;
    int64_t sum = 0;
    uint32_t cnt = 0;
    const XcdRange xr = xcd_range(a.page_begin, a.page_end);
    uint64_t p = xr.first;
    // the next tile's catalog is loaded while the current tile runs
    Q1Page nd{}, nq{}, ns{}, np{};
    if (p < xr.end) {
        nd = q1_page(a.date, p);
        nq = q1_page(a.qty, p);
        ns = q1_page(a.disc, p);
        np = q1_page(a.price, p);
    }
    for (; p < xr.end; p += xr.stride) {
        const Q1Page pd = nd, pq = nq, ps = ns, pp = np;
        const uint64_t pn = p + xr.stride;
        if (pn < xr.end) {
            nd = q1_page(a.date, pn);
            nq = q1_page(a.qty, pn);
            ns = q1_page(a.disc, pn);
            np = q1_page(a.price, pn);
        }
        if (a.chunk_mask && a.chunk_mask[p / a.pages_per_chunk] == 0) continue;
        if (page_direct(pd.code) && page_direct(ps.code) && page_direct(pq.code) && page_direct(pp.code))
            continue;  // k_q1_direct's tile
        const uint32_t rows = (uint32_t)min<uint64_t>(P, a.n_rows - p * P);
        const bool full = rows == (uint32_t)(256 * G4);
        Q1Lanes<G4> st;
        st.reset();
        if (!full)
            for (int b = 0; b < 4 * G4; ++b)
                if (256 * (b / 4) + 4 * __lane_id() + (b % 4) >= rows) st.clear(b);
        // every column's bytes are requested before the first is decoded; the price page
        // reuses the date page's registers once those are in LDS
        RawPage<G4 + 1> big;
        RawPage<G4 / 4 + 1> rq, rs;
        {
            const uint8_t *f;
            uint32_t n;
            q1_ahead(pd, rows * 4, f, n);
            big.load(f, n);
            q1_ahead(pq, rows, f, n);
            rq.load(f, n);
            q1_ahead(ps, rows, f, n);
            rs.load(f, n);
        }
        bool bad = false;
        // the date column: staged, the price page requested into the same registers (in flight
        // while the date page decodes), decoded, tested
#define Q1_DATE_STEP()                                                                           \
    do {                                                                                         \
        const int how = q1_stage(a, pd, rows * 4, big, win);                                     \
        {                                                                                        \
            const uint8_t *f;                                                                    \
            uint32_t n;                                                                          \
            q1_ahead(pp, rows * 4, f, n);                                                        \
            big.load(f, n);                                                                      \
        }                                                                                        \
        bad = !q1_decode(a, how, pd, rows * 4, 4, win, tok);                                     \
        if (!bad) {                                                                              \
            if (full) q1_date_pass<WAVE, G4, true>(a, tab, win, rows, st, HASH);                       \
            else q1_date_pass<WAVE, G4, false>(a, tab, win, rows, st, HASH);                           \
        }                                                                                        \
    } while (0)
        if (!HASH) Q1_DATE_STEP();
        if (!bad && __any(st.any())) {
            bad = !q1_decode(a, q1_stage(a, pq, rows, rq, win), pq, rows, 1, win, tok);
            if (!bad) {
                if (full) q1_byte_pass<WAVE, G4, true, false>(win, rows, a.qty_lo, a.qty_span, st);
                else q1_byte_pass<WAVE, G4, false, false>(win, rows, a.qty_lo, a.qty_span, st);
            }
        }
        // the discount bytes are kept in registers, so the window is free for the next column
        if (!bad && __any(st.any())) {
            __builtin_amdgcn_wave_barrier();
            bad = !q1_decode(a, q1_stage(a, ps, rows, rs, win), ps, rows, 1, win, tok);
            if (!bad) {
                if (full) q1_byte_pass<WAVE, G4, true, true>(win, rows, a.disc_lo, a.disc_span, st);
                else q1_byte_pass<WAVE, G4, false, true>(win, rows, a.disc_lo, a.disc_span, st);
            }
        }
        if (HASH && !bad && __any(st.any())) {
            __builtin_amdgcn_wave_barrier();
            Q1_DATE_STEP();
        }
#undef Q1_DATE_STEP
        if (!bad && __any(st.any())) {
            __builtin_amdgcn_wave_barrier();
            bad = !q1_decode(a, q1_stage(a, pp, rows * 4, big, win), pp, rows * 4, 4, win, tok);
            if (!bad) {
                if (full) q1_price_pass<WAVE, G4, true>(a, win, rows, st, sum, cnt);
                else q1_price_pass<WAVE, G4, false>(a, win, rows, st, sum, cnt);
            }
        }
        if (bad && __lane_id() == 0) raise_err(a.err, DEVERR_CORRUPT);
        __builtin_amdgcn_wave_barrier();  // the window is rewritten by the next tile
    }
    q1_flush(a, sum, cnt);
}

template <typename K>
static hipError_t launch_persistent(golap_ctx *ctx, K kernel, const Q1Args &a, size_t lds, int block = BLOCK) {
    hipError_t e = prepare_lds(ctx, kernel, lds);
    if (e != hipSuccess) return e;
    int occ = 0;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kernel, block, lds);
    if (e != hipSuccess) return e;
    occ = std::max(occ, 1);
    const uint64_t tiles = a.page_end - a.page_begin;
    uint64_t grid = std::max<uint64_t>(1, std::min<uint64_t>(tiles, (uint64_t)ctx->cus * occ));
    if (grid >= XCDS) grid -= grid % XCDS;  // whole rounds over the XCDs (see xcd_range)
    hipLaunchKernelGGL(kernel, dim3((uint32_t)grid), dim3(block), lds, ctx->stream, a);
    return hipGetLastError();
}

template <int PROBE>
static hipError_t launch_q1_mode(golap_ctx *ctx, Q1Args a, size_t tab, bool direct) {
    if (!direct) {
        // one wave per tile where the page size allows, else the four-wave general kernel
        a.tab_bytes = (uint32_t)((tab + 15) & ~(size_t)15);
        const size_t lds = (size_t)a.tab_bytes + a.win_cap + lz4i::scratch_bytes(a.win_cap, a.tok_cap);
        if (a.win_cap && lds <= (size_t)ctx->lds_max) {
            switch (a.page_rows) {
            case 1024: return launch_persistent(ctx, k_q1_wave<PROBE, 4>, a, lds, WAVE);
            case 2048: return launch_persistent(ctx, k_q1_wave<PROBE, 8>, a, lds, WAVE);
            case 4096: return launch_persistent(ctx, k_q1_wave<PROBE, 16>, a, lds, WAVE);
            case 8192: return launch_persistent(ctx, k_q1_wave<PROBE, 32>, a, lds, WAVE);
            default: break;
            }
        }
        return launch_persistent(ctx, k_q1_general<PROBE>, a, tab + (size_t)a.page_rows * 10);
    }
    switch (a.page_rows) {
    case 1024: return launch_persistent(ctx, k_q1_direct<PROBE, 1>, a, tab);
    case 2048: return launch_persistent(ctx, k_q1_direct<PROBE, 2>, a, tab);
    case 4096: return launch_persistent(ctx, k_q1_direct<PROBE, 4>, a, tab);
    case 8192: return launch_persistent(ctx, k_q1_direct<PROBE, 8>, a, tab);
    default: return launch_persistent(ctx, k_q1_direct<PROBE, 0>, a, tab);
    }
}

hipError_t launch_q1(golap_ctx *ctx, const Q1Launch &l, bool direct) {
    Q1Args a{};
    a.date = l.date;
    a.disc = l.disc;
    a.qty = l.qty;
    a.price = l.price;
    a.page_rows = l.page_rows;
    a.n_rows = l.n_rows;
    a.page_begin = l.page_begin;
    a.page_end = l.page_end;
    a.chunk_mask = l.chunk_mask;
    a.pages_per_chunk = l.pages_per_chunk;
    a.disc_lo = l.disc_lo;
    a.disc_span = l.disc_span;
    a.qty_lo = l.qty_lo;
    a.qty_span = l.qty_span;
    a.price_lo = l.price_lo;
    a.price_span = l.price_span;
    a.slots = l.slots;
    a.log2cap = l.log2cap;
    a.bitmap = l.bitmap;
    a.kmin = l.kmin;
    a.krange = l.krange;
    a.sum = l.sum;
    a.count = l.count;
    a.err = l.err;
    a.win_cap = l.win_cap;
    a.tok_cap = l.tok_cap;
    switch (l.probe) {
    case PROBE_BITMAP:
        return launch_q1_mode<PROBE_BITMAP>(ctx, a, (((size_t)l.krange + 31) / 32 * 4 + 15) & ~(size_t)15, direct);
    case PROBE_HASH_LDS:
        return launch_q1_mode<PROBE_HASH_LDS>(ctx, a, 8ull << l.log2cap, direct);
    default:
        return launch_q1_mode<PROBE_HASH_GLOBAL>(ctx, a, 0, direct);
    }
}

}  // namespace golap_impl
