// gfx950 kernels of the compressed-chunk -> filter -> join-probe -> int64-sum path.
//
// Launch shape everywhere: 256-thread workgroups (4 wave64s). Scan kernels are persistent
// (grid = CUs x resident workgroups) and walk pages with a grid stride, so per-workgroup set-up
// (copying a dimension table into LDS) is paid once per workgroup, not once per page.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "common.hpp"
#include "device_codecs.hpp"

namespace golap_impl {
using namespace golap_dev;
using golap_fmt::PAGE_CODEC_MASK;
using golap_fmt::PAGE_LITERAL_ONLY;

constexpr int BLOCK = 256;
constexpr int WAVES_PER_BLOCK = BLOCK / WAVE;

// Dynamic LDS above the default 64 KiB must be requested per kernel; sizes the device cannot
// give are refused before launch.
template <typename K>
static hipError_t prepare_lds(golap_ctx *ctx, K kernel, size_t lds) {
    if (lds > (size_t)ctx->lds_max) return hipErrorInvalidConfiguration;
    if (lds > 65536)
        return hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    return hipSuccess;
}

__device__ __forceinline__ void raise_err(uint32_t *err, uint32_t bit) {
    if (err) atomicOr(err, bit);
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int d = WAVE / 2; d > 0; d >>= 1) v += __shfl_xor(v, d, WAVE);
    return v;
}

__device__ __forceinline__ int64_t load_typed(const void *col, uint32_t type, uint64_t row) {
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

// True when a page can be read in place: a RAW page, or an LZ4 page that is one literal run.
__device__ __host__ __forceinline__ bool page_direct(uint8_t codec) {
    return (codec & PAGE_CODEC_MASK) == golap_fmt::RAW || (codec & PAGE_LITERAL_ONLY);
}
__device__ __forceinline__ uint32_t direct_header(uint8_t codec, uint32_t nbytes) {
    return (codec & PAGE_LITERAL_ONLY) ? golap_fmt::lz4_literal_header_bytes(nbytes) : 0u;
}

// ------------------------------------------------------------------------ page decompression
struct DecodeArgs {
    ColView c;
    uint64_t page_begin, page_end;
    uint32_t page_rows;
    uint64_t n_rows;
    uint8_t *out;
    uint64_t out_row0;
    uint32_t window;  // LDS bytes per wave
    uint32_t *err;
};

// One wave per page. In-place pages are copied global->global with 16-byte accesses; coded
// pages are decoded into the wave's LDS window, then written out with 16-byte stores.
__global__ __launch_bounds__(BLOCK) void k_decode_pages(DecodeArgs a) {
    extern __shared__ __align__(16) uint8_t lds[];
    const uint32_t wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    uint8_t *win = lds + (size_t)wave * a.window;
    const uint64_t n = a.page_end - a.page_begin;
    for (uint64_t i = (uint64_t)blockIdx.x * WAVES_PER_BLOCK + wave; i < n;
         i += (uint64_t)gridDim.x * WAVES_PER_BLOCK) {
        const uint64_t p = a.page_begin + i;
        const uint64_t r0 = p * a.page_rows;
        const uint32_t rows = (uint32_t)min<uint64_t>(a.page_rows, a.n_rows - r0);
        const uint32_t nbytes = rows * a.c.elem;
        const uint8_t codec = a.c.page_codec[p];
        const uint8_t *src = a.c.comp + a.c.page_off[p];
        const uint32_t csize = a.c.page_csize[p];
        uint8_t *dst = a.out + (r0 - a.out_row0) * a.c.elem;
        const uint8_t *from;
        if (page_direct(codec)) {
            const uint32_t hdr = direct_header(codec, nbytes);
            if ((uint64_t)hdr + nbytes != csize) {
                if (lane == 0) raise_err(a.err, DEVERR_CORRUPT);
                continue;
            }
            from = src + hdr;
        } else {
            if (nbytes > a.window) {
                if (lane == 0) raise_err(a.err, DEVERR_CORRUPT);
                continue;
            }
            const int64_t got = page_decode_wave(codec, src, csize, win, nbytes, a.c.elem);
            if (got != (int64_t)nbytes) {
                if (lane == 0) raise_err(a.err, DEVERR_CORRUPT);
                continue;
            }
            from = win;
        }
        const bool aligned = ((reinterpret_cast<uintptr_t>(from) | reinterpret_cast<uintptr_t>(dst)) & 15) == 0;
        uint32_t done = 0;
        if (aligned) {
            const uint32_t nv = nbytes / 16;
            const uint4 *s4 = reinterpret_cast<const uint4 *>(from);
            uint4 *d4 = reinterpret_cast<uint4 *>(dst);
            for (uint32_t k = lane; k < nv; k += WAVE) d4[k] = s4[k];
            done = nv * 16;
        }
        for (uint32_t k = done + lane; k < nbytes; k += WAVE) dst[k] = from[k];
    }
}

hipError_t launch_decode_pages(golap_ctx *ctx, const golap_column *col, uint64_t page_begin,
                               uint64_t page_end, uint8_t *dev_out, uint64_t out_row0) {
    if (page_end <= page_begin) return hipSuccess;
    DecodeArgs a;
    a.c = view(col);
    a.page_begin = page_begin;
    a.page_end = page_end;
    a.page_rows = col->page_rows;
    a.n_rows = col->n_rows;
    a.out = dev_out;
    a.out_row0 = out_row0;
    a.window = (col->page_rows * col->elem + 15) & ~15u;
    a.err = reinterpret_cast<uint32_t *>(ctx->d_scratch);
    const size_t lds = (size_t)a.window * WAVES_PER_BLOCK;
    hipError_t e = prepare_lds(ctx, k_decode_pages, lds);
    if (e != hipSuccess) return e;
    int occ = 0;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_decode_pages, BLOCK, lds);
    if (e != hipSuccess) return e;
    occ = std::max(occ, 1);
    const uint64_t need = (page_end - page_begin + WAVES_PER_BLOCK - 1) / WAVES_PER_BLOCK;
    const uint32_t grid = (uint32_t)std::min<uint64_t>(need, (uint64_t)ctx->cus * occ);
    hipLaunchKernelGGL(k_decode_pages, dim3(grid), dim3(BLOCK), lds, ctx->stream, a);
    return hipGetLastError();
}

// ------------------------------------------------------------------------ hash table
__device__ __host__ __forceinline__ uint32_t ht_hash(int32_t key, uint32_t log2cap) {
    return ((uint32_t)key * 0x9E3779B1u) >> (32 - log2cap);
}
__device__ __forceinline__ uint32_t ht_stored_key(int32_t key) { return (uint32_t)key ^ 0x80000000u; }

// Returns true and the payload if key is in the table.
__device__ __forceinline__ bool ht_probe(const uint64_t *slots, uint32_t log2cap, int32_t key,
                                         uint32_t &payload) {
    const uint32_t mask = (1u << log2cap) - 1, sk = ht_stored_key(key);
    uint32_t h = ht_hash(key, log2cap);
    for (uint32_t i = 0; i <= mask; ++i) {
        const uint64_t s = slots[h];
        if (s == 0) return false;
        if ((uint32_t)s == sk) {
            payload = (uint32_t)(s >> 32);
            return true;
        }
        h = (h + 1) & mask;
    }
    return false;
}

__device__ __forceinline__ bool ht_insert(uint64_t *slots, uint32_t log2cap, int32_t key,
                                          uint32_t payload) {
    const uint32_t mask = (1u << log2cap) - 1;
    const uint64_t slot = ((uint64_t)payload << 32) | ht_stored_key(key);
    uint32_t h = ht_hash(key, log2cap);
    for (uint32_t i = 0; i <= mask; ++i) {
        if (atomicCAS(reinterpret_cast<unsigned long long *>(&slots[h]), 0ull,
                      (unsigned long long)slot) == 0ull)
            return true;
        h = (h + 1) & mask;
    }
    return false;
}

__global__ void k_fill_zero(uint64_t *p, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = 0;
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
    uint32_t *err;
};

// hash_map_build_pred_other (core/join.cuh:107-135): one thread per build row; the row is
// inserted iff every predicate passes. Insert is one 64-bit CAS on an empty slot (the key is
// the occupancy marker), so no thread ever waits on another.
__global__ __launch_bounds__(BLOCK) void k_hash_build(BuildArgs a) {
    uint32_t inserted = 0;
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
        if (ht_insert(a.slots, a.log2cap, (int32_t)key, pay)) ++inserted;
        else raise_err(a.err, DEVERR_CAPACITY);
    }
    const uint32_t tot = wave_sum(inserted);
    if (__lane_id() == 0 && tot) atomicAdd(reinterpret_cast<unsigned long long *>(a.filled), (unsigned long long)tot);
}

hipError_t launch_hash_build(golap_ctx *ctx, golap_hashmap *m, uint64_t n_rows,
                             const PredView *preds, uint32_t n_preds, const int32_t *payload,
                             uint64_t *d_filled) {
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

// ------------------------------------------------------------------------ fused Q1 plan
// pipeline_q1_lineorder (usecases/ssb/queries/01.cuh:26-54) fused with the decompression the
// reference runs beforehand per chunk. One workgroup handles one page-row tile of all four
// columns: DIRECT tiles (all four pages readable in place) are evaluated straight from the
// compressed buffers with 16-byte loads; other tiles are first decoded into LDS windows, one
// wave per column. Each thread evaluates rows in groups of four; a row that passes the
// discount/quantity/price ranges probes the date table (LDS-pinned when it fits); matches add
// price*discount to a per-thread int64, reduced per wave into one global atomic at the end.
struct Q1Args {
    ColView date, disc, qty, price;
    uint32_t page_rows;
    uint64_t n_rows;
    uint64_t page_begin, page_end;
    const uint16_t *chunk_mask;  // indexed by page / pages_per_chunk, may be null
    uint32_t pages_per_chunk;
    int32_t disc_lo, disc_hi, qty_lo, qty_hi, price_lo, price_hi;
    const uint64_t *slots;
    uint32_t log2cap;
    int64_t *sum;
    uint64_t *count;
    uint32_t *err;
};

template <bool LDS_TABLE>
__device__ __forceinline__ void q1_row(const Q1Args &a, const uint64_t *tab, int32_t dk, uint32_t d,
                                       uint32_t q, int32_t pr, int64_t &sum, uint32_t &cnt) {
    if ((int32_t)d >= a.disc_lo && (int32_t)d <= a.disc_hi && (int32_t)q >= a.qty_lo &&
        (int32_t)q <= a.qty_hi && pr >= a.price_lo && pr <= a.price_hi) {
        uint32_t pay;
        if (ht_probe(tab, a.log2cap, dk, pay)) {
            sum += (int64_t)pr * (int64_t)d;
            cnt += 1;
        }
    }
}

template <bool DIRECT, bool LDS_TABLE>
__global__ __launch_bounds__(BLOCK) void k_q1(Q1Args a) {
    extern __shared__ __align__(16) uint8_t lds[];
    __shared__ uint32_t tile_bad;
    const uint64_t *tab = a.slots;
    uint32_t tab_bytes = 0;
    if (LDS_TABLE) {
        const uint32_t n = 1u << a.log2cap;
        uint64_t *t = reinterpret_cast<uint64_t *>(lds);
        for (uint32_t i = threadIdx.x; i < n; i += BLOCK) t[i] = a.slots[i];
        tab = t;
        tab_bytes = n * 8;
        __syncthreads();
    }
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
        const uint8_t cd = a.date.page_codec[p], cs = a.disc.page_codec[p],
                      cq = a.qty.page_codec[p], cp = a.price.page_codec[p];
        const bool direct = page_direct(cd) && page_direct(cs) && page_direct(cq) && page_direct(cp);
        if (direct != DIRECT) continue;
        const uint64_t r0 = p * P;
        const uint32_t rows = (uint32_t)min<uint64_t>(P, a.n_rows - r0);
        const uint8_t *dp, *sp, *qp, *pp;
        bool bad = false;
        if (DIRECT) {
            const uint32_t hd = direct_header(cd, rows * 4), hs = direct_header(cs, rows),
                           hq = direct_header(cq, rows), hp = direct_header(cp, rows * 4);
            dp = a.date.comp + a.date.page_off[p] + hd;
            sp = a.disc.comp + a.disc.page_off[p] + hs;
            qp = a.qty.comp + a.qty.page_off[p] + hq;
            pp = a.price.comp + a.price.page_off[p] + hp;
            // the catalog must say exactly "header + values", at 16-byte aligned payloads
            if (a.date.page_csize[p] != hd + rows * 4 || a.disc.page_csize[p] != hs + rows ||
                a.qty.page_csize[p] != hq + rows || a.price.page_csize[p] != hp + rows * 4 ||
                ((reinterpret_cast<uintptr_t>(dp) | reinterpret_cast<uintptr_t>(sp) |
                  reinterpret_cast<uintptr_t>(qp) | reinterpret_cast<uintptr_t>(pp)) & 15)) {
                if (threadIdx.x == 0) raise_err(a.err, DEVERR_CORRUPT);
                continue;  // uniform: every thread evaluates the same test
            }
        } else {
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
            bad = tile_bad != 0;
            dp = w_date;
            sp = w_disc;
            qp = w_qty;
            pp = w_price;
        }
        for (uint32_t g = threadIdx.x; !bad && g * 4 < rows; g += BLOCK) {
            const uint32_t rr = g * 4;
            if (rr + 4 <= rows) {
                const int4 dv = *reinterpret_cast<const int4 *>(dp + 16 * (size_t)g);
                const int4 pv = *reinterpret_cast<const int4 *>(pp + 16 * (size_t)g);
                const uint32_t sv = *reinterpret_cast<const uint32_t *>(sp + 4 * (size_t)g);
                const uint32_t qv = *reinterpret_cast<const uint32_t *>(qp + 4 * (size_t)g);
                q1_row<LDS_TABLE>(a, tab, dv.x, sv & 0xff, qv & 0xff, pv.x, sum, cnt);
                q1_row<LDS_TABLE>(a, tab, dv.y, (sv >> 8) & 0xff, (qv >> 8) & 0xff, pv.y, sum, cnt);
                q1_row<LDS_TABLE>(a, tab, dv.z, (sv >> 16) & 0xff, (qv >> 16) & 0xff, pv.z, sum, cnt);
                q1_row<LDS_TABLE>(a, tab, dv.w, sv >> 24, qv >> 24, pv.w, sum, cnt);
            } else {
                for (uint32_t r = rr; r < rows; ++r) {
                    int32_t dk, pr;
                    __builtin_memcpy(&dk, dp + 4 * (size_t)r, 4);
                    __builtin_memcpy(&pr, pp + 4 * (size_t)r, 4);
                    q1_row<LDS_TABLE>(a, tab, dk, sp[r], qp[r], pr, sum, cnt);
                }
            }
        }
        if (!DIRECT) __syncthreads();  // windows and tile_bad are reused by the next tile
    }
    const int64_t ws = wave_sum(sum);
    const uint32_t wc = wave_sum(cnt);
    if (__lane_id() == 0 && wc) {
        atomicAdd(reinterpret_cast<unsigned long long *>(a.sum), (unsigned long long)ws);
        atomicAdd(reinterpret_cast<unsigned long long *>(a.count), (unsigned long long)wc);
    }
}

template <bool DIRECT, bool LDS_TABLE>
static hipError_t launch_q1_t(golap_ctx *ctx, const Q1Args &a, size_t lds) {
    hipError_t e = prepare_lds(ctx, k_q1<DIRECT, LDS_TABLE>, lds);
    if (e != hipSuccess) return e;
    int occ = 0;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_q1<DIRECT, LDS_TABLE>, BLOCK, lds);
    if (e != hipSuccess) return e;
    occ = std::max(occ, 1);
    const uint64_t tiles = a.page_end - a.page_begin;
    const uint32_t grid = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(tiles, (uint64_t)ctx->cus * occ));
    hipLaunchKernelGGL((k_q1<DIRECT, LDS_TABLE>), dim3(grid), dim3(BLOCK), lds, ctx->stream, a);
    return hipGetLastError();
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
    a.disc_hi = l.disc_hi;
    a.qty_lo = l.qty_lo;
    a.qty_hi = l.qty_hi;
    a.price_lo = l.price_lo;
    a.price_hi = l.price_hi;
    a.slots = l.slots;
    a.log2cap = l.log2cap;
    a.sum = l.sum;
    a.count = l.count;
    a.err = l.err;
    const size_t tab = l.in_lds ? (8ull << l.log2cap) : 0;
    const size_t win = direct ? 0 : (size_t)l.page_rows * 10;
    const size_t lds = tab + win;
    if (direct) return l.in_lds ? launch_q1_t<true, true>(ctx, a, lds) : launch_q1_t<true, false>(ctx, a, lds);
    return l.in_lds ? launch_q1_t<false, true>(ctx, a, lds) : launch_q1_t<false, false>(ctx, a, lds);
}

}  // namespace golap_impl
