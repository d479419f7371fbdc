// Chunk metadata beyond min/max (gfx950): per-chunk equal-width histograms (EqHistogram,
// core/metadata.cuh:225-382) and per-chunk Bloom filters (BloomMeta, :386-537), built on the
// device from a column's values and checked per chunk into the same 0/1 flags the zone-map
// check writes.
//
// The reference fills them chunk by chunk during prepare_compressed (core/apps.cuh:212-224):
// a cub reduction for min/max, then fill_hist / fill_bloom with one global atomic per value.
// Here one workgroup takes a chunk: min/max by block reduction, counts and filter bits in LDS
// (LDS atomics), one coalesced write of the chunk's histogram or filter. The bucket and hash
// arithmetic is the reference's (double-precision bucket index, MurmurHash3's 64-bit
// finalizer modulo the filter bits), so the arrays are identical.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>

#include "common.hpp"
#include "star_dev.hpp"

namespace golap_impl {
namespace {

constexpr int BLOCK = 256;
constexpr uint32_t HIST_MAX_BUCKETS = 8192;   // u32 LDS counters: 32 KiB
constexpr uint64_t BLOOM_MAX_BYTES = 65536;  // one chunk's filter in LDS

enum : uint32_t { META_HIST = 1, META_BLOOM = 2, META_MINMAX = 3 };

// Value of row r of a typed column, sign- or zero-extended.
__device__ __forceinline__ int64_t load_typed(const void *p, uint32_t type, uint64_t r) {
    switch (type) {
    case GOLAP_I8: return static_cast<const int8_t *>(p)[r];
    case GOLAP_U8: return static_cast<const uint8_t *>(p)[r];
    case GOLAP_I16: return static_cast<const int16_t *>(p)[r];
    case GOLAP_U16: return static_cast<const uint16_t *>(p)[r];
    case GOLAP_I32: return static_cast<const int32_t *>(p)[r];
    case GOLAP_U32: return static_cast<const uint32_t *>(p)[r];
    default: return static_cast<const int64_t *>(p)[r];
    }
}

// Block-wide min and max of the chunk's values (every thread gets them).
__device__ void chunk_minmax(const void *v, uint32_t type, uint64_t r0, uint64_t r1, int64_t *red, int64_t &mn,
                             int64_t &mx) {
    int64_t a = INT64_MAX, b = INT64_MIN;
    for (uint64_t r = r0 + threadIdx.x; r < r1; r += BLOCK) {
        const int64_t x = load_typed(v, type, r);
        a = min(a, x);
        b = max(b, x);
    }
    red[threadIdx.x] = a;
    red[BLOCK + threadIdx.x] = b;
    __syncthreads();
    for (int s = BLOCK / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            red[threadIdx.x] = min(red[threadIdx.x], red[threadIdx.x + s]);
            red[BLOCK + threadIdx.x] = max(red[BLOCK + threadIdx.x], red[BLOCK + threadIdx.x + s]);
        }
        __syncthreads();
    }
    mn = red[0];
    mx = red[BLOCK];
    __syncthreads();
}

// EqHistogram::init_chunk + fill_hist (metadata.cuh:281-285, 333-347): workgroup per chunk.
__global__ __launch_bounds__(BLOCK) void k_hist_build(const void *v, uint32_t type, uint64_t n, uint64_t chunk_rows,
                                                      uint64_t n_chunks, uint32_t buckets, int64_t *mins,
                                                      int64_t *maxs, uint64_t *hist) {
    __shared__ int64_t red[2 * BLOCK];
    __shared__ uint32_t cnt[HIST_MAX_BUCKETS];
    for (uint64_t c = blockIdx.x; c < n_chunks; c += gridDim.x) {
        const uint64_t r0 = c * chunk_rows, r1 = min(n, r0 + chunk_rows);
        int64_t mn, mx;
        chunk_minmax(v, type, r0, r1, red, mn, mx);
        for (uint32_t b = threadIdx.x; b < buckets; b += BLOCK) cnt[b] = 0;
        __syncthreads();
        const double inv = mx != mn ? (double)buckets / (double)(mx - mn) : 0.0;
        for (uint64_t r = r0 + threadIdx.x; r < r1; r += BLOCK) {
            uint64_t b = 0;
            if (mx != mn) {
                b = (uint64_t)((double)(load_typed(v, type, r) - mn) * inv);
                b = min<uint64_t>(buckets - 1, b);  // the maximum maps to bucket `buckets`
            }
            atomicAdd(&cnt[b], 1u);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            mins[c] = mn;
            maxs[c] = mx;
        }
        for (uint32_t b = threadIdx.x; b < buckets; b += BLOCK) hist[c * buckets + b] = cnt[b];
        __syncthreads();
    }
}

// check_hist / EqHistogram::check_pred (metadata.cuh:349-372, 573-585): one thread per chunk,
// the buckets between first and last (the reference's truncating conversions) searched for a
// non-empty one.
__global__ __launch_bounds__(BLOCK) void k_hist_check(const int64_t *mins, const int64_t *maxs, const uint64_t *hist,
                                                      uint64_t n_chunks, uint32_t buckets, int64_t lo, int64_t hi,
                                                      uint16_t *flags) {
    for (uint64_t c = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; c < n_chunks; c += (uint64_t)gridDim.x * BLOCK) {
        const int64_t mn = mins[c], mx = maxs[c];
        bool keep = false;
        if (mn == mx) {
            keep = lo <= mn && hi >= mx && hist[c * buckets] != 0;
        } else {
            const double inv = (double)buckets / (double)(mx - mn);
            int64_t first = (int64_t)(((double)lo - (double)mn) * inv);
            const int64_t last = (int64_t)(((double)hi - (double)mn) * inv);
            if (lo == mx) first = (int64_t)buckets - 1;
            const int64_t b0 = max<int64_t>(first, 0), b1 = min<int64_t>(last, (int64_t)buckets - 1);
            for (int64_t b = b0; b <= b1 && !keep; ++b) keep = hist[c * buckets + b] != 0;
        }
        flags[c] = keep ? 1 : 0;
    }
}

// BloomMeta::hash (metadata.cuh:457-465): MurmurHash3's 64-bit finalizer, modulo the bits.
__device__ __forceinline__ uint64_t bloom_hash(uint64_t x, uint64_t mbits) {
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdull;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ull;
    x ^= x >> 33;
    return x % mbits;
}

// fill_bloom / BloomMeta::add (metadata.cuh:467-481): workgroup per chunk, the chunk's filter
// in LDS.
__global__ __launch_bounds__(BLOCK) void k_bloom_build(const void *v, uint32_t type, uint64_t n, uint64_t chunk_rows,
                                                       uint64_t n_chunks, uint64_t m, uint32_t k, uint32_t *state) {
    extern __shared__ uint32_t f[];
    const uint64_t words = m / 4, mbits = m << 3;
    for (uint64_t c = blockIdx.x; c < n_chunks; c += gridDim.x) {
        for (uint64_t w = threadIdx.x; w < words; w += BLOCK) f[w] = 0;
        __syncthreads();
        const uint64_t r0 = c * chunk_rows, r1 = min(n, r0 + chunk_rows);
        for (uint64_t r = r0 + threadIdx.x; r < r1; r += BLOCK) {
            const uint64_t x = (uint64_t)load_typed(v, type, r);
            for (uint32_t i = 0; i < k; ++i) {
                const uint64_t bit = bloom_hash(x + i, mbits);
                atomicOr(&f[bit >> 5], 1u << (bit & 31));
            }
        }
        __syncthreads();
        for (uint64_t w = threadIdx.x; w < words; w += BLOCK) state[c * words + w] = f[w];
        __syncthreads();
    }
}

// check_bloom / BloomMeta::check_pred (metadata.cuh:483-501): one thread per chunk.
__global__ __launch_bounds__(BLOCK) void k_bloom_check(const uint32_t *state, uint64_t n_chunks, uint64_t m, uint32_t k,
                                                       uint64_t x, uint16_t *flags) {
    const uint64_t words = m / 4, mbits = m << 3;
    for (uint64_t c = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; c < n_chunks; c += (uint64_t)gridDim.x * BLOCK) {
        bool keep = true;
        for (uint32_t i = 0; i < k; ++i) {
            const uint64_t bit = bloom_hash(x + i, mbits);
            keep &= ((state[c * words + (bit >> 5)] >> (bit & 31)) & 1u) != 0;
        }
        flags[c] = keep ? 1 : 0;
    }
}

bool int_type(uint32_t t) { return t >= GOLAP_I8 && t <= GOLAP_I64; }

// SortHelper::prejoin_tables (usecases/ssb/helper.hpp:115-152) + MinMaxMeta of the prejoined
// column: workgroup per chunk; each row's foreign key (decoded) is probed into the dimension map,
// whose payload is the attribute; min and max over the rows that join (none: min > max).
__global__ __launch_bounds__(BLOCK) void k_prejoin_minmax(const int32_t *fk, uint64_t n, uint64_t chunk_rows,
                                                          uint64_t n_chunks, const uint64_t *slots, uint32_t log2cap,
                                                          int64_t *mins, int64_t *maxs) {
    __shared__ int64_t red[2 * BLOCK];
    for (uint64_t c = blockIdx.x; c < n_chunks; c += gridDim.x) {
        const uint64_t r0 = c * chunk_rows, r1 = min(n, r0 + chunk_rows);
        int64_t a = INT64_MAX, b = INT64_MIN;
        for (uint64_t r = r0 + threadIdx.x; r < r1; r += BLOCK) {
            const int32_t key = fk[r];
            const uint32_t h = stardev::hash32((uint32_t)key) >> (32 - log2cap);
            uint32_t pay;
            if (stardev::probe_exact(slots, log2cap, key, slots[h], h, pay)) {
                a = min(a, (int64_t)(int32_t)pay);
                b = max(b, (int64_t)(int32_t)pay);
            }
        }
        red[threadIdx.x] = a;
        red[BLOCK + threadIdx.x] = b;
        __syncthreads();
        for (int s = BLOCK / 2; s > 0; s >>= 1) {
            if ((int)threadIdx.x < s) {
                red[threadIdx.x] = min(red[threadIdx.x], red[threadIdx.x + s]);
                red[BLOCK + threadIdx.x] = max(red[BLOCK + threadIdx.x], red[BLOCK + threadIdx.x + s]);
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            mins[c] = red[0];
            maxs[c] = red[BLOCK];
        }
        __syncthreads();
    }
}

// check_mmmeta (core/metadata.cuh:539-547) over a min/max meta: flag = lo <= max && hi >= min.
__global__ __launch_bounds__(BLOCK) void k_minmax_meta_check(const int64_t *mins, const int64_t *maxs, uint64_t n,
                                                             int64_t lo, int64_t hi, uint16_t *flags) {
    for (uint64_t c = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; c < n; c += (uint64_t)gridDim.x * BLOCK)
        flags[c] = (lo <= maxs[c] && hi >= mins[c]) ? 1 : 0;
}

}  // namespace
}  // namespace golap_impl

using namespace golap_impl;
using Lock = std::lock_guard<std::recursive_mutex>;

struct golap_chunkmeta {
    golap_ctx *ctx = nullptr;
    uint32_t kind = 0, type = 0;
    uint64_t n_rows = 0, chunk_rows = 0, n_chunks = 0;
    uint32_t buckets = 0;         // histogram
    uint64_t m = 0, k = 0, n = 0;  // Bloom: bytes per chunk, hashes, design items
    int64_t *d_mins = nullptr, *d_maxs = nullptr;
    uint64_t *d_hist = nullptr;
    uint32_t *d_state = nullptr;
};

namespace {

void meta_release(golap_chunkmeta *h) {
    if (!h) return;
    dev_release(h->ctx, h->d_mins);
    dev_release(h->ctx, h->d_maxs);
    dev_release(h->ctx, h->d_hist);
    dev_release(h->ctx, h->d_state);
    delete h;
}

int meta_new(golap_ctx *ctx, const void *v, uint32_t type, uint64_t n, uint64_t chunk_rows, golap_chunkmeta **out,
             golap_chunkmeta **h) {
    if (!ctx || !out || (n && !v) || !chunk_rows || !int_type(type)) return GOLAP_EINVAL;
    *out = nullptr;
    *h = new golap_chunkmeta();
    (*h)->ctx = ctx;
    (*h)->type = type;
    (*h)->n_rows = n;
    (*h)->chunk_rows = chunk_rows;
    (*h)->n_chunks = (n + chunk_rows - 1) / chunk_rows;
    return GOLAP_OK;
}

uint32_t grid_for(golap_ctx *ctx, uint64_t items) {
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(items, (uint64_t)ctx->cus * 8));
}

}  // namespace

extern "C" {

int golap_eqhist_build(golap_ctx *ctx, const void *dev_values, uint32_t type, uint64_t n_rows, uint64_t chunk_rows,
                       uint32_t bucket_num, golap_chunkmeta **out) {
    golap_chunkmeta *h = nullptr;
    if (int r = meta_new(ctx, dev_values, type, n_rows, chunk_rows, out, &h)) return r;
    Lock l(ctx->mu);
    if (int r = set_device(ctx)) return meta_release(h), r;
    if (!bucket_num || bucket_num > HIST_MAX_BUCKETS) {
        meta_release(h);
        return fail(ctx, GOLAP_EINVAL, "eqhist: bucket_num must be 1..8192");
    }
    h->kind = META_HIST;
    h->buckets = bucket_num;
    const uint64_t nc = std::max<uint64_t>(h->n_chunks, 1);
    void *a = nullptr, *b = nullptr, *c = nullptr;
    if (int r = dev_alloc(ctx, nc * 8, &a)) return meta_release(h), r;
    h->d_mins = static_cast<int64_t *>(a);
    if (int r = dev_alloc(ctx, nc * 8, &b)) return meta_release(h), r;
    h->d_maxs = static_cast<int64_t *>(b);
    if (int r = dev_alloc(ctx, nc * bucket_num * 8, &c)) return meta_release(h), r;
    h->d_hist = static_cast<uint64_t *>(c);
    if (h->n_chunks)
        hipLaunchKernelGGL(k_hist_build, dim3(grid_for(ctx, h->n_chunks)), dim3(BLOCK), 0, ctx->stream, dev_values,
                           type, n_rows, chunk_rows, h->n_chunks, bucket_num, h->d_mins, h->d_maxs, h->d_hist);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        meta_release(h);
        return fail(ctx, GOLAP_EHIP, std::string("eqhist_build: ") + hipGetErrorString(e));
    }
    *out = h;
    return GOLAP_OK;
}

int golap_eqhist_check(golap_ctx *ctx, const golap_chunkmeta *h, int64_t lo, int64_t hi, uint16_t *dev_flags) {
    if (!ctx || !h || h->ctx != ctx || h->kind != META_HIST || (h->n_chunks && !dev_flags)) return GOLAP_EINVAL;
    Lock l(ctx->mu);
    if (int r = set_device(ctx)) return r;
    if (!h->n_chunks) return GOLAP_OK;
    hipLaunchKernelGGL(k_hist_check, dim3(grid_for(ctx, (h->n_chunks + BLOCK - 1) / BLOCK)), dim3(BLOCK), 0,
                       ctx->stream, h->d_mins, h->d_maxs, h->d_hist, h->n_chunks, h->buckets, lo, hi, dev_flags);
    GOLAP_HIP_CHECK(ctx, hipGetLastError());
    return GOLAP_OK;
}

int golap_bloommeta_build(golap_ctx *ctx, const void *dev_values, uint32_t type, uint64_t n_rows, uint64_t chunk_rows,
                          double p, uint64_t m_bytes, golap_chunkmeta **out) {
    golap_chunkmeta *h = nullptr;
    if (int r = meta_new(ctx, dev_values, type, n_rows, chunk_rows, out, &h)) return r;
    Lock l(ctx->mu);
    if (int r = set_device(ctx)) return meta_release(h), r;
    if (!(p > 0.0 && p < 1.0) || !m_bytes || m_bytes % 4 || m_bytes > BLOOM_MAX_BYTES) {
        meta_release(h);
        return fail(ctx, GOLAP_EINVAL, "bloommeta: p in (0, 1) and m a multiple of 4 up to 65536 bytes");
    }
    // BloomMeta's (p, m) parameters (metadata.cuh:398-401)
    const uint64_t mbits = m_bytes << 3;
    h->kind = META_BLOOM;
    h->m = m_bytes;
    h->n = (uint64_t)std::ceil((double)mbits * std::log(std::pow(0.5, std::log(2.0))) / std::log(p));
    h->k = (uint64_t)std::ceil(std::log(2.0) * (double)mbits / (double)std::max<uint64_t>(h->n, 1));
    void *s = nullptr;
    if (int r = dev_alloc(ctx, std::max<uint64_t>(h->n_chunks, 1) * m_bytes, &s)) return meta_release(h), r;
    h->d_state = static_cast<uint32_t *>(s);
    if (h->n_chunks)  // the chunk's filter in dynamic LDS (<= 64 KiB)
        hipLaunchKernelGGL(k_bloom_build, dim3(grid_for(ctx, h->n_chunks)), dim3(BLOCK), m_bytes, ctx->stream,
                           dev_values, type, n_rows, chunk_rows, h->n_chunks, m_bytes, (uint32_t)h->k, h->d_state);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        meta_release(h);
        return fail(ctx, GOLAP_EHIP, std::string("bloommeta_build: ") + hipGetErrorString(e));
    }
    *out = h;
    return GOLAP_OK;
}

int golap_bloommeta_check(golap_ctx *ctx, const golap_chunkmeta *h, int64_t value, uint16_t *dev_flags) {
    if (!ctx || !h || h->ctx != ctx || h->kind != META_BLOOM || (h->n_chunks && !dev_flags)) return GOLAP_EINVAL;
    Lock l(ctx->mu);
    if (int r = set_device(ctx)) return r;
    if (!h->n_chunks) return GOLAP_OK;
    // the value as the column type holds it, widened as BloomMeta::_to_uint64_t does (:514-520)
    uint64_t x = 0;
    switch (h->type) {
    case GOLAP_I8: x = (uint64_t)(int64_t)(int8_t)value; break;
    case GOLAP_U8: x = (uint64_t)(uint8_t)value; break;
    case GOLAP_I16: x = (uint64_t)(int64_t)(int16_t)value; break;
    case GOLAP_U16: x = (uint64_t)(uint16_t)value; break;
    case GOLAP_I32: x = (uint64_t)(int64_t)(int32_t)value; break;
    case GOLAP_U32: x = (uint64_t)(uint32_t)value; break;
    default: x = (uint64_t)value; break;
    }
    hipLaunchKernelGGL(k_bloom_check, dim3(grid_for(ctx, (h->n_chunks + BLOCK - 1) / BLOCK)), dim3(BLOCK), 0,
                       ctx->stream, h->d_state, h->n_chunks, h->m, (uint32_t)h->k, x, dev_flags);
    GOLAP_HIP_CHECK(ctx, hipGetLastError());
    return GOLAP_OK;
}

int golap_prejoin_minmax(golap_ctx *ctx, const golap_column *fk, const golap_hashmap *map, golap_chunkmeta **out) {
    if (!ctx || !fk || !map || !out || fk->ctx != ctx || map->ctx != ctx) return GOLAP_EINVAL;
    if (fk->type != GOLAP_I32 && fk->type != GOLAP_U32) return fail(ctx, GOLAP_EINVAL, "prejoin: 32-bit keys only");
    if (!map->built) return fail(ctx, GOLAP_EINVAL, "prejoin: map not built");
    golap_chunkmeta *h = nullptr;
    if (int r = meta_new(ctx, fk, GOLAP_I32, fk->n_rows, fk->chunk_rows, out, &h)) return r;
    Lock l(ctx->mu);
    if (int r = set_device(ctx)) return meta_release(h), r;
    h->kind = META_MINMAX;
    const uint64_t nc = std::max<uint64_t>(h->n_chunks, 1);
    void *a = nullptr, *b = nullptr, *keys = nullptr;
    if (int r = dev_alloc(ctx, nc * 8, &a)) return meta_release(h), r;
    h->d_mins = static_cast<int64_t *>(a);
    if (int r = dev_alloc(ctx, nc * 8, &b)) return meta_release(h), r;
    h->d_maxs = static_cast<int64_t *>(b);
    if (h->n_chunks) {
        // the keys decoded chunk range by chunk range (at most 64 chunks at a time)
        const uint64_t span = std::min<uint64_t>(h->n_chunks, 64);
        if (int r = dev_alloc(ctx, span * fk->chunk_rows * 4, &keys)) return meta_release(h), r;
        const uint64_t ppc = fk->chunk_rows / fk->page_rows;
        const uint64_t *slots = map->small_log2 ? map->d_small : map->d_slots;
        const uint32_t log2cap = map->small_log2 ? map->small_log2 : map->log2cap;
        hipError_t e = hipSuccess;
        for (uint64_t c0 = 0; c0 < h->n_chunks && e == hipSuccess; c0 += span) {
            const uint64_t c1 = std::min(h->n_chunks, c0 + span);
            const uint64_t rows = std::min(fk->n_rows, c1 * fk->chunk_rows) - c0 * fk->chunk_rows;
            e = launch_decode_pages(ctx, fk, c0 * ppc, std::min<uint64_t>(fk->n_pages, c1 * ppc),
                                    static_cast<uint8_t *>(keys), c0 * fk->chunk_rows);
            if (e == hipSuccess) {
                hipLaunchKernelGGL(k_prejoin_minmax, dim3(grid_for(ctx, c1 - c0)), dim3(BLOCK), 0, ctx->stream,
                                   static_cast<const int32_t *>(keys), rows, fk->chunk_rows, c1 - c0, slots, log2cap,
                                   h->d_mins + c0, h->d_maxs + c0);
                e = hipGetLastError();
            }
        }
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        dev_release(ctx, keys);
        if (e != hipSuccess) {
            meta_release(h);
            return fail(ctx, GOLAP_EHIP, std::string("prejoin_minmax: ") + hipGetErrorString(e));
        }
        if (int r = check_device_errors(ctx, "prejoin_minmax")) return meta_release(h), r;
    }
    *out = h;
    return GOLAP_OK;
}

int golap_chunkmeta_minmax_check(golap_ctx *ctx, const golap_chunkmeta *h, int64_t lo, int64_t hi,
                                 uint16_t *dev_flags) {
    if (!ctx || !h || h->ctx != ctx || !h->d_mins || (h->n_chunks && !dev_flags)) return GOLAP_EINVAL;
    Lock l(ctx->mu);
    if (int r = set_device(ctx)) return r;
    if (!h->n_chunks) return GOLAP_OK;
    hipLaunchKernelGGL(k_minmax_meta_check, dim3(grid_for(ctx, (h->n_chunks + BLOCK - 1) / BLOCK)), dim3(BLOCK), 0,
                       ctx->stream, h->d_mins, h->d_maxs, h->n_chunks, lo, hi, dev_flags);
    GOLAP_HIP_CHECK(ctx, hipGetLastError());
    return GOLAP_OK;
}

int golap_chunkmeta_info(const golap_chunkmeta *h, golap_chunkmeta_desc *out) {
    if (!h || !out) return GOLAP_EINVAL;
    out->n_chunks = h->n_chunks;
    out->buckets = h->buckets;
    out->bloom_bytes = h->m;
    out->bloom_k = h->k;
    out->bloom_n = h->n;
    out->mins = h->d_mins;
    out->maxs = h->d_maxs;
    out->hist = h->d_hist;
    out->bloom_state = h->d_state;
    return GOLAP_OK;
}

int golap_chunkmeta_free(golap_ctx *ctx, golap_chunkmeta *h) {
    if (!ctx || !h || h->ctx != ctx) return GOLAP_EINVAL;
    Lock l(ctx->mu);
    if (int r = set_device(ctx)) return r;
    meta_release(h);
    return GOLAP_OK;
}

}  // extern "C"
