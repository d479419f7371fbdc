// Device code shared by the fused Q1 kernels (kernels.hip, q1wg.hip): page access, the date
// probe structures, the plan's arguments, XCD-aware tile order and the per-thread row passes.
#pragma once
#include <hip/hip_runtime.h>

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
    // predicates as lo <= v <= lo + span (unsigned arithmetic)
    uint32_t disc_lo, disc_span, qty_lo, qty_span, price_lo, price_span;
    const uint64_t *slots;   // PROBE_HASH_*: open-addressing table, 2^log2cap slots
    uint32_t log2cap;
    const uint32_t *bitmap;  // PROBE_BITMAP: membership bits of keys in [kmin, kmin+krange)
    int32_t kmin;
    uint32_t krange;
    int64_t *sum;
    uint64_t *count;
    uint32_t *err;
    // k_q1_wave: LDS bytes of the probe table, of the decode window and token offsets it holds
    uint32_t tab_bytes, win_cap, tok_cap;
};

// How the date probe is answered: the open-addressing table in HBM/L2, the same table pinned
// in LDS, or (dense key range) an exact membership bitmap pinned in LDS.
enum : int { PROBE_HASH_GLOBAL = 0, PROBE_HASH_LDS = 1, PROBE_BITMAP = 2 };

template <int PROBE>
__device__ __forceinline__ void q1_row(const Q1Args &a, const void *tab, int32_t dk, uint32_t d,
                                       uint32_t q, int32_t pr, int64_t &sum, uint32_t &cnt) {
    const bool pass = (d - a.disc_lo <= a.disc_span) & (q - a.qty_lo <= a.qty_span) &
                      ((uint32_t)pr - a.price_lo <= a.price_span);
    if (PROBE == PROBE_BITMAP) {
        // branch-free: one LDS word per row
        const uint32_t off = (uint32_t)dk - (uint32_t)a.kmin;
        const bool in = off < a.krange;
        const uint32_t w = static_cast<const uint32_t *>(tab)[in ? (off >> 5) : 0u];
        const bool hit = pass & in & (((w >> (off & 31)) & 1u) != 0);
        sum += (int64_t)pr * (int64_t)(hit ? d : 0u);
        cnt += hit ? 1u : 0u;
    } else if (pass) {
        uint32_t pay;
        if (ht_probe(static_cast<const uint64_t *>(tab), a.log2cap, dk, pay)) {
            sum += (int64_t)pr * (int64_t)d;
            cnt += 1;
        }
    }
}

template <int PROBE>
__device__ __forceinline__ void q1_group(const Q1Args &a, const void *tab, int4 dv, int4 pv,
                                         uint32_t sv, uint32_t qv, int64_t &sum, uint32_t &cnt) {
    q1_row<PROBE>(a, tab, dv.x, sv & 0xff, qv & 0xff, pv.x, sum, cnt);
    q1_row<PROBE>(a, tab, dv.y, (sv >> 8) & 0xff, (qv >> 8) & 0xff, pv.y, sum, cnt);
    q1_row<PROBE>(a, tab, dv.z, (sv >> 16) & 0xff, (qv >> 16) & 0xff, pv.z, sum, cnt);
    q1_row<PROBE>(a, tab, dv.w, sv >> 24, qv >> 24, pv.w, sum, cnt);
}

// Rows [0, rows) of one tile whose four columns are readable at dp/sp/qp/pp (in HBM or LDS).
template <int PROBE>
__device__ __forceinline__ void q1_rows_generic(const Q1Args &a, const void *tab,
                                                const uint8_t *dp, const uint8_t *sp,
                                                const uint8_t *qp, const uint8_t *pp, uint32_t rows,
                                                int64_t &sum, uint32_t &cnt) {
    for (uint32_t g = threadIdx.x; g * 4 < rows; g += BLOCK) {
        const uint32_t rr = g * 4;
        if (rr + 4 <= rows) {
            q1_group<PROBE>(a, tab, *reinterpret_cast<const int4 *>(dp + 16 * (size_t)g),
                     *reinterpret_cast<const int4 *>(pp + 16 * (size_t)g),
                     *reinterpret_cast<const uint32_t *>(sp + 4 * (size_t)g),
                     *reinterpret_cast<const uint32_t *>(qp + 4 * (size_t)g), sum, cnt);
        } else {
            for (uint32_t r = rr; r < rows; ++r) {
                int32_t dk, pr;
                __builtin_memcpy(&dk, dp + 4 * (size_t)r, 4);
                __builtin_memcpy(&pr, pp + 4 * (size_t)r, 4);
                q1_row<PROBE>(a, tab, dk, sp[r], qp[r], pr, sum, cnt);
            }
        }
    }
}

// Copies the probe structure into LDS (PROBE_HASH_LDS, PROBE_BITMAP); returns what to probe.
template <int PROBE>
__device__ __forceinline__ const void *q1_table(const Q1Args &a, uint8_t *lds) {
    if (PROBE == PROBE_HASH_GLOBAL) return a.slots;
    uint32_t *t = reinterpret_cast<uint32_t *>(lds);
    const uint32_t *src = PROBE == PROBE_BITMAP ? a.bitmap : reinterpret_cast<const uint32_t *>(a.slots);
    const uint32_t words = PROBE == PROBE_BITMAP ? (a.krange + 31) / 32 : (2u << a.log2cap);
    for (uint32_t i = threadIdx.x; i < words; i += BLOCK) t[i] = src[i];
    __syncthreads();
    return t;
}

__device__ __forceinline__ void q1_flush(const Q1Args &a, int64_t sum, uint32_t cnt) {
    const int64_t ws = wave_sum(sum);
    const uint32_t wc = wave_sum(cnt);
    if (__lane_id() == 0 && wc) {
        atomicAdd(reinterpret_cast<unsigned long long *>(a.sum), (unsigned long long)ws);
        atomicAdd(reinterpret_cast<unsigned long long *>(a.count), (unsigned long long)wc);
    }
}

// XCD-aware tile order. Workgroups are dispatched round-robin over the XCDs (workgroup i runs
// on XCD i mod XCDS), and each XCD has its own L2. Giving each XCD one contiguous range of pages
// makes neighbouring pages, which share the cache lines at their boundaries, meet in the same L2.
constexpr uint32_t XCDS = 8;
struct XcdRange {
    uint64_t first, end, stride;
};
__device__ __forceinline__ XcdRange xcd_range(uint64_t begin, uint64_t end) {
    if (gridDim.x % XCDS != 0) return XcdRange{begin + blockIdx.x, end, gridDim.x};
    const uint64_t n = end - begin;
    const uint32_t x = blockIdx.x % XCDS;
    return XcdRange{begin + n * x / XCDS + blockIdx.x / XCDS, begin + n * (x + 1) / XCDS,
                    gridDim.x / XCDS};
}

template <int G4>
struct Q1Lanes {
    static constexpr int NW = (4 * G4 + 63) / 64;
    uint64_t m[NW];     // bit 4 i + j: row 256 i + 4 lane + j is a candidate
    uint32_t disc[G4];  // discount bytes of the four rows of group i
    __device__ __forceinline__ bool bit(int b) const { return (m[b >> 6] >> (b & 63)) & 1ull; }
    __device__ __forceinline__ void clear(int b) { m[b >> 6] &= ~(1ull << (b & 63)); }
    // every row a candidate (bits past 4 G4 stay clear)
    __device__ __forceinline__ void reset() {
#pragma unroll
        for (int w = 0; w < NW; ++w) {
            const int n = 4 * G4 - 64 * w;
            m[w] = n >= 64 ? ~0ull : (1ull << n) - 1;
        }
    }
    __device__ __forceinline__ bool any() const {
        uint64_t o = 0;
#pragma unroll
        for (int w = 0; w < NW; ++w) o |= m[w];
        return o != 0;
    }
};

template <int NT, int G4, bool FULL>
__device__ __forceinline__ void q1_date_pass(const Q1Args &a, const uint32_t *tab, const uint8_t *dp, uint32_t rows,
                                             Q1Lanes<G4> &st, bool hash) {
    const uint32_t t = threadIdx.x;
#pragma unroll
    for (int i = 0; i < G4; ++i) {
        const uint32_t r0 = 4 * NT * i + 4 * t;
        int32_t v[4];
        if (FULL || r0 + 4 <= rows) {
            const int4 x = *reinterpret_cast<const int4 *>(dp + 4 * r0);
            v[0] = x.x; v[1] = x.y; v[2] = x.z; v[3] = x.w;
        } else {
#pragma unroll
            for (int j = 0; j < 4; ++j) v[j] = r0 + j < rows ? reinterpret_cast<const int32_t *>(dp)[r0 + j] : 0;
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int b = 4 * i + j;
            bool hit;
            if (!hash) {
                const uint32_t off = (uint32_t)v[j] - (uint32_t)a.kmin;
                const bool in = off < a.krange;
                const uint32_t w = tab[in ? (off >> 5) : 0u];
                hit = in && ((w >> (off & 31)) & 1u);
            } else {
                uint32_t pay;
                hit = st.bit(b) && ht_probe(reinterpret_cast<const uint64_t *>(tab), a.log2cap, v[j], pay);
            }
            if (!hit) st.clear(b);
        }
    }
}

// quantity (KEEP = false) or discount (KEEP = true: the bytes are kept for the sum)
template <int NT, int G4, bool FULL, bool KEEP>
__device__ __forceinline__ void q1_byte_pass(const uint8_t *cp, uint32_t rows, uint32_t lo, uint32_t span,
                                             Q1Lanes<G4> &st) {
    const uint32_t t = threadIdx.x;
#pragma unroll
    for (int i = 0; i < G4; ++i) {
        const uint32_t r0 = 4 * NT * i + 4 * t;
        uint32_t x;
        if (FULL || r0 + 4 <= rows) {
            x = *reinterpret_cast<const uint32_t *>(cp + r0);
        } else {
            x = 0;
#pragma unroll
            for (int j = 0; j < 4; ++j) x |= r0 + j < rows ? (uint32_t)cp[r0 + j] << (8 * j) : 0u;
        }
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (((x >> (8 * j)) & 0xffu) - lo > span) st.clear(4 * i + j);
        if (KEEP) st.disc[i] = x;
    }
}

template <int NT, int G4, bool FULL>
__device__ __forceinline__ void q1_price_pass(const Q1Args &a, const uint8_t *pp, uint32_t rows, const Q1Lanes<G4> &st,
                                              int64_t &sum, uint32_t &cnt) {
    const uint32_t t = threadIdx.x;
#pragma unroll
    for (int i = 0; i < G4; ++i) {
        const uint32_t r0 = 4 * NT * i + 4 * t;
        int32_t v[4];
        if (FULL || r0 + 4 <= rows) {
            const int4 x = *reinterpret_cast<const int4 *>(pp + 4 * r0);
            v[0] = x.x; v[1] = x.y; v[2] = x.z; v[3] = x.w;
        } else {
#pragma unroll
            for (int j = 0; j < 4; ++j) v[j] = r0 + j < rows ? reinterpret_cast<const int32_t *>(pp)[r0 + j] : 0;
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const bool hit = st.bit(4 * i + j) && ((uint32_t)v[j] - a.price_lo <= a.price_span);
            sum += (int64_t)v[j] * (int64_t)(hit ? (st.disc[i] >> (8 * j)) & 0xffu : 0u);
            cnt += hit ? 1u : 0u;
        }
    }
}

// One column's page of the current tile (uniform values).
struct Q1Page {
    const uint8_t *src;
    uint32_t csize;
    uint8_t code;
};
__device__ __forceinline__ Q1Page q1_page(const ColView &c, uint64_t p) {
    return Q1Page{c.comp + c.page_off[p], c.page_csize[p], c.page_codec[p]};
}

}  // namespace golap_impl
