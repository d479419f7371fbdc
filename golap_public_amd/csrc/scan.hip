// Fused scan + filter + group-by over paged typed columns (gfx950): the trip-table plans.
//
// The reference groups with group_by_agg_2col / 4col over decompressed attribute columns and
// HashAggregate::add (core/group.cuh:14-36, core/dev_structs.cuh:209-244) after its range
// predicates (usecases/ssb/queries/01.cuh:26-54). Here one persistent kernel reads the
// compressed pages of every column a query uses:
//   - in place (raw and literal-only LZ4 pages), by point access (frame-of-reference and RLE
//     pages, delta pages with bw 0), in registers (delta pages), or decoded into an LDS window
//     (other coded pages, by one wave);
//   - thread t owns rows [t*PER, t*PER + PER) of the page and works column by column over
//     four of them at a time, with wave-uniform column plans, so the loads of one column are
//     independent and issue back to back; delta-bitpack pages of up to two columns are decoded
//     in registers (running sums started from a block scan of the per-thread delta totals);
//   - range predicates on any column; group keys computed per row as (v - base) / div with a
//     magic-number division (hour of a timestamp, zone id), mixed-radix group index;
//   - SUM/COUNT into an LDS-private hashed table keyed by (query, group) that spills rows to
//     global int64 atomics when a slot run is full, and is flushed to the global dense result
//     when it fills up at a work-item end (sorted time columns give every stretch of rows its
//     own small set of groups); when all passing rows of a wave share one group, the wave adds
//     their sum with a single update.
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

// How one typed column is read in the current tile. Offsets, not pointers: global addresses
// are rebuilt from the kernel argument's column base (so loads stay global loads) and window
// addresses from the LDS array.
enum : uint32_t { TA_DIRECT = 0, TA_WINDOW = 1, TA_AFFINE = 2, TA_RLE = 3, TA_FOR = 4, TA_DBP = 5, TA_BAD = 6 };
struct TPlan {
    uint64_t off;  // DIRECT: payload; FOR: fields; RLE, DBP: page start (byte offsets into comp); WINDOW: LDS offset
    int64_t a, b;  // AFFINE: v[r] = a + r*b; FOR: a = frame of reference; RLE: b = value bytes (padded);
                   // DBP: a = v[0], b = minimum delta
    uint32_t mode;
    uint32_t aux;  // RLE: runs; FOR, DBP: bits per field
};

struct ScanShared {
    ScanQueryDev q;
    ScanColView cols[SCAN_MAX_COLS];  // the launch's columns: indexed at run time from LDS, not kernarg
    TPlan tp[SCAN_MAX_COLS];
    uint64_t wtot[SCAN_DBP_SLOTS][BLOCK / WAVE];
    uint32_t tile_bad, fill, item;
};

__device__ __forceinline__ uint32_t uni(uint32_t x) { return __builtin_amdgcn_readfirstlane(x); }

// A column base read back from LDS is a generic pointer; it always points to device memory, and
// loads through the global address space are plain global loads.
typedef const __attribute__((address_space(1))) uint8_t gbyte;
__device__ __forceinline__ gbyte *gmem(const uint8_t *p) { return (gbyte *)p; }
__device__ __forceinline__ uint64_t uni64(uint64_t x) {
    return ((uint64_t)uni((uint32_t)(x >> 32)) << 32) | uni((uint32_t)x);
}

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
__device__ __forceinline__ uint64_t gload_elem(const __attribute__((address_space(1))) uint8_t *p, uint32_t elem) {
    switch (elem) {
    case 1: return *p;
    case 2: return *reinterpret_cast<const __attribute__((address_space(1))) uint16_t *>(p);
    case 4: return *reinterpret_cast<const __attribute__((address_space(1))) uint32_t *>(p);
    default: return *reinterpret_cast<const __attribute__((address_space(1))) uint64_t *>(p);
    }
}

// Access plan of column v for page p; woff: its LDS window offset or -1.
__device__ inline TPlan tcol_plan(const ScanColView &v, uint64_t p, uint32_t rows) {
    TPlan x{};
    const uint32_t elem = v.c.elem;
    const uint8_t codec = v.c.page_codec[p];
    const uint64_t off = v.c.page_off[p];
    const uint8_t *src = v.c.comp + off;
    const uint32_t csize = v.c.page_csize[p];
    const uint32_t cd = codec & PAGE_CODEC_MASK;
    if (is_direct(codec)) {
        const uint32_t hdr = direct_hdr(codec, rows * elem);
        x.off = off + hdr;
        x.mode = (csize == hdr + rows * elem && !(reinterpret_cast<uintptr_t>(src + hdr) % elem)) ? TA_DIRECT : TA_BAD;
    } else if (cd == golap_fmt::DBP && (codec & golap_fmt::PAGE_POINT)) {
        if (csize < 32 || src[20] != 0 || src[21] != elem || ld_u32_unaligned(src + 16) != rows) {
            x.mode = TA_BAD;
        } else {
            x.a = (int64_t)ld_u64_unaligned(src);
            x.b = (int64_t)ld_u64_unaligned(src + 8);
            x.mode = TA_AFFINE;
        }
    } else if (cd == golap_fmt::FOR) {
        // fields read as aligned 8-byte words: the payload must be 8-byte aligned
        if (for_check(src, csize, elem) != (int64_t)rows || (reinterpret_cast<uintptr_t>(src) & 7)) {
            x.mode = TA_BAD;
        } else {
            x.off = off + 32;
            x.a = (int64_t)ld_u64_unaligned(src);
            x.aux = src[20];
            x.mode = TA_FOR;
        }
    } else if (cd == golap_fmt::RLE) {
        const uint32_t nruns = csize >= 16 ? ld_u32_unaligned(src) : 0u;
        const uint64_t vbytes = ((uint64_t)nruns * elem + 3) & ~3ull;
        const bool ok = csize >= 16 && src[8] == elem && ld_u32_unaligned(src + 4) == rows && nruns >= 1 &&
                        16ull + vbytes + 4ull * nruns <= csize;
        x.off = off;
        x.aux = nruns;
        x.b = (int64_t)vbytes;
        x.mode = ok ? TA_RLE : TA_BAD;
    } else if (cd == golap_fmt::DBP && v.dslot >= 0) {
        const uint32_t n = ld_u32_unaligned(src + 16), bw = src[20];
        const uint64_t words = n ? (((uint64_t)(n - 1) * bw + 31) / 32) : 0;
        if (csize < 32 || n != rows || src[21] != elem || bw > 32 || 32 + words * 4 > csize ||
            (reinterpret_cast<uintptr_t>(src) & 3)) {
            x.mode = TA_BAD;
        } else {
            x.off = off;
            x.a = (int64_t)ld_u64_unaligned(src);
            x.b = (int64_t)ld_u64_unaligned(src + 8);
            x.aux = bw;
            x.mode = TA_DBP;
        }
    } else if (v.woff >= 0) {
        x.off = (uint64_t)v.woff;
        x.mode = TA_WINDOW;
    } else {
        x.mode = TA_BAD;
    }
    return x;
}

constexpr int NR = 4;  // rows of a thread processed together (independent loads per column)

// Uniform view of a register-decoded delta-bitpack column in the current tile.
struct DbpReg {
    const __attribute__((address_space(1))) uint32_t *pk;  // packed fields
    uint64_t mind, fmask, words;
    uint32_t bw;
};
// v[i] - v[i-1] of row i >= 1.
__device__ __forceinline__ uint64_t dbp_delta(const DbpReg &d, uint32_t i) {
    uint64_t p = 0;
    if (d.bw) {
        const uint64_t bit = (uint64_t)(i - 1) * d.bw, w = bit >> 5;
        const uint64_t lo = d.pk[w], hi = (w + 1 < d.words) ? d.pk[w + 1] : 0;
        p = ((lo | (hi << 32)) >> (bit & 31)) & d.fmask;
    }
    return d.mind + p;
}

// Raw values of column c (wave-uniform) at the thread's rows r[j] whose bit is set in valid;
// dv holds the current group's values of the register-decoded columns.
__device__ __forceinline__ void fetch(const ScanArgs &a, const ScanShared &S, const uint8_t *lds, uint32_t c,
                                      const uint32_t (&r)[NR], uint32_t valid, uint64_t (&raw)[NR],
                                      const uint64_t (&dv)[SCAN_DBP_SLOTS][NR]) {
    const uint32_t mode = uni(S.tp[c].mode);
    if (mode == TA_DBP) {
        const uint32_t slot = uni((uint32_t)S.cols[c].dslot);
#pragma unroll
        for (int j = 0; j < NR; ++j) raw[j] = slot ? dv[1][j] : dv[0][j];
        return;
    }
    const uint32_t elem = uni(S.cols[c].c.elem);
    const uint64_t off = uni64(S.tp[c].off);
    const uint8_t *comp = reinterpret_cast<const uint8_t *>(uni64(reinterpret_cast<uint64_t>(S.cols[c].c.comp)));
    switch (mode) {
    case TA_DIRECT: {
        gbyte *base = gmem(comp) + off;
#pragma unroll
        for (int j = 0; j < NR; ++j) raw[j] = ((valid >> j) & 1u) ? gload_elem(base + (size_t)r[j] * elem, elem) : 0;
        break;
    }
    case TA_WINDOW: {
        const uint8_t *base = lds + off;
#pragma unroll
        for (int j = 0; j < NR; ++j) raw[j] = ((valid >> j) & 1u) ? load_elem(base + (size_t)r[j] * elem, elem) : 0;
        break;
    }
    case TA_FOR: {
        const __attribute__((address_space(1))) uint64_t *w =
            reinterpret_cast<const __attribute__((address_space(1))) uint64_t *>(gmem(comp) + off);
        const uint32_t bw = uni(S.tp[c].aux);
        const uint64_t ref = uni64((uint64_t)S.tp[c].a);
        const uint64_t mask = bw ? (~0ull >> (64 - bw)) : 0ull;
#pragma unroll
        for (int j = 0; j < NR; ++j) {
            uint64_t x = 0;
            if (((valid >> j) & 1u) && bw) {
                const uint64_t bit = (uint64_t)r[j] * bw;
                const uint32_t sh = (uint32_t)(bit & 63);
                x = w[bit >> 6] >> sh;
                if (sh + bw > 64) x |= w[(bit >> 6) + 1] << (64 - sh);  // the field ends in the next word
            }
            raw[j] = ref + (x & mask);
        }
        break;
    }
    case TA_AFFINE: {
        const uint64_t a0 = uni64((uint64_t)S.tp[c].a), b0 = uni64((uint64_t)S.tp[c].b);
#pragma unroll
        for (int j = 0; j < NR; ++j) raw[j] = a0 + (uint64_t)r[j] * b0;
        break;
    }
    default: {  // RLE: first run whose exclusive end is above the row
        const uint8_t *pg = comp + off;
        const uint32_t nruns = uni(S.tp[c].aux);
        const uint64_t vbytes = uni64((uint64_t)S.tp[c].b);
#pragma unroll
        for (int j = 0; j < NR; ++j) {
            raw[j] = 0;
            if (!((valid >> j) & 1u)) continue;
            uint32_t lo = 0, hi = nruns - 1;
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (ld_u32_unaligned(pg + 16 + vbytes + 4ull * mid) > r[j]) hi = mid;
                else lo = mid + 1;
            }
            uint64_t v = 0;
            for (uint32_t b = 0; b < elem; ++b) v |= (uint64_t)pg[16 + (uint64_t)lo * elem + b] << (8 * b);
            raw[j] = v;
        }
        break;
    }
    }
}

// Decodes a delta-bitpack page (rows rows of elem bytes) into the LDS window with the whole
// block: thread t sums the deltas of rows [t*k, t*k + k), a block scan of the per-thread totals
// gives every thread its start value. Uniform; false if the header does not fit.
__device__ __forceinline__ bool dbp_decode_block(const uint8_t *src, uint32_t csize, uint8_t *dst, uint32_t rows, uint32_t elem,
                                 ScanShared &S) {
    const uint32_t n = ld_u32_unaligned(src + 16), bw = src[20], e = src[21];
    const uint64_t words = n ? (((uint64_t)(n - 1) * bw + 31) / 32) : 0;
    if (csize < 32 || n != rows || e != elem || bw > 32 || 32 + words * 4 > csize ||
        (reinterpret_cast<uintptr_t>(src) & 3))
        return false;
    const uint64_t base = ld_u64_unaligned(src), mind = ld_u64_unaligned(src + 8);
    const uint32_t *pk = reinterpret_cast<const uint32_t *>(src + 32);
    const uint64_t fmask = bw == 32 ? 0xffffffffull : ((1ull << bw) - 1);
    const uint32_t k = (rows + BLOCK - 1) / BLOCK, r0 = threadIdx.x * k, r1 = min(rows, r0 + k);
    auto delta = [&](uint32_t i) -> uint64_t {  // v[i] - v[i-1], i >= 1
        uint64_t p = 0;
        if (bw) {
            const uint64_t bit = (uint64_t)(i - 1) * bw, w = bit >> 5;
            const uint64_t lo = pk[w], hi = (w + 1 < words) ? pk[w + 1] : 0;
            p = ((lo | (hi << 32)) >> (bit & 31)) & fmask;
        }
        return mind + p;
    };
    uint64_t tot = 0;
    for (uint32_t i = max(r0, 1u); i < r1; ++i) tot += delta(i);
    const uint64_t incl = wave_incl_scan_u64(tot);
    const uint32_t wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    if (lane == WAVE - 1) S.wtot[0][wave] = incl;
    __syncthreads();
    uint64_t run = base + incl - tot;
    for (uint32_t w = 0; w < wave; ++w) run += S.wtot[0][w];
    for (uint32_t i = r0; i < r1; ++i) {
        if (i > 0) run += delta(i);
        store_elem(dst + (size_t)i * elem, run, elem);
    }
    return true;
}

// group_add_lds for n rows of one key with sum v, counting the keys the table holds (fill) for
// the flush decision.
__device__ __forceinline__ bool group_add_counted(const GroupTable &T, uint32_t key, int64_t v, uint32_t n,
                                                  uint32_t *fill) {
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
            atomicAdd(&T.gcnt[h], n);
            return true;
        }
        h = (h + 1) & mask;
    }
    return false;
}

// Moves every LDS entry into its query's dense global result and empties the table.
__device__ __forceinline__ void scan_flush(const ScanArgs &a, const GroupTable &T, ScanShared &S) {
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

__device__ __forceinline__ void scan_set_query(const ScanArgs &a, ScanShared &S, uint32_t q) {
    __syncthreads();
    const uint32_t *src = reinterpret_cast<const uint32_t *>(a.queries ? a.queries + q : &a.q0);
    uint32_t *dst = reinterpret_cast<uint32_t *>(&S.q);
    for (uint32_t i = threadIdx.x; i < SQ_WORDS; i += BLOCK) dst[i] = src[i];
    __syncthreads();
}

template <int PER>
__device__ __forceinline__ void scan_tile(const ScanArgs &a, ScanShared &S, uint32_t q, uint64_t p, uint8_t *lds,
                                          const GroupTable &T) {
    const uint32_t wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    const uint32_t rows = (uint32_t)min<uint64_t>(a.page_rows, a.n_rows - p * a.page_rows);
    const uint32_t used = S.q.colmask;
    if (threadIdx.x == 0) S.tile_bad = 0;
    if (threadIdx.x < a.n_cols && ((used >> threadIdx.x) & 1u)) S.tp[threadIdx.x] = tcol_plan(S.cols[threadIdx.x], p, rows);
    __syncthreads();
    bool bad = false;
    for (uint32_t c = 0; c < a.n_cols; ++c) bad |= ((used >> c) & 1u) && S.tp[c].mode == TA_BAD;
    if (bad) {  // uniform
        if (threadIdx.x == 0) raise_err(a.err, DEVERR_CORRUPT);
        __syncthreads();
        return;
    }
    // thread t owns rows [t*PER, t*PER + PER): delta-bitpack columns with a register slot are
    // decoded by running sums; every thread's start value comes from a block scan of the
    // per-thread delta totals
    const uint32_t row0 = threadIdx.x * PER, row1 = min(rows, row0 + PER);
    DbpReg dr[SCAN_DBP_SLOTS];
    uint64_t run[SCAN_DBP_SLOTS], tot[SCAN_DBP_SLOTS], incl[SCAN_DBP_SLOTS];
    uint32_t dmask = 0;  // active slots (uniform)
#pragma unroll
    for (uint32_t sl = 0; sl < SCAN_DBP_SLOTS; ++sl) {
        run[sl] = tot[sl] = incl[sl] = 0;
        dr[sl] = DbpReg{nullptr, 0, 0, 0, 0};
    }
    for (uint32_t c = 0; c < a.n_cols; ++c) {
        if (!((used >> c) & 1u) || uni(S.tp[c].mode) != TA_DBP) continue;
        const uint32_t sl = uni((uint32_t)S.cols[c].dslot);
        const uint8_t *comp = reinterpret_cast<const uint8_t *>(uni64(reinterpret_cast<uint64_t>(S.cols[c].c.comp)));
        DbpReg d;
        d.pk = reinterpret_cast<const __attribute__((address_space(1))) uint32_t *>(gmem(comp) + uni64(S.tp[c].off) + 32);
        d.bw = uni(S.tp[c].aux);
        d.mind = uni64((uint64_t)S.tp[c].b);
        d.fmask = d.bw == 32 ? 0xffffffffull : ((1ull << d.bw) - 1);
        d.words = rows ? (((uint64_t)(rows - 1) * d.bw + 31) / 32) : 0;
        uint64_t t = 0;
        for (uint32_t i = max(row0, 1u); i < row1; ++i) t += dbp_delta(d, i);
        const uint64_t in = wave_incl_scan_u64(t);
        if (lane == WAVE - 1) S.wtot[sl][wave] = in;
        if (sl == 0) {
            dr[0] = d;
            tot[0] = t;
            incl[0] = in;
            run[0] = uni64((uint64_t)S.tp[c].a);
        } else {
            dr[1] = d;
            tot[1] = t;
            incl[1] = in;
            run[1] = uni64((uint64_t)S.tp[c].a);
        }
        dmask |= 1u << sl;
    }
    // coded pages without a slot: delta pages by the whole block, others by one wave each
    bool any_wave = false;
    for (uint32_t c = 0; c < a.n_cols; ++c) {
        if (!((used >> c) & 1u) || S.tp[c].mode != TA_WINDOW) continue;
        const ColView &cv = S.cols[c].c;
        if ((cv.page_codec[p] & PAGE_CODEC_MASK) == golap_fmt::DBP) {
            __syncthreads();  // slot 0's totals are read before a block decode reuses wtot[0]
            if ((dmask & 1u) && !(dmask & 4u)) {
                run[0] += incl[0] - tot[0];
                for (uint32_t w = 0; w < wave; ++w) run[0] += S.wtot[0][w];
                incl[0] = tot[0] = 0;
                dmask |= 4u;  // slot 0 start value taken
            }
            __syncthreads();
            if (!dbp_decode_block(cv.comp + cv.page_off[p], cv.page_csize[p], lds + S.tp[c].off, rows, cv.elem, S) &&
                threadIdx.x == 0) {
                S.tile_bad = 1;
                raise_err(a.err, DEVERR_CORRUPT);
            }
        } else {
            any_wave = true;
        }
    }
    if (any_wave) {
        uint32_t k = 0;
        for (uint32_t c = 0; c < a.n_cols; ++c) {
            if (!((used >> c) & 1u) || S.tp[c].mode != TA_WINDOW) continue;
            const ColView &cv = S.cols[c].c;
            if ((cv.page_codec[p] & PAGE_CODEC_MASK) == golap_fmt::DBP) continue;
            if (k++ % (BLOCK / WAVE) != wave) continue;
            const uint32_t want = rows * cv.elem;
            const int64_t got = page_decode_wave(cv.page_codec[p], cv.comp + cv.page_off[p], cv.page_csize[p],
                                                 lds + S.tp[c].off, want, cv.elem);
            if (got != (int64_t)want && lane == 0) {
                S.tile_bad = 1;
                raise_err(a.err, DEVERR_CORRUPT);
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (uint32_t sl = 0; sl < SCAN_DBP_SLOTS; ++sl) {
        if (!((dmask >> sl) & 1u) || (sl == 0 && (dmask & 4u))) continue;
        run[sl] += incl[sl] - tot[sl];
        for (uint32_t w = 0; w < wave; ++w) run[sl] += S.wtot[sl][w];
    }
    if (!S.tile_bad) {
        const uint32_t np = uni(S.q.n_preds), nk = uni(S.q.n_keys);
        const int32_t vc = (int32_t)uni((uint32_t)S.q.val_col);
        const bool use_lds = uni(S.q.use_lds) && T.slots;
        const uint32_t qkey = q << a.group_bits;
        int64_t *g_sum = reinterpret_cast<int64_t *>(uni64(reinterpret_cast<uint64_t>(S.q.g_sum)));
        uint64_t *g_cnt = reinterpret_cast<uint64_t *>(uni64(reinterpret_cast<uint64_t>(S.q.g_cnt)));
        for (int j0 = 0; j0 < PER; j0 += NR) {
            uint32_t r[NR];
            uint32_t valid = 0;
#pragma unroll
            for (int j = 0; j < NR; ++j) {
                r[j] = row0 + (uint32_t)(j0 + j);
                valid |= (uint32_t)(r[j] < rows) << j;
            }
            // this group's values of the register-decoded columns
            uint64_t dv[SCAN_DBP_SLOTS][NR];
#pragma unroll
            for (uint32_t sl = 0; sl < SCAN_DBP_SLOTS; ++sl) {
                const bool on = (dmask >> sl) & 1u;
#pragma unroll
                for (int j = 0; j < NR; ++j) {
                    if (on && ((valid >> j) & 1u) && r[j] > 0) run[sl] += dbp_delta(dr[sl], r[j]);
                    dv[sl][j] = run[sl];
                }
            }
            uint32_t pass = valid;
            uint64_t raw[NR];
            for (uint32_t i = 0; i < np; ++i) {
                const uint32_t c = uni(S.q.pred[i].col);
                const uint32_t ty = uni(S.cols[c].type);
                const int64_t lo = (int64_t)uni64((uint64_t)S.q.pred[i].lo), hi = (int64_t)uni64((uint64_t)S.q.pred[i].hi);
                fetch(a, S, lds, c, r, pass, raw, dv);
#pragma unroll
                for (int j = 0; j < NR; ++j) {
                    const int64_t v = typed(raw[j], ty);
                    if (v < lo || v > hi) pass &= ~(1u << j);
                }
            }
            if (__ballot(pass != 0) == 0) continue;  // wave-uniform
            uint32_t g[NR];
#pragma unroll
            for (int j = 0; j < NR; ++j) g[j] = 0;
            for (uint32_t i = 0; i < nk; ++i) {
                const uint32_t c = uni(S.q.key[i].col);
                const uint32_t ty = uni(S.cols[c].type);
                const uint64_t base = uni64((uint64_t)S.q.key[i].base), span = uni64(S.q.key[i].span);
                golap_fmt::UDiv32 dvz;
                dvz.m = uni(S.q.key[i].div.m);
                dvz.s1 = uni(S.q.key[i].div.s1);
                dvz.s2 = uni(S.q.key[i].div.s2);
                const uint32_t stride = uni((uint32_t)S.q.key[i].stride);
                fetch(a, S, lds, c, r, pass, raw, dv);
#pragma unroll
                for (int j = 0; j < NR; ++j) {
                    const uint64_t u = (uint64_t)typed(raw[j], ty) - base;
                    if (((pass >> j) & 1u) && u >= span) {
                        raise_err(a.err, DEVERR_KEYRANGE);
                        pass &= ~(1u << j);
                    }
                    g[j] += golap_fmt::udiv32((uint32_t)u, dvz) * stride;
                }
            }
            if (vc >= 0) {
                const uint32_t ty = uni(S.cols[vc].type);
                fetch(a, S, lds, (uint32_t)vc, r, pass, raw, dv);
#pragma unroll
                for (int j = 0; j < NR; ++j) raw[j] = (uint64_t)typed(raw[j], ty);
            } else {
#pragma unroll
                for (int j = 0; j < NR; ++j) raw[j] = 0;
            }
#pragma unroll
            for (int j = 0; j < NR; ++j) {
                const bool act = (pass >> j) & 1u;
                const uint64_t am = __ballot(act);
                if (!am) continue;
                const uint32_t leader = (uint32_t)__ffsll((long long)am) - 1;
                const uint32_t g0 = __shfl(g[j], leader, WAVE);
                if (__ballot(act && g[j] != g0) == 0) {
                    // every passing row of the wave is in one group (sorted keys): one update
                    int64_t sum = act ? (int64_t)raw[j] : 0;
#pragma unroll
                    for (int o = WAVE / 2; o > 0; o >>= 1) sum += __shfl_xor(sum, o, WAVE);
                    if (lane == leader) {
                        const uint32_t cnt = (uint32_t)__popcll(am);
                        if (!use_lds || !group_add_counted(T, qkey | g0, sum, cnt, &S.fill)) {
                            atomicAdd(reinterpret_cast<unsigned long long *>(&g_sum[g0]), (unsigned long long)sum);
                            atomicAdd(reinterpret_cast<unsigned long long *>(&g_cnt[g0]), (unsigned long long)cnt);
                        }
                    }
                } else if (act) {
                    if (!use_lds || !group_add_counted(T, qkey | g[j], (int64_t)raw[j], 1u, &S.fill))
                        group_add_global(g_sum, g_cnt, g[j], (int64_t)raw[j]);
                }
            }
        }
    }
    __syncthreads();
}

// Persistent scan kernel. Static schedule: the queries' pages flattened in query order, workgroup
// b takes [b*n/G, (b+1)*n/G). Queue schedule: device-built (chunk, query, part) items taken with
// one atomic each.
template <int PER, bool QUEUE>
__global__ __launch_bounds__(BLOCK) void k_scan(ScanArgs a) {
    extern __shared__ __align__(16) uint8_t lds[];
    __shared__ ScanShared S;
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
#pragma unroll
    for (uint32_t i = 0; i < SCAN_MAX_COLS; ++i)
        if (threadIdx.x == i) S.cols[i] = a.col[i];
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
            scan_tile<PER>(a, S, q, p, lds, T);
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
            for (uint64_t p = pb; p < pe; ++p) scan_tile<PER>(a, S, q, p, lds, T);
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
    const uint32_t mb = l.max_blocks_per_cu;
#define GOLAP_SCAN_CASE(PER)                                                                        \
    return l.items ? launch_scan_t(ctx, k_scan<PER, true>, a, lds, stream, mb, max_grid, prepare_only) \
                   : launch_scan_t(ctx, k_scan<PER, false>, a, lds, stream, mb, max_grid, prepare_only);
    // rows per thread: page_rows / BLOCK rounded up to a multiple of NR
    if (l.page_rows <= 8 * BLOCK) { GOLAP_SCAN_CASE(8) }
    if (l.page_rows <= 16 * BLOCK) { GOLAP_SCAN_CASE(16) }
    GOLAP_SCAN_CASE(32)
#undef GOLAP_SCAN_CASE
}

}  // namespace golap_impl
