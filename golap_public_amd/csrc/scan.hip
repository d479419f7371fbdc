// Fused shared scan + filter + group-by over paged typed columns (gfx950): the trip-table plans.
//
// The reference groups with group_by_agg_2col / 4col over decompressed attribute columns and
// HashAggregate::add (core/group.cuh:14-36, core/dev_structs.cuh:209-244) after its range
// predicates (usecases/ssb/queries/01.cuh:26-54), one plan per launch. Here one persistent
// kernel serves any number of plans over one table:
//   - a work item is a chunk part with a 64-bit mask of the plans (of one word of 64) that
//     need the chunk; the workgroup reads the item's pages once for all of them;
//   - per (plan, item), the zone maps of the chunk give every key its local range (a sorted
//     timestamp's hour spans one or two values in a chunk, a zone id all 265) and tell which
//     predicates hold for the whole chunk (skipped) and which keys are constant (hoisted);
//   - SUM/COUNT go into a dense LDS array over the local key ranges (no hashing, no probing),
//     flushed to the plan's dense global result once per item; a plan whose local domain does
//     not fit, or that asks for global aggregation, adds straight to global int64 atomics;
//     when all passing rows of a wave share one group, the wave adds their sum in one update;
//   - plans that evaluate the chunk's rows identically (same tested predicates, same local key
//     functions with every key in range, same value column: typically one request template
//     over different time ranges) share one set of partials, computed once and flushed into
//     each plan's result with its own group mapping;
//   - plans whose partials do not fit the LDS array together run in several passes over the
//     item (the pages are then re-read from L2 / MALL);
//   - the waves of a workgroup take the item's pages round robin and work on a page alone:
//     column plans, window decodes and delta-bitpack prefix sums are per wave, so no workgroup
//     barrier stands between pages (only the item's set-up and flush synchronise the block);
//   - a page is read in place (raw and literal-only LZ4 pages), by point access (frame-of-
//     reference and RLE pages, delta pages with bw 0), by running sums in registers (delta
//     pages; start values from a wave prefix scan of per-lane delta totals), or from an LDS
//     window the wave decodes it into (other coded pages); lane l owns HH consecutive rows of
//     every sub-tile of WAVE * HH rows, and every column use reads all HH of them at once.
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
    uint32_t win_bytes, agg_entries;  // windows of one wave; dense partial entries
    uint32_t page_waves;              // waves that take pages (each with its own windows)
    const ScanQueryDev *queries;
    uint32_t n_queries;
    const uint4 *items;
    uint32_t *queue;
    uint32_t pages_per_item;
    uint32_t *err;
    int32_t tok_off;   // LZ4I token offsets + sink in the wave's region, -1 = none
    uint32_t tok_cap;
};

// How one typed column is read in the current page. off: byte offset of the data in the
// column buffer (windows: in the wave's LDS windows).
enum : uint32_t { TA_DIRECT = 0, TA_WINDOW = 1, TA_AFFINE = 2, TA_RLE = 3, TA_FOR = 4, TA_DBP = 5, TA_BAD = 6 };
struct TPlan {
    uint64_t off;  // DIRECT: payload; FOR: fields; RLE, DBP: page start; WINDOW: LDS offset
    int64_t a, b;  // AFFINE: v[r] = a + r*b; FOR: a = frame of reference; RLE: b = value bytes (padded);
                   // DBP: a = v[0], b = minimum delta
    uint32_t mode;
    uint32_t aux;  // RLE: runs; FOR, DBP: bits per field; DIRECT: payload 16-byte aligned
};

constexpr uint32_t PASS_MAX = 16;  // plans evaluated together (their item views in LDS)
constexpr uint32_t NO_LDS = 0xffffffffu;
constexpr uint32_t NWAVE = BLOCK / WAVE;
// HH (template parameter): rows of a lane per sub-tile of WAVE * HH rows; 8 for short launches
// (fewer sub-tiles per page), 4 for long ones (fewer registers: four workgroups per CU).

// A plan's view of the current work item. Predicates that hold for the whole chunk are left
// out; keys constant over the chunk contribute cg to the global group index; the others are
// listed first with their local range [klo, klo + lcard) and local stride.
struct ItemQuery {
    int64_t plo[SCAN_MAX_PREDS], phi[SCAN_MAX_PREDS];
    int64_t kbase[SCAN_MAX_KEYS];
    int64_t *g_sum;
    uint64_t *g_cnt;
    uint32_t kspan[SCAN_MAX_KEYS];
    uint32_t km[SCAN_MAX_KEYS], ks[SCAN_MAX_KEYS];  // division magic: m; s1 | s2 << 8
    uint32_t kgs[SCAN_MAX_KEYS];                     // global stride
    uint32_t klo[SCAN_MAX_KEYS], lcard[SCAN_MAX_KEYS], lstride[SCAN_MAX_KEYS];
    uint32_t pcol, kcol;  // column of predicate / key i in bits [8i, 8i + 8)
    uint32_t np, nk;      // predicates tested and keys computed per row
    int32_t vcol;         // -1: COUNT only
    uint32_t cg;
    uint32_t lbase, L;    // dense partials at [lbase, lbase + L); lbase NO_LDS: global atomics
    uint32_t colmask;     // columns read per row
    // Sharing: plans whose rows evaluate identically on the chunk (same tested predicates, same
    // local key functions, every key inside its range, same value column) add into one set of
    // partials, computed by the first of them (leader) and flushed into every one's result.
    int64_t korg[SCAN_MAX_KEYS];  // value where local key index 0 starts (kbase + klo * div)
    uint32_t sig;                 // hash of the evaluation signature
    uint32_t shareable;
    uint32_t leader;              // pass index of the plan whose partials this plan uses
    uint32_t members;             // leaders: other plans use these partials
};

struct ScanShared {
    ScanColView cols[SCAN_MAX_COLS];  // the launch's columns: indexed at run time from LDS, not kernarg
    TPlan tp[NWAVE][SCAN_MAX_COLS];   // each wave's plans of its current page
    ItemQuery iq[PASS_MAX];
    uint32_t qidx[WAVE], qlen[WAVE];  // the item's live plans and their local domains (0: global)
    uint32_t nlive, pass_end, pass_used, item;
};

__device__ __forceinline__ uint32_t uni(uint32_t x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ uint64_t uni64(uint64_t x) {
    return ((uint64_t)uni((uint32_t)(x >> 32)) << 32) | uni((uint32_t)x);
}

// Address-space views: column buffers are global memory, windows are LDS. Loads at a uniform
// base plus a 32-bit byte offset compile to the scalar-base form of global loads and to plain
// LDS reads.
typedef const __attribute__((address_space(1))) uint8_t gbyte;
typedef const __attribute__((address_space(3))) uint8_t lbyte;
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ gbyte *gmem(const uint8_t *p) { return (gbyte *)p; }

template <typename T>
__device__ __forceinline__ T ldt(gbyte *base, uint32_t off) {
    return *reinterpret_cast<const __attribute__((address_space(1))) T *>(base + off);
}
template <typename T>
__device__ __forceinline__ T ldt(lbyte *base, uint32_t off) {
    return *reinterpret_cast<const __attribute__((address_space(3))) T *>(base + off);
}
template <typename B>
__device__ __forceinline__ uint64_t ld_elem(B *base, uint32_t off, uint32_t elem) {
    switch (elem) {
    case 1: return ldt<uint8_t>(base, off);
    case 2: return ldt<uint16_t>(base, off);
    case 4: return ldt<uint32_t>(base, off);
    default: return ldt<uint64_t>(base, off);
    }
}

// Shift that sign-extends a value of `type` held in its low bytes (0: none needed).
__device__ __forceinline__ uint32_t sext_shift(uint32_t type, uint32_t elem) {
    const bool sgn = type == GOLAP_I8 || type == GOLAP_I16 || type == GOLAP_I32;
    return sgn ? 64u - 8u * elem : 0u;
}
__device__ __forceinline__ int64_t sext(uint64_t raw, uint32_t sh) { return sh ? ((int64_t)(raw << sh) >> sh) : (int64_t)raw; }

// Access plan of column v for a page whose bytes start at src (global or LDS, generic pointer)
// at offset base_off from the column's base, with the catalog's codec and size.
__device__ inline TPlan plan_page(const ScanColView &v, const uint8_t *src, uint64_t base_off, uint8_t codec, uint32_t csize,
                                  uint32_t rows) {
    TPlan x{};
    const uint32_t elem = v.c.elem;
    const uint32_t cd = codec & PAGE_CODEC_MASK;
    if (is_direct(codec)) {
        const uint32_t hdr = direct_hdr(codec, rows * elem);
        x.off = base_off + hdr;
        x.mode = (csize == hdr + rows * elem && !(reinterpret_cast<uintptr_t>(src + hdr) % elem)) ? TA_DIRECT : TA_BAD;
        x.aux = !(reinterpret_cast<uintptr_t>(src + hdr) % 16);  // vector loads of a thread's rows
    } else if (cd == golap_fmt::DBP && (codec & golap_fmt::PAGE_POINT)) {
        if (csize < 32 || src[20] != 0 || src[21] != elem || ld_u32_unaligned(src + 16) != rows) {
            x.mode = TA_BAD;
        } else {
            x.a = (int64_t)ld_u64_unaligned(src);
            x.b = (int64_t)ld_u64_unaligned(src + 8);
            x.mode = TA_AFFINE;
        }
    } else if (cd == golap_fmt::FOR) {
        // fields read as aligned 4- or 8-byte words: the payload must be 8-byte aligned
        if (for_check(src, csize, elem) != (int64_t)rows || (reinterpret_cast<uintptr_t>(src) & 7)) {
            x.mode = TA_BAD;
        } else {
            x.off = base_off + 32;
            x.a = (int64_t)ld_u64_unaligned(src);
            x.aux = src[20];
            x.mode = TA_FOR;
        }
    } else if (cd == golap_fmt::RLE) {
        const uint32_t nruns = csize >= 16 ? ld_u32_unaligned(src) : 0u;
        const uint64_t vbytes = ((uint64_t)nruns * elem + 3) & ~3ull;
        const bool ok = csize >= 16 && src[8] == elem && ld_u32_unaligned(src + 4) == rows && nruns >= 1 &&
                        16ull + vbytes + 4ull * nruns <= csize;
        x.off = base_off;
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
            x.off = base_off;
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

// Register-decoded delta-bitpack column of the current page (uniform); B: its address space.
template <typename B>
struct DbpReg {
    B *pk;  // packed fields
    uint64_t mind;
    uint32_t bw;
};
// The delta fields of a lane's HH rows r0 + j of a sub-tile, decoded once per sub-tile and
// shared by every use of the column: p[j] = packed field of row r0 + j (the delta v[r] - v[r-1]
// is mind + p[j]); valid bit j: the row has a delta (1 <= r0 + j <= rmax).
template <int HH>
struct DbpRows {
    uint32_t p[SCAN_DBP_SLOTS][HH];
    uint32_t valid;
};
// Field extraction at bit offset `bit` (fields of bw <= 32 bits; the column buffer extends at
// least 8 bytes past any page, so the second dword is always readable).
template <typename B>
__device__ __forceinline__ uint32_t packed_field(B *pk, uint32_t bit, uint32_t mask) {
    const u32x2 w = ldt<u32x2>(pk, (bit >> 5) << 2);
    return __builtin_amdgcn_alignbit(w.y, w.x, bit & 31u) & mask;
}

// Values (as int64) of column c (wave-uniform) at the thread's rows r0 + j, j < HH; a row past
// rmax (the page's last row) reads row rmax instead (its value is never used). base: the data
// base the plans' offsets refer to; windows are in the wave's LDS windows (lds). Frame-of-reference, delta and
// affine pages give the exact int64 values; in-place, window and run-length reads are sign- or
// zero-extended from their element width. drs: the delta slots' values at row r0 - 1 (at row
// 0: the page's first value).
template <int HH, typename B>
__device__ __forceinline__ void fetch_col(const TPlan *tp, const ScanColView *cols, B *base, lbyte *lds, uint32_t c,
                                          uint32_t r0, uint32_t rmax, int64_t (&out)[HH],
                                          const DbpReg<B> (&dr)[SCAN_DBP_SLOTS], const DbpRows<HH> &dp,
                                          const uint64_t (&drs)[SCAN_DBP_SLOTS]) {
    const uint32_t mode = uni(tp[c].mode);
    if (mode == TA_DBP) {  // running sums of the sub-tile's decoded deltas
        const uint32_t slot = uni((uint32_t)cols[c].dslot);
        uint64_t mind = dr[0].mind, run = drs[0];
#pragma unroll
        for (uint32_t t = 1; t < SCAN_DBP_SLOTS; ++t)
            if (slot == t) {
                mind = dr[t].mind;
                run = drs[t];
            }
#pragma unroll
        for (int j = 0; j < HH; ++j) {
            uint32_t p = dp.p[0][j];
#pragma unroll
            for (uint32_t t = 1; t < SCAN_DBP_SLOTS; ++t)
                if (slot == t) p = dp.p[t][j];
            if ((dp.valid >> j) & 1u) run += mind + p;
            out[j] = (int64_t)run;
        }
        return;
    }
    const uint32_t elem = uni(cols[c].c.elem);
    const uint32_t sh = sext_shift(uni(cols[c].type), elem);
    const uint32_t off = (uint32_t)uni64(tp[c].off);
    B *b = base + off;
    switch (mode) {
    case TA_DIRECT: {
        if (r0 + HH - 1 <= rmax && uni(tp[c].aux)) {
            // HH consecutive rows of a 16-byte aligned payload (r0 % HH == 0): HH * elem bytes
            // in dwords w, loaded as 16-byte vectors where they fill one
            static_assert(HH == 4 || HH == 8, "row vectors of 4 or 8 rows");
            const uint32_t vo = r0 * elem;
            uint32_t w[2 * HH];
            auto load = [&](int nw) {
                if (nw == 1) {
                    w[0] = ldt<uint32_t>(b, vo);
                } else if (nw == 2) {
                    const u32x2 v = ldt<u32x2>(b, vo);
                    w[0] = v.x;
                    w[1] = v.y;
                } else {
#pragma unroll
                    for (int q = 0; q < nw / 4; ++q) {
                        const u32x4 v = ldt<u32x4>(b, vo + 16 * q);
                        w[4 * q] = v.x;
                        w[4 * q + 1] = v.y;
                        w[4 * q + 2] = v.z;
                        w[4 * q + 3] = v.w;
                    }
                }
            };
            if (elem == 1) {
                load(HH / 4);
#pragma unroll
                for (int j = 0; j < HH; ++j) out[j] = sext((w[j >> 2] >> (8 * (j & 3))) & 0xffu, sh);
            } else if (elem == 2) {
                load(HH / 2);
#pragma unroll
                for (int j = 0; j < HH; ++j) out[j] = sext((w[j >> 1] >> (16 * (j & 1))) & 0xffffu, sh);
            } else if (elem == 4) {
                load(HH);
#pragma unroll
                for (int j = 0; j < HH; ++j) out[j] = sext(w[j], sh);
            } else {
                load(2 * HH);
#pragma unroll
                for (int j = 0; j < HH; ++j) out[j] = (int64_t)(((uint64_t)w[2 * j + 1] << 32) | w[2 * j]);
            }
        } else {
#pragma unroll
            for (int j = 0; j < HH; ++j) out[j] = sext(ld_elem(b, min(r0 + j, rmax) * elem, elem), sh);
        }
        break;
    }
    case TA_WINDOW: {
        lbyte *w = lds + off;
#pragma unroll
        for (int j = 0; j < HH; ++j) out[j] = sext(ld_elem(w, min(r0 + j, rmax) * elem, elem), sh);
        break;
    }
    case TA_FOR: {
        const uint32_t bw = uni(tp[c].aux);
        const int64_t ref = (int64_t)uni64((uint64_t)tp[c].a);
        if (bw <= 32) {
            // field r at bits [r*bw, r*bw + bw): two dwords and a funnel shift (the payload is
            // followed by 8 zero bytes, so the second dword is always inside the page)
            const uint32_t mask = bw == 32 ? 0xffffffffu : ((1u << bw) - 1u);
            const uint32_t b0 = __umul24(r0, bw), bmax = rmax * bw;  // rows < 2^24
#pragma unroll
            for (int j = 0; j < HH; ++j) out[j] = ref + (int64_t)packed_field(b, min(b0 + j * bw, bmax), mask);
        } else {
            const uint64_t mask = ~0ull >> (64 - bw);
#pragma unroll
            for (int j = 0; j < HH; ++j) {
                const uint32_t bit = min(r0 + j, rmax) * bw;  // < 2^32: rows <= 8192, bw <= 56
                const uint32_t s = bit & 63;
                uint64_t x = ldt<uint64_t>(b, (bit >> 6) << 3) >> s;
                if (s + bw > 64) x |= ldt<uint64_t>(b, ((bit >> 6) << 3) + 8) << (64 - s);  // the field ends in the next word
                out[j] = ref + (int64_t)(x & mask);
            }
        }
        break;
    }
    case TA_AFFINE: {
        const uint64_t a0 = uni64((uint64_t)tp[c].a), b0 = uni64((uint64_t)tp[c].b);
#pragma unroll
        for (int j = 0; j < HH; ++j) out[j] = (int64_t)(a0 + (uint64_t)(r0 + j) * b0);
        break;
    }
    default: {  // RLE: first run whose exclusive end is above the row
        const uint32_t nruns = uni(tp[c].aux);
        const uint32_t vbytes = (uint32_t)uni64((uint64_t)tp[c].b);
#pragma unroll
        for (int j = 0; j < HH; ++j) {
            const uint32_t r = min(r0 + j, rmax);
            uint32_t lo = 0, hi = nruns - 1;
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (ldt<uint32_t>(b, 16 + vbytes + 4 * mid) > r) hi = mid;
                else lo = mid + 1;
            }
            uint64_t v = 0;
            for (uint32_t k = 0; k < elem; ++k) v |= (uint64_t)ldt<uint8_t>(b, 16 + lo * elem + k) << (8 * k);
            out[j] = sext(v, sh);
        }
        break;
    }
    }
}

// Predicate P on chunk c: 0 fails for every row, 1 holds for every row, 2 must be tested.
__device__ __forceinline__ uint32_t pred_state(const ScanPredDev &P, uint64_t c) {
    const int64_t zmin = P.zmin[c], zmax = P.zmax[c];
    if (zmax < P.lo || zmin > P.hi) return 0;
    return (zmin >= P.lo && zmax <= P.hi) ? 1 : 2;
}

// Key K on chunk c: its values' key indexes lie in [klo, khi]; inside: every row's key is in
// range; false when no row's key is (a passing row is then an error, found per row).
__device__ __forceinline__ bool key_range(const ScanKeyDev &K, uint64_t c, uint32_t &klo, uint32_t &khi, bool &inside) {
    const int64_t zmin = K.zmin[c], zmax = K.zmax[c];
    // offsets of the chunk's extreme values from base, clamped at 0
    const uint64_t ulo = zmin <= K.base ? 0 : (uint64_t)zmin - (uint64_t)K.base;
    const uint64_t uhi = zmax < K.base ? 0 : (uint64_t)zmax - (uint64_t)K.base;
    inside = zmin >= K.base && uhi < K.span;
    klo = khi = 0;
    if (zmax < K.base || ulo >= K.span) return false;
    klo = golap_fmt::udiv32((uint32_t)ulo, K.div);
    khi = golap_fmt::udiv32((uint32_t)min<uint64_t>(uhi, K.span - 1), K.div);
    return true;
}

// Plan Q on chunk c: false when a predicate fails for the whole chunk; *L = the local domain
// of the keys that are not constant over the chunk, 0 when the plan aggregates in global
// memory (it asks for that, a key's chunk range leaves its span, or the domain exceeds
// agg_entries).
__device__ bool item_domain(const ScanQueryDev &Q, uint64_t c, uint32_t agg_entries, uint32_t *L) {
    for (uint32_t i = 0; i < Q.n_preds; ++i)
        if (!pred_state(Q.pred[i], c)) return false;
    uint64_t n = 1;
    bool local = Q.use_lds != 0;
    for (uint32_t i = 0; i < Q.n_keys; ++i) {
        uint32_t klo, khi;
        bool inside;
        local &= key_range(Q.key[i], c, klo, khi, inside);
        if (!(inside && klo == khi)) n *= khi - klo + 1;
    }
    *L = (local && n <= agg_entries) ? (uint32_t)n : 0u;
    return true;
}

// Plan Q's view of chunk c (written by one thread; lbase is set by the caller). Predicates
// that hold for the whole chunk are left out; keys constant over the chunk go into cg.
__device__ void item_query(const ScanQueryDev &Q, uint64_t c, uint32_t agg_entries, ItemQuery &o) {
    o.g_sum = Q.g_sum;
    o.g_cnt = Q.g_cnt;
    o.vcol = Q.val_col;
    o.colmask = Q.val_col >= 0 ? 1u << Q.val_col : 0u;
    uint32_t np = 0, pcol = 0;
    for (uint32_t i = 0; i < Q.n_preds; ++i) {
        const ScanPredDev &P = Q.pred[i];
        if (pred_state(P, c) != 2) continue;
        o.plo[np] = P.lo;
        o.phi[np] = P.hi;
        pcol |= P.col << (8 * np);
        o.colmask |= 1u << P.col;
        ++np;
    }
    o.np = np;
    o.pcol = pcol;
    uint32_t nk = 0, kcol = 0, cg = 0;
    bool shareable = true;
    for (uint32_t i = 0; i < Q.n_keys; ++i) {
        const ScanKeyDev &K = Q.key[i];
        uint32_t klo, khi;
        bool inside;
        (void)key_range(K, c, klo, khi, inside);
        if (inside && klo == khi) {  // constant over the chunk
            cg += klo * (uint32_t)K.stride;
            continue;
        }
        o.kbase[nk] = K.base;
        o.korg[nk] = K.base + (int64_t)klo * (int64_t)(K.span / K.card);
        shareable &= inside;
        o.kspan[nk] = (uint32_t)K.span;
        o.km[nk] = K.div.m;
        o.ks[nk] = K.div.s1 | (K.div.s2 << 8);
        o.kgs[nk] = (uint32_t)K.stride;
        o.klo[nk] = klo;
        o.lcard[nk] = khi - klo + 1;
        kcol |= K.col << (8 * nk);
        o.colmask |= 1u << K.col;
        ++nk;
    }
    o.nk = nk;
    o.kcol = kcol;
    o.cg = cg;
    uint32_t st = 1;
    for (int n = (int)nk - 1; n >= 0; --n) {
        o.lstride[n] = st;
        st *= o.lcard[n];
    }
    uint32_t L = 0;
    (void)item_domain(Q, c, agg_entries, &L);
    o.L = L;
    o.shareable = shareable && L;
    // FNV-1a over the signature fields
    uint32_t hsh = 2166136261u;
    auto mix = [&](uint64_t v) {
        for (int b = 0; b < 8; ++b) hsh = (hsh ^ (uint32_t)((v >> (8 * b)) & 0xffu)) * 16777619u;
    };
    mix(np | ((uint64_t)nk << 8) | ((uint64_t)(uint32_t)o.vcol << 16) | ((uint64_t)L << 32));
    mix(pcol | ((uint64_t)kcol << 32));
    for (uint32_t i = 0; i < np; ++i) {
        mix((uint64_t)o.plo[i]);
        mix((uint64_t)o.phi[i]);
    }
    for (uint32_t i = 0; i < nk; ++i) {
        mix((uint64_t)o.korg[i]);
        mix(o.km[i] | ((uint64_t)o.ks[i] << 32));
        mix(o.lcard[i]);
    }
    o.sig = hsh;
}

// Whether plans a and b (both shareable) evaluate every row of the chunk identically.
__device__ bool same_rows(const ItemQuery &x, const ItemQuery &y) {
    if (x.sig != y.sig || x.np != y.np || x.nk != y.nk || x.pcol != y.pcol || x.kcol != y.kcol || x.vcol != y.vcol ||
        x.L != y.L)
        return false;
    for (uint32_t i = 0; i < x.np; ++i)
        if (x.plo[i] != y.plo[i] || x.phi[i] != y.phi[i]) return false;
    for (uint32_t i = 0; i < x.nk; ++i)
        if (x.korg[i] != y.korg[i] || x.km[i] != y.km[i] || x.ks[i] != y.ks[i] || x.lcard[i] != y.lcard[i])
            return false;
    return true;
}

// Adds the sum and count of rows of one group to dense partial e, or to the global result g.
__device__ __forceinline__ void agg_lds(unsigned long long *lsum, uint32_t *lcnt, uint32_t e, int64_t v, uint32_t n,
                                        bool sum) {
    if (sum) atomicAdd(&lsum[e], (unsigned long long)v);
    atomicAdd(&lcnt[e], n);
}
__device__ __forceinline__ void agg_global(const ItemQuery &Q, uint32_t g, int64_t v, uint32_t n, bool sum) {
    if (sum) atomicAdd(reinterpret_cast<unsigned long long *>(&Q.g_sum[g]), (unsigned long long)v);
    atomicAdd(reinterpret_cast<unsigned long long *>(&Q.g_cnt[g]), (unsigned long long)n);
}

// Key i of plan Q at the lane's HH rows with values x: kk[j] = (x[j] - base) / div; returns
// the rows whose value lies outside [base, base + span) (their kk is meaningless).
template <int HH>
__device__ __forceinline__ uint32_t key_index(const ItemQuery &Q, uint32_t i, const int64_t (&x)[HH], uint32_t (&kk)[HH]) {
    const uint64_t kb = uni64((uint64_t)Q.kbase[i]);
    const uint32_t span = uni(Q.kspan[i]), ks = uni(Q.ks[i]);
    golap_fmt::UDiv32 dv;
    dv.m = uni(Q.km[i]);
    dv.s1 = ks & 0xffu;
    dv.s2 = ks >> 8;
    uint32_t bad = 0;
    if (dv.m == 1 && ks == 0) {  // divisor 1
#pragma unroll
        for (int j = 0; j < HH; ++j) {
            const uint64_t uo = (uint64_t)x[j] - kb;
            bad |= (uint32_t)(uo >= span) << j;
            kk[j] = (uint32_t)uo;
        }
    } else {
#pragma unroll
        for (int j = 0; j < HH; ++j) {
            const uint64_t uo = (uint64_t)x[j] - kb;
            bad |= (uint32_t)(uo >= span) << j;
            kk[j] = golap_fmt::udiv32((uint32_t)uo, dv);
        }
    }
    return bad;
}

// Page p of the item for the plans of the pass, processed by one wave alone: column plans
// (wave-private), window decodes into the wave's windows (wlds), delta-bitpack start values
// from wave prefix scans, then plan by plan and column use by column use over the lane's HH
// rows of every sub-tile: the plan's and the column's parameters stay uniform across them,
// every use issues HH independent loads, and rows are carried between uses as bit masks (pass)
// and partial group indexes (g global, e local).
template <int HH>
__device__ void wave_page(const ScanArgs &a, ScanShared &S, uint64_t p, uint32_t used, uint32_t npass, uint8_t *wlds,
                          unsigned long long *lsum, uint32_t *lcnt) {
    const uint32_t wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    const uint32_t rows = (uint32_t)min<uint64_t>(a.page_rows, a.n_rows - p * a.page_rows);
    TPlan *tp = S.tp[wave];
    if (lane < a.n_cols && ((used >> lane) & 1u)) {
        const ScanColView &v = S.cols[lane];
        if (v.decoded) {  // decoded ahead of the launch: read in place
            TPlan x{};
            x.off = (p * a.page_rows - v.dec_row0) * v.c.elem;
            x.mode = TA_DIRECT;
            x.aux = !((reinterpret_cast<uintptr_t>(v.c.comp) + x.off) % 16);
            tp[lane] = x;
        } else {
            const uint64_t off = v.c.page_off[p];
            tp[lane] = plan_page(v, v.c.comp + off, off, v.c.page_codec[p], v.c.page_csize[p], rows);
        }
    }
    // the wave's own LDS writes are visible to its later LDS reads (in order within a wave)
    bool bad = false;
    for (uint32_t c = 0; c < a.n_cols; ++c) bad |= ((used >> c) & 1u) && tp[c].mode == TA_BAD;
    if (bad) {  // wave-uniform
        if (lane == 0) raise_err(a.err, DEVERR_CORRUPT);
        return;
    }
    for (uint32_t c = 0; c < a.n_cols; ++c) {
        if (!((used >> c) & 1u) || tp[c].mode != TA_WINDOW) continue;
        const ColView &cv = S.cols[c].c;
        const uint32_t want = rows * cv.elem;
        // LZ4I pages of a column with a staging window decode in parallel (token offsets and
        // sink behind the windows), others serially
        uint16_t *tok = a.tok_off >= 0 ? reinterpret_cast<uint16_t *>(wlds + a.tok_off) : nullptr;
        const int64_t got = page_decode_wave(cv.page_codec[p], cv.comp + cv.page_off[p], cv.page_csize[p],
                                             wlds + tp[c].off, want, cv.elem, S.cols[c].wcap, tok, a.tok_cap);
        if (got != (int64_t)want) {  // wave-uniform
            if (lane == 0) raise_err(a.err, DEVERR_CORRUPT);
            return;
        }
    }
    // register-decoded delta columns of the page
    DbpReg<gbyte> dr[SCAN_DBP_SLOTS];
    uint64_t carry[SCAN_DBP_SLOTS];  // value before the sub-tile's first row (uniform)
    uint32_t dmask = 0;
#pragma unroll
    for (uint32_t sl = 0; sl < SCAN_DBP_SLOTS; ++sl) {
        dr[sl] = DbpReg<gbyte>{nullptr, 0, 0};
        carry[sl] = 0;
    }
    for (uint32_t c = 0; c < a.n_cols; ++c) {
        if (!((used >> c) & 1u) || uni(tp[c].mode) != TA_DBP) continue;
        const uint32_t sl = uni((uint32_t)S.cols[c].dslot);
        DbpReg<gbyte> d;
        d.pk = gmem(reinterpret_cast<const uint8_t *>(uni64(reinterpret_cast<uint64_t>(S.cols[c].c.comp)))) +
               uni64(tp[c].off) + 32;
        d.bw = uni(tp[c].aux);
        d.mind = uni64((uint64_t)tp[c].b);
#pragma unroll
        for (uint32_t t = 0; t < SCAN_DBP_SLOTS; ++t)
            if (sl == t) {
                dr[t] = d;
                carry[t] = uni64((uint64_t)tp[c].a);
            }
        dmask |= 1u << sl;
    }
    auto col_base = [&](uint32_t c) -> gbyte * {
        return gmem(reinterpret_cast<const uint8_t *>(uni64(reinterpret_cast<uint64_t>(S.cols[c].c.comp))));
    };
    const uint32_t rmax = rows - 1;
    for (uint32_t sub = 0; sub * (WAVE * HH) < rows; ++sub) {
        const uint32_t row0 = sub * (WAVE * HH) + lane * HH;
        // delta fields of the lane's rows, once per sub-tile; start values by a wave prefix scan
        DbpRows<HH> dp;
        dp.valid = 0;
#pragma unroll
        for (int j = 0; j < HH; ++j) dp.valid |= (uint32_t)(row0 + j >= 1 && row0 + j <= rmax) << j;
        uint64_t run[SCAN_DBP_SLOTS];
#pragma unroll
        for (uint32_t sl = 0; sl < SCAN_DBP_SLOTS; ++sl) {
            run[sl] = carry[sl];
#pragma unroll
            for (int j = 0; j < HH; ++j) dp.p[sl][j] = 0;
            if (!((dmask >> sl) & 1u)) continue;  // uniform
            const uint32_t bw = dr[sl].bw;
            uint64_t tot = 0;
            if (bw) {
                const uint32_t mask = bw == 32 ? 0xffffffffu : ((1u << bw) - 1u);
                // field of row r >= 1 at bit (r - 1) * bw
                const uint32_t b0 = __umul24(row0, bw) - bw;
#pragma unroll
                for (int j = 0; j < HH; ++j)
                    if ((dp.valid >> j) & 1u) {
                        dp.p[sl][j] = packed_field(dr[sl].pk, b0 + j * bw, mask);
                        tot += dp.p[sl][j];
                    }
            }
            tot += (uint64_t)__popc(dp.valid) * dr[sl].mind;
            const uint64_t incl = wave_incl_scan_u64(tot);
            run[sl] += incl - tot;
            carry[sl] += uni64(__shfl(incl, WAVE - 1, WAVE));
        }
        for (uint32_t k = 0; k < npass; ++k) {
            const ItemQuery &Q = S.iq[k];
            if (uni(Q.leader) != k) continue;  // adds into its leader's partials
            const uint32_t np = uni(Q.np), nk = uni(Q.nk), pcol = uni(Q.pcol), kcol = uni(Q.kcol);
            const int32_t vc = (int32_t)uni((uint32_t)Q.vcol);
            const bool with_sum = vc >= 0;
            const uint32_t nu = np + nk + 1;  // the last use is the value (0 for COUNT only)
            const uint32_t lbase = uni(Q.lbase), cg = uni(Q.cg);
            const bool glob = lbase == NO_LDS;
            const bool one = !glob && uni(Q.L) == 1u;  // every row of the item in one group
            uint32_t pass = 0;
#pragma unroll
            for (int j = 0; j < HH; ++j) pass |= (uint32_t)(row0 + j <= rmax) << j;
            uint32_t inl = pass;
            // group index of the row: global (g, from cg) for global plans, local (e, from 0)
            // for plans with dense partials
            uint32_t ge[HH];
#pragma unroll
            for (int j = 0; j < HH; ++j) ge[j] = glob ? cg : 0u;
            for (uint32_t u = 0; u < nu; ++u) {
                const bool last = u + 1 == nu;
                if (__ballot(pass != 0) == 0) break;  // wave-uniform: no row passes
                int64_t x[HH];
                if (!last || with_sum) {
                    const uint32_t c = u < np ? (pcol >> (8 * u)) & 0xffu
                                              : u < np + nk ? (kcol >> (8 * (u - np))) & 0xffu : (uint32_t)vc;
                    fetch_col<HH, gbyte>(tp, S.cols, col_base(c), (lbyte *)wlds, c, row0, rmax, x, dr, dp, run);
                } else {
#pragma unroll
                    for (int j = 0; j < HH; ++j) x[j] = 0;
                }
                if (u < np) {
                    const int64_t lo = (int64_t)uni64((uint64_t)Q.plo[u]), hi = (int64_t)uni64((uint64_t)Q.phi[u]);
                    const uint64_t w = (uint64_t)hi - (uint64_t)lo;
#pragma unroll
                    for (int j = 0; j < HH; ++j)
                        if ((uint64_t)x[j] - (uint64_t)lo > w) pass &= ~(1u << j);
                } else if (!last) {
                    const uint32_t i = u - np;
                    uint32_t kk[HH];
                    const uint32_t bad_rows = key_index(Q, i, x, kk);
                    if (glob) {
                        const uint32_t gs = uni(Q.kgs[i]);
#pragma unroll
                        for (int j = 0; j < HH; ++j) ge[j] += kk[j] * gs;
                    } else {
                        // local index: lcard, lstride <= the LDS entries (< 2^24)
                        const uint32_t klo = uni(Q.klo[i]), lcard = uni(Q.lcard[i]), ls = uni(Q.lstride[i]);
#pragma unroll
                        for (int j = 0; j < HH; ++j) {
                            const uint32_t lk = kk[j] - klo;
                            inl &= ~((uint32_t)(lk >= lcard) << j);
                            ge[j] += __umul24(lk, ls);
                        }
                    }
                    if (bad_rows & pass) {  // a qualifying row's key outside its declared range
                        raise_err(a.err, DEVERR_KEYRANGE);
                        pass &= ~bad_rows;
                    }
                } else {
                    // aggregate the passing rows: x holds their values
                    inl &= pass;
                    const uint32_t outside = pass & ~inl;  // key outside the local range: global
                    if (one) {  // the wave's rows in one update
                        int64_t tsum = 0;
                        uint32_t tcnt = 0;
#pragma unroll
                        for (int j = 0; j < HH; ++j)
                            if ((inl >> j) & 1u) {
                                tsum += x[j];
                                ++tcnt;
                            }
                        if (with_sum) {
#pragma unroll
                            for (int o = WAVE / 2; o > 0; o >>= 1) tsum += __shfl_xor(tsum, o, WAVE);
                        }
#pragma unroll
                        for (int o = WAVE / 2; o > 0; o >>= 1) tcnt += __shfl_xor(tcnt, o, WAVE);
                        if (lane == 0 && tcnt) agg_lds(lsum, lcnt, lbase, tsum, tcnt, with_sum);
                    } else if (!glob) {
#pragma unroll
                        for (int j = 0; j < HH; ++j)
                            if ((inl >> j) & 1u) agg_lds(lsum, lcnt, lbase + ge[j], x[j], 1u, with_sum);
                    }
                    if (!glob) {
                        if (__ballot(outside != 0)) {
                            // zone maps bound every row's key; a plan that shares its partials
                            // cannot send rows elsewhere
                            if (uni(Q.members) && lane == 0) raise_err(a.err, DEVERR_CORRUPT);
                            // global group index of these rows: the key columns once more
#pragma unroll
                            for (int j = 0; j < HH; ++j) ge[j] = cg;
                            for (uint32_t i = 0; i < nk; ++i) {
                                const uint32_t c = (kcol >> (8 * i)) & 0xffu;
                                int64_t xk[HH];
                                fetch_col<HH, gbyte>(tp, S.cols, col_base(c), (lbyte *)wlds, c, row0, rmax, xk, dr, dp, run);
                                uint32_t kk[HH];
                                (void)key_index(Q, i, xk, kk);
                                const uint32_t gs = uni(Q.kgs[i]);
#pragma unroll
                                for (int j = 0; j < HH; ++j) ge[j] += kk[j] * gs;
                            }
#pragma unroll
                            for (int j = 0; j < HH; ++j)
                                if ((outside >> j) & 1u) agg_global(Q, ge[j], x[j], 1u, with_sum);
                        }
                    } else {
                        // global aggregation: rows of a wave that share one group (sorted keys) add once
#pragma unroll
                        for (int j = 0; j < HH; ++j) {
                            const bool act = (pass >> j) & 1u;
                            const uint64_t am = __ballot(act);
                            if (!am) continue;
                            const uint32_t leader = (uint32_t)__ffsll((long long)am) - 1;
                            const uint32_t g0 = __builtin_amdgcn_readlane(ge[j], leader);
                            if (__ballot(act && ge[j] != g0) == 0) {
                                int64_t sum = act ? x[j] : 0;
                                if (with_sum) {
#pragma unroll
                                    for (int o = WAVE / 2; o > 0; o >>= 1) sum += __shfl_xor(sum, o, WAVE);
                                }
                                if (lane == leader) agg_global(Q, g0, sum, (uint32_t)__popcll(am), with_sum);
                            } else if (act) {
                                agg_global(Q, ge[j], x[j], 1u, with_sum);
                            }
                        }
                    }
                }
            }
        }
    }
}

// Moves the pass's dense partials into their plans' global results.
__device__ __forceinline__ void pass_flush(ScanShared &S, uint32_t npass, const unsigned long long *lsum,
                                           const uint32_t *lcnt) {
    __syncthreads();
    for (uint32_t k = 0; k < npass; ++k) {
        const ItemQuery &Q = S.iq[k];
        if (Q.lbase == NO_LDS) continue;
        const bool with_sum = Q.vcol >= 0;
        for (uint32_t e = threadIdx.x; e < Q.L; e += BLOCK) {
            const uint32_t n = lcnt[Q.lbase + e];
            if (!n) continue;
            uint32_t rem = e, g = Q.cg;
            for (uint32_t i = 0; i < Q.nk; ++i) {
                const uint32_t d = rem / Q.lstride[i];
                rem -= d * Q.lstride[i];
                g += (Q.klo[i] + d) * Q.kgs[i];
            }
            agg_global(Q, g, (int64_t)lsum[Q.lbase + e], n, with_sum);
        }
    }
    __syncthreads();
}

// Persistent shared-scan kernel: workgroups take (chunk part, plan word, plan mask) items with
// one atomic each; the waves share the item's dense partials and split its pages.
template <int HH>
__global__ __launch_bounds__(BLOCK, HH == 4 ? 4 : 2) void k_scan(ScanArgs a) {
    extern __shared__ __align__(16) uint8_t lds[];
    __shared__ ScanShared S;
    const uint32_t wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    uint8_t *wlds = lds + wave * a.win_bytes;
    unsigned long long *lsum = reinterpret_cast<unsigned long long *>(lds + a.page_waves * a.win_bytes);
    uint32_t *lcnt = reinterpret_cast<uint32_t *>(lds + a.page_waves * a.win_bytes + (size_t)a.agg_entries * 8);
#pragma unroll
    for (uint32_t i = 0; i < SCAN_MAX_COLS; ++i)
        if (threadIdx.x == i) S.cols[i] = a.col[i];
    const uint32_t n_items = a.queue[1];
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) S.item = atomicAdd(&a.queue[0], 1u);
        __syncthreads();
        const uint32_t it = S.item;
        if (it >= n_items) break;
        const uint4 w = a.items[it];
        const uint64_t c = w.x;
        const uint32_t part = w.y & 0xffu, word = w.y >> 8;
        const uint64_t mask = (uint64_t)w.z | ((uint64_t)w.w << 32);
        // the item's live plans (wave 0, one lane per plan of the word)
        if (threadIdx.x < WAVE) {
            const uint32_t q = word * WAVE + lane;
            bool live = false;
            uint32_t L = 0;
            if (((mask >> lane) & 1u) && q < a.n_queries) live = item_domain(a.queries[q], c, a.agg_entries, &L);
            const uint64_t b = __ballot(live);
            if (live) {
                const uint32_t rank = (uint32_t)__popcll(lane ? (b << (64 - lane)) : 0ull);
                S.qidx[rank] = q;
                S.qlen[rank] = L;
            }
            if (lane == 0) S.nlive = (uint32_t)__popcll(b);
        }
        __syncthreads();
        const uint32_t nlive = S.nlive;
        const uint64_t c0 = c * a.pages_per_chunk;
        const uint64_t pb = c0 + (uint64_t)part * a.pages_per_item;
        const uint64_t pe = min(min(pb + a.pages_per_item, c0 + a.pages_per_chunk), a.n_pages);
        for (uint32_t k0 = 0; k0 < nlive;) {
            // a pass: the views of the next plans, up to PASS_MAX; plans that share a leader's
            // partials take no LDS, the others are admitted while their partials fit
            const uint32_t nview = min(nlive - k0, PASS_MAX);
            if (threadIdx.x < nview) item_query(a.queries[S.qidx[k0 + threadIdx.x]], c, a.agg_entries, S.iq[threadIdx.x]);
            __syncthreads();
            if (threadIdx.x == 0) {
                uint32_t k = 0, used = 0;
                for (; k < nview; ++k) {
                    ItemQuery &Q = S.iq[k];
                    Q.leader = k;
                    Q.members = 0;
                    if (Q.shareable)
                        for (uint32_t j = 0; j < k; ++j)
                            if (S.iq[j].leader == j && S.iq[j].shareable && same_rows(S.iq[j], Q)) {
                                Q.leader = j;
                                break;
                            }
                    if (Q.leader != k) {
                        Q.lbase = S.iq[Q.leader].lbase;
                        S.iq[Q.leader].members = 1;
                        continue;
                    }
                    if (Q.L && used + Q.L > a.agg_entries) break;
                    Q.lbase = Q.L ? used : NO_LDS;
                    used += Q.L;
                }
                S.pass_end = k0 + max(k, 1u);
                S.pass_used = used;
            }
            __syncthreads();
            const uint32_t k1 = S.pass_end, npass = k1 - k0, entries = S.pass_used;
            for (uint32_t e = threadIdx.x; e < entries; e += BLOCK) {
                lsum[e] = 0;
                lcnt[e] = 0;
            }
            __syncthreads();
            uint32_t used = 0;
            for (uint32_t k = 0; k < npass; ++k)
                if (S.iq[k].leader == k) used |= S.iq[k].colmask;
            if (wave < a.page_waves)
                for (uint64_t p = pb + wave; p < pe; p += a.page_waves) wave_page<HH>(a, S, p, used, npass, wlds, lsum, lcnt);
            pass_flush(S, npass, lsum, lcnt);
            k0 = k1;
        }
    }
}

// Scan work list: wave w handles (chunk, word) pairs [64w, 64w + 64); pairs with a plan that
// keeps the chunk are compacted with a wave ballot and one global atomic per wave.
__global__ __launch_bounds__(BLOCK) void k_mask_items(const uint8_t *qbase, uint32_t qstride, uint32_t nq, uint64_t c0,
                                                      uint64_t nch, uint32_t parts, uint4 *items, uint32_t *queue) {
    const uint32_t lane = threadIdx.x % WAVE;
    const uint32_t nw = (nq + WAVE - 1) / WAVE;
    const uint64_t total = nch * nw;
    const uint64_t waves = (uint64_t)gridDim.x * (BLOCK / WAVE);
    for (uint64_t wv = (uint64_t)blockIdx.x * (BLOCK / WAVE) + threadIdx.x / WAVE; wv * WAVE < total; wv += waves) {
        const uint64_t t = wv * WAVE + lane;
        uint64_t mask = 0;
        uint64_t c = 0;
        uint32_t word = 0;
        if (t < total) {
            c = c0 + t / nw;
            word = (uint32_t)(t % nw);
            const uint32_t qe = min(nq, (word + 1) * (uint32_t)WAVE);
            for (uint32_t q = word * WAVE; q < qe; ++q) {
                const SliceDev &s = *reinterpret_cast<const SliceDev *>(qbase + (size_t)q * qstride);
                if (c >= s.chunk_begin && c < s.chunk_end && slice_keeps(s, c)) mask |= 1ull << (q - word * WAVE);
            }
        }
        const uint64_t ball = __ballot(mask != 0);
        uint32_t base = 0;
        if (lane == 0 && ball) base = atomicAdd(&queue[1], (uint32_t)__popcll(ball) * parts);
        base = __shfl(base, 0, WAVE);
        if (mask) {
            const uint64_t below = lane ? (ball << (64 - lane)) : 0ull;
            const uint32_t slot = base + (uint32_t)__popcll(below) * parts;
            for (uint32_t t2 = 0; t2 < parts; ++t2)
                items[slot + t2] = make_uint4((uint32_t)c, t2 | (word << 8), (uint32_t)mask, (uint32_t)(mask >> 32));
        }
    }
}

template <typename K>
hipError_t launch_scan_t(golap_ctx *ctx, K kernel, const ScanArgs &a, size_t lds, hipStream_t stream,
                         uint32_t max_blocks_per_cu, bool prepare_only) {
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
    const uint64_t grid = std::max<uint64_t>(1, (uint64_t)ctx->cus * occ);
    if (prepare_only) return hipSuccess;
    hipLaunchKernelGGL(kernel, dim3((uint32_t)grid), dim3(BLOCK), lds, stream, a);
    return hipGetLastError();
}

}  // namespace

// The kernel's own LDS (ScanShared), the same on every device.
static size_t scan_static_lds() {
    static size_t static_lds = ~(size_t)0;
    if (static_lds == ~(size_t)0) {
        hipFuncAttributes fa{};
        static_lds = hipFuncGetAttributes(&fa, reinterpret_cast<const void *>(k_scan<4>)) == hipSuccess ? fa.sharedSizeBytes
                                                                                                      : 16384;
    }
    return static_lds;
}

size_t scan_lds_budget(golap_ctx *ctx) {
    const size_t st = scan_static_lds();
    return (size_t)ctx->lds_max > st ? (size_t)ctx->lds_max - st : 0;
}


// 8 rows per lane when the pages do not give every wave of four workgroups per CU a page (each
// wave's steps per page then bound the time) or when the LDS leaves no more workgroups per CU
// for the 4-row kernel (its fewer registers buy nothing); else 4.
uint32_t scan_rows_per_lane(golap_ctx *ctx, uint64_t pages, size_t lds) {
    if (pages < (uint64_t)ctx->cus * 16) return 8;
    int o4 = 0, o8 = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&o4, k_scan<4>, BLOCK, lds) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&o8, k_scan<8>, BLOCK, lds) != hipSuccess)
        return 8;
    return o4 > o8 ? 4u : 8u;
}

size_t scan_lds_bytes(uint32_t win_bytes, uint32_t agg_entries, uint32_t page_waves) {
    return (size_t)page_waves * win_bytes + (size_t)agg_entries * 12;
}

hipError_t launch_scan(golap_ctx *ctx, const ScanLaunch &l, hipStream_t stream, bool prepare_only) {
    if (l.n_cols > SCAN_MAX_COLS || !l.queries || !l.n_queries || !l.items || !l.queue || !l.pages_per_item ||
        l.pages_per_item > l.pages_per_chunk || l.win_bytes % 16 || l.agg_entries > SCAN_AGG_MAX || l.page_rows > 8192 ||
        !l.page_waves || l.page_waves > NWAVE || (l.rows_per_lane != 4 && l.rows_per_lane != 8))
        return hipErrorInvalidValue;
    ScanArgs a{};
    for (uint32_t c = 0; c < l.n_cols; ++c) a.col[c] = l.col[c];
    a.n_cols = l.n_cols;
    a.page_rows = l.page_rows;
    a.pages_per_chunk = l.pages_per_chunk;
    a.n_rows = l.n_rows;
    a.n_pages = l.n_pages;
    a.win_bytes = l.win_bytes;
    a.agg_entries = l.agg_entries;
    a.page_waves = l.page_waves;
    a.queries = l.queries;
    a.n_queries = l.n_queries;
    a.items = l.items;
    a.queue = l.queue;
    a.pages_per_item = l.pages_per_item;
    a.err = l.err;
    a.tok_off = l.tok_off;
    a.tok_cap = l.tok_cap;
    if (l.tok_off >= 0 && ((uint32_t)l.tok_off % 16 || (uint32_t)l.tok_off + lz4i::scratch_bytes(0, l.tok_cap) > l.win_bytes))
        return hipErrorInvalidValue;
    for (uint32_t c = 0; c < l.n_cols; ++c)
        if (l.col[c].wcap && (l.col[c].woff < 0 || l.col[c].wcap % 16 || (uint32_t)l.col[c].woff + l.col[c].wcap > l.win_bytes ||
                              l.col[c].wcap < l.page_rows * l.col[c].c.elem))
            return hipErrorInvalidValue;
    const size_t lds = scan_lds_bytes(l.win_bytes, l.agg_entries, l.page_waves);
    if (lds > scan_lds_budget(ctx)) return hipErrorInvalidConfiguration;
    return l.rows_per_lane == 8 ? launch_scan_t(ctx, k_scan<8>, a, lds, stream, l.max_blocks_per_cu, prepare_only)
                                : launch_scan_t(ctx, k_scan<4>, a, lds, stream, l.max_blocks_per_cu, prepare_only);
}

hipError_t launch_mask_items(const void *qbase, uint32_t qstride, uint32_t n_queries, uint64_t c0, uint64_t nch,
                             uint32_t parts, uint4 *items, uint32_t *queue, hipStream_t stream) {
    if (!n_queries || !parts || parts > 256 || n_queries > (1u << 24)) return hipErrorInvalidValue;
    const uint64_t total = nch * ((n_queries + WAVE - 1) / WAVE);
    if (!total) return hipSuccess;
    const uint64_t waves = (total + WAVE - 1) / WAVE;
    const uint32_t grid = (uint32_t)std::min<uint64_t>((waves + BLOCK / WAVE - 1) / (BLOCK / WAVE), 4096);
    hipLaunchKernelGGL(k_mask_items, dim3(grid), dim3(BLOCK), 0, stream, static_cast<const uint8_t *>(qbase), qstride,
                       n_queries, c0, nch, parts, items, queue);
    return hipGetLastError();
}

hipError_t launch_scan_items(const ScanQueryDev *queries, uint32_t n_queries, uint64_t c0, uint64_t nch,
                             uint32_t parts, uint4 *items, uint32_t *queue, hipStream_t stream) {
    return launch_mask_items(queries, (uint32_t)sizeof(ScanQueryDev), n_queries, c0, nch, parts, items, queue, stream);
}

}  // namespace golap_impl
