// Fused Q1 plan (pipeline_q1_lineorder, usecases/ssb/queries/01.cuh:26-54, fused with the
// per-chunk decompression the reference runs before it, core/comp.cuh:171-195) for tiles with
// coded pages, one 256-thread workgroup per tile.
//
// The tile's columns are taken one after another. An in-place page (RAW, or LZ4 made of one
// literal run) is read straight from HBM; an LZ4I page is staged in LDS and decoded by the whole
// workgroup (lz4wg::decode); any other coded page (and an LZ4I page the cooperative decode hands
// back) is decoded by the first wave in order. Thread t owns rows 1024 i + 4 t + {0..3} (i < G4),
// so every column access is one 16-byte (int32) or 4-byte (uint8) access per four rows,
// consecutive across the workgroup, and keeps their state in registers between columns: a
// candidate bit per row and the discount bytes. The date probe goes first when it is the
// membership bitmap (cheap and the most selective test); a hash probe goes after the quantity and discount
// tests, for the rows still candidates. Once no row of the tile is a candidate, the remaining
// columns are not touched.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "lz4i_wg.hpp"
#include "q1_dev.hpp"

namespace golap_impl {
namespace {

// The in-order decoders (any codec), kept out of line: the rare path, whose registers would
// otherwise add to the cooperative decode's.
__device__ __attribute__((noinline)) int64_t wg_decode_in_order(uint8_t codec, const uint8_t *src, uint32_t csize,
                                                                uint8_t *dst, uint32_t cap, uint32_t elem) {
    return page_decode_wave(codec, src, csize, dst, cap, elem);
}

// The tile's nb bytes of one column, readable at the returned address: an in-place page in HBM,
// any other page decoded into the LDS output window. nullptr: a corrupt page. Every thread of the
// workgroup calls it with the same page; the result is the same in every thread.
__device__ __forceinline__ const uint8_t *wg_rows(uint8_t *base, const lz4wg::Layout &l, const Q1Page &pg, uint32_t nb,
                                                  uint32_t elem) {
    if (page_direct(pg.code)) {
        const uint32_t hd = direct_header(pg.code, nb);
        const uint8_t *p = pg.src + hd;
        return (pg.csize == hd + nb && (reinterpret_cast<uintptr_t>(p) & 15) == 0) ? p : nullptr;
    }
    __syncthreads();  // the previous column's window and the decode's reduction words are read
    int64_t got = -2;
    const uint8_t *in_order = pg.src;
    if ((pg.code & PAGE_CODEC_MASK) == golap_fmt::LZ4I && pg.csize <= l.stage_cap && nb <= l.out_cap &&
        (reinterpret_cast<uintptr_t>(pg.src) & 15) == 0) {
        lz4wg::stage<BLOCK>(base, l, pg.src, pg.csize);
        __syncthreads();
        got = lz4wg::decode<BLOCK>(base, l, pg.csize, nb);
        in_order = base + l.stage;
    }
    if (got == -2) {
        uint32_t *res = reinterpret_cast<uint32_t *>(base + l.misc) + 1;
        if (threadIdx.x < WAVE) {
            const int64_t r =
                nb <= l.out_cap ? wg_decode_in_order(pg.code, in_order, pg.csize, base + l.out, nb, elem) : -1;
            if (threadIdx.x == 0) *res = r == (int64_t)nb ? 1u : 0u;
        }
        __syncthreads();
        got = *res ? (int64_t)nb : -1;
    }
    return got == (int64_t)nb ? base + l.out : nullptr;
}

// LDS: the decode regions (l) and three words for the workgroup OR after them. The date probe
// (bitmap or hash table) is read through the vector L1 rather than pinned in LDS: its gathers run
// on the texture path beside other workgroups' LDS work, and the LDS left is enough for four
// workgroups per CU (a 4,096-row tile of the workload needs about 37 KiB). Registers are capped
// to match (four waves per SIMD).
template <int PROBE, int G4>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(4, 4))) void k_q1_wg(Q1Args a,
                                                                                           lz4wg::Layout l) {
    static_assert(4 * G4 <= 64, "one 64-bit candidate word per thread");
    extern __shared__ __align__(16) uint8_t lds[];
    constexpr bool HASH = PROBE != PROBE_BITMAP;
    uint8_t *base = lds;
    lz4wg::BlockAny any{reinterpret_cast<uint32_t *>(base + l.total), 0};
    if (threadIdx.x < 3) any.red[threadIdx.x] = 0;
    const uint32_t *tab = HASH ? reinterpret_cast<const uint32_t *>(a.slots) : a.bitmap;
    __syncthreads();
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
        const bool full = rows == (uint32_t)(1024 * G4);
        Q1Lanes<G4> st;
        st.reset();
        if (!full)
            for (int b = 0; b < 4 * G4; ++b)
                if (1024 * (b / 4) + 4 * threadIdx.x + (b % 4) >= rows) st.clear(b);
        // Columns in the order of the tests (0 date, 1 quantity, 2 discount, 3 price), one decode
        // site for all of them: the decode is large, and four inlined copies would not stay in
        // the instruction cache.
        bool bad = false;
#pragma nounroll
        for (int step = 0; step < 4; ++step) {
            const int col = HASH ? (step == 2 ? 0 : step == 3 ? 3 : step + 1) : step;
            if (step > 0 && !any(st.any())) break;  // no candidate left in the tile
            const Q1Page pg = col == 0 ? pd : col == 1 ? pq : col == 2 ? ps : pp;
            const uint32_t elem = (col == 0 || col == 3) ? 4u : 1u;
            const uint8_t *c = wg_rows(base, l, pg, rows * elem, elem);
            if (c == nullptr) {
                bad = true;
                break;
            }
            if (col == 0) {
                if (full) q1_date_pass<BLOCK, G4, true>(a, tab, c, rows, st, HASH);
                else q1_date_pass<BLOCK, G4, false>(a, tab, c, rows, st, HASH);
            } else if (col == 1) {
                if (full) q1_byte_pass<BLOCK, G4, true, false>(c, rows, a.qty_lo, a.qty_span, st);
                else q1_byte_pass<BLOCK, G4, false, false>(c, rows, a.qty_lo, a.qty_span, st);
            } else if (col == 2) {  // the discount bytes stay in registers for the sum
                if (full) q1_byte_pass<BLOCK, G4, true, true>(c, rows, a.disc_lo, a.disc_span, st);
                else q1_byte_pass<BLOCK, G4, false, true>(c, rows, a.disc_lo, a.disc_span, st);
            } else {
                if (full) q1_price_pass<BLOCK, G4, true>(a, c, rows, st, sum, cnt);
                else q1_price_pass<BLOCK, G4, false>(a, c, rows, st, sum, cnt);
            }
        }
        if (bad && threadIdx.x == 0) raise_err(a.err, DEVERR_CORRUPT);
    }
    q1_flush(a, sum, cnt);
}

template <int PROBE, int G4>
hipError_t launch_wg(golap_ctx *ctx, const Q1Args &a, const lz4wg::Layout &l, size_t lds) {
    auto kernel = k_q1_wg<PROBE, G4>;
    hipError_t e = prepare_lds(ctx, kernel, lds);
    if (e != hipSuccess) return e;
    int occ = 0;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kernel, BLOCK, lds);
    if (e != hipSuccess) return e;
    occ = std::max(occ, 1);
    const uint64_t tiles = a.page_end - a.page_begin;
    uint64_t grid = std::max<uint64_t>(1, std::min<uint64_t>(tiles, (uint64_t)ctx->cus * occ));
    if (grid >= XCDS) grid -= grid % XCDS;  // whole rounds over the XCDs (xcd_range)
    hipLaunchKernelGGL(kernel, dim3((uint32_t)grid), dim3(BLOCK), lds, ctx->stream, a, l);
    return hipGetLastError();
}

template <int PROBE>
hipError_t launch_wg_mode(golap_ctx *ctx, const Q1Args &a, const lz4wg::Layout &l, size_t lds) {
    switch (a.page_rows) {
    case 1024: return launch_wg<PROBE, 1>(ctx, a, l, lds);
    case 2048: return launch_wg<PROBE, 2>(ctx, a, l, lds);
    default: return launch_wg<PROBE, 4>(ctx, a, l, lds);
    }
}

}  // namespace

hipError_t launch_q1_wg(golap_ctx *ctx, const Q1Launch &q) {
    // page sizes whose LZ4I pages the cooperative decode takes whole (at most NT * MAXJ
    // sequences: 4096 rows of 4-byte values at four bytes per sequence, with room)
    if (q.page_rows != 1024 && q.page_rows != 2048 && q.page_rows != 4096) return hipErrorNotSupported;
    Q1Args a{};
    a.date = q.date;
    a.disc = q.disc;
    a.qty = q.qty;
    a.price = q.price;
    a.page_rows = q.page_rows;
    a.n_rows = q.n_rows;
    a.page_begin = q.page_begin;
    a.page_end = q.page_end;
    a.chunk_mask = q.chunk_mask;
    a.pages_per_chunk = q.pages_per_chunk;
    a.disc_lo = q.disc_lo;
    a.disc_span = q.disc_span;
    a.qty_lo = q.qty_lo;
    a.qty_span = q.qty_span;
    a.price_lo = q.price_lo;
    a.price_span = q.price_span;
    a.slots = q.slots;
    a.log2cap = q.log2cap;
    a.bitmap = q.bitmap;
    a.kmin = q.kmin;
    a.krange = q.krange;
    a.sum = q.sum;
    a.count = q.count;
    a.err = q.err;
    const uint32_t page_bytes = q.page_rows * 4;  // the widest (4-byte) column
    const uint32_t tok_cap = q.max_nseq ? q.max_nseq : lz4wg::max_sequences(page_bytes);
    const lz4wg::Layout l = lz4wg::layout(q.stage_cap, page_bytes, tok_cap, BLOCK);
    const size_t lds = (size_t)l.total + 16;
    if (lds > (size_t)ctx->lds_max) return hipErrorNotSupported;
    // the small table (PROBE_HASH_LDS) is probed where it lies, like the large one
    return q.probe == PROBE_BITMAP ? launch_wg_mode<PROBE_BITMAP>(ctx, a, l, lds)
                                   : launch_wg_mode<PROBE_HASH_GLOBAL>(ctx, a, l, lds);
}

}  // namespace golap_impl
