// gfx950 page decompression (ChunkDecompressor's decode, core/comp.cuh:171-195): every page of
// a chunk range of one column decoded into a device buffer.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "common.hpp"
#include "device_codecs.hpp"
#include "lz4i_wg.hpp"

namespace golap_impl {
using namespace golap_dev;
using golap_fmt::PAGE_CODEC_MASK;
using golap_fmt::PAGE_LITERAL_ONLY;

namespace {
constexpr int BLOCK = 256;

template <typename K>
hipError_t prepare_lds(golap_ctx *ctx, K kernel, size_t lds) {
    if (lds > (size_t)ctx->lds_max) return hipErrorInvalidConfiguration;
    if (lds > 65536)
        return hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    return hipSuccess;
}

// Workgroups of 256 threads per CU that LDS allows (at most 8), for kernels whose registers
// allow four waves per SIMD.
uint32_t wg_per_cu(const golap_ctx *ctx, uint32_t lds) {
    const uint32_t lds_cu = std::max<uint32_t>((uint32_t)ctx->lds_max, 65536);
    return std::max<uint32_t>(1, std::min<uint32_t>(4, lds_cu / std::max<uint32_t>(lds, 1)));
}

__device__ __forceinline__ void raise_err(uint32_t *err, uint32_t bit) {
    if (err) atomicOr(err, bit);
}
// True when a page can be read in place: a RAW page, or an LZ4 page that is one literal run.
__device__ __forceinline__ bool page_direct(uint8_t codec) {
    return (codec & PAGE_CODEC_MASK) == golap_fmt::RAW || (codec & PAGE_LITERAL_ONLY);
}
__device__ __forceinline__ uint32_t direct_header(uint8_t codec, uint32_t nbytes) {
    return (codec & PAGE_LITERAL_ONLY) ? golap_fmt::lz4_literal_header_bytes(nbytes) : 0u;
}
// The in-order decoders, run by one wave: any page, from global memory or (an LZ4I page the
// cooperative decode hands back) from its staged copy in LDS. Kept out of line: it is the rare
// path and would otherwise add its registers to the cooperative decode's.
__device__ __attribute__((noinline)) int64_t decode_in_order(uint8_t codec, const uint8_t *src, uint32_t csize,
                                                             uint8_t *dst, uint32_t cap, uint32_t elem) {
    return page_decode_wave(codec, src, csize, dst, cap, elem);
}
}  // namespace

// ------------------------------------------------------------------------ page decompression
struct DecodeArgs {
    ColView c;
    uint64_t page_begin, page_end;
    uint32_t page_rows;
    uint64_t n_rows;
    uint8_t *out;
    uint64_t out_row0;
    uint32_t window;   // LDS decode window (a multiple of 16 bytes, the page plus LZ4I slack)
    uint32_t tok_cap;  // LZ4I token offsets after the window
    uint32_t *err;
};

// One wave (workgroup) per page. In-place pages are copied global->global with 16-byte
// accesses; coded pages are decoded into the wave's LDS window (LZ4I pages by the parallel
// parse), then written out with 16-byte stores.
namespace {
__global__ __launch_bounds__(WAVE) void k_decode_pages(DecodeArgs a) {
    extern __shared__ __align__(16) uint8_t lds[];
    const uint32_t lane = threadIdx.x;
    uint8_t *win = lds;
    uint16_t *tok = reinterpret_cast<uint16_t *>(lds + a.window);
    const uint64_t n = a.page_end - a.page_begin;
    for (uint64_t i = blockIdx.x; i < n; i += gridDim.x) {
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
            const int64_t got = page_decode_wave(codec, src, csize, win, nbytes, a.c.elem, a.window, tok, a.tok_cap);
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

// One workgroup (256 threads) per page, for columns with LZ4I pages: an LZ4I page is staged in
// LDS and decoded by the whole workgroup (lz4wg::decode); in-place pages are copied
// global->global; any other coded page (and an LZ4I page the cooperative decode does not take)
// is decoded by the first wave. The decoded page is written out with 16-byte stores. JUMP: the
// layout has a source map, and matches left pending by the ordered pass are resolved by pointer
// jumping (pages whose matches copy from matches, in long chains). NT threads take up to 8 NT
// sequences a page: 256, or 1,024 for columns whose pages have more than 2,048.
template <bool JUMP, int NT>
__global__ __launch_bounds__(NT) void k_decode_wg(DecodeArgs a, lz4wg::Layout l) {
    extern __shared__ __align__(16) uint8_t lds[];
    const uint32_t tid = threadIdx.x;
    uint8_t *win = lds + l.out;
    uint32_t *res = reinterpret_cast<uint32_t *>(lds + l.misc) + 1;
    const uint64_t n = a.page_end - a.page_begin;
    for (uint64_t i = blockIdx.x; i < n; i += gridDim.x) {
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
                if (tid == 0) raise_err(a.err, DEVERR_CORRUPT);
                continue;
            }
            from = src + hdr;
        } else {
            int64_t got = -2;
            const uint8_t *in_order = src;
            if ((codec & PAGE_CODEC_MASK) == golap_fmt::LZ4I && csize <= l.stage_cap && nbytes <= l.out_cap &&
                (reinterpret_cast<uintptr_t>(src) & 15) == 0) {
                lz4wg::stage<NT>(lds, l, src, csize);
                __syncthreads();
                got = lz4wg::decode<NT, JUMP>(lds, l, csize, nbytes);
                in_order = lds + l.stage;
            }
            if (got == -2) {
                if (tid < WAVE) {
                    const int64_t r = nbytes <= l.out_cap ? decode_in_order(codec, in_order, csize, win, nbytes, a.c.elem)
                                                          : -1;
                    if (tid == 0) *res = r == (int64_t)nbytes ? 1u : 0u;
                }
                __syncthreads();
                got = *res ? (int64_t)nbytes : -1;
            }
            if (got != (int64_t)nbytes) {
                if (tid == 0) raise_err(a.err, DEVERR_CORRUPT);
                __syncthreads();  // the next page's staging rewrites the LDS this one read
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
            for (uint32_t k = tid; k < nv; k += NT) d4[k] = s4[k];
            done = nv * 16;
        }
        for (uint32_t k = done + tid; k < nbytes; k += NT) dst[k] = from[k];
        __syncthreads();
    }
}

}  // namespace

hipError_t launch_decode_pages(golap_ctx *ctx, const golap_column *col, uint64_t page_begin,
                               uint64_t page_end, uint8_t *dev_out, uint64_t out_row0, hipStream_t stream,
                               uint32_t *err) {
    if (page_end <= page_begin) return hipSuccess;
    if (!stream) stream = ctx->stream;
    if (!err) err = reinterpret_cast<uint32_t *>(ctx->d_scratch);
    DecodeArgs a{};
    a.c = view(col);
    a.page_begin = page_begin;
    a.page_end = page_end;
    a.page_rows = col->page_rows;
    a.n_rows = col->n_rows;
    a.out = dev_out;
    a.out_row0 = out_row0;
    const uint32_t page_bytes = col->page_rows * col->elem;
    a.err = err;
    if (col->extra.has_lz4i && page_bytes < 65536) {
        const uint32_t tok_cap = col->extra.lz4i_max_nseq ? col->extra.lz4i_max_nseq : lz4wg::max_sequences(page_bytes);
        const uint32_t nt = tok_cap > lz4wg::MAXJ * BLOCK ? 4 * BLOCK : BLOCK;
        const lz4wg::Layout lj = lz4wg::layout(lz4i_stage_bytes(col->extra), page_bytes, tok_cap, nt, true, true);
        const lz4wg::Layout l = lz4wg::layout(lz4i_stage_bytes(col->extra), page_bytes, tok_cap, nt, true);
        // the source map when it keeps at least three quarters of the workgroups per CU
        auto per_cu = [&](uint32_t lds) {
            return std::min<uint64_t>((uint64_t)ctx->lds_max / std::max<uint32_t>(lds, 1), 2048 / nt);
        };
        const bool jump = lj.total <= (uint32_t)ctx->lds_max && 4 * per_cu(lj.total) >= 3 * per_cu(l.total);
        if (l.total <= (uint32_t)ctx->lds_max) {
            const lz4wg::Layout &u = jump ? lj : l;
            auto kernel = nt == BLOCK ? (jump ? k_decode_wg<true, BLOCK> : k_decode_wg<false, BLOCK>)
                                      : (jump ? k_decode_wg<true, 4 * BLOCK> : k_decode_wg<false, 4 * BLOCK>);
            hipError_t e = prepare_lds(ctx, kernel, u.total);
            if (e != hipSuccess) return e;
            int occ = 0;
            e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kernel, (int)nt, u.total);
            if (e != hipSuccess) return e;
            const uint32_t wgs = std::max<uint32_t>(1, std::min<uint32_t>((uint32_t)occ, wg_per_cu(ctx, u.total)));
            const uint32_t grid = (uint32_t)std::min<uint64_t>(page_end - page_begin, (uint64_t)ctx->cus * wgs);
            hipLaunchKernelGGL(kernel, dim3(grid), dim3(nt), u.total, stream, a, u);
            return hipGetLastError();
        }
    }
    if (col->extra.has_lz4i && page_bytes <= 65536) {
        lz4i_window(page_bytes, col->extra.lz4i_max_nseq, a.window, a.tok_cap);
    } else {
        a.window = (page_bytes + 15) & ~15u;
        a.tok_cap = 0;
    }
    const size_t lds = (size_t)a.window + (a.tok_cap ? lz4i::scratch_bytes(a.window, a.tok_cap) : 0);
    hipError_t e = prepare_lds(ctx, k_decode_pages, lds);
    if (e != hipSuccess) return e;
    int occ = 0;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_decode_pages, WAVE, lds);
    if (e != hipSuccess) return e;
    occ = std::max(occ, 1);
    const uint32_t grid = (uint32_t)std::min<uint64_t>(page_end - page_begin, (uint64_t)ctx->cus * occ);
    hipLaunchKernelGGL(k_decode_pages, dim3(grid), dim3(WAVE), lds, stream, a);
    return hipGetLastError();
}

}  // namespace golap_impl
