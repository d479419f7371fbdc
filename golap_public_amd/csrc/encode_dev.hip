// Device write path (gfx950): prepare_compressed's compression step on the GPU.
//
// The reference compresses chunk by chunk on the device with nvcomp (ChunkCompressor,
// core/comp.cuh:81-149, chained DataCopy -> ChunkCompressor -> DenseWriter by
// prepare_compressed, core/apps.cuh:148-255) and fills the zone maps with a cub reduction per
// chunk. Here the values already in HBM are coded into the paged column format of the host
// encoder (format.hpp), byte for byte: same codec choice per page, same LZ4 parse, same page
// layout, so a column coded here equals golap_encode + golap_column_upload of the same values.
//
//   k_encode_pages   one wave per page. RAW copies; DBP / FOR reduce the page's delta or value
//                    range across the wave and assemble every 32-bit output word from the
//                    fields that overlap it (no atomics); RLE counts run starts per lane and
//                    places runs by a wave prefix sum; LZ4 runs the host encoder's greedy
//                    single-probe parse wave-uniformly (hash table of page positions in LDS,
//                    matches extended 64 bytes per step by a ballot, literals copied by all
//                    lanes) and, as the host does, puts a block with matches behind its parse
//                    index (LZ4I: the token offset of every 16th sequence kept in LDS). AUTO
//                    keeps the smallest of LZ4 / DBP / RLE / FOR, ties to the earlier. Each
//                    page is coded into its own slot of a scratch buffer.
//   k_chunk_minmax   one workgroup per chunk: the zone map.
//   k_place_pages    after the host has laid the pages out (compute_page_offsets, the host
//                    encoder's own layout), copies every slot to its place in the column.
#include <hip/hip_runtime.h>

#include <chrono>
#include <memory>
#include <string>
#include <vector>

#include "common.hpp"
#include "format.hpp"
#include "host_column.hpp"

namespace golap_impl {
namespace {

using namespace golap_fmt;

constexpr int PW = 64;                          // threads of k_encode_pages: one wave per page
constexpr int BLOCK = 256;                      // threads of the other kernels
constexpr uint32_t HT_LOG = golap_host::LZ4_HASH_LOG;
constexpr uint32_t HT_SIZE = 1u << HT_LOG;
constexpr uint32_t HT_EMPTY = 0xffffu;          // never accepted: a stored position is <= n - 12
constexpr uint64_t LZ4_DEV_MAX_PAGE = 65536;    // page positions fit 16 bits
// The LZ4 block is coded this far into its slot, room for the largest LZ4I header in front of
// it (4 bytes + 255 index entries of 2 bytes, rounded up to 16).
constexpr uint32_t LZ4I_HDR_MAX = 528;
static_assert(LZ4I_HDR_MAX >= ((4 + 2 * (golap_fmt::LZ4I_MAX_SEGS - 1) + 15) & ~15u), "LZ4I header room");

struct EncodeArgs {
    const uint8_t *src;
    uint32_t type, elem, page_rows, codec, min_match, unit;
    uint64_t n_rows, n_pages;
    uint8_t *slots;
    uint64_t slot_bytes;
    uint32_t *csize;
    uint32_t *shift;     // where the page's bytes start in its slot
    uint8_t *codecs;
    uint32_t *max_nseq;  // most sequences of an LZ4I page (atomic maximum)
};

__device__ __forceinline__ int64_t load_typed(const uint8_t *p, uint32_t type, uint64_t i) {
    switch (type) {
    case GOLAP_I8: return reinterpret_cast<const int8_t *>(p)[i];
    case GOLAP_U8: return p[i];
    case GOLAP_I16: return reinterpret_cast<const int16_t *>(p)[i];
    case GOLAP_U16: return reinterpret_cast<const uint16_t *>(p)[i];
    case GOLAP_I32: return reinterpret_cast<const int32_t *>(p)[i];
    case GOLAP_U32: return reinterpret_cast<const uint32_t *>(p)[i];
    default: return reinterpret_cast<const int64_t *>(p)[i];
    }
}

__device__ __forceinline__ void store_typed(uint8_t *p, uint32_t elem, uint32_t i, int64_t v) {
    switch (elem) {
    case 1: p[i] = (uint8_t)v; break;
    case 2: reinterpret_cast<uint16_t *>(p)[i] = (uint16_t)v; break;
    case 4: reinterpret_cast<uint32_t *>(p)[i] = (uint32_t)v; break;
    default: reinterpret_cast<uint64_t *>(p)[i] = (uint64_t)v; break;
    }
}

// Four bytes at byte x of a 4-byte aligned page (staged in LDS): two aligned dwords funnelled by
// x & 3 bytes (v_alignbyte_b32 shifts the 64-bit pair right by 8 * its third operand).
__device__ __forceinline__ uint32_t rd32(const uint8_t *p, uint32_t x) {
    const uint32_t *w = reinterpret_cast<const uint32_t *>(p + (x & ~3u));
    return __builtin_amdgcn_alignbyte(w[1], w[0], x & 3u);
}

__device__ __forceinline__ int64_t wave_min(int64_t v) {
    for (int o = PW / 2; o; o >>= 1) v = min(v, (int64_t)__shfl_xor((long long)v, o));
    return v;
}
__device__ __forceinline__ int64_t wave_max(int64_t v) {
    for (int o = PW / 2; o; o >>= 1) v = max(v, (int64_t)__shfl_xor((long long)v, o));
    return v;
}
__device__ __forceinline__ __int128 shfl_xor128(__int128 v, int o) {
    const long long hi = __shfl_xor((long long)(v >> 64), o);
    const unsigned long long lo = __shfl_xor((unsigned long long)v, o);
    return (__int128)(((unsigned __int128)(unsigned long long)hi << 64) | lo);
}
// Inclusive prefix sum over the wave.
__device__ __forceinline__ uint32_t wave_incl_sum(uint32_t v) {
    const uint32_t lane = threadIdx.x;
    for (int o = 1; o < PW; o <<= 1) {
        const uint32_t t = __shfl_up(v, o);
        if (lane >= (uint32_t)o) v += t;
    }
    return v;
}

__device__ __forceinline__ uint32_t bits_of(uint64_t range) { return range ? 64u - (uint32_t)__clzll(range) : 0u; }

// ------------------------------------------------------------------ LZ4 (lz4_encode_block)
__device__ __forceinline__ uint32_t len_bytes(uint32_t len) { return len >= 15 ? 1u + (len - 15) / 255u : 0u; }

// Token plus length extension of a sequence at out + op (lane 0 writes); returns the bytes.
__device__ __forceinline__ uint32_t put_token(uint8_t *out, uint32_t op, uint32_t lit, uint32_t ml, bool has_match) {
    if (threadIdx.x == 0) {
        out[op] = (uint8_t)(((lit >= 15 ? 15u : lit) << 4) | (has_match ? (ml >= 15 ? 15u : ml) : 0u));
        if (lit >= 15) {
            uint32_t r = lit - 15, q = op + 1;
            while (r >= 255) {
                out[q++] = 255;
                r -= 255;
            }
            out[q] = (uint8_t)r;
        }
    }
    return 1 + len_bytes(lit);
}

__device__ __forceinline__ void copy_bytes(uint8_t *dst, const uint8_t *src, uint32_t n) {
    for (uint32_t x = threadIdx.x; x < n; x += PW) dst[x] = src[x];
}

// The host encoder's greedy parse (one hash probe per unit-aligned position, acceleration after 64
// misses, match lengths rounded down to whole units, matches shorter than min_match not emitted), run by the whole wave with identical control
// flow over the page staged in LDS (src). The hash table holds page positions; HT_EMPTY
// entries and entries not below ip are rejected exactly as the host rejects entries of earlier
// pages. The workgroup is one wave, whose LDS accesses complete in issue order, so the probe
// needs no barrier between reading and rewriting a table entry. A literal-only block is written
// `shift` bytes into the slot so that its payload is 16-byte aligned there, as it will be in the
// column. The block offset of every 8th sequence's token goes to rec[seq / 8] (lane 0).
// Returns the block size; literal_only, lit_shift (the literal-only block's offset in out) and
// nseq (the block's sequences) set.
__device__ __forceinline__ uint32_t lz4_page(const uint8_t *src, uint32_t n, uint8_t *out, uint32_t min_match,
                                             uint32_t unit, uint16_t *ht, uint16_t *rec, bool &literal_only,
                                             uint32_t &lit_shift, uint32_t &nseq) {
    const uint32_t lane = threadIdx.x;
    if (min_match < LZ4_MINMATCH) min_match = LZ4_MINMATCH;
    if (unit != 2 && unit != 4) unit = 1;
    uint32_t op = 0, anchor = 0, ip = 0;
    nseq = 0;
    if (n >= LZ4_MFLIMIT + 1) {
        const uint32_t msl = n - LZ4_MFLIMIT, mel = n - LZ4_LASTLITERALS;
        uint32_t misses = 0;
        while (ip <= msl) {
            const uint32_t seq = rd32(src, ip);
            const uint32_t h = (seq * 2654435761u) >> (32 - HT_LOG);
            const uint32_t ref = ht[h];
            ht[h] = (uint16_t)ip;  // every lane writes the same value
            if (ref != HT_EMPTY && ref < ip && ip - ref <= LZ4_MAX_OFFSET && rd32(src, ref) == seq) {
                uint32_t len = 4;
                for (;;) {  // first position where the bytes differ or the match must end
                    const uint32_t q = ip + len + lane;
                    const bool same = q < mel && src[ref + len + lane] == src[q];
                    const uint64_t stop = __ballot(!same);
                    if (stop) {
                        len += (uint32_t)__ffsll((unsigned long long)stop) - 1;
                        break;
                    }
                    len += PW;
                }
                len -= len % unit;
                if (len >= min_match) {
                    const uint32_t lit = ip - anchor, ml = len - LZ4_MINMATCH;
                    if (lane == 0 && (nseq & 7) == 0) rec[nseq >> 3] = (uint16_t)op;
                    ++nseq;
                    op += put_token(out, op, lit, ml, true);
                    copy_bytes(out + op, src + anchor, lit);
                    op += lit;
                    const uint32_t off = ip - ref;
                    if (lane == 0) {
                        out[op] = (uint8_t)(off & 0xff);
                        out[op + 1] = (uint8_t)(off >> 8);
                        if (ml >= 15) {
                            uint32_t r = ml - 15, q = op + 2;
                            while (r >= 255) {
                                out[q++] = 255;
                                r -= 255;
                            }
                            out[q] = (uint8_t)r;
                        }
                    }
                    op += 2 + len_bytes(ml);
                    ip += len;
                    anchor = ip;
                    misses = 0;
                    continue;
                }
            }
            ip += unit * (1 + (misses++ >> 6));
        }
    }
    const uint32_t lit = n - anchor;
    literal_only = op == 0;
    lit_shift = 0;
    if (literal_only) {
        const uint32_t hdr = 1 + len_bytes(lit);
        lit_shift = (PAGE_ALIGN - hdr % PAGE_ALIGN) % PAGE_ALIGN;
        out += lit_shift;
    }
    if (lane == 0 && (nseq & 7) == 0) rec[nseq >> 3] = (uint16_t)op;
    ++nseq;
    op += put_token(out, op, lit, 0, false);
    copy_bytes(out + op, src + anchor, lit);
    return op + lit;
}

// ------------------------------------------------------------------ DBP / FOR / RLE
struct PageStats {
    bool dbp_ok, for_ok;
    uint32_t dbp_bw, for_bw, nruns;
    int64_t dbp_mind, for_min;
    uint64_t dbp_size, for_size, rle_size;
};

__device__ void page_stats(const uint8_t *src, uint32_t type, uint32_t elem, uint32_t rows, bool need_dbp,
                           bool need_for, bool need_rle, PageStats &s) {
    const uint32_t lane = threadIdx.x;
    if (need_dbp) {
        __int128 mind = 0, maxd = 0;
        bool any = false;
        for (uint32_t i = 1 + lane; i < rows; i += PW) {
            const __int128 d = (__int128)load_typed(src, type, i) - load_typed(src, type, i - 1);
            if (!any || d < mind) mind = d;
            if (!any || d > maxd) maxd = d;
            any = true;
        }
        if (!any) {  // lanes without deltas take an existing one (row 1's), or 0 for one-row pages
            const __int128 d = rows > 1 ? (__int128)load_typed(src, type, 1) - load_typed(src, type, 0) : 0;
            mind = maxd = d;
        }
        for (int o = PW / 2; o; o >>= 1) {
            const __int128 a = shfl_xor128(mind, o), b = shfl_xor128(maxd, o);
            mind = a < mind ? a : mind;
            maxd = b > maxd ? b : maxd;
        }
        const __int128 lo64 = (__int128)INT64_MIN, hi64 = (__int128)INT64_MAX;
        const unsigned __int128 range = (unsigned __int128)(maxd - mind);
        s.dbp_ok = mind >= lo64 && mind <= hi64 && (range >> 64) == 0;
        s.dbp_bw = s.dbp_ok ? bits_of((uint64_t)range) : 64;
        s.dbp_ok = s.dbp_ok && s.dbp_bw <= 32;
        s.dbp_mind = s.dbp_ok ? (int64_t)mind : 0;
        const uint64_t words = rows > 1 ? ((uint64_t)(rows - 1) * s.dbp_bw + 31) / 32 : 0;
        s.dbp_size = s.dbp_ok ? sizeof(DbpHeader) + words * 4 : 0;
    }
    if (need_for) {
        int64_t mn = INT64_MAX, mx = INT64_MIN;
        for (uint32_t i = lane; i < rows; i += PW) {
            const int64_t v = load_typed(src, type, i);
            mn = min(mn, v);
            mx = max(mx, v);
        }
        mn = wave_min(mn);
        mx = wave_max(mx);
        s.for_bw = bits_of((uint64_t)mx - (uint64_t)mn);
        s.for_ok = s.for_bw <= FOR_MAX_BW;
        s.for_min = mn;
        s.for_size = s.for_ok ? for_page_bytes(rows, s.for_bw) : 0;
    }
    if (need_rle) {
        uint32_t cnt = 0;
        for (uint32_t i = lane; i < rows; i += PW)
            cnt += (i == 0 || load_typed(src, type, i) != load_typed(src, type, i - 1)) ? 1u : 0u;
        for (int o = PW / 2; o; o >>= 1) cnt += __shfl_xor(cnt, o);
        s.nruns = cnt;
        s.rle_size = sizeof(RleHeader) + (((uint64_t)cnt * elem + 3) & ~3ull) + 4ull * cnt;
    }
}

// DbpHeader / FOR header: base, min_delta, n, bw, elem, zero padding (32 bytes).
__device__ __forceinline__ void put_header32(uint8_t *out, int64_t base, int64_t min_delta, uint32_t n, uint32_t bw,
                                             uint32_t elem) {
    if (threadIdx.x == 0) {
        uint32_t *w = reinterpret_cast<uint32_t *>(out);
        w[0] = (uint32_t)(uint64_t)base;
        w[1] = (uint32_t)((uint64_t)base >> 32);
        w[2] = (uint32_t)(uint64_t)min_delta;
        w[3] = (uint32_t)((uint64_t)min_delta >> 32);
        w[4] = n;
        w[5] = bw | (elem << 8);
        w[6] = 0;
        w[7] = 0;
    }
}

__device__ void dbp_write(const uint8_t *src, uint32_t type, uint32_t elem, uint32_t rows, const PageStats &s,
                          uint8_t *out) {
    put_header32(out, load_typed(src, type, 0), s.dbp_mind, rows, s.dbp_bw, elem);
    const uint32_t bw = s.dbp_bw;
    const uint64_t words = rows > 1 ? ((uint64_t)(rows - 1) * bw + 31) / 32 : 0;
    uint32_t *pk = reinterpret_cast<uint32_t *>(out + sizeof(DbpHeader));
    for (uint64_t w = threadIdx.x; w < words; w += PW) {
        uint32_t word = 0;
        // field f holds row f+1's delta at bits [f*bw, f*bw + bw)
        for (uint64_t f = w * 32 / bw; f + 1 < rows && f * bw < w * 32 + 32; ++f) {
            const uint64_t p = (uint64_t)((__int128)load_typed(src, type, (uint32_t)f + 1) -
                                          load_typed(src, type, (uint32_t)f) - s.dbp_mind);
            const int64_t sh = (int64_t)(f * bw) - (int64_t)(w * 32);
            word |= sh >= 0 ? (uint32_t)(p << sh) : (uint32_t)(p >> -sh);
        }
        pk[w] = word;
    }
}

__device__ void for_write(const uint8_t *src, uint32_t type, uint32_t elem, uint32_t rows, const PageStats &s,
                          uint8_t *out) {
    put_header32(out, s.for_min, 0, rows, s.for_bw, elem);
    const uint32_t bw = s.for_bw;
    const uint64_t payload = for_page_bytes(rows, bw) - 32;  // fields + 8 zero bytes
    const uint64_t words = (payload + 3) / 4;
    uint32_t *pk = reinterpret_cast<uint32_t *>(out + 32);
    for (uint64_t w = threadIdx.x; w < words; w += PW) {
        uint32_t word = 0;
        if (bw)
            for (uint64_t f = w * 32 / bw; f < rows && f * bw < w * 32 + 32; ++f) {
                const uint64_t off = (uint64_t)load_typed(src, type, (uint32_t)f) - (uint64_t)s.for_min;
                const int64_t sh = (int64_t)(f * bw) - (int64_t)(w * 32);
                word |= sh >= 0 ? (uint32_t)(off << sh) : (uint32_t)(off >> -sh);
            }
        pk[w] = word;
    }
}

__device__ void rle_write(const uint8_t *src, uint32_t type, uint32_t elem, uint32_t rows, const PageStats &s,
                          uint8_t *out) {
    const uint32_t lane = threadIdx.x;
    const uint32_t nruns = s.nruns;
    const uint64_t vbytes = ((uint64_t)nruns * elem + 3) & ~3ull;
    uint8_t *vals = out + sizeof(RleHeader);
    uint32_t *ends = reinterpret_cast<uint32_t *>(vals + vbytes);
    if (lane == 0) {
        uint32_t *h = reinterpret_cast<uint32_t *>(out);
        h[0] = nruns;
        h[1] = rows;
        h[2] = elem;
        h[3] = 0;
        for (uint64_t b = (uint64_t)nruns * elem; b < vbytes; ++b) vals[b] = 0;
        ends[nruns - 1] = rows;
    }
    // contiguous row segments per lane, run indices by a wave prefix sum
    const uint32_t seg = (rows + PW - 1) / PW, r0 = min(rows, lane * seg), r1 = min(rows, r0 + seg);
    uint32_t cnt = 0;
    for (uint32_t i = r0; i < r1; ++i)
        cnt += (i == 0 || load_typed(src, type, i) != load_typed(src, type, i - 1)) ? 1u : 0u;
    uint32_t k = wave_incl_sum(cnt) - cnt;
    for (uint32_t i = r0; i < r1; ++i) {
        const int64_t v = load_typed(src, type, i);
        if (i == 0 || v != load_typed(src, type, i - 1)) {
            if (k > 0) ends[k - 1] = i;
            store_typed(vals, elem, k, v);
            ++k;
        }
    }
}

__device__ void raw_write(const uint8_t *src, uint32_t n, uint8_t *out) {
    const uint32_t n16 = n / 16;
    const uint4 *s = reinterpret_cast<const uint4 *>(src);
    uint4 *d = reinterpret_cast<uint4 *>(out);
    for (uint32_t i = threadIdx.x; i < n16; i += PW) d[i] = s[i];
    for (uint32_t i = n16 * 16 + threadIdx.x; i < n; i += PW) out[i] = src[i];
}

// Dynamic LDS of the LZ4 variant: the hash table, the page being parsed, then the token offset
// of every 8th sequence. Every sequence but the last ends in a match of at least 4 bytes, so a
// page of n bytes has at most n / 4 + 1 sequences.
constexpr uint32_t HT_BYTES = HT_SIZE * 2;
__host__ __device__ inline uint32_t page_lds_bytes(uint32_t page_bytes) { return (page_bytes + 15) / 16 * 16; }
__host__ __device__ inline uint32_t rec_lds_bytes(uint32_t page_bytes) {
    return ((page_bytes / 4 + 1) / 8 + 1) * 2;
}
uint32_t encode_lds_bytes(uint64_t page_bytes) {
    return HT_BYTES + page_lds_bytes((uint32_t)page_bytes) + rec_lds_bytes((uint32_t)page_bytes);
}

// The LZ4I header (format.hpp) of a block of nseq sequences into hdr (lz4i_header_bytes bytes,
// ending where the block starts), from the token offsets of every 8th sequence.
__device__ __forceinline__ uint32_t lz4i_put_header(uint8_t *hdr_end, uint32_t nseq, const uint16_t *rec, uint32_t unit) {
    const uint32_t sl = lz4i_seg_log2(nseq), nseg = lz4i_segments(nseq, sl), hb = lz4i_header_bytes(nseq, sl);
    uint8_t *h = hdr_end - hb;
    for (uint32_t x = threadIdx.x; x < hb; x += PW) {
        uint32_t b = 0;
        if (x == 0) b = nseq & 0xff;
        else if (x == 1) b = nseq >> 8;
        else if (x == 2) b = sl;
        else if (x == 3) b = unit == 4 ? 2u : unit == 2 ? 1u : 0u;
        else if (x >= 4) {
            const uint32_t s = (x - 4) / 2 + 1;  // entry s: the token of sequence s << sl
            if (s < nseg) {
                const uint32_t t = rec[s << (sl - LZ4I_MIN_SEG_LOG2)];
                b = (x & 1) ? t >> 8 : t & 0xff;
            }
        }
        h[x] = (uint8_t)b;
    }
    return hb;
}

// The page's bytes into LDS and the hash table emptied (the wave's loads and stores complete
// before the barrier's wait).
__device__ __forceinline__ void stage_page(const uint8_t *src, uint32_t n, uint8_t *page, uint16_t *ht) {
    const uint4 empty = make_uint4(0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu);
    uint4 *h16 = reinterpret_cast<uint4 *>(ht);
    for (uint32_t i = threadIdx.x; i < HT_BYTES / 16; i += PW) h16[i] = empty;
    const uint32_t n16 = n / 16;
    const uint4 *s16 = reinterpret_cast<const uint4 *>(src);
    uint4 *p16 = reinterpret_cast<uint4 *>(page);
    for (uint32_t i = threadIdx.x; i < n16; i += PW) p16[i] = s16[i];
    for (uint32_t i = n16 * 16 + threadIdx.x; i < n; i += PW) page[i] = src[i];
    __syncthreads();
}

template <bool LZ>
__global__ __launch_bounds__(PW) void k_encode_pages(EncodeArgs a) {
    extern __shared__ __align__(16) uint8_t lds[];
    uint16_t *ht = reinterpret_cast<uint16_t *>(lds);
    uint8_t *page = lds + HT_BYTES;
    uint16_t *rec = reinterpret_cast<uint16_t *>(page + page_lds_bytes(a.page_rows * a.elem));
    for (uint64_t p = blockIdx.x; p < a.n_pages; p += gridDim.x) {
        const uint64_t r0 = p * a.page_rows;
        const uint32_t rows = (uint32_t)min<uint64_t>(a.page_rows, a.n_rows - r0);
        const uint32_t nbytes = rows * a.elem;
        const uint8_t *src = a.src + r0 * a.elem;
        uint8_t *out = a.slots + p * a.slot_bytes;
        const uint32_t c = a.codec;
        uint32_t size = 0, shift = 0, lz4i_nseq = 0;
        uint8_t code = RAW;
        if (c == RAW) {
            raw_write(src, nbytes, out);
            size = nbytes;
        } else {
            const bool auto_ = c == GOLAP_CODEC_AUTO;
            PageStats s{};
            page_stats(src, a.type, a.elem, rows, auto_ || c == DBP, auto_ || c == FOR, auto_ || c == RLE, s);
            uint64_t best = UINT64_MAX;
            uint32_t pick = 0xff;  // codec whose bytes must be written (LZ4 is written as it is coded)
            // LZ4 as the host encoder keeps it: one literal run stays a plain LZ4 page, any other
            // block of up to LZ4I_MAX_BLOCK bytes goes behind its parse index (LZ4I)
            if (LZ && (c == LZ4 || auto_ || (c == DBP && !s.dbp_ok) || (c == FOR && !s.for_ok))) {
                bool lit_only = false;
                uint32_t lit_shift = 0, nseq = 0;
                stage_page(src, nbytes, page, ht);
                const uint32_t bsize =
                    lz4_page(page, nbytes, out + LZ4I_HDR_MAX, a.min_match, a.unit, ht, rec, lit_only, lit_shift, nseq);
                best = bsize;
                shift = LZ4I_HDR_MAX + lit_shift;
                code = (uint8_t)(LZ4 | (lit_only ? PAGE_LITERAL_ONLY : 0));
                if (!lit_only && bsize <= LZ4I_MAX_BLOCK) {
                    __syncthreads();  // lane 0's token offsets
                    const uint32_t hb = lz4i_put_header(out + LZ4I_HDR_MAX, nseq, rec, a.unit);
                    best = hb + bsize;
                    shift = LZ4I_HDR_MAX - hb;
                    code = LZ4I;
                    lz4i_nseq = nseq;
                }
                pick = LZ4;
            }
            if ((auto_ || c == DBP) && s.dbp_ok && s.dbp_size < best) {
                best = s.dbp_size;
                pick = DBP;
                code = (uint8_t)(DBP | (s.dbp_bw == 0 ? PAGE_POINT : 0));
            }
            if ((auto_ || c == RLE) && s.rle_size < best) {
                best = s.rle_size;
                pick = RLE;
                code = (uint8_t)(RLE | PAGE_POINT);
            }
            if ((auto_ || c == FOR) && s.for_ok && s.for_size < best) {
                best = s.for_size;
                pick = FOR;
                code = (uint8_t)(FOR | PAGE_POINT);
            }
            if (pick != LZ4) shift = 0;
            if (pick == DBP) dbp_write(src, a.type, a.elem, rows, s, out);
            else if (pick == RLE) rle_write(src, a.type, a.elem, rows, s, out);
            else if (pick == FOR) for_write(src, a.type, a.elem, rows, s, out);
            size = (uint32_t)best;
        }
        if (threadIdx.x == 0) {
            a.csize[p] = size;
            a.shift[p] = shift;
            a.codecs[p] = code;
            if (code == LZ4I) atomicMax(a.max_nseq, lz4i_nseq);
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(BLOCK) void k_chunk_minmax(const uint8_t *v, uint32_t type, uint64_t n, uint64_t chunk_rows,
                                                        uint64_t n_chunks, int64_t *mins, int64_t *maxs) {
    __shared__ int64_t red[2][BLOCK / PW];
    const uint32_t lane = threadIdx.x % PW, wave = threadIdx.x / PW;
    for (uint64_t c = blockIdx.x; c < n_chunks; c += gridDim.x) {
        const uint64_t r0 = c * chunk_rows, r1 = min(n, r0 + chunk_rows);
        int64_t mn = INT64_MAX, mx = INT64_MIN;
        for (uint64_t r = r0 + threadIdx.x; r < r1; r += BLOCK) {
            const int64_t x = load_typed(v, type, r);
            mn = min(mn, x);
            mx = max(mx, x);
        }
        mn = wave_min(mn);
        mx = wave_max(mx);
        if (lane == 0) {
            red[0][wave] = mn;
            red[1][wave] = mx;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int w = 1; w < BLOCK / PW; ++w) {
                mn = min(mn, red[0][w]);
                mx = max(mx, red[1][w]);
            }
            mins[c] = mn;
            maxs[c] = mx;
        }
        __syncthreads();
    }
}

// Slot p (its bytes from shift[p] on) -> comp + page_off[p]. Source and destination agree
// modulo 16 (pages start 16-byte aligned, literal-only payloads are aligned in both), so the
// bulk moves in 16-byte units.
__global__ __launch_bounds__(BLOCK) void k_place_pages(const uint8_t *slots, uint64_t slot_bytes, const uint64_t *page_off,
                                                       const uint32_t *csize, const uint32_t *shift, uint64_t n_pages,
                                                       uint8_t *comp) {
    for (uint64_t p = blockIdx.x; p < n_pages; p += gridDim.x) {
        const uint8_t *s = slots + p * slot_bytes + shift[p];
        uint8_t *d = comp + page_off[p];
        const uint32_t n = csize[p];
        const uint32_t head = min<uint32_t>(n, (uint32_t)((PAGE_ALIGN - (reinterpret_cast<uintptr_t>(d) & 15)) & 15));
        const uint32_t n16 = (n - head) / 16;
        for (uint32_t i = threadIdx.x; i < head; i += BLOCK) d[i] = s[i];
        const uint4 *s16 = reinterpret_cast<const uint4 *>(s + head);
        uint4 *d16 = reinterpret_cast<uint4 *>(d + head);
        for (uint32_t i = threadIdx.x; i < n16; i += BLOCK) d16[i] = s16[i];
        for (uint32_t i = head + n16 * 16 + threadIdx.x; i < n; i += BLOCK) d[i] = s[i];
    }
}

uint64_t slot_bytes_for(uint32_t codec, uint32_t page_rows, uint32_t elem) {
    const uint64_t nb = (uint64_t)page_rows * elem;
    uint64_t b = nb;
    if (codec != RAW) b = std::max<uint64_t>(b, LZ4I_HDR_MAX + golap_host::lz4_bound(nb) + PAGE_ALIGN);
    if (codec == DBP || codec == GOLAP_CODEC_AUTO) b = std::max<uint64_t>(b, golap_host::dbp_bound(page_rows));
    if (codec == RLE || codec == GOLAP_CODEC_AUTO) b = std::max<uint64_t>(b, golap_host::rle_bound(page_rows, elem));
    if (codec == FOR || codec == GOLAP_CODEC_AUTO)
        b = std::max<uint64_t>(b, for_page_bytes(page_rows, FOR_MAX_BW) + 4);  // whole words
    return (b + 2 * PAGE_ALIGN - 1) / PAGE_ALIGN * PAGE_ALIGN;
}

}  // namespace
}  // namespace golap_impl

using namespace golap_impl;
using Lock = std::lock_guard<std::recursive_mutex>;

extern "C" int golap_encode_device(golap_ctx *ctx, const void *dev_values, uint64_t n_rows, uint32_t type,
                                   const golap_encode_opts *opts, golap_column **out, double *kernel_ms) {
    if (!ctx || !out || !valid_type(type)) return GOLAP_EINVAL;
    *out = nullptr;
    if (kernel_ms) *kernel_ms = 0;
    golap_encode_opts o;
    golap_encode_opts_default(&o);
    if (opts) o = *opts;
    const uint32_t elem = golap_host::type_size(type);
    if (!o.page_rows || o.page_rows % 64 || !o.chunk_rows || o.chunk_rows % o.page_rows)
        return fail(ctx, GOLAP_EINVAL, "encode_device: page_rows must be a positive multiple of 64 and divide chunk_rows");
    if (o.codec != RAW && o.codec != LZ4 && o.codec != DBP && o.codec != RLE && o.codec != FOR &&
        o.codec != GOLAP_CODEC_AUTO)
        return fail(ctx, GOLAP_EINVAL, "encode_device: unknown codec");
    const uint64_t page_bytes = (uint64_t)o.page_rows * elem;
    const bool lz = o.codec != RAW && o.codec != RLE;  // DBP and FOR fall back to LZ4 per page
    if (page_bytes > (1u << 20) || (lz && page_bytes > LZ4_DEV_MAX_PAGE))
        return fail(ctx, GOLAP_EINVAL, "encode_device: pages up to 64 KiB (1 MiB for RAW and RLE)");
    if (lz && encode_lds_bytes(page_bytes) > (uint64_t)ctx->lds_max)
        return fail(ctx, GOLAP_EINVAL, "encode_device: page and hash table exceed the device's LDS");
    if (n_rows && (!dev_values || reinterpret_cast<uintptr_t>(dev_values) % 16))
        return fail(ctx, GOLAP_EINVAL, "encode_device: values must be 16-byte aligned device memory");
    Lock l(ctx->mu);
    if (int r = set_device(ctx)) return r;

    golap_host_column h;
    h.type = type;
    h.elem = elem;
    h.page_rows = o.page_rows;
    h.chunk_rows = o.chunk_rows;
    h.n_rows = n_rows;
    h.n_pages = (n_rows + o.page_rows - 1) / o.page_rows;
    h.n_chunks = (n_rows + o.chunk_rows - 1) / o.chunk_rows;
    h.page_off.resize(h.n_pages);
    h.page_csize.resize(h.n_pages);
    h.page_codec.resize(h.n_pages);
    h.chunk_min.resize(h.n_chunks);
    h.chunk_max.resize(h.n_chunks);
    const uint64_t slot = slot_bytes_for(o.codec, o.page_rows, elem);
    // scratch: page slots | csize | shift | codec | chunk min | chunk max | LZ4I sequence maximum
    const uint64_t off_cs = h.n_pages * slot, off_sh = off_cs + h.n_pages * 4, off_cd = off_sh + h.n_pages * 4;
    const uint64_t off_mn = (off_cd + h.n_pages + 15) / 16 * 16, off_mx = off_mn + h.n_chunks * 8;
    const uint64_t off_nq = off_mx + h.n_chunks * 8, total = off_nq + 16;
    void *scratch = nullptr;
    if (total)
        if (int r = dev_alloc(ctx, total, &scratch)) return r;
    uint8_t *sb = static_cast<uint8_t *>(scratch);
    auto release = [&] {
        (void)hipStreamSynchronize(ctx->stream);
        if (scratch) dev_release(ctx, scratch);
    };
    float ms1 = 0, ms2 = 0;
    uint32_t max_nseq = 0;
    hipError_t e = hipSuccess;
    if (h.n_pages) {
        EncodeArgs a{};
        a.src = static_cast<const uint8_t *>(dev_values);
        a.type = type;
        a.elem = elem;
        a.page_rows = o.page_rows;
        a.codec = o.codec;
        a.min_match = o.lz4_min_match;
        a.unit = golap_host::lz4_unit_of(o, elem);
        a.n_rows = n_rows;
        a.n_pages = h.n_pages;
        a.slots = sb;
        a.slot_bytes = slot;
        a.csize = reinterpret_cast<uint32_t *>(sb + off_cs);
        a.shift = reinterpret_cast<uint32_t *>(sb + off_sh);
        a.codecs = sb + off_cd;
        a.max_nseq = reinterpret_cast<uint32_t *>(sb + off_nq);
        const uint32_t grid = (uint32_t)std::min<uint64_t>(h.n_pages, (uint64_t)ctx->cus * 16);
        const uint32_t cgrid = (uint32_t)std::min<uint64_t>(h.n_chunks, (uint64_t)ctx->cus * 8);
        e = hipMemsetAsync(a.max_nseq, 0, 4, ctx->stream);
        if (e == hipSuccess) e = hipEventRecord(ctx->ev0, ctx->stream);
        if (e == hipSuccess) {
            const uint32_t lds = lz ? encode_lds_bytes(page_bytes) : 0;
            if (lds > 65536)
                e = hipFuncSetAttribute(reinterpret_cast<const void *>(k_encode_pages<true>),
                                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            if (e == hipSuccess) {
                if (lz) hipLaunchKernelGGL(k_encode_pages<true>, dim3(grid), dim3(PW), lds, ctx->stream, a);
                else hipLaunchKernelGGL(k_encode_pages<false>, dim3(grid), dim3(PW), 0, ctx->stream, a);
                e = hipGetLastError();
            }
        }
        if (e == hipSuccess) {
            hipLaunchKernelGGL(k_chunk_minmax, dim3(cgrid), dim3(BLOCK), 0, ctx->stream, a.src, type, n_rows,
                               (uint64_t)o.chunk_rows, h.n_chunks, reinterpret_cast<int64_t *>(sb + off_mn),
                               reinterpret_cast<int64_t *>(sb + off_mx));
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipEventRecord(ctx->ev1, ctx->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(h.page_csize.data(), a.csize, h.n_pages * 4, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(h.page_codec.data(), a.codecs, h.n_pages, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(h.chunk_min.data(), sb + off_mn, h.n_chunks * 8, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(h.chunk_max.data(), sb + off_mx, h.n_chunks * 8, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(&max_nseq, a.max_nseq, 4, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e == hipSuccess) e = hipEventElapsedTime(&ms1, ctx->ev0, ctx->ev1);
    }
    if (e != hipSuccess) {
        release();
        return fail(ctx, GOLAP_EHIP, std::string("encode_device: ") + hipGetErrorString(e));
    }
    // the host encoder's layout: 16-byte aligned pages, literal-only payloads aligned
    const uint64_t comp_bytes = golap_host::compute_page_offsets(h);
    golap_column *col = nullptr;
    if (int r = create_device_column(ctx, &h, comp_bytes, nullptr, &col)) {
        release();
        return r;
    }
    e = hipMemsetAsync(col->d_comp, 0, comp_bytes, ctx->stream);  // gaps between pages are zero
    if (e == hipSuccess && h.n_pages) {
        const uint32_t grid = (uint32_t)std::min<uint64_t>(h.n_pages, (uint64_t)ctx->cus * 8);
        e = hipEventRecord(ctx->ev0, ctx->stream);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(k_place_pages, dim3(grid), dim3(BLOCK), 0, ctx->stream, sb, slot, col->d_page_off,
                               reinterpret_cast<const uint32_t *>(sb + off_cs),
                               reinterpret_cast<const uint32_t *>(sb + off_sh), h.n_pages, col->d_comp);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipEventRecord(ctx->ev1, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e == hipSuccess) e = hipEventElapsedTime(&ms2, ctx->ev0, ctx->ev1);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    release();
    if (e != hipSuccess) {
        golap_column_free(ctx, col);
        return fail(ctx, GOLAP_EHIP, std::string("encode_device: ") + hipGetErrorString(e));
    }
    col->extra.lz4i_max_nseq = max_nseq;
    if (kernel_ms) *kernel_ms = (double)ms1 + (double)ms2;
    *out = col;
    return GOLAP_OK;
}
