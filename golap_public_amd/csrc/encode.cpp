// Host write path: prepare_compressed (core/apps.cuh:148-255) re-designed for HBM-resident
// paged columns. Pages are coded in parallel on the host; the column buffer places every page at
// a 16-byte boundary (literal-only LZ4 pages: their payload), so device kernels read pages with
// aligned 16-byte loads.
#include <algorithm>
#include <cstring>
#include <limits>
#include <thread>

#include "format.hpp"
#include "host_column.hpp"

namespace golap_host {
using namespace golap_fmt;

uint32_t type_size(uint32_t type) {
    switch (type) {
    case GOLAP_I8: case GOLAP_U8: return 1;
    case GOLAP_I16: case GOLAP_U16: return 2;
    case GOLAP_I32: case GOLAP_U32: return 4;
    case GOLAP_I64: return 8;
    default: return 0;
    }
}
bool type_signed(uint32_t type) {
    return type == GOLAP_I8 || type == GOLAP_I16 || type == GOLAP_I32 || type == GOLAP_I64;
}
int64_t load_value(const uint8_t *p, uint32_t type) {
    switch (type) {
    case GOLAP_I8: return *(const int8_t *)p;
    case GOLAP_U8: return *p;
    case GOLAP_I16: { int16_t v; memcpy(&v, p, 2); return v; }
    case GOLAP_U16: { uint16_t v; memcpy(&v, p, 2); return v; }
    case GOLAP_I32: { int32_t v; memcpy(&v, p, 4); return v; }
    case GOLAP_U32: { uint32_t v; memcpy(&v, p, 4); return v; }
    case GOLAP_I64: { int64_t v; memcpy(&v, p, 8); return v; }
    default: return 0;
    }
}

// ------------------------------------------------------------------------------------ LZ4
uint64_t lz4_bound(uint64_t n) { return n + n / 255 + 16; }

static inline uint32_t rd32(const uint8_t *p) { uint32_t v; memcpy(&v, p, 4); return v; }
static inline uint32_t lz4_hash(uint32_t seq) { return (seq * 2654435761u) >> (32 - LZ4_HASH_LOG); }

static inline uint8_t *put_len(uint8_t *op, uint64_t len) {  // length extension bytes
    while (len >= 255) { *op++ = 255; len -= 255; }
    *op++ = (uint8_t)len;
    return op;
}

// Greedy single-probe LZ4 block encoder (format: lz4_Block_format.md). The hash table holds
// absolute positions (base + offset into src); entries older than this block are < base and
// rejected, so the table needs no reset between pages. Matches shorter than min_match are not
// emitted: a short match saves at most a byte or two but costs the decoder a full sequence.
// unit (1, 2 or 4): every sequence starts at a multiple of unit bytes and every literal run,
// match and offset is a multiple of unit bytes (the parse visits unit-aligned positions only and
// rounds match lengths down), so a decoder copies whole units; the block is standard LZ4.
uint64_t lz4_encode_block(const uint8_t *src, uint64_t n, uint8_t *dst, uint32_t min_match,
                          int64_t *ht, int64_t base, std::vector<uint32_t> *tokens, uint32_t unit) {
    uint8_t *op = dst;
    if (tokens) tokens->clear();
    uint64_t anchor = 0, ip = 0;
    if (min_match < LZ4_MINMATCH) min_match = LZ4_MINMATCH;
    if (unit != 2 && unit != 4) unit = 1;
    if (n >= LZ4_MFLIMIT + 1) {
        const uint64_t match_start_limit = n - LZ4_MFLIMIT;  // a match starts at ip <= n-12
        const uint64_t match_end_limit = n - LZ4_LASTLITERALS;
        uint32_t misses = 0;
        while (ip <= match_start_limit) {
            uint32_t seq = rd32(src + ip);
            uint32_t h = lz4_hash(seq);
            int64_t ref = ht[h] - base;
            ht[h] = base + (int64_t)ip;
            if (ref >= 0 && (uint64_t)ref < ip && ip - (uint64_t)ref <= LZ4_MAX_OFFSET &&
                rd32(src + ref) == seq) {
                uint64_t len = 4;
                while (ip + len < match_end_limit && src[ref + len] == src[ip + len]) ++len;
                len -= len % unit;
                if (len >= min_match) {
                    uint64_t lit = ip - anchor, ml = len - LZ4_MINMATCH;
                    uint8_t *token = op++;
                    if (tokens) tokens->push_back((uint32_t)(token - dst));
                    *token = (uint8_t)(((lit >= 15 ? 15 : lit) << 4) | (ml >= 15 ? 15 : ml));
                    if (lit >= 15) op = put_len(op, lit - 15);
                    memcpy(op, src + anchor, lit);
                    op += lit;
                    uint64_t off = ip - (uint64_t)ref;
                    *op++ = (uint8_t)(off & 0xff);
                    *op++ = (uint8_t)(off >> 8);
                    if (ml >= 15) op = put_len(op, ml - 15);
                    ip += len;
                    anchor = ip;
                    misses = 0;
                    continue;
                }
            }
            // liblz4-style acceleration: step grows with consecutive misses.
            ip += unit * (1 + (misses++ >> 6));
        }
    }
    uint64_t lit = n - anchor;
    if (tokens) tokens->push_back((uint32_t)(op - dst));
    *op++ = (uint8_t)((lit >= 15 ? 15 : lit) << 4);
    if (lit >= 15) op = put_len(op, lit - 15);
    memcpy(op, src + anchor, lit);
    op += lit;
    return (uint64_t)(op - dst);
}

uint64_t lz4i_make_page(const uint8_t *blk, uint64_t bsize, const std::vector<uint32_t> &tokens, uint8_t *out,
                        uint32_t unit) {
    const uint64_t nseq = tokens.size();
    if (!nseq || bsize > LZ4I_MAX_BLOCK) return 0;
    const uint32_t sl = lz4i_seg_log2((uint32_t)nseq);
    const uint32_t nseg = lz4i_segments((uint32_t)nseq, sl), hb = lz4i_header_bytes((uint32_t)nseq, sl);
    memmove(out + hb, blk, bsize);  // the block may already sit in out, behind the header
    memset(out, 0, hb);
    out[0] = (uint8_t)(nseq & 0xff);
    out[1] = (uint8_t)(nseq >> 8);
    out[2] = (uint8_t)sl;
    out[3] = (uint8_t)(unit == 4 ? 2 : unit == 2 ? 1 : 0);
    for (uint32_t s = 1; s < nseg; ++s) {
        const uint32_t t = tokens[(uint64_t)s << sl];
        out[4 + 2 * (s - 1)] = (uint8_t)(t & 0xff);
        out[5 + 2 * (s - 1)] = (uint8_t)(t >> 8);
    }
    return hb + bsize;
}

// ------------------------------------------------------------------------------------ DBP
uint64_t dbp_bound(uint64_t rows) { return sizeof(DbpHeader) + 4 * ((rows * 32 + 31) / 32) + 4; }

uint64_t dbp_encode_page(const uint8_t *src, uint64_t rows, uint32_t type, uint8_t *dst) {
    const uint32_t elem = type_size(type);
    DbpHeader h{};
    h.n = (uint32_t)rows;
    h.elem = (uint8_t)elem;
    h.base = rows ? load_value(src, type) : 0;
    __int128 mind = 0, maxd = 0;
    for (uint64_t i = 1; i < rows; ++i) {
        __int128 d = (__int128)load_value(src + i * elem, type) - load_value(src + (i - 1) * elem, type);
        if (i == 1 || d < mind) mind = d;
        if (i == 1 || d > maxd) maxd = d;
    }
    if (mind < std::numeric_limits<int64_t>::min() || mind > std::numeric_limits<int64_t>::max())
        return 0;
    unsigned __int128 range = (unsigned __int128)(maxd - mind);
    uint32_t bw = 0;
    while (bw < 64 && (range >> bw) != 0) ++bw;
    if (bw > 32) return 0;
    h.min_delta = (int64_t)mind;
    h.bw = (uint8_t)bw;
    memcpy(dst, &h, sizeof(h));
    uint64_t words = rows > 1 ? ((rows - 1) * bw + 31) / 32 : 0;
    uint32_t *pk = reinterpret_cast<uint32_t *>(dst + sizeof(h));
    memset(pk, 0, words * 4);
    if (bw) {
        for (uint64_t i = 1; i < rows; ++i) {
            uint64_t p = (uint64_t)((__int128)load_value(src + i * elem, type) -
                                    load_value(src + (i - 1) * elem, type) - mind);
            uint64_t bit = (i - 1) * bw, w = bit / 32, s = bit % 32;
            pk[w] |= (uint32_t)(p << s);
            if (s + bw > 32) pk[w + 1] |= (uint32_t)(p >> (32 - s));
        }
    }
    return sizeof(h) + words * 4;
}

// ------------------------------------------------------------------------------------ FOR
uint64_t for_bound(uint64_t rows) { return golap_fmt::for_page_bytes(rows, golap_fmt::FOR_MAX_BW); }

// Frame of reference: the page minimum plus bw-bit offsets; 0 if the range needs more than
// FOR_MAX_BW bits.
uint64_t for_encode_page(const uint8_t *src, uint64_t rows, uint32_t type, uint8_t *dst) {
    const uint32_t elem = type_size(type);
    if (!rows) return 0;
    int64_t mn = load_value(src, type), mx = mn;
    for (uint64_t i = 1; i < rows; ++i) {
        const int64_t v = load_value(src + i * elem, type);
        mn = std::min(mn, v);
        mx = std::max(mx, v);
    }
    const unsigned __int128 range = (unsigned __int128)((__int128)mx - mn);
    uint32_t bw = 0;
    while (bw < 64 && (range >> bw) != 0) ++bw;
    if (bw > golap_fmt::FOR_MAX_BW) return 0;
    DbpHeader h{};
    h.base = mn;
    h.n = (uint32_t)rows;
    h.bw = (uint8_t)bw;
    h.elem = (uint8_t)elem;
    memcpy(dst, &h, sizeof(h));
    const uint64_t total = golap_fmt::for_page_bytes(rows, bw);
    uint8_t *pk = dst + sizeof(h);
    memset(pk, 0, total - sizeof(h));
    if (bw)
        for (uint64_t i = 0; i < rows; ++i) {
            const uint64_t off = (uint64_t)((__int128)load_value(src + i * elem, type) - mn);
            const uint64_t bit = i * bw;
            for (uint32_t b = 0; b < bw;) {  // byte-wise, LSB first
                const uint64_t pos = bit + b;
                const uint32_t take = std::min<uint32_t>(bw - b, 8 - (uint32_t)(pos & 7));
                pk[pos >> 3] |= (uint8_t)(((off >> b) & ((1u << take) - 1)) << (pos & 7));
                b += take;
            }
        }
    return total;
}

// ------------------------------------------------------------------------------------ RLE
uint64_t rle_bound(uint64_t rows, uint32_t elem) {
    return sizeof(RleHeader) + ((rows * elem + 3) & ~3ull) + 4 * rows;
}

uint64_t rle_encode_page(const uint8_t *src, uint64_t rows, uint32_t elem, uint8_t *dst) {
    uint32_t nruns = 0;
    for (uint64_t i = 0; i < rows; ++i)
        if (i == 0 || memcmp(src + i * elem, src + (i - 1) * elem, elem)) ++nruns;
    RleHeader h{};
    h.nruns = nruns;
    h.n = (uint32_t)rows;
    h.elem = (uint8_t)elem;
    memcpy(dst, &h, sizeof(h));
    uint64_t vbytes = ((uint64_t)nruns * elem + 3) & ~3ull;
    uint8_t *vals = dst + sizeof(h);
    memset(vals, 0, vbytes);
    uint32_t *ends = reinterpret_cast<uint32_t *>(vals + vbytes);
    uint32_t k = 0;
    for (uint64_t i = 0; i < rows; ++i) {
        if (i > 0 && memcmp(src + i * elem, src + (i - 1) * elem, elem)) {
            ends[k] = (uint32_t)i;
            ++k;
        }
        if (i == 0 || memcmp(src + i * elem, src + (i - 1) * elem, elem))
            memcpy(vals + (uint64_t)k * elem, src + i * elem, elem);
    }
    if (rows) ends[k] = (uint32_t)rows;
    return sizeof(h) + vbytes + 4ull * nruns;
}

// ------------------------------------------------------------------------------------ column
uint32_t lz4_unit_of(const golap_encode_opts &o, uint32_t elem) {
    if (o.lz4_unit == 1 || o.lz4_unit == 2 || o.lz4_unit == 4) return std::min(o.lz4_unit, std::max(elem, 1u));
    return std::min(elem, 4u);  // the element width; 8-byte elements in 4-byte units
}
namespace {
struct PageOut {
    uint64_t pos;    // offset in the worker's buffer
    uint32_t csize;
    uint8_t codec;   // codec | PAGE_LITERAL_ONLY
};

// Codes one page; for AUTO the smallest of LZ4 / DBP / RLE / FOR wins (ties keep the earlier).
void code_page(const uint8_t *src, uint64_t rows, uint32_t type, uint32_t elem, uint32_t codec,
               uint32_t min_match, uint32_t unit, int64_t *ht, int64_t base, std::vector<uint8_t> &tmp,
               std::vector<uint8_t> &buf, PageOut &po) {
    const uint64_t nbytes = rows * elem;
    auto emit = [&](const uint8_t *p, uint64_t sz, uint8_t c) {
        po.pos = buf.size();
        po.csize = (uint32_t)sz;
        po.codec = c;
        buf.insert(buf.end(), p, p + sz);
    };
    if (codec == RAW) { emit(src, nbytes, RAW); return; }
    uint64_t best = UINT64_MAX;
    uint8_t best_codec = RAW;
    std::vector<uint8_t> &cand = tmp;
    std::vector<uint8_t> keep;
    auto consider = [&](uint64_t sz, uint8_t c) {
        if (sz && sz < best) {
            best = sz;
            best_codec = c;
            keep.assign(cand.begin(), cand.begin() + sz);
        }
    };
    // LZ4: a block of one literal run stays a plain LZ4 page (read in place); any other block
    // becomes an LZ4I page, the block behind its parse index, unless it is too large for one.
    std::vector<uint32_t> toks;
    auto lz4_candidate = [&]() {
        constexpr uint64_t HDR_ROOM = (4 + 2 * (LZ4I_MAX_SEGS - 1) + 15) & ~15u;  // the largest LZ4I header
        cand.resize(HDR_ROOM + lz4_bound(nbytes) + 16);
        uint8_t *blk = cand.data() + HDR_ROOM;
        const uint64_t sz = lz4_encode_block(src, nbytes, blk, min_match, ht, base, &toks, unit);
        // a block with a match can have exactly the literal-only size: parse, do not compare sizes
        if (lz4_is_literal_only(blk, sz, nbytes)) {
            memmove(cand.data(), blk, sz);
            consider(sz, (uint8_t)(LZ4 | PAGE_LITERAL_ONLY));
            return;
        }
        uint64_t pz = lz4i_make_page(blk, sz, toks, cand.data(), unit);
        if (pz) {
            consider(pz, LZ4I);
        } else {
            memmove(cand.data(), blk, sz);
            consider(sz, LZ4);
        }
    };
    if (codec == LZ4 || codec == GOLAP_CODEC_AUTO) lz4_candidate();
    if (codec == DBP || codec == GOLAP_CODEC_AUTO) {
        cand.resize(dbp_bound(rows));
        const uint64_t sz = dbp_encode_page(src, rows, type, cand.data());
        // bw == 0: every row is base + r * min_delta
        consider(sz, sz && cand[20] == 0 ? (uint8_t)(DBP | PAGE_POINT) : (uint8_t)DBP);
    }
    if (codec == RLE || codec == GOLAP_CODEC_AUTO) {
        cand.resize(rle_bound(rows, elem));
        consider(rle_encode_page(src, rows, elem, cand.data()), (uint8_t)(RLE | PAGE_POINT));
    }
    if (codec == golap_fmt::FOR || codec == GOLAP_CODEC_AUTO) {
        cand.resize(for_bound(rows));
        consider(for_encode_page(src, rows, type, cand.data()), (uint8_t)(golap_fmt::FOR | PAGE_POINT));
    }
    if (best == UINT64_MAX) lz4_candidate();  // DBP / FOR requested but the page's deltas or range are too wide
    emit(keep.data(), best, best_codec);
}
}  // namespace

int encode_column(const void *values, uint64_t n_rows, uint32_t type, const golap_encode_opts &o,
                  golap_host_column &out, std::string &err) {
    const uint32_t elem = type_size(type);
    if (!elem) { err = "encode: unknown column type"; return GOLAP_EINVAL; }
    if (!o.page_rows || o.page_rows % 64 || !o.chunk_rows || o.chunk_rows % o.page_rows) {
        err = "encode: page_rows must be a positive multiple of 64 and divide chunk_rows";
        return GOLAP_EINVAL;
    }
    if ((uint64_t)o.page_rows * elem > (1u << 20)) { err = "encode: page larger than 1 MiB"; return GOLAP_EINVAL; }
    if (o.codec != RAW && o.codec != LZ4 && o.codec != DBP && o.codec != RLE && o.codec != golap_fmt::FOR &&
        o.codec != GOLAP_CODEC_AUTO) { err = "encode: unknown codec"; return GOLAP_EINVAL; }
    if (n_rows && !values) { err = "encode: null values"; return GOLAP_EINVAL; }
    const uint8_t *src = static_cast<const uint8_t *>(values);

    out = golap_host_column{};
    out.type = type;
    out.elem = elem;
    out.page_rows = o.page_rows;
    out.chunk_rows = o.chunk_rows;
    out.n_rows = n_rows;
    out.n_pages = (n_rows + o.page_rows - 1) / o.page_rows;
    out.n_chunks = (n_rows + o.chunk_rows - 1) / o.chunk_rows;
    out.page_off.resize(out.n_pages);
    out.page_csize.resize(out.n_pages);
    out.page_codec.resize(out.n_pages);
    out.chunk_min.assign(out.n_chunks, std::numeric_limits<int64_t>::max());
    out.chunk_max.assign(out.n_chunks, std::numeric_limits<int64_t>::min());

    uint32_t threads = o.threads ? o.threads : std::max(1u, std::thread::hardware_concurrency());
    threads = (uint32_t)std::min<uint64_t>(threads, std::max<uint64_t>(1, out.n_pages));
    std::vector<std::vector<uint8_t>> bufs(threads);
    std::vector<std::vector<PageOut>> pos(threads);
    std::vector<uint64_t> first(threads + 1);
    for (uint32_t t = 0; t <= threads; ++t) first[t] = out.n_pages * t / threads;

    auto work = [&](uint32_t t) {
        std::vector<int64_t> ht(1u << LZ4_HASH_LOG, -1);
        std::vector<uint8_t> tmp;
        auto &buf = bufs[t];
        buf.reserve((first[t + 1] - first[t]) * (uint64_t)o.page_rows * elem * 101 / 100 + 64);
        for (uint64_t p = first[t]; p < first[t + 1]; ++p) {
            uint64_t r0 = p * o.page_rows;
            uint64_t rows = std::min<uint64_t>(o.page_rows, n_rows - r0);
            PageOut po;
            code_page(src + r0 * elem, rows, type, elem, o.codec, o.lz4_min_match, lz4_unit_of(o, elem), ht.data(),
                      (int64_t)(r0 * elem), tmp, buf, po);
            pos[t].push_back(po);
        }
        // zone maps of the chunks whose first row lies in this worker's pages
        for (uint64_t c = first[t] * o.page_rows / o.chunk_rows; c < out.n_chunks; ++c) {
            uint64_t r0 = c * o.chunk_rows;
            if (r0 < first[t] * o.page_rows) continue;
            if (r0 >= first[t + 1] * o.page_rows) break;
            uint64_t r1 = std::min<uint64_t>(n_rows, r0 + o.chunk_rows);
            int64_t mn = std::numeric_limits<int64_t>::max(), mx = std::numeric_limits<int64_t>::min();
            for (uint64_t r = r0; r < r1; ++r) {
                int64_t v = load_value(src + r * elem, type);
                mn = std::min(mn, v);
                mx = std::max(mx, v);
            }
            out.chunk_min[c] = mn;
            out.chunk_max[c] = mx;
        }
    };
    std::vector<std::thread> ts;
    for (uint32_t t = 1; t < threads; ++t) ts.emplace_back(work, t);
    work(0);
    for (auto &th : ts) th.join();

    for (uint32_t t = 0; t < threads; ++t) {
        for (uint64_t i = 0; i < pos[t].size(); ++i) {
            out.page_csize[first[t] + i] = pos[t][i].csize;
            out.page_codec[first[t] + i] = pos[t][i].codec;
        }
    }
    out.comp.assign(compute_page_offsets(out), 0);
    auto place = [&](uint32_t t) {
        for (uint64_t i = 0; i < pos[t].size(); ++i) {
            uint64_t p = first[t] + i;
            memcpy(out.comp.data() + out.page_off[p], bufs[t].data() + pos[t][i].pos, pos[t][i].csize);
        }
    };
    ts.clear();
    for (uint32_t t = 1; t < threads; ++t) ts.emplace_back(place, t);
    place(0);
    for (auto &th : ts) th.join();
    return GOLAP_OK;
}

bool lz4_is_literal_only(const uint8_t *blk, uint64_t csize, uint64_t n) {
    const uint64_t hdr = lz4_literal_header_bytes((uint32_t)n);
    if (n > 0xffffffffull || csize != hdr + n) return false;
    uint64_t lit = blk[0] >> 4, ip = 1;
    if (lit == 15) {
        uint8_t b;
        do {
            if (ip >= hdr) return false;
            b = blk[ip++];
            lit += b;
        } while (b == 255);
    }
    return ip == hdr && lit == n;
}

uint64_t compute_page_offsets(golap_host_column &col) {
    uint64_t cur = 0;
    col.payload_bytes = 0;
    for (uint64_t p = 0; p < col.n_pages; ++p) {
        uint64_t lead = 0;
        if (col.page_codec[p] & PAGE_LITERAL_ONLY) {
            uint64_t rows = std::min<uint64_t>(col.page_rows, col.n_rows - p * col.page_rows);
            lead = lz4_literal_header_bytes((uint32_t)(rows * col.elem));
        }
        // each page at a 16-byte boundary; a literal-only LZ4 page shifted so that its payload
        // (after the token and length bytes) is 16-byte aligned
        uint64_t start = ((cur + lead + PAGE_ALIGN - 1) / PAGE_ALIGN) * PAGE_ALIGN - lead;
        if (start < cur) start += PAGE_ALIGN;
        col.page_off[p] = start;
        col.payload_bytes += col.page_csize[p];
        cur = start + col.page_csize[p];
    }
    // one spare 16-byte block at the end: vector loads of a page tail stay inside the buffer
    return ((cur + PAGE_ALIGN - 1) / PAGE_ALIGN) * PAGE_ALIGN + PAGE_ALIGN;
}

void layout_pages(golap_host_column &col, const std::vector<const uint8_t *> &pages) {
    col.comp.assign(compute_page_offsets(col), 0);
    for (uint64_t p = 0; p < col.n_pages; ++p)
        if (col.page_csize[p]) memcpy(col.comp.data() + col.page_off[p], pages[p], col.page_csize[p]);
}

}  // namespace golap_host
