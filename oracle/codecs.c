/*
 * CPU page decoders (TEST INFRASTRUCTURE: the oracle, never the product).
 *
 * The reference's CPU codec calls LZ4_decompress_safe_partial(in, out, comp, chunk, chunk)
 * (core/comp_cpu.hpp:22-32) per chunk; the LZ4 library is an un-vendored dependency
 * (lz4 @ce8ee024, scripts/init.sh:32). oracle_lz4_decode restates the published LZ4 block
 * format (lz4/doc/lz4_Block_format.md): sequences of token | literal-length bytes | literals |
 * 2-byte little-endian offset | match-length bytes, the last sequence carrying literals only.
 * Every read and write is bounds checked, as in the library's "safe" decoder.
 */
#include "oracle.h"
#include <stdlib.h>
#include <string.h>

int64_t oracle_lz4_decode(const uint8_t *src, uint64_t csize, uint8_t *dst, uint64_t cap) {
    uint64_t ip = 0, op = 0;
    if (csize == 0) return cap == 0 ? 0 : -1;
    for (;;) {
        if (ip >= csize) return -1;
        uint32_t token = src[ip++];
        uint64_t lit = token >> 4;
        if (lit == 15) {
            uint32_t b;
            do {
                if (ip >= csize) return -1;
                b = src[ip++];
                lit += b;
            } while (b == 255);
        }
        if (ip + lit > csize || op + lit > cap) return -1;
        memcpy(dst + op, src + ip, lit);
        ip += lit;
        op += lit;
        if (ip == csize) break; /* last sequence: literals only */
        if (ip + 2 > csize) return -1;
        uint64_t off = (uint64_t)src[ip] | ((uint64_t)src[ip + 1] << 8);
        ip += 2;
        if (off == 0 || off > op) return -1;
        uint64_t ml = token & 15;
        if (ml == 15) {
            uint32_t b;
            do {
                if (ip >= csize) return -1;
                b = src[ip++];
                ml += b;
            } while (b == 255);
        }
        ml += 4; /* MINMATCH */
        if (op + ml > cap) return -1;
        for (uint64_t k = 0; k < ml; ++k) dst[op + k] = dst[op - off + k]; /* overlap-safe */
        op += ml;
    }
    return (int64_t)op;
}

/* Token offsets of an LZ4 block's sequences by a serial parse (the parse of
 * oracle_lz4_decode): returns the number of sequences, -1 if the block does not parse. */
static int64_t lz4_token_offsets(const uint8_t *src, uint64_t csize, uint32_t *tok, uint64_t cap_tok) {
    uint64_t ip = 0, n = 0;
    if (csize == 0) return -1;
    for (;;) {
        if (ip >= csize) return -1;
        if (n < cap_tok) tok[n] = (uint32_t)ip;
        ++n;
        uint32_t token = src[ip++];
        uint64_t lit = token >> 4;
        if (lit == 15) {
            uint32_t b;
            do {
                if (ip >= csize) return -1;
                b = src[ip++];
                lit += b;
            } while (b == 255);
        }
        if (ip + lit > csize) return -1;
        ip += lit;
        if (ip == csize) break;
        if (ip + 2 > csize) return -1;
        ip += 2;
        if ((token & 15) == 15) {
            uint32_t b;
            do {
                if (ip >= csize) return -1;
                b = src[ip++];
            } while (b == 255);
        }
    }
    return (int64_t)n;
}

/* Does every sequence of an LZ4 block (parsed as above) measure its literal run, its offset and
 * its match in whole units of `unit` bytes? */
static int lz4_in_units(const uint8_t *src, uint64_t csize, uint32_t unit) {
    uint64_t ip = 0;
    for (;;) {
        uint32_t token = src[ip++];
        uint64_t lit = token >> 4;
        if (lit == 15) {
            uint32_t b;
            do {
                b = src[ip++];
                lit += b;
            } while (b == 255);
        }
        if (lit % unit) return 0;
        ip += lit;
        if (ip == csize) return 1;
        const uint32_t off = (uint32_t)src[ip] | ((uint32_t)src[ip + 1] << 8);
        ip += 2;
        uint64_t ml = token & 15;
        if (ml == 15) {
            uint32_t b;
            do {
                b = src[ip++];
                ml += b;
            } while (b == 255);
        }
        if (off % unit || (ml + 4) % unit) return 0;
    }
}

/* LZ4I page (this build's page format, golap_public_amd/csrc/format.hpp): header {u16 nseq,
 * u8 seg_log2, u8 unit_log2}, u16 tok[nseg - 1] (block offset of the token of sequence
 * s << seg_log2), zero bytes to a multiple of 16, then a standard LZ4 block below 64 KiB. The page
 * is valid iff the block decodes (oracle_lz4_decode) and the header agrees with the block's
 * serial parse: nseq sequences, every indexed offset at the token of its sequence, and with
 * unit_log2 = u in 0..2 every literal run, offset and match a multiple of 2^u bytes. */
int64_t oracle_lz4i_decode(const uint8_t *src, uint64_t csize, uint8_t *dst, uint64_t cap) {
    if (csize < 4) return -1;
    const uint32_t nseq = (uint32_t)src[0] | ((uint32_t)src[1] << 8), sl = src[2];
    if (nseq == 0 || sl > 15 || src[3] > 2) return -1;
    const uint32_t nseg = (uint32_t)(((uint64_t)nseq + (1ull << sl) - 1) >> sl);
    const uint64_t hb = (4 + 2 * (uint64_t)(nseg - 1) + 15) & ~15ull;
    if (hb > csize || csize - hb > 65535) return -1;
    const uint8_t *blk = src + hb;
    const uint64_t bs = csize - hb;
    uint32_t *tok = (uint32_t *)malloc(((uint64_t)nseq + 1) * sizeof(uint32_t));
    if (!tok) return -1;
    int64_t got = lz4_token_offsets(blk, bs, tok, (uint64_t)nseq + 1);
    int ok = got == (int64_t)nseq;
    for (uint32_t s = 1; ok && s < nseg; ++s)
        ok = ((uint32_t)src[4 + 2 * (s - 1)] | ((uint32_t)src[5 + 2 * (s - 1)] << 8)) == tok[(uint64_t)s << sl];
    for (uint64_t i = 4 + 2 * (uint64_t)(nseg - 1); ok && i < hb; ++i) ok = src[i] == 0;
    free(tok);
    if (!ok || (src[3] && !lz4_in_units(blk, bs, 1u << src[3]))) return -1;
    return oracle_lz4_decode(blk, bs, dst, cap);
}

static inline uint64_t rd_le(const uint8_t *p, uint32_t n) {
    uint64_t v = 0;
    for (uint32_t i = 0; i < n; ++i) v |= (uint64_t)p[i] << (8 * i);
    return v;
}
static inline void wr_le(uint8_t *p, uint64_t v, uint32_t n) {
    for (uint32_t i = 0; i < n; ++i) p[i] = (uint8_t)(v >> (8 * i));
}

/* Delta-bitpack page: 32-byte header {i64 base, i64 min_delta, u32 n, u8 bw, u8 elem, 6 pad},
 * then (n-1) packed values of bw bits (little-endian bit stream in u32 words);
 * v[0] = base, v[i] = v[i-1] + min_delta + packed[i-1] (mod 2^64), low elem bytes stored. */
int64_t oracle_dbp_decode(const uint8_t *src, uint64_t csize, uint8_t *dst, uint64_t cap) {
    if (csize < 32) return -1;
    uint64_t base = rd_le(src, 8), mind = rd_le(src + 8, 8);
    uint32_t n = (uint32_t)rd_le(src + 16, 4), bw = src[20], elem = src[21];
    if (bw > 32 || (elem != 1 && elem != 2 && elem != 4 && elem != 8)) return -1;
    if ((uint64_t)n * elem > cap) return -1;
    uint64_t words = n ? (((uint64_t)(n - 1) * bw + 31) / 32) : 0;
    if (32 + words * 4 > csize) return -1;
    const uint8_t *pk = src + 32;
    uint64_t v = base;
    for (uint32_t i = 0; i < n; ++i) {
        if (i > 0) {
            uint64_t bit = (uint64_t)(i - 1) * bw, p = 0;
            if (bw) {
                uint64_t w = bit / 32;
                uint64_t lo = rd_le(pk + 4 * w, 4);
                uint64_t hi = (w + 1 < words) ? rd_le(pk + 4 * (w + 1), 4) : 0;
                uint64_t both = lo | (hi << 32);
                p = (both >> (bit % 32)) & ((bw == 64) ? ~0ull : ((1ull << bw) - 1));
            }
            v = v + mind + p;
        }
        wr_le(dst + (uint64_t)i * elem, v, elem);
    }
    return n;
}

/* FOR page: the 32-byte header of a delta-bitpack page (base = frame of reference, min_delta
 * unused), then n fields of bw bits (bw <= 56), field r at bits [r*bw, r*bw+bw) of a
 * little-endian byte stream, then 8 zero bytes; v[r] = base + field r (mod 2^64). */
int64_t oracle_for_decode(const uint8_t *src, uint64_t csize, uint8_t *dst, uint64_t cap) {
    if (csize < 32) return -1;
    uint64_t base = rd_le(src, 8);
    uint32_t n = (uint32_t)rd_le(src + 16, 4), bw = src[20], elem = src[21];
    if (bw > 56 || (elem != 1 && elem != 2 && elem != 4 && elem != 8)) return -1;
    if ((uint64_t)n * elem > cap || 32 + ((uint64_t)n * bw + 7) / 8 + 8 > csize) return -1;
    const uint8_t *pk = src + 32;
    for (uint32_t r = 0; r < n; ++r) {
        uint64_t v = 0;
        for (uint32_t b = 0; b < bw; ++b) {  /* bit by bit: the plainest reading of the format */
            uint64_t pos = (uint64_t)r * bw + b;
            v |= (uint64_t)((pk[pos >> 3] >> (pos & 7)) & 1u) << b;
        }
        wr_le(dst + (uint64_t)r * elem, base + v, elem);
    }
    return n;
}

/* RLE page: 16-byte header {u32 nruns, u32 n, u8 elem, 7 pad}, values[nruns] of elem bytes
 * padded to 4, then u32 ends[nruns] (exclusive end row of each run, ends[nruns-1] == n). */
int64_t oracle_rle_decode(const uint8_t *src, uint64_t csize, uint8_t *dst, uint64_t cap) {
    if (csize < 16) return -1;
    uint32_t nruns = (uint32_t)rd_le(src, 4), n = (uint32_t)rd_le(src + 4, 4), elem = src[8];
    if (elem != 1 && elem != 2 && elem != 4 && elem != 8) return -1;
    uint64_t vbytes = ((uint64_t)nruns * elem + 3) & ~3ull;
    if (16 + vbytes + 4ull * nruns > csize || (uint64_t)n * elem > cap) return -1;
    const uint8_t *vals = src + 16, *ends = src + 16 + vbytes;
    uint32_t row = 0;
    for (uint32_t k = 0; k < nruns; ++k) {
        uint32_t e = (uint32_t)rd_le(ends + 4 * k, 4);
        if (e < row || e > n) return -1;
        for (; row < e; ++row) memcpy(dst + (uint64_t)row * elem, vals + (uint64_t)k * elem, elem);
    }
    if (row != n) return -1;
    return n;
}

/* codec: low 4 bits are the codec id (0 RAW, 1 LZ4, 2 DBP, 3 RLE, 4 FOR, 5 LZ4I); bit 7
 * (literal-only LZ4) and bit 6 (point-readable) are layout hints only. */
int64_t oracle_page_decode(uint32_t codec, const uint8_t *src, uint64_t csize, uint8_t *dst,
                           uint64_t cap) {
    switch (codec & 0x0f) {
    case 0:
        if (csize > cap) return -1;
        memcpy(dst, src, csize);
        return (int64_t)csize;
    case 1: return oracle_lz4_decode(src, csize, dst, cap);
    case 2: return oracle_dbp_decode(src, csize, dst, cap);
    case 3: return oracle_rle_decode(src, csize, dst, cap);
    case 4: return oracle_for_decode(src, csize, dst, cap);
    case 5: return oracle_lz4i_decode(src, csize, dst, cap);
    default: return -1;
    }
}

int oracle_decode_column(const uint8_t *comp, const uint64_t *page_off, const uint32_t *page_csize,
                         const uint8_t *page_codec, uint64_t n_rows, uint32_t elem,
                         uint32_t page_rows, uint32_t chunk_rows, uint8_t *dst) {
    if (!page_rows || !chunk_rows || chunk_rows % page_rows) return -1;
    uint64_t n_pages = (n_rows + page_rows - 1) / page_rows;
    uint64_t n_chunks = (n_rows + chunk_rows - 1) / chunk_rows;
    memset(dst, 0, n_chunks * (uint64_t)chunk_rows * elem);
    for (uint64_t p = 0; p < n_pages; ++p) {
        uint64_t r0 = p * page_rows;
        uint64_t rows = (n_rows - r0 < page_rows) ? n_rows - r0 : page_rows;
        int64_t got = oracle_page_decode(page_codec[p], comp + page_off[p], page_csize[p],
                                         dst + r0 * elem, rows * elem);
        uint32_t cd = page_codec[p] & 0x0f;
        uint64_t want = cd == 0 || cd == 1 || cd == 5 ? rows * elem : rows;
        if (got < 0 || (uint64_t)got != want) return -1;
    }
    return 0;
}
