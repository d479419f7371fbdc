// Paged compressed column format shared by the host encoder and the gfx950 decoders.
//
// A column of n_rows values of `elem` bytes is cut into chunks of chunk_rows rows (the
// reference's BlockInfo unit, core/comp.cuh:23-32, and its zone-map unit, core/metadata.cuh:
// 110-160). Each chunk is cut into pages of page_rows rows; page p covers rows
// [p*page_rows, min(n_rows, (p+1)*page_rows)) and is coded independently. Pages play the role of
// nvcomp's 64 KiB sub-blocks (nvchunk, core/comp.cuh:40), but are sized in rows, so page p of
// every column of a table covers the same rows and one workgroup can consume a row tile of all
// columns without a second pass.
//
// Page codecs:
//   RAW  (0)  the values themselves.
//   LZ4  (1)  one LZ4 block (lz4_Block_format.md): token | lit-len bytes | literals | offset |
//             match-len bytes; the last sequence has literals only.
//   DBP  (2)  frame-of-reference delta bit-packing (sorted or slowly varying integers).
//   RLE  (3)  run-length encoding (value, exclusive end row) pairs.
//   FOR  (4)  frame of reference: v[r] = base + the r-th bw-bit field of a little-endian bit
//             stream (unsorted low-range integers); every row is readable on its own.
//   LZ4I (5)  one LZ4 block behind a parse index, so that the 64 lanes of a wave can walk the
//             block's sequence list from 64 places at once: a 4-byte header {u16 nseq, the
//             block's sequences (the last one literals only); u8 seg_log2; u8 0}, then
//             u16 tok[nseg - 1] with nseg = ceil(nseq / 2^seg_log2), tok[s - 1] being the block
//             offset of the token of sequence s * 2^seg_log2, zero bytes up to a multiple of 16,
//             then the block (below 64 KiB). The block is standard LZ4 on its own; the index
//             only says where the sequences start, and a reader rejects a page whose index
//             disagrees with the block's own parse.
// Flag bits above the codec: PAGE_POINT (0x40, see below) and PAGE_LITERAL_ONLY (0x80), which
// marks an LZ4 page made of a single literal run; such a page's payload starts lz4_literal_header_bytes(n_bytes) bytes after the page start and the
// encoder places it 16-byte aligned in the column buffer, so a scan reads it in place.
#pragma once
#include <cstdint>

namespace golap_fmt {

enum Codec : uint8_t { RAW = 0, LZ4 = 1, DBP = 2, RLE = 3, FOR = 4, LZ4I = 5 };
constexpr uint8_t PAGE_CODEC_MASK = 0x0f;
constexpr uint8_t PAGE_LITERAL_ONLY = 0x80;
// Bit 6 (PAGE_POINT): any row of the page can be read without decoding the page: every RLE
// page (binary search over the run ends), every FOR page (one unaligned 8-byte read) and DBP
// pages with bw == 0 (v[r] = base + r*min_delta).
constexpr uint8_t PAGE_POINT = 0x40;

constexpr uint32_t LZ4_MINMATCH = 4;
constexpr uint32_t LZ4_LASTLITERALS = 5;  // the last 5 bytes of a block are literals
constexpr uint32_t LZ4_MFLIMIT = 12;      // the last match starts >= 12 bytes before the end
constexpr uint32_t LZ4_MAX_OFFSET = 65535;

// Bytes of token + literal-length extension for a literal-only block of n bytes.
__host__ __device__ inline uint32_t lz4_literal_header_bytes(uint32_t n) {
    return n < 15 ? 1u : 2u + (n - 15) / 255u;
}

// LZ4I page header bytes for nseq sequences indexed every 2^seg_log2 sequences.
constexpr uint32_t LZ4I_MAX_BLOCK = 65535;  // token offsets fit 16 bits
constexpr uint32_t LZ4I_MIN_SEG_LOG2 = 3;
constexpr uint32_t LZ4I_MAX_SEGS = 256;
// Pages of at most this many sequences are decoded in order by one wave (few, long sequences:
// cheapest that way); more go to the workgroup-cooperative decode.
constexpr uint32_t LZ4I_SERIAL_MAX = 32;     // the encoder's choice: one segment per thread of a workgroup
__host__ __device__ inline uint32_t lz4i_segments(uint32_t nseq, uint32_t seg_log2) {
    return (uint32_t)(((uint64_t)nseq + (1ull << seg_log2) - 1) >> seg_log2);
}
__host__ __device__ inline uint32_t lz4i_header_bytes(uint32_t nseq, uint32_t seg_log2) {
    const uint32_t nseg = lz4i_segments(nseq, seg_log2);
    return (4 + 2 * (nseg ? nseg - 1 : 0) + 15) & ~15u;
}
// The encoder's segment length: the shortest power of two, at least 8 sequences, that cuts
// the block into at most LZ4I_MAX_SEGS segments.
__host__ __device__ inline uint32_t lz4i_seg_log2(uint32_t nseq) {
    uint32_t l = LZ4I_MIN_SEG_LOG2;
    while (((uint64_t)LZ4I_MAX_SEGS << l) < nseq) ++l;
    return l;
}

// DBP page header (32 bytes, little endian) followed by (n-1) packed deltas.
struct DbpHeader {
    int64_t base;       // v[0]
    int64_t min_delta;  // min over i>=1 of v[i]-v[i-1]
    uint32_t n;         // rows in the page
    uint8_t bw;         // bits per packed delta, 0..32
    uint8_t elem;       // bytes per value
    uint8_t pad[6];
};
static_assert(sizeof(DbpHeader) == 32, "DbpHeader layout");

// RLE page header (16 bytes) followed by values[nruns] (elem bytes each, padded to 4) and
// uint32 ends[nruns].
struct RleHeader {
    uint32_t nruns;
    uint32_t n;
    uint8_t elem;
    uint8_t pad[7];
};
static_assert(sizeof(RleHeader) == 16, "RleHeader layout");

// FOR page header (32 bytes, the DbpHeader layout with min_delta unused) followed by
// ceil(n*bw/8) bytes of bw-bit fields (field r at bits [r*bw, r*bw + bw) of the byte stream,
// LSB first) and 8 zero bytes, so an 8-byte read at byte (r*bw)/8 never leaves the page.
constexpr uint32_t FOR_MAX_BW = 56;
__host__ __device__ inline uint64_t for_page_bytes(uint64_t n, uint32_t bw) { return 32 + (n * bw + 7) / 8 + 8; }

constexpr uint32_t PAGE_ALIGN = 16;

}  // namespace golap_fmt
