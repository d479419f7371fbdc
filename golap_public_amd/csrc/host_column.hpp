// Host-side encoded column (the CompInfo/BlockInfo catalog of core/comp.cuh:23-78 plus the
// MinMaxMeta zone map of core/metadata.cuh:110-221, in one object).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/golap_hip.h"

struct golap_host_column {
    uint32_t type = 0, elem = 0, page_rows = 0, chunk_rows = 0;
    uint64_t n_rows = 0, n_pages = 0, n_chunks = 0, payload_bytes = 0;
    std::vector<uint8_t> comp;
    std::vector<uint64_t> page_off;
    std::vector<uint32_t> page_csize;
    std::vector<uint8_t> page_codec;
    std::vector<int64_t> chunk_min, chunk_max;
};

namespace golap_host {

uint32_t type_size(uint32_t type);
bool type_signed(uint32_t type);
int64_t load_value(const uint8_t *p, uint32_t type);

// LZ4 block encoder; returns the block size. dst must hold lz4_bound(n) bytes.
uint64_t lz4_bound(uint64_t n);
// hash_table: 1<<LZ4_HASH_LOG absolute positions; base: absolute position of src (must grow
// from block to block when the table is reused, so stale entries fall below base).
// tokens, when given, receives the block offset of every sequence's token, in order.
// unit (1, 2 or 4): sequences start, and literal runs, matches and offsets measure, in whole units.
uint64_t lz4_encode_block(const uint8_t *src, uint64_t n, uint8_t *dst, uint32_t min_match,
                          int64_t *hash_table, int64_t base, std::vector<uint32_t> *tokens = nullptr,
                          uint32_t unit = 1);
// The LZ4 unit of a column of elem-byte values under opts (lz4_unit 0: the element width, 8-byte
// elements in 4-byte units).
uint32_t lz4_unit_of(const golap_encode_opts &opts, uint32_t elem);
// An LZ4I page (format.hpp) from a block of bsize bytes and its token offsets; returns the
// page size, 0 if the block does not fit the index (64 KiB or more). out must hold
// lz4i_header_bytes(tokens.size(), lz4i_seg_log2(tokens.size())) + bsize bytes.
// unit: the block's LZ4 unit, recorded in the header.
uint64_t lz4i_make_page(const uint8_t *blk, uint64_t bsize, const std::vector<uint32_t> &tokens, uint8_t *out,
                        uint32_t unit = 1);
constexpr uint32_t LZ4_HASH_LOG = 14;

// Delta-bitpack page; returns 0 if the value range needs more than 32 bits per delta.
uint64_t dbp_bound(uint64_t rows);
uint64_t dbp_encode_page(const uint8_t *src, uint64_t rows, uint32_t type, uint8_t *dst);
// RLE page; returns the page size (always succeeds).
uint64_t rle_bound(uint64_t rows, uint32_t elem);
uint64_t rle_encode_page(const uint8_t *src, uint64_t rows, uint32_t elem, uint8_t *dst);

int encode_column(const void *values, uint64_t n_rows, uint32_t type, const golap_encode_opts &o,
                  golap_host_column &out, std::string &err);

}  // namespace golap_host

namespace golap_host {
// Places coded pages into the column buffer (16-byte aligned pages; literal-only LZ4 payloads
// aligned) and fills page_off / payload_bytes. pages[p] holds page_csize[p] bytes.
void layout_pages(golap_host_column &col, const std::vector<const uint8_t *> &pages);
// Fills page_off / payload_bytes from page_csize / page_codec; returns the buffer size.
uint64_t compute_page_offsets(golap_host_column &col);
// Is this LZ4 block exactly one literal run of n bytes (a single sequence whose token and
// length bytes announce n literals)? Size alone does not tell: a block with one short match
// can have the same size.
bool lz4_is_literal_only(const uint8_t *blk, uint64_t csize, uint64_t n);
}  // namespace golap_host
