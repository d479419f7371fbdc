// Wavefront-cooperative page decoders for gfx950 (wave64).
//
// These replace the nvcomp decompression the reference calls per chunk
// (ChunkDecompressor::step -> nvcomp decompress, core/comp.cuh:171-195). One wavefront decodes
// one page: the sequence headers are parsed redundantly by all 64 lanes (every lane reads the
// same byte, a broadcast load), and every copy is spread over the lanes. The destination is an
// LDS window of the page's decoded size, so matches copy LDS to LDS. Every read is checked
// against the page size and every write against the window size; a malformed page returns -1
// instead of touching memory outside its buffers.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "format.hpp"

namespace golap_dev {

constexpr int WAVE = 64;

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// Decode one LZ4 block src[0, csize) into dst[0, cap). Returns bytes written or -1.
// All lanes of the wave must call it with the same arguments.
__device__ inline int64_t lz4_decode_wave(const uint8_t *__restrict__ src, uint32_t csize,
                                          uint8_t *__restrict__ dst, uint32_t cap) {
    const uint32_t lane = lane_id();
    uint32_t ip = 0, op = 0;
    if (csize == 0) return cap == 0 ? 0 : -1;
    for (;;) {
        if (ip >= csize) return -1;
        const uint32_t token = src[ip++];
        uint32_t lit = token >> 4;
        if (lit == 15) {
            uint32_t b;
            do {
                if (ip >= csize) return -1;
                b = src[ip++];
                lit += b;
            } while (b == 255 && lit < (1u << 30));
        }
        if (lit > csize - ip || lit > cap - op) return -1;
        // literals: global -> window
        uint32_t k = lane;
        for (; k + 3 * WAVE < lit; k += 4 * WAVE) {
            uint8_t a0 = src[ip + k], a1 = src[ip + k + WAVE], a2 = src[ip + k + 2 * WAVE],
                    a3 = src[ip + k + 3 * WAVE];
            dst[op + k] = a0;
            dst[op + k + WAVE] = a1;
            dst[op + k + 2 * WAVE] = a2;
            dst[op + k + 3 * WAVE] = a3;
        }
        for (; k < lit; k += WAVE) dst[op + k] = src[ip + k];
        ip += lit;
        op += lit;
        if (ip == csize) break;  // last sequence carries literals only
        if (csize - ip < 2) return -1;
        const uint32_t off = (uint32_t)src[ip] | ((uint32_t)src[ip + 1] << 8);
        ip += 2;
        if (off == 0 || off > op) return -1;
        uint32_t ml = token & 15;
        if (ml == 15) {
            uint32_t b;
            do {
                if (ip >= csize) return -1;
                b = src[ip++];
                ml += b;
            } while (b == 255 && ml < (1u << 30));
        }
        ml += golap_fmt::LZ4_MINMATCH;
        if (ml > cap - op) return -1;
        // match: byte j of the match equals byte (j mod off) of dst[op-off, op), which is final
        // before the copy starts, so every lane may copy independently.
        const uint32_t from = op - off;
        if (off >= WAVE) {
            // rounds of 64 bytes never read bytes written in the same round
            for (uint32_t j = lane; j < ml; j += WAVE) dst[op + j] = dst[from + j];
        } else {
            uint32_t j = lane, r = lane % off;
            const uint32_t step = WAVE % off;
            for (; j < ml; j += WAVE) {
                dst[op + j] = dst[from + r];
                r += step;
                if (r >= off) r -= off;
            }
        }
        op += ml;
    }
    return op;
}

__device__ __forceinline__ uint64_t wave_incl_scan_u64(uint64_t v) {
    const uint32_t lane = lane_id();
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) {
        uint64_t t = __shfl_up(v, d, WAVE);
        if (lane >= (uint32_t)d) v += t;
    }
    return v;
}

__device__ __forceinline__ void store_elem(uint8_t *dst, uint64_t v, uint32_t elem) {
    switch (elem) {
    case 1: *dst = (uint8_t)v; break;
    case 2: *reinterpret_cast<uint16_t *>(dst) = (uint16_t)v; break;
    case 4: *reinterpret_cast<uint32_t *>(dst) = (uint32_t)v; break;
    default: *reinterpret_cast<uint64_t *>(dst) = v; break;
    }
}

__device__ __forceinline__ uint32_t ld_u32_unaligned(const uint8_t *p) {
    return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}
__device__ __forceinline__ uint64_t ld_u64_unaligned(const uint8_t *p) {
    return (uint64_t)ld_u32_unaligned(p) | ((uint64_t)ld_u32_unaligned(p + 4) << 32);
}

// Delta-bitpack page -> dst (elem-aligned). Returns rows or -1.
__device__ inline int64_t dbp_decode_wave(const uint8_t *__restrict__ src, uint32_t csize,
                                          uint8_t *__restrict__ dst, uint32_t cap) {
    if (csize < sizeof(golap_fmt::DbpHeader)) return -1;
    const uint64_t base = ld_u64_unaligned(src), mind = ld_u64_unaligned(src + 8);
    const uint32_t n = ld_u32_unaligned(src + 16), bw = src[20], elem = src[21];
    if (bw > 32 || (elem != 1 && elem != 2 && elem != 4 && elem != 8)) return -1;
    if ((uint64_t)n * elem > cap) return -1;
    const uint64_t words = n ? (((uint64_t)(n - 1) * bw + 31) / 32) : 0;
    if (32 + words * 4 > csize) return -1;
    const uint8_t *pk = src + 32;
    const uint32_t lane = lane_id();
    const uint64_t mask = bw == 32 ? 0xffffffffull : ((1ull << bw) - 1);
    uint64_t carry = base;  // value of the row before this round (row 0 adds nothing)
    for (uint32_t r0 = 0; r0 < n; r0 += WAVE) {
        const uint32_t i = r0 + lane;
        uint64_t d = 0;
        if (i > 0 && i < n) {
            uint64_t p = 0;
            if (bw) {
                const uint64_t bit = (uint64_t)(i - 1) * bw, w = bit >> 5;
                const uint64_t lo = ld_u32_unaligned(pk + 4 * w);
                const uint64_t hi = (w + 1 < words) ? ld_u32_unaligned(pk + 4 * (w + 1)) : 0;
                p = ((lo | (hi << 32)) >> (bit & 31)) & mask;
            }
            d = mind + p;
        }
        const uint64_t s = wave_incl_scan_u64(d);
        if (i < n) store_elem(dst + (uint64_t)i * elem, carry + s, elem);
        carry += __shfl(s, WAVE - 1, WAVE);
    }
    return n;
}

// FOR page header check; returns rows or -1.
__device__ __forceinline__ int64_t for_check(const uint8_t *src, uint32_t csize, uint32_t elem_want) {
    if (csize < 32) return -1;
    const uint32_t n = ld_u32_unaligned(src + 16), bw = src[20], elem = src[21];
    if (bw > golap_fmt::FOR_MAX_BW || (elem_want ? elem != elem_want : (elem != 1 && elem != 2 && elem != 4 && elem != 8)))
        return -1;
    if (golap_fmt::for_page_bytes(n, bw) > csize) return -1;
    return n;
}
// Row r of a checked FOR page (payload = src + 32): one unaligned 8-byte read.
__device__ __forceinline__ uint64_t for_value(const uint8_t *payload, uint64_t base, uint32_t bw, uint32_t r) {
    const uint64_t bit = (uint64_t)r * bw;
    const uint64_t x = ld_u64_unaligned(payload + (bit >> 3)) >> (bit & 7);
    return base + (bw ? (x & (~0ull >> (64 - bw))) : 0ull);
}
// FOR page -> dst. Returns rows or -1.
__device__ inline int64_t for_decode_wave(const uint8_t *__restrict__ src, uint32_t csize,
                                          uint8_t *__restrict__ dst, uint32_t cap) {
    const int64_t n = for_check(src, csize, 0);
    if (n < 0) return -1;
    const uint32_t bw = src[20], elem = src[21];
    if ((uint64_t)n * elem > cap) return -1;
    const uint64_t base = ld_u64_unaligned(src);
    for (uint32_t r = lane_id(); r < (uint32_t)n; r += WAVE)
        store_elem(dst + (uint64_t)r * elem, for_value(src + 32, base, bw, r), elem);
    return n;
}

// RLE page -> dst. Returns rows or -1.
__device__ inline int64_t rle_decode_wave(const uint8_t *__restrict__ src, uint32_t csize,
                                          uint8_t *__restrict__ dst, uint32_t cap) {
    if (csize < sizeof(golap_fmt::RleHeader)) return -1;
    const uint32_t nruns = ld_u32_unaligned(src), n = ld_u32_unaligned(src + 4), elem = src[8];
    if (elem != 1 && elem != 2 && elem != 4 && elem != 8) return -1;
    const uint64_t vbytes = ((uint64_t)nruns * elem + 3) & ~3ull;
    if (16 + vbytes + 4ull * nruns > csize || (uint64_t)n * elem > cap) return -1;
    const uint8_t *vals = src + 16, *ends = src + 16 + vbytes;
    const uint32_t lane = lane_id();
    uint32_t row = 0;
    for (uint32_t k = 0; k < nruns; ++k) {
        const uint32_t e = ld_u32_unaligned(ends + 4 * k);
        if (e < row || e > n) return -1;
        uint64_t v = 0;
        for (uint32_t b = 0; b < elem; ++b) v |= (uint64_t)vals[(uint64_t)k * elem + b] << (8 * b);
        for (uint32_t r = row + lane; r < e; r += WAVE) store_elem(dst + (uint64_t)r * elem, v, elem);
        row = e;
    }
    return row == n ? (int64_t)n : -1;
}

// Decodes one page of any codec into dst[0, cap). Returns bytes written or -1.
__device__ inline int64_t page_decode_wave(uint8_t codec, const uint8_t *__restrict__ src,
                                           uint32_t csize, uint8_t *__restrict__ dst, uint32_t cap,
                                           uint32_t elem) {
    switch (codec & golap_fmt::PAGE_CODEC_MASK) {
    case golap_fmt::RAW: {
        if (csize > cap) return -1;
        for (uint32_t k = lane_id(); k < csize; k += WAVE) dst[k] = src[k];
        return csize;
    }
    case golap_fmt::LZ4: return lz4_decode_wave(src, csize, dst, cap);
    case golap_fmt::DBP: {
        int64_t r = dbp_decode_wave(src, csize, dst, cap);
        return r < 0 ? -1 : r * elem;
    }
    case golap_fmt::RLE: {
        int64_t r = rle_decode_wave(src, csize, dst, cap);
        return r < 0 ? -1 : r * elem;
    }
    case golap_fmt::FOR: {
        int64_t r = for_decode_wave(src, csize, dst, cap);
        return r < 0 ? -1 : r * elem;
    }
    default: return -1;
    }
}

}  // namespace golap_dev
