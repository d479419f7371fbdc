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
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));  // a native vector type: kept in registers

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// Decode one LZ4 block src[0, csize) into dst[0, cap). Returns bytes written or -1.
// All lanes of the wave must call it with the same arguments. With idx (the u16 token offsets of
// an LZ4I header, format.hpp), the block's own parse must agree with the index: nseq sequences,
// sequence s << seg_log2 starting at idx[s - 1].
__device__ __forceinline__ int64_t lz4_decode_wave(const uint8_t *__restrict__ src, uint32_t csize,
                                          uint8_t *__restrict__ dst, uint32_t cap,
                                          const uint8_t *idx = nullptr, uint32_t nseq = 0,
                                          uint32_t seg_log2 = 0, uint32_t unit_log2 = 0) {
    const uint32_t umask = (1u << unit_log2) - 1;  // an LZ4I page's lengths and offsets in whole units
    const uint32_t lane = lane_id();
    uint32_t ip = 0, op = 0, seq = 0;
    if (csize == 0) return cap == 0 ? 0 : -1;
    for (;;) {
        if (ip >= csize) return -1;
        if (idx) {
            if (seq >= nseq) return -1;
            if (seq && (seq & ((1u << seg_log2) - 1)) == 0) {
                const uint32_t s = seq >> seg_log2;
                if (((uint32_t)idx[2 * (s - 1)] | ((uint32_t)idx[2 * (s - 1) + 1] << 8)) != ip) return -1;
            }
            ++seq;
        }
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
        if (lit > csize - ip || lit > cap - op || (lit & umask)) return -1;
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
        if (ml > cap - op || ((off | ml) & umask)) return -1;
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
    if (idx && seq != nseq) return -1;
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

// ------------------------------------------------------------------ LZ4I: parallel parse
// An LZ4I page (format.hpp) decoded by one wave with every lane busy:
//   stage   the page is copied into the tail of the LDS window with 16-byte loads; the output
//           grows from the window's start towards it.
//   walk    lane s walks segment s of the sequence list from the offset the index gives,
//           storing each sequence's token offset in tok[]; every segment must end where the next
//           one starts (the last one at the end of the block), so the index is checked against
//           the block's own parse.
//   copy    sequences are taken 64 at a time in block order, one per lane: a wave prefix sum of
//           (literals + match) gives every lane its output position; literals are copied lane by
//           lane with unaligned 4-byte LDS accesses; a match is copied once every byte it reads
//           is final: all bytes below the output position of the first lane whose match is not
//           done yet are final, so each round copies at least that lane's match and usually all.
// Lengths of 32 bytes or more are copied by the whole wave. The staged page is never overwritten
// before it is read: each group's output must end below the group's first input byte, or the
// decoder hands the page to the serial decoder.
namespace lz4i {
constexpr uint32_t LONG = 64;        // per-lane copies below, whole-wave copies from here
constexpr uint32_t PAR_MIN = 8;      // fewer sequences: the serial decoder is cheaper
constexpr uint32_t SINK_BYTES = 4 * 64;  // per-lane 4-byte slots that take the stores of idle lanes
// LDS after a decode window for pages of up to tok_cap sequences: the token offsets, the sink.
__host__ __device__ inline uint32_t tok_bytes(uint32_t tok_cap) { return (2 * tok_cap + 15) & ~15u; }
__host__ __device__ inline uint32_t scratch_bytes(uint32_t win_cap, uint32_t tok_cap) {
    (void)win_cap;
    return tok_bytes(tok_cap) + SINK_BYTES;
}

struct Hdr {
    uint32_t nseq, sl, nseg, hb, ul;  // ul: log2 of the unit the header claims (0..2)
};

// Header of an LZ4I page (in global memory or LDS); false if malformed.
__device__ __forceinline__ bool header(const uint8_t *src, uint32_t csize, Hdr &h) {
    if (csize < 16) return false;
    uint32_t w;
    __builtin_memcpy(&w, src, 4);
    h.nseq = w & 0xffffu;
    h.sl = (w >> 16) & 0xffu;
    h.ul = w >> 24;
    if (h.nseq == 0 || h.sl > 15 || h.ul > 2) return false;
    h.nseg = golap_fmt::lz4i_segments(h.nseq, h.sl);
    h.hb = (4 + 2 * (h.nseg - 1) + 15) & ~15u;
    return h.hb <= csize && csize - h.hb <= golap_fmt::LZ4I_MAX_BLOCK;
}

// Header bytes between the index and the block must be zero (lanes check 64 at a time).
__device__ __forceinline__ bool pad_zero(const uint8_t *src, const Hdr &h) {
    bool ok = true;
    for (uint32_t i = 4 + 2 * (h.nseg - 1) + lane_id(); i < h.hb; i += WAVE) ok &= src[i] == 0;
    return __all(ok);
}

__device__ __forceinline__ uint32_t rd16(const uint8_t *p) { return (uint32_t)p[0] | ((uint32_t)p[1] << 8); }
__device__ __forceinline__ uint32_t ld32u(const uint8_t *p) {
    uint32_t v;
    __builtin_memcpy(&v, p, 4);
    return v;
}
__device__ __forceinline__ void st32u(uint8_t *p, uint32_t v) { __builtin_memcpy(p, &v, 4); }
__device__ __forceinline__ void st16u(uint8_t *p, uint32_t v) {
    const uint16_t x = (uint16_t)v;
    __builtin_memcpy(p, &x, 2);
}

// Inclusive prefix sum over the wave by DPP: row shifts by 1, 2, 4, 8 within each row of 16
// lanes, then lane 15 of rows 0 and 2 into rows 1 and 3 and lane 31 into rows 2 and 3.
__device__ __forceinline__ uint32_t wave_scan_dpp(uint32_t v) {
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false);
    return v;
}
// Maximum and minimum over the wave by the same DPP steps (lane 63 ends with the whole wave's),
// read back as a uniform value.
__device__ __forceinline__ uint32_t wave_max_dpp(uint32_t v) {
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, false));
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, false));
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, false));
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, false));
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false));
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false));
    return __builtin_amdgcn_readlane(v, WAVE - 1);
}
__device__ __forceinline__ uint32_t wave_min_dpp(uint32_t v) {
    return ~wave_max_dpp(~v);
}

// Exactly n <= 16 bytes from s to d when on; lanes that are off (and the unused stores of the
// others) write their own 4-byte sink slot instead, so the copy has no branch. All 16 source
// bytes are loaded before any store: s + 16 <= d, or only the first n bytes of s matter and
// they lie below d, or s and d are different buffers. Reads up to 16 bytes past s + n.
__device__ __forceinline__ void copy16(uint8_t *d, const uint8_t *s, uint32_t n, bool on, uint8_t *sink) {
    const uint32_t v0 = ld32u(s), v1 = ld32u(s + 4), v2 = ld32u(s + 8), v3 = ld32u(s + 12);
    const uint32_t k = n & ~3u, r = n & 3u;
    const uint32_t vt = k == 0 ? v0 : k == 4 ? v1 : k == 8 ? v2 : v3;
    st32u(on && n >= 4 ? d : sink, v0);
    st32u(on && n >= 8 ? d + 4 : sink, v1);
    st32u(on && n >= 12 ? d + 8 : sink, v2);
    st32u(on && n >= 16 ? d + 12 : sink, v3);
    st16u(on && (r & 2) ? d + k : sink, vt);
    *(on && (r & 1) ? d + k + (r & 2) : sink) = (uint8_t)(vt >> (8 * (r & 2)));
}

// n < LONG bytes by chunks of 16 in order (each lane its own n; the loop runs while any lane
// has bytes left). Valid for copies between different buffers and for matches whose offset is
// at least 16 or at least n.
__device__ __forceinline__ void copy_chunks(uint8_t *d, const uint8_t *s, uint32_t n, bool on, uint8_t *sink) {
    for (uint32_t done = 0; __any(on && n > done); done += 16) {
        const uint32_t c = n > done ? min(n - done, 16u) : 0u;
        copy16(d + done, s + done, c, on && c > 0, sink);
        __builtin_amdgcn_wave_barrier();
    }
}

// n bytes from s to d, lane by lane; d below s or at least 4 bytes above it (s + j + 4 <= d + j
// for a match, so every 4-byte read sees only final bytes or this lane's earlier writes).
__device__ __forceinline__ void copy_lane(uint8_t *d, const uint8_t *s, uint32_t n) {
    uint32_t j = 0;
    for (; j + 4 <= n; j += 4) {
        uint32_t v;
        __builtin_memcpy(&v, s + j, 4);
        __builtin_memcpy(d + j, &v, 4);
    }
    for (; j < n; ++j) d[j] = s[j];
}

// A match of n bytes at d whose source starts off bytes back, by one lane.
__device__ __forceinline__ void match_lane(uint8_t *d, uint32_t off, uint32_t n) {
    const uint8_t *s = d - off;  // off <= the output position: s stays in the window
    if (off >= 4) {
        copy_lane(d, s, n);
    } else {
        for (uint32_t j = 0; j < n; ++j) d[j] = s[j];  // byte by byte: replicates the pattern
    }
}

// n bytes to d by the whole wave (arguments uniform). off == 0: a literal run from s, which does
// not overlap d; otherwise a match with source d - off: rounds of 256 bytes when off >= 256,
// else byte j of the match is byte j mod off of [d - off, d), final before the copy.
__device__ __forceinline__ void copy_wave(uint8_t *d, const uint8_t *s, uint32_t n, uint32_t off) {
    const uint32_t lane = lane_id();
    if (off == 0 || off >= 4 * WAVE) {
        const uint8_t *from = off ? d - off : s;
        for (uint32_t b = 0; b < n; b += 4 * WAVE) {
            const uint32_t j = b + 4 * lane;
            if (j + 4 <= n) {
                uint32_t v;
                __builtin_memcpy(&v, from + j, 4);
                __builtin_memcpy(d + j, &v, 4);
            } else {
                for (uint32_t x = j; x < n; ++x) d[x] = from[x];
            }
            __builtin_amdgcn_wave_barrier();
        }
    } else {
        const uint8_t *pat = d - off;
        for (uint32_t j = 4 * lane; j < n; j += 4 * WAVE) {
            const uint32_t e = min(n - j, 4u);
            uint32_t r = j % off;
            for (uint32_t x = 0; x < e; ++x) {
                d[j + x] = pat[r];
                r = r + 1 == off ? 0 : r + 1;
            }
        }
    }
}

// Walks one segment: sequences [seq, seq_end) starting at block offset ip; must end at end_ip.
// Stores every sequence's token offset. Returns false on any disagreement or overrun.
__device__ __forceinline__ bool walk(const uint8_t *blk, uint32_t bs, uint32_t ip, uint32_t seq, uint32_t seq_end,
                                     uint32_t end_ip, uint32_t nseq, uint16_t *tok) {
    for (; seq < seq_end; ++seq) {
        if (ip >= bs) return false;
        tok[seq] = (uint16_t)ip;
        const uint32_t t = blk[ip++];
        uint32_t lit = t >> 4;
        if (lit == 15) {
            uint32_t b;
            do {
                if (ip >= bs) return false;
                b = blk[ip++];
                lit += b;
            } while (b == 255);
        }
        ip += lit;
        if (ip > bs) return false;
        if (ip == bs) {  // literals reach the end: the last sequence
            if (seq + 1 != nseq) return false;
            ++seq;
            break;
        }
        if (seq + 1 == nseq) return false;  // the last sequence must end the block
        ip += 2;
        if (ip > bs) return false;
        if ((t & 15) == 15) {
            uint32_t b;
            do {
                if (ip >= bs) return false;
                b = blk[ip++];
            } while (b == 255);
        }
    }
    return seq == seq_end && ip == end_ip;
}

// The parallel decode of a checked header h: the page src[0, csize) (16-byte aligned, readable
// up to its next 16-byte boundary) into win[0, osize). win_cap: LDS window bytes (a multiple of
// 16, at least csize rounded up to 16), followed in LDS by scratch_bytes(win_cap, tok_cap)
// bytes (token offsets, sink) that tok points at; nseq <= tok_cap. Returns
// osize, -1 if the page is malformed, -2 if its groups cannot be decoded in place (the caller
// decodes it serially).
__device__ __forceinline__ int64_t decode_par(const uint8_t *__restrict__ src, uint32_t csize, const Hdr &h, uint8_t *win,
                                     uint32_t win_cap, uint32_t osize, uint16_t *tok, uint32_t tok_cap,
                                     bool staged = false) {
    const uint32_t lane = lane_id();
    uint8_t *sink = reinterpret_cast<uint8_t *>(tok) + tok_bytes(tok_cap);
    const uint32_t c16 = (csize + 15) & ~15u;
    const uint32_t P = win_cap - c16;  // staged page
    if (!staged) {
        // every load in flight before the first store: one memory latency per 16 KiB of page
        const u32x4 *s4 = reinterpret_cast<const u32x4 *>(src);
        u32x4 *d4 = reinterpret_cast<u32x4 *>(win + P);
        const uint32_t n16 = c16 / 16;
        for (uint32_t i0 = 0; i0 < n16; i0 += 16 * WAVE) {
            u32x4 v[16];
#pragma unroll
            for (int j = 0; j < 16; ++j) v[j] = s4[min(i0 + j * WAVE + lane, n16 - 1)];  // unconditional: in flight together
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                const uint32_t i = i0 + j * WAVE + lane;
                if (i < n16) d4[i] = v[j];
            }
        }
    }
    __builtin_amdgcn_wave_barrier();
    const uint8_t *hdr = win + P;
    const uint8_t *blk = hdr + h.hb;
    const uint32_t bs = csize - h.hb;
    bool ok = true;
    for (uint32_t sgm = lane; sgm < h.nseg; sgm += WAVE) {
        const uint32_t ip = sgm ? rd16(hdr + 4 + 2 * (sgm - 1)) : 0u;
        const uint32_t end_ip = sgm + 1 < h.nseg ? rd16(hdr + 4 + 2 * sgm) : bs;
        const uint32_t s0 = sgm << h.sl, s1 = min(h.nseq, (sgm + 1) << h.sl);
        ok &= walk(blk, bs, ip, s0, s1, end_ip, h.nseq, tok);
    }
    if (!__all(ok)) return -1;
    __builtin_amdgcn_wave_barrier();
    uint32_t base = 0;
    uint8_t *slot = sink + 4 * lane;
    // the headers of group g + 1 are loaded before group g's copies: its input lies above
    // everything group g writes (the in-place check), so the loads can run ahead
    uint32_t nip = lane < h.nseq ? tok[lane] : 0u;
    uint32_t nt = ld32u(blk + nip) & 0xffu;
    uint32_t now = ld32u(blk + nip + 1 + (nt >> 4));  // the offset word unless the literal length extends
    for (uint32_t g0 = 0; g0 < h.nseq; g0 += WAVE) {
        const uint32_t k = g0 + lane;
        const bool valid = k < h.nseq;
        const bool last = k + 1 == h.nseq;
        // the sequence header: token, literal length, offset, match length (the walk has
        // checked every byte of it lies in the block)
        const uint32_t ip = nip, t = nt;
        uint32_t lit = valid ? t >> 4 : 0u, ls = ip + 1, ow = now;
        if (__any(lit == 15)) {
            if (lit == 15) {
                uint32_t b;
                do {
                    b = blk[ls++];
                    lit += b;
                } while (b == 255);
                ow = ld32u(blk + ls + lit);
            }
        }
        const bool has_match = valid && !last;
        const uint32_t mp = ls + lit;  // offset bytes
        uint32_t ml = has_match ? (t & 15) : 0u;
        const uint32_t off = has_match ? (ow & 0xffffu) : 0u;
        if (__any(ml == 15)) {
            if (ml == 15) {
                uint32_t q = mp + 2, b;
                do {
                    b = blk[q++];
                    ml += b;
                } while (b == 255);
            }
        }
        if (has_match) ml += golap_fmt::LZ4_MINMATCH;
        const uint32_t len = lit + ml;
        const uint32_t incl = wave_scan_dpp(len);
        const uint32_t dst = base + incl - len, mdst = dst + lit;
        const uint32_t total = __builtin_amdgcn_readlane(incl, WAVE - 1);
        const uint32_t in0 = __builtin_amdgcn_readfirstlane(ip);
        // the walk bounds every length by the block size, so these sums do not wrap
        const uint32_t umask = (1u << h.ul) - 1;  // lengths and offsets in whole units
        const bool bad = valid && ((uint64_t)dst + len > osize || (ml && (off == 0 || off > mdst)) ||
                                   ((lit | off | ml) & umask) != 0);
        if (__any(bad)) return -1;
        // In place, the group writes [base, base + total) while its input still lies in the
        // staged page. When that range ends below the group's first input byte, literals are
        // copied straight from the page. Otherwise short literals are first read into registers
        // and long ones (LONG bytes or more) copied first, each by the whole wave from its start
        // forwards: then the writes must only stay below the next group's input, and a long
        // literal must not start above its own staged bytes.
        const uint32_t kn = k + WAVE;
        const bool direct = base + total <= P + h.hb + in0;
        if (!direct) {
            const uint32_t next_in = g0 + WAVE < h.nseq ? __builtin_amdgcn_readfirstlane(tok[g0 + WAVE]) : bs;
            if (base + total > P + h.hb + next_in || __any(lit >= LONG && dst > P + h.hb + ls)) return -2;
        }
        nip = kn < h.nseq ? tok[kn] : 0u;
        nt = ld32u(blk + nip) & 0xffu;
        now = ld32u(blk + nip + 1 + (nt >> 4));
        if (direct) {
            copy_chunks(win + dst, blk + ls, lit, lit < LONG, slot);
            for (uint64_t lw = __ballot(lit >= LONG); lw; lw &= lw - 1) {
                const int l = __builtin_ctzll(lw);
                copy_wave(win + __builtin_amdgcn_readlane(dst, l), blk + __builtin_amdgcn_readlane(ls, l),
                          __builtin_amdgcn_readlane(lit, l), 0);
            }
        } else {
            const uint32_t sl = lit < LONG ? lit : 0u;
            const uint32_t nd = wave_max_dpp((sl + 3) >> 2);  // dwords, at most LONG / 4
            uint32_t lv[LONG / 4];
#pragma unroll
            for (int j = 0; j < (int)(LONG / 4); ++j)
                if ((uint32_t)j < nd) lv[j] = ld32u(blk + ls + 4 * j);
            __builtin_amdgcn_wave_barrier();
            for (uint64_t lw = __ballot(lit >= LONG); lw; lw &= lw - 1) {
                const int l = __builtin_ctzll(lw);
                copy_wave(win + __builtin_amdgcn_readlane(dst, l), blk + __builtin_amdgcn_readlane(ls, l),
                          __builtin_amdgcn_readlane(lit, l), 0);
            }
#pragma unroll
            for (int j = 0; j < (int)(LONG / 4); ++j) {
                if ((uint32_t)j < nd) {
                    const uint32_t b = 4 * j;
                    uint8_t *d = win + dst + b;
                    st32u(b + 4 <= sl ? d : slot, lv[j]);
                    const uint32_t r = sl - b;  // 1..3 bytes of a last, partial dword
                    st16u(b < sl && r < 4 && (r & 2) ? d : slot, lv[j]);
                    *(b < sl && r < 4 && (r & 1) ? d + (r & 2) : slot) = (uint8_t)(lv[j] >> (8 * (r & 2)));
                }
            }
        }
        __builtin_amdgcn_wave_barrier();
        bool pend = ml != 0;
        const uint32_t need = off >= ml ? mdst - off + ml : mdst;
        const bool chunked = ml < LONG && (off >= 16 || off >= ml);
        for (uint64_t pm = __ballot(pend); pm; pm = __ballot(pend)) {
            const uint32_t frontier = __builtin_amdgcn_readlane(mdst, __builtin_ctzll(pm));
            const bool ready = pend && need <= frontier;
            copy_chunks(win + mdst, win + mdst - off, ml, ready && chunked, slot);
            if (ready && !chunked && ml < LONG) match_lane(win + mdst, off, ml);
            for (uint64_t lw = __ballot(ready && ml >= LONG); lw; lw &= lw - 1) {
                const int l = __builtin_ctzll(lw);
                copy_wave(win + __builtin_amdgcn_readlane(mdst, l), nullptr, __builtin_amdgcn_readlane(ml, l),
                          __builtin_amdgcn_readlane(off, l));
            }
            pend = pend && !ready;
            __builtin_amdgcn_wave_barrier();
        }
        base += total;
    }
    return base;
}
}  // namespace lz4i

// Where decode_par stages a page of csize bytes in a window of win_cap bytes.
__device__ __forceinline__ uint32_t lz4i_stage_offset(uint32_t win_cap, uint32_t csize) {
    return win_cap - ((csize + 15) & ~15u);
}

// An LZ4I page already staged by the caller at win + stage_offset(win_cap, csize) (its bytes in
// LDS, from registers loaded ahead): header and padding read from LDS, then the parallel parse;
// the serial fallback reads the page from global memory (src). Returns bytes written or -1.
__device__ __forceinline__ int64_t lz4i_decode_staged(const uint8_t *__restrict__ src, uint32_t csize, uint8_t *win,
                                                      uint32_t win_cap, uint32_t osize, uint16_t *tok,
                                                      uint32_t tok_cap) {
    const uint8_t *pg = win + lz4i_stage_offset(win_cap, csize);
    lz4i::Hdr h;
    if (!lz4i::header(pg, csize, h) || !lz4i::pad_zero(pg, h)) return -1;
    if (osize > win_cap) return -1;
    if (h.nseq >= lz4i::PAR_MIN && h.nseq <= tok_cap) {
        const int64_t r = lz4i::decode_par(src, csize, h, win, win_cap, osize, tok, tok_cap, true);
        if (r != -2) return r;
    }
    return lz4_decode_wave(src + h.hb, csize - h.hb, win, osize, src + 4, h.nseq, h.sl, h.ul);
}

// Decodes an LZ4I page src[0, csize) into win[0, osize): in parallel when the page has enough
// sequences, tok points at lz4i::scratch_bytes(win_cap, tok_cap) bytes of LDS for pages of up to
// tok_cap sequences and the window (win_cap bytes, a multiple of 16) holds the staged page;
// otherwise, and when its groups do not fit in place, serially from global memory. Returns the
// bytes written or -1.
__device__ __forceinline__ int64_t lz4i_decode_wave(const uint8_t *__restrict__ src, uint32_t csize, uint8_t *win,
                                           uint32_t win_cap, uint32_t osize, uint16_t *tok, uint32_t tok_cap) {
    lz4i::Hdr h;
    if (!lz4i::header(src, csize, h) || !lz4i::pad_zero(src, h)) return -1;
    if (osize > win_cap) return -1;
    if (h.nseq >= lz4i::PAR_MIN && h.nseq <= tok_cap && tok && ((csize + 15) & ~15u) <= win_cap &&
        (reinterpret_cast<uintptr_t>(src) & 15) == 0 && (win_cap & 15) == 0) {
        const int64_t r = lz4i::decode_par(src, csize, h, win, win_cap, osize, tok, tok_cap);
        if (r != -2) return r;
    }
    return lz4_decode_wave(src + h.hb, csize - h.hb, win, osize, src + 4, h.nseq, h.sl, h.ul);
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

// Decodes one page of any codec into dst[0, cap). Returns bytes written or -1. LZ4I pages are
// decoded in parallel when the caller gives a window of win_cap >= cap bytes (a multiple of 16)
// and room for tok_cap token offsets, serially otherwise.
__device__ __forceinline__ int64_t page_decode_wave(uint8_t codec, const uint8_t *__restrict__ src,
                                           uint32_t csize, uint8_t *__restrict__ dst, uint32_t cap,
                                           uint32_t elem, uint32_t win_cap = 0, uint16_t *tok = nullptr,
                                           uint32_t tok_cap = 0) {
    switch (codec & golap_fmt::PAGE_CODEC_MASK) {
    case golap_fmt::RAW: {
        if (csize > cap) return -1;
        for (uint32_t k = lane_id(); k < csize; k += WAVE) dst[k] = src[k];
        return csize;
    }
    case golap_fmt::LZ4: return lz4_decode_wave(src, csize, dst, cap);
    case golap_fmt::LZ4I: return lz4i_decode_wave(src, csize, dst, win_cap >= cap ? win_cap : cap, cap, tok, tok_cap);
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
