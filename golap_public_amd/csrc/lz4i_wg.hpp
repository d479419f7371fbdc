// Workgroup-cooperative decode of one LZ4I page (format.hpp) by all NT threads of a workgroup.
//
// The reference decodes each 64 KiB LZ4 sub-block with nvcomp, one warp per sub-block
// (ChunkDecompressor::step, core/comp.cuh:171-195). Here the page is staged in LDS and every
// thread of the workgroup works on it at once; the decoded page lands in a separate LDS output
// window, so staged bytes are never overwritten:
//   walk     one thread per index segment parses its sequences from the token offset the index
//            gives and records every sequence's token offset; the block's own parse must agree
//            with the index.
//   parse    thread t takes sequences t, t + NT, t + 2 NT, ... (at most MAXJ), each parsed on its
//            own from its token offset, all reads of a thread in flight together. A workgroup
//            prefix sum of the decoded lengths, in sequence order, gives every sequence's output
//            position; the total must equal the page's decoded size.
//   literals are copied and their bytes marked in a bitmap of final output units; each match
//            (destination, offset, length) stays in registers.
//   matches  a match is copied once every unit of its source is marked final, then its
//            destination is marked. One pass over the registers, in sequence order with a
//            barrier per step, copies most of them (their sources lie further back); the rest are
//            gathered into an LDS list that the whole workgroup works through in passes. The
//            earliest pending match always has a final source, so every pass copies at least one
//            and the loop ends.
// Every LDS access is aligned to its own width: on gfx950 a wave instruction with a misaligned
// lane costs about 64 cycles of LDS time against 4 to 11 aligned. Bytes at any offset are read
// as the two aligned 16-byte blocks that hold them and shifted in registers; bytes at any offset
// are written by atomic ORs of aligned words into the zeroed output window. Pages whose header
// claims 4-byte units (typed LZ4, format.hpp) are copied as whole aligned words.
// Every length and offset is checked against the page and the output window before use; a
// malformed page returns -1 and never reads or writes outside its LDS regions.
#pragma once
#include "device_codecs.hpp"

namespace golap_dev {
namespace lz4wg {

constexpr uint32_t MAXJ = 8;         // sequences per thread (match descriptors kept in registers)
constexpr uint32_t SERIAL_MAX = golap_fmt::LZ4I_SERIAL_MAX;  // pages with at most this many sequences: in order
constexpr uint32_t SHORT = 16;       // literals and matches up to this long are copied in the main passes
constexpr uint32_t LIST_MIN = 256;   // entries the lists hold at least
constexpr uint32_t SPIN = 32;        // barrier-free tries of a thread's list entries in the first pass
constexpr uint32_t STEP_TRIES = 4;   // tries of a short match in its step of the ordered pass

// Byte offsets of the decode's LDS regions from the caller's base, for pages of up to stage_cap
// coded bytes and tok_cap sequences decoding to at most out_cap bytes, with nt threads. The
// token array is reused for the lists of long literals and pending matches (8 bytes an entry):
// at least LIST_MIN entries, or one per sequence with full_list (a page whose pending matches
// overflow the list is handed back for the in-order decode).
// With jump, a source map of two bytes per output byte follows for decode<NT, true>.
struct Layout {
    uint32_t stage, out, bits, tok, wtot, misc, jump, total;
    uint32_t stage_cap, out_cap, tok_cap, list_cap;
};
__host__ __device__ inline Layout layout(uint32_t stage_cap, uint32_t out_cap, uint32_t tok_cap, uint32_t nt,
                                         bool full_list = false, bool jump = false) {
    Layout l;
    l.stage_cap = stage_cap;
    l.out_cap = out_cap;
    l.tok_cap = tok_cap;
    l.stage = 0;
    l.out = (stage_cap + 32 + 15) & ~15u;           // two 16-byte blocks may be read past the page end
    l.bits = l.out + ((out_cap + 32 + 15) & ~15u);  // and past the output end
    l.tok = l.bits + ((((out_cap + 31) / 32 + 2) * 4 + 15) & ~15u);
    const uint32_t list_bytes = max(2 * tok_cap, 8u * max(LIST_MIN, full_list ? tok_cap : 0u));
    l.list_cap = list_bytes / 8;
    l.wtot = l.tok + ((list_bytes + 15) & ~15u);  // the token array, later the lists
    l.misc = l.wtot + MAXJ * (nt / 64) * 4;
    l.jump = jump ? l.misc + 32 : 0u;
    l.total = jump ? l.jump + ((2 * out_cap + 15) & ~15u) : l.misc + 32;
    return l;
}
// The most sequences a page decoding to osize bytes can have: every sequence but the last
// decodes to at least 4 bytes.
__host__ __device__ inline uint32_t max_sequences(uint32_t osize) { return osize / 4 + 1; }

// ---- aligned LDS access
// Bytes [p, p + 16) of base (4-byte aligned, readable 20 bytes past p & ~3): the five aligned
// words from p & ~3, funnel-shifted by p % 4 bytes.
__device__ __forceinline__ u32x4 ld_any16(const uint8_t *base, uint32_t p) {
    const uint32_t *w = reinterpret_cast<const uint32_t *>(base + (p & ~3u));
    const uint32_t d0 = w[0], d1 = w[1], d2 = w[2], d3 = w[3], d4 = w[4];
    const uint32_t sh = 8 * (p & 3);
    return u32x4{__builtin_amdgcn_alignbit(d1, d0, sh), __builtin_amdgcn_alignbit(d2, d1, sh),
                 __builtin_amdgcn_alignbit(d3, d2, sh), __builtin_amdgcn_alignbit(d4, d3, sh)};
}
// Bytes [p, p + 4) of base (readable 8 bytes past p & ~3): two aligned words.
__device__ __forceinline__ uint32_t ld_any4(const uint8_t *base, uint32_t p) {
    const uint32_t *w = reinterpret_cast<const uint32_t *>(base + (p & ~3u));
    return (uint32_t)((((uint64_t)w[1] << 32) | w[0]) >> (8 * (p & 3)));
}
__device__ __forceinline__ uint32_t ld32(const uint8_t *p) { return *reinterpret_cast<const uint32_t *>(p); }
__device__ __forceinline__ uint32_t ld16b(const uint8_t *p) { return *reinterpret_cast<const uint16_t *>(p); }
// The byte mask of a word from its four byte enables (bit i: byte i): the enables spread to bits
// 0, 8, 16 and 24 by one 24-bit multiply, then each widened to its byte.
__device__ __forceinline__ uint32_t byte_mask(uint32_t e) {
    const uint32_t x = (e * 0x204081u) & 0x01010101u;  // e < 16: no carries between the copies
    return (x << 8) - x;
}
// Bytes 0..n-1 of v (n <= 16) ORed into out[op, op + n): the output bytes there are zero.
__device__ __forceinline__ void or_bytes(uint8_t *out, uint32_t op, const u32x4 &v, uint32_t n) {
    uint32_t *o32 = reinterpret_cast<uint32_t *>(out) + (op >> 2);
    const int o = (int)(op & 3), sh = 32 - 8 * o;
    const uint32_t L[6] = {0u, v.x, v.y, v.z, v.w, 0u};
    const uint32_t en = ((1u << n) - 1u) << o;  // byte enables of the five words, four bits each
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        const uint32_t w = (uint32_t)((((uint64_t)L[k + 1] << 32) | L[k]) >> sh);
        const uint32_t m = byte_mask((en >> (4 * k)) & 15u);
        if (m) atomicOr(o32 + k, w & m);
    }
}
// Words 0..n/4-1 of v (n a multiple of 4, <= 16) to out[op, op + n), op a multiple of 4.
__device__ __forceinline__ void st_words(uint8_t *out, uint32_t op, const u32x4 &v, uint32_t n) {
    uint32_t *o32 = reinterpret_cast<uint32_t *>(out) + (op >> 2);
    if (n >= 4) o32[0] = v.x;
    if (n >= 8) o32[1] = v.y;
    if (n >= 12) o32[2] = v.z;
    if (n >= 16) o32[3] = v.w;
}
// n <= 16 bytes of v to out + op: whole words for 4-byte units (typed), else ORs.
__device__ __forceinline__ void put16(uint8_t *out, uint32_t op, const u32x4 &v, uint32_t n, bool words) {
    if (words) st_words(out, op, v, n);
    else or_bytes(out, op, v, n);
}

// ---- the bitmap of final output units (bit i: unit i, units of 2^ub bytes)
__device__ __forceinline__ uint32_t bmask(uint32_t lo, uint32_t hi) {
    return (0xffffffffu >> (31u - hi)) & (0xffffffffu << lo);
}
// Masks of units [a, a + n) in words a >> 5 and (a >> 5) + 1, for n <= 33 (both zero for n == 0):
// the halves of one 64-bit mask.
__device__ __forceinline__ void masks2(uint32_t a, uint32_t n, uint32_t &m0, uint32_t &m1) {
    const uint64_t m = ((1ull << n) - 1) << (a & 31);
    m0 = (uint32_t)m;
    m1 = (uint32_t)(m >> 32);
}
// Are units [a, a + n) final (1 <= n <= 33)? The bitmap has a spare word past its end.
__device__ __forceinline__ bool bits_all(const uint32_t *bits, uint32_t a, uint32_t n) {
    uint32_t m0, m1;
    masks2(a, n, m0, m1);
    const uint32_t b0 = bits[a >> 5], b1 = bits[(a >> 5) + 1];
    return (b0 & m0) == m0 && (b1 & m1) == m1;
}
__device__ __forceinline__ bool bits_all_any(const uint32_t *bits, uint32_t a, uint32_t n) {
    if (n <= 33) return bits_all(bits, a, n);
    const uint32_t e = a + n - 1, w0 = a >> 5, w1 = e >> 5;
    bool ok = (bits[w0] | ~bmask(a & 31, 31)) == 0xffffffffu && (bits[w1] | ~bmask(0, e & 31)) == 0xffffffffu;
    for (uint32_t w = w0 + 1; w < w1; ++w) ok = ok && bits[w] == 0xffffffffu;
    return ok;
}
// Marks units [a, a + n) final; n <= 33 (nothing for n == 0).
__device__ __forceinline__ void bits_set(uint32_t *bits, uint32_t a, uint32_t n) {
    uint32_t m0, m1;
    masks2(a, n, m0, m1);
    if (m0) atomicOr(bits + (a >> 5), m0);
    if (m1) atomicOr(bits + (a >> 5) + 1, m1);
}
__device__ __forceinline__ void bits_set_any(uint32_t *bits, uint32_t a, uint32_t n) {
    if (n <= 33) {
        bits_set(bits, a, n);
        return;
    }
    const uint32_t e = a + n - 1, w0 = a >> 5, w1 = e >> 5;
    atomicOr(bits + w0, bmask(a & 31, 31));
    for (uint32_t w = w0 + 1; w < w1; ++w) atomicOr(bits + w, 0xffffffffu);
    atomicOr(bits + w1, bmask(0, e & 31));
}

// Compiler-only ordering point: LDS accesses of one wave are performed in issue order, so a
// thread that sees a mark also sees the bytes written before it was set.
__device__ __forceinline__ void order() { asm volatile("" ::: "memory"); }

// Workgroup OR of pred (every thread calls it): one LDS atomic per wave and one barrier; the
// runtime's reduction barrier costs about ten plain ones. Call c ORs into word c % 3 and then
// thread 0 clears word (c + 2) % 3, the word of calls c - 1 and c + 2: every reader of call
// c - 1 read it before arriving at call c's barrier, and no thread can OR into it for call c + 2
// before passing call c + 1's barrier, where thread 0 arrives only after the clear.
struct BlockAny {
    uint32_t *red;  // three LDS words, zero before the first call
    uint32_t c;
    __device__ __forceinline__ bool operator()(bool pred) {
        const uint32_t w = c % 3;
        if (__ballot(pred) != 0 && (threadIdx.x % WAVE) == 0) atomicOr(red + w, 1u);
        __syncthreads();
        const bool r = red[w] != 0;
        if (threadIdx.x == 0) red[(c + 2) % 3] = 0;
        ++c;
        return r;
    }
};

// A match of n bytes at out + dst from dst - off whose source is final, by one thread: in 16-byte
// pieces when the source lies at least 16 bytes (or n) back, else in order by words (typed) or
// bytes.
__device__ __forceinline__ void match_thread(uint8_t *out, uint32_t dst, uint32_t off, uint32_t n, bool words) {
    if (off >= 16 || off >= n) {
        for (uint32_t c = 0; c < n; c += 16) put16(out, dst + c, ld_any16(out, dst - off + c), min(n - c, 16u), words);
    } else if (words) {
        uint32_t *o32 = reinterpret_cast<uint32_t *>(out);
        for (uint32_t c = 0; c < n; c += 4) o32[(dst + c) >> 2] = o32[(dst - off + c) >> 2];
    } else {
        for (uint32_t c = 0; c < n; ++c) {
            const uint32_t b = out[dst - off + c];
            atomicOr(reinterpret_cast<uint32_t *>(out) + ((dst + c) >> 2), b << (8 * ((dst + c) & 3)));
        }
    }
}
// A literal run of n bytes from stage + s to out + op, by one thread, in 16-byte pieces.
__device__ __forceinline__ void literal_thread(uint8_t *out, uint32_t op, const uint8_t *blk, uint32_t s, uint32_t n,
                                               bool words) {
    for (uint32_t c = 0; c < n; c += 16) put16(out, op + c, ld_any16(blk, s + c), min(n - c, 16u), words);
}

// Bytes of the sequence whose token lies at blk[ip] (token, length bytes, literals, offset, and
// for a sequence with a match the match length bytes); 0 when it does not fit before bs.
__device__ __forceinline__ uint32_t seq_bytes(const uint8_t *blk, uint32_t bs, uint32_t ip, bool has_match) {
    const uint32_t ip0 = ip;
    const uint32_t t = blk[ip++];
    uint32_t lit = t >> 4;
    if (lit == 15) {
        uint32_t b;
        do {
            if (ip >= bs) return 0;
            b = blk[ip++];
            lit += b;
        } while (b == 255);
    }
    if (lit > bs - ip) return 0;
    ip += lit;
    if (!has_match) return ip - ip0;
    if (bs - ip < 2) return 0;
    ip += 2;
    if ((t & 15) == 15) {
        uint32_t b;
        do {
            if (ip >= bs) return 0;
            b = blk[ip++];
        } while (b == 255);
    }
    return ip - ip0;
}

// Walks sequences [k, k_end) of the block from offset ip, which must end at end_ip, storing
// each sequence's token offset in tok; the block's last sequence (nseq - 1) must end the block.
// False on any disagreement or overrun. A step reads two bytes at ip <= bs (the staged page has
// spare bytes past its end); length extensions beyond one byte take the slow path.
__device__ __forceinline__ bool walk(const uint8_t *blk, uint32_t bs, uint32_t ip, uint32_t k, uint32_t k_end,
                                     uint32_t end_ip, uint32_t nseq, uint16_t *tok) {
    bool ok = ip < bs;
    const uint32_t k_stop = k_end == nseq ? k_end - 1 : k_end;  // sequences with a match
    for (; k < k_stop; ++k) {
        tok[k] = (uint16_t)ip;
        const uint32_t t = blk[ip], e1 = blk[ip + 1];
        // token, one literal-length byte when the literal length is 15 or more, literals, offset
        uint32_t adv = (t >> 4) < 15 ? 3 + (t >> 4) : 4 + 15 + e1;
        if (__any((t & 15) == 15)) {  // one match-length byte, read where the offset ends
            if ((t & 15) == 15) adv += 1 + (blk[min(ip + adv, bs)] == 255 ? 0x10000u : 0u);
        }
        const bool slow = ((t >> 4) == 15 && e1 == 255) || adv >= 0x10000u;  // more length bytes
        if (__any(slow)) {
            if (slow) adv = seq_bytes(blk, bs, ip, true);
        }
        ok = ok && adv != 0 && adv <= bs - ip;
        ip = min(ip + adv, bs);
    }
    if (k_end == nseq && k < k_end) {  // the last sequence: literals up to the block's end
        ok = ok && ip < bs;
        tok[k] = (uint16_t)(ok ? ip : 0u);
        const uint32_t adv = ok ? seq_bytes(blk, bs, ip, false) : 0u;
        ok = ok && adv != 0 && ip + adv == bs;
        ip = bs;
    }
    return ok && ip == end_ip;
}

// What the parse leaves for the later phases: per thread, sequences k = tid + NT j (j < J):
// lt[j] = literal start in the block | literal length << 16, mo[j] = match offset | match
// length << 16, pos[j] = output position.
struct Parsed {
    uint32_t lt[MAXJ], mo[MAXJ], pos[MAXJ];
    uint32_t J, nseq, ub, bs;
    bool words;
    const uint8_t *blk;
};

// Header, walk, parse and output positions of the LZ4I page staged at lds + l.stage (csize bytes,
// staged before a __syncthreads()): 0, -1 (malformed) or -2 (not taken: pages of at most
// SERIAL_MAX sequences, which are cheapest in order, more sequences than MAXJ per thread or than
// l.tok_cap, more segments than threads, a decoded size of 64 KiB or more). FULL: the output
// window is zeroed and the bitmap cleared for decode(). The result is the same in every thread.
template <int NT, bool FULL>
__device__ __forceinline__ int parse(uint8_t *lds, const Layout &l, uint32_t csize, uint32_t osize, BlockAny &any,
                                     Parsed &ps) {
    constexpr uint32_t NW = NT / WAVE;
    const uint32_t tid = threadIdx.x, lane = tid % WAVE, wave = tid / WAVE;
    const uint8_t *hdr = lds + l.stage;
    uint8_t *out = lds + l.out;
    uint32_t *bits = reinterpret_cast<uint32_t *>(lds + l.bits);
    uint16_t *tok = reinterpret_cast<uint16_t *>(lds + l.tok);
    uint32_t *wtot = reinterpret_cast<uint32_t *>(lds + l.wtot);
    // header (uniform): {u16 nseq, u8 seg_log2, u8 unit_log2}, nseg - 1 token offsets, zero padding
    if (csize < 16 || csize > l.stage_cap || osize > l.out_cap) return -1;
    const uint32_t w = ld32(hdr);
    const uint32_t nseq = w & 0xffffu, sl = (w >> 16) & 0xffu, ul = w >> 24;
    if (nseq == 0 || sl > 15 || ul > 2) return -1;
    const uint32_t nseg = golap_fmt::lz4i_segments(nseq, sl);
    const uint32_t ib = 4 + 2 * (nseg - 1), hb = (ib + 15) & ~15u;
    if (hb > csize || csize - hb > golap_fmt::LZ4I_MAX_BLOCK) return -1;
    const uint8_t *blk = hdr + hb;
    const uint32_t bs = csize - hb;
    if (nseq <= SERIAL_MAX || nseg > (uint32_t)NT || osize > 0xffffu || nseq > l.tok_cap ||
        nseq > (uint32_t)NT * MAXJ)
        return -2;
    const uint32_t J = (nseq + NT - 1) / NT;
    const uint32_t umask = (1u << ul) - 1;  // lengths and offsets are multiples of the unit
    const bool words = ul == 2 && (osize & 3) == 0;  // whole aligned words
    const uint32_t ub = words ? 2u : 0u;             // bitmap unit: a word or a byte
    ps.J = J;
    ps.nseq = nseq;
    ps.words = words;
    ps.ub = ub;
    ps.blk = blk;
    ps.bs = bs;

    // walk; the output window zeroed (bytes are ORed in) and the bitmap cleared
    if (FULL) {
        const uint32_t nz = words ? 0u : (osize + 15) / 16;
        u32x4 *o16 = reinterpret_cast<u32x4 *>(out);
        for (uint32_t i = tid; i < nz; i += NT) o16[i] = u32x4{0, 0, 0, 0};
        for (uint32_t i = tid; i < ((osize >> ub) + 31) / 32 + 2; i += NT) bits[i] = 0;
    }
    bool bad = tid < hb - ib && hdr[ib + tid] != 0;
    if (tid < nseg) {
        const uint32_t ip = tid ? ld16b(hdr + 4 + 2 * (tid - 1)) : 0u;
        const uint32_t end_ip = tid + 1 < nseg ? ld16b(hdr + 4 + 2 * tid) : bs;
        bad |= !walk(blk, bs, ip, tid << sl, min(nseq, (tid + 1) << sl), end_ip, nseq, tok);
    }
    if (any(bad)) return -1;

    // parse: sequence k = tid + NT j (the walk has checked every byte of every header; reads
    // past a short last sequence stay inside the staged page's spare bytes). All token offsets
    // are read together, then the two 16-byte blocks at every token, then every offset.
    // lt[j] = literal start | literal length << 16, mo[j] = match offset | match length << 16.
    uint32_t *lt = ps.lt, *mo = ps.mo, *pos = ps.pos;
    uint32_t rd[MAXJ], tk[MAXJ];
#pragma unroll
    for (uint32_t j = 0; j < MAXJ; ++j) {
        const uint32_t k = tid + NT * j;
        tk[j] = j < J ? tok[k < nseq ? k : 0u] : 0u;
    }
    uint32_t slow = 0;
#pragma unroll
    for (uint32_t j = 0; j < MAXJ; ++j) {
        lt[j] = 1;
        mo[j] = 0;
        pos[j] = 0;
        rd[j] = 0;
        if (j >= J) continue;  // uniform
        const uint32_t k = tid + NT * j;
        const bool v = k < nseq, has = v && k + 1 < nseq;
        const uint32_t x = ld_any4(blk, tk[j]);
        const uint32_t t = v ? x & 0xffu : 0u, e1 = (x >> 8) & 0xffu;
        const bool ext = (t >> 4) == 15;
        const uint32_t n = ext ? 15 + e1 : t >> 4, s0 = tk[j] + (ext ? 2u : 1u);
        slow |= (ext && e1 == 255) ? 1u << j : 0u;
        lt[j] = s0 | (n << 16);
        mo[j] = has ? (t & 15) | 0x100u : 0u;  // token's match length, 0x100: has a match
    }
    if (__any(slow != 0)) {  // literal lengths of 270 bytes or more: parsed byte by byte
#pragma unroll
        for (uint32_t j = 0; j < MAXJ; ++j) {
            if ((slow >> j) & 1u) {
                uint32_t s1 = tk[j] + 1, n = 15, c;
                do {
                    c = blk[s1++];
                    n += c;
                } while (c == 255);
                lt[j] = s1 | (n << 16);
            }
        }
    }
#pragma unroll
    for (uint32_t j = 0; j < MAXJ; ++j)
        if (j < J) rd[j] = ld_any4(blk, (lt[j] & 0xffffu) + (lt[j] >> 16));  // the offset and a length byte
    uint32_t mext = 0;
#pragma unroll
    for (uint32_t j = 0; j < MAXJ; ++j) mext |= (mo[j] & 0xffu) == 15 && ((rd[j] >> 16) & 0xffu) == 255 ? 1u << j : 0u;
#pragma unroll
    for (uint32_t j = 0; j < MAXJ; ++j) {
        if (j >= J) continue;
        const uint32_t n = lt[j] >> 16, has = mo[j] >> 8;
        uint32_t ml = mo[j] & 15;
        if (ml == 15) ml += (rd[j] >> 16) & 0xffu;  // the first match-length byte
        if (__any((mext >> j) & 1u)) {              // more of them: byte by byte
            if ((mext >> j) & 1u) {
                uint32_t q = (lt[j] & 0xffffu) + n + 3, c;
                do {
                    c = blk[q++];
                    ml += c;
                } while (c == 255 && ml < 0x10000u);
            }
        }
        ml = has ? ml + golap_fmt::LZ4_MINMATCH : 0u;
        const uint32_t off = has ? rd[j] & 0xffffu : 0u;
        bad |= ml > 0xffffu || ((n | off | ml) & umask) != 0;  // no longer than a page; whole units
        mo[j] = off | (min(ml, 0xffffu) << 16);
        pos[j] = n + ml;  // decoded length, turned into the output position below
    }
    // output positions: per j a wave prefix sum, then the wave totals of every j in sequence order
#pragma unroll
    for (uint32_t j = 0; j < MAXJ; ++j) {
        if (j >= J) continue;
        const uint32_t len = min(pos[j], 0x1ffffu);
        const uint32_t incl = lz4i::wave_scan_dpp(len);
        if (lane == WAVE - 1) wtot[j * NW + wave] = incl;
        pos[j] = incl - len;
    }
    if (any(bad)) return -1;
    {
        uint32_t carry = 0;
#pragma unroll
        for (uint32_t j = 0; j < MAXJ; ++j) {
            if (j >= J) continue;
            uint32_t before = 0, all = 0;
#pragma unroll
            for (uint32_t q = 0; q < NW; ++q) {
                const uint32_t t = wtot[j * NW + q];
                before += q < wave ? t : 0u;
                all += t;
            }
            pos[j] += carry + before;
            carry += all;
        }
        if (carry != osize) return -1;  // every thread computes the same total
    }
    return 0;
}

// Resolves the pending matches of the threads (pend, mo = destination | offset << 16, ml =
// length) by pointer jumping over a source map of the output units: a pending unit maps to the
// unit it copies, every other unit (final: literal or copied match) to itself. Each round
// replaces a unit's source by its source's source, halving every chain (a source always lies
// before its unit, so chains end at a final unit); then every pending unit takes its final
// source's value. Rounds: about log2 of the longest chain, against one list pass per link.
template <int NT>
__device__ __forceinline__ int64_t resolve_by_jumping(uint8_t *lds, const Layout &l, uint32_t osize, uint32_t ub,
                                                      uint32_t pend, const uint32_t (&mo)[MAXJ],
                                                      const uint32_t (&ml)[MAXJ], BlockAny &any) {
    const uint32_t tid = threadIdx.x;
    uint8_t *out = lds + l.out;
    uint16_t *map = reinterpret_cast<uint16_t *>(lds + l.jump);
    const uint32_t nu = osize >> ub;
    for (uint32_t i = tid; i < nu; i += NT) map[i] = (uint16_t)i;
    __syncthreads();
#pragma unroll
    for (uint32_t j = 0; j < MAXJ; ++j) {
        if (!((pend >> j) & 1u)) continue;
        const uint32_t d = (mo[j] & 0xffffu) >> ub, s = ((mo[j] & 0xffffu) - (mo[j] >> 16)) >> ub, n = ml[j] >> ub;
        for (uint32_t k = 0; k < n; ++k) map[d + k] = (uint16_t)(s + k);
    }
    __syncthreads();
    for (uint32_t round = 0;; ++round) {
        bool changed = false;
        for (uint32_t i = tid; i < nu; i += NT) {
            const uint32_t m = map[i];
            if (m != i) {
                const uint32_t mm = map[m];
                if (mm != m) {
                    map[i] = (uint16_t)mm;
                    changed = true;
                }
            }
        }
        if (!any(changed)) break;
        if (round > 20) return -1;  // chains of at most 2^16 units end within 17 rounds
    }
    // every source is final now and no final unit is written below
    uint32_t *o32 = reinterpret_cast<uint32_t *>(out);
    if (ub == 2) {
        for (uint32_t i = tid; i < nu; i += NT) {
            const uint32_t m = map[i];
            if (m != i) o32[i] = o32[m];
        }
    } else {
        for (uint32_t w = tid; w < (osize + 3) / 4; w += NT) {
            uint32_t add = 0;
#pragma unroll
            for (uint32_t b = 0; b < 4; ++b) {
                const uint32_t i = 4 * w + b;
                if (i < osize) {
                    const uint32_t m = map[i];
                    if (m != i) add |= (uint32_t)out[m] << (8 * b);
                }
            }
            if (add) o32[w] |= add;  // the word's other bytes are final; this thread owns it here
        }
    }
    __syncthreads();
    return osize;
}

// Decodes the LZ4I page staged at lds + l.stage (csize bytes, staged before a __syncthreads())
// into lds + l.out[0, osize). All NT threads call it. Returns osize, -1 if the page is malformed,
// -2 if this path does not take it (parse's cases, or more long literals or pending matches than
// a list holds). The caller then decodes the page in order. The result is the same in every thread.
// JUMP (l.jump set): matches the ordered pass leaves pending are resolved by pointer jumping
// instead of list passes when they are at least 1/16 of the output units.
template <int NT, bool JUMP = false>
__device__ __forceinline__ int64_t decode(uint8_t *lds, const Layout &l, uint32_t csize, uint32_t osize) {
    const uint32_t tid = threadIdx.x;
    uint8_t *out = lds + l.out;
    uint32_t *bits = reinterpret_cast<uint32_t *>(lds + l.bits);
    uint32_t *misc = reinterpret_cast<uint32_t *>(lds + l.misc);
    BlockAny any{misc + 4, 0};  // zeroed by stage()
    Parsed ps;
    if (const int r = parse<NT, true>(lds, l, csize, osize, any, ps)) return r;
    const uint32_t J = ps.J, ub = ps.ub;
    const bool words = ps.words;
    const uint8_t *blk = ps.blk;
    uint32_t *lt = ps.lt, *mo = ps.mo, *pos = ps.pos;
    bool bad = false;


    // literals up to SHORT bytes; mo[j] becomes the match destination | offset << 16, ml[] its length
    uint32_t ml[MAXJ], pend = 0, longs = 0;
#pragma unroll
    for (uint32_t j = 0; j < MAXJ; ++j) {
        ml[j] = 0;
        if (j >= J) continue;
        const uint32_t op = pos[j], s = lt[j] & 0xffffu, n = lt[j] >> 16;
        const uint32_t ns = n <= SHORT ? n : 0u;
        if (ns) put16(out, op, ld_any16(blk, s), ns, words);
        if (ns) bits_set(bits, op >> ub, ns >> ub);
        longs |= n > SHORT ? 1u << j : 0u;  // copied below, by the whole workgroup
        const uint32_t dst = op + n, off = mo[j] & 0xffffu;
        ml[j] = mo[j] >> 16;
        bad |= ml[j] && (off == 0 || off > dst);
        pend |= ml[j] ? 1u << j : 0u;
        mo[j] = dst | (off << 16);
    }
    // longer literals: gathered into a list (in the token array's place, no longer read) and
    // copied one per thread
    uint32_t *list = reinterpret_cast<uint32_t *>(lds + l.tok);
    const uint32_t list_cap = l.list_cap;
    if (tid == 0) misc[0] = 0;
    if (any(bad)) return -1;
    if (any(longs != 0)) {
        uint32_t pieces = 0;  // 16-byte pieces of the thread's long literals, one list entry each
#pragma unroll
        for (uint32_t j = 0; j < MAXJ; ++j) pieces += (longs >> j) & 1u ? ((lt[j] >> 16) + 15) / 16 : 0u;
        const uint32_t at = pieces ? atomicAdd(misc, pieces) : 0u;
        __syncthreads();
        const uint32_t total = misc[0];
        if (total > list_cap) {  // more pieces than the list holds: every thread copies its own
#pragma unroll
            for (uint32_t j = 0; j < MAXJ; ++j)
                if ((longs >> j) & 1u) {
                    literal_thread(out, pos[j], blk, lt[j] & 0xffffu, lt[j] >> 16, words);
                    bits_set_any(bits, pos[j] >> ub, (lt[j] >> 16) >> ub);
                }
        } else {
            uint32_t i = at;
#pragma unroll
            for (uint32_t j = 0; j < MAXJ; ++j)
                if ((longs >> j) & 1u) {
                    const uint32_t s = lt[j] & 0xffffu, n = lt[j] >> 16;
                    for (uint32_t c = 0; c < n; c += 16, ++i) {
                        list[2 * i] = (pos[j] + c) | (min(n - c, 16u) << 16);
                        list[2 * i + 1] = s + c;
                    }
                }
        }
        __syncthreads();
        for (uint32_t e = tid; e < (total <= list_cap ? total : 0u); e += NT) {
            const uint32_t op = list[2 * e] & 0xffffu, n = list[2 * e] >> 16, s = list[2 * e + 1];
            put16(out, op, ld_any16(blk, s), n, words);
            bits_set(bits, op >> ub, n >> ub);
        }
        __syncthreads();  // the list region is reused below
        if (tid == 0) misc[0] = 0;
    }

    // matches: one pass over the registers in sequence order (a barrier per step), short matches
    // whose source lies at least their length (or 16 bytes) back; a match whose source another
    // thread of the same step is still copying is tried again a few times before it goes to the
    // list
#pragma unroll
    for (uint32_t j = 0; j < MAXJ; ++j) {
        if (j >= J) continue;
        __syncthreads();  // every earlier sequence's match is done or pending
        if (!__any((pend >> j) & 1u)) continue;  // wave-uniform
        const uint32_t dst = mo[j] & 0xffffu, off = mo[j] >> 16, n = ml[j];
        const bool fast = ((pend >> j) & 1u) && n <= SHORT && (off >= 16 || off >= n);
        const uint32_t src = fast ? dst - off : 0u;
        bool todo = fast;
        for (uint32_t it = 0; it < STEP_TRIES && __any(todo); ++it) {
            const bool ready = todo && bits_all(bits, src >> ub, n >> ub);
            order();
            const u32x4 v = ld_any16(out, src);
            if (ready) put16(out, dst, v, n, words);
            order();
            if (ready) bits_set(bits, dst >> ub, n >> ub);
            pend &= ready ? ~(1u << j) : ~0u;
            todo = todo && !ready;
        }
    }
    // the rest, gathered into the list, worked through by the whole workgroup
    __syncthreads();
    const uint32_t mine = __builtin_popcount(pend);
    const uint32_t at = mine ? atomicAdd(misc, mine) : 0u;
    __syncthreads();
    const uint32_t total = misc[0];
    if (total == 0) return osize;
    // pointer jumping scans every output unit a round: worth it when pending matches are dense
    // (long copy chains), the list passes otherwise
    if constexpr (JUMP)
        if (16 * total > (osize >> ub)) return resolve_by_jumping<NT>(lds, l, osize, ub, pend, ps.mo, ml, any);
    if (total > list_cap) return -2;
    {
        uint32_t i = at;
#pragma unroll
        for (uint32_t j = 0; j < MAXJ; ++j)
            if ((pend >> j) & 1u) {
                list[2 * i] = mo[j];
                list[2 * i + 1] = ml[j];
                ++i;
            }
    }
    __syncthreads();
    // In the first pass a thread retries its entries without a barrier, up to SPIN times, while
    // other threads and waves complete the sources they wait for (the earliest pending match
    // always has a final source); every later pass tries each entry once and ends in a barrier.
    for (uint32_t pass = 0;; ++pass) {
        bool left = false;
        for (uint32_t it = 0; it < (pass ? 1u : SPIN); ++it) {
            left = false;
            for (uint32_t e = tid; e < total; e += NT) {
                const uint32_t n = list[2 * e + 1];
                if (!n) continue;
                const uint32_t dst = list[2 * e] & 0xffffu, off = list[2 * e] >> 16;
                const bool ready = bits_all_any(bits, (dst - off) >> ub, min(n, off) >> ub);
                order();
                if (ready) {
                    match_thread(out, dst, off, n, words);
                    order();
                    bits_set_any(bits, dst >> ub, n >> ub);
                    list[2 * e + 1] = 0;
                } else {
                    left = true;
                }
            }
            if (!left) break;
            order();  // the next try reads the bitmap again
        }
        if (!any(left)) return osize;
        if (pass > total) return -1;  // a pass that copies nothing: a broken invariant
    }
}

// Copies a page of csize bytes (16-byte aligned source) to lds + l.stage with 16-byte loads, all
// loads of a thread in flight together, and clears decode's reduction words. The caller
// synchronises before decode.
template <int NT>
__device__ __forceinline__ void stage(uint8_t *lds, const Layout &l, const uint8_t *src, uint32_t csize) {
    const u32x4 *s4 = reinterpret_cast<const u32x4 *>(src);
    u32x4 *d4 = reinterpret_cast<u32x4 *>(lds + l.stage);
    const uint32_t n16 = (csize + 15) / 16;
    constexpr int U = 8;
    if (threadIdx.x < 3) reinterpret_cast<uint32_t *>(lds + l.misc)[4 + threadIdx.x] = 0;  // decode's reductions
    for (uint32_t i0 = 0; i0 < n16; i0 += U * NT) {
        u32x4 v[U];
#pragma unroll
        for (int j = 0; j < U; ++j) {
            const uint32_t i = i0 + j * NT + threadIdx.x;
            if (i < n16) v[j] = s4[i];
        }
#pragma unroll
        for (int j = 0; j < U; ++j) {
            const uint32_t i = i0 + j * NT + threadIdx.x;
            if (i < n16) d4[i] = v[j];
        }
    }
}

}  // namespace lz4wg
}  // namespace golap_dev
