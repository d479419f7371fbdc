// Device helpers shared by the star-join kernels (single query: star.hip, many queries in one
// persistent launch: batch.hip): column access per page tile, the driver filter test, and the
// LDS-private group table with spill to global int64 atomics.
#pragma once
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "device_codecs.hpp"

namespace golap_impl {
namespace stardev {
using namespace golap_dev;
using golap_fmt::PAGE_CODEC_MASK;
using golap_fmt::PAGE_LITERAL_ONLY;

constexpr int BLOCK = 256;
constexpr uint32_t XCDS = 8;
constexpr uint32_t GROUP_EMPTY = 0xffffffffu;
constexpr int GROUP_PROBES = 8;

__device__ __forceinline__ void raise_err(uint32_t *err, uint32_t bit) {
    if (err) atomicOr(err, bit);
}
__device__ __forceinline__ bool is_direct(uint8_t codec) {
    return (codec & PAGE_CODEC_MASK) == golap_fmt::RAW || (codec & PAGE_LITERAL_ONLY);
}
__device__ __forceinline__ uint32_t direct_hdr(uint8_t codec, uint32_t nbytes) {
    return (codec & PAGE_LITERAL_ONLY) ? golap_fmt::lz4_literal_header_bytes(nbytes) : 0u;
}
__device__ __forceinline__ uint32_t hash32(uint32_t x) { return x * 0x9E3779B1u; }
// MurmurHash3 finaliser: all output bits depend on all key bits (the filter uses low and high).
__device__ __forceinline__ uint32_t fmix32(uint32_t h) {
    h ^= h >> 16;
    h *= 0x85ebca6bu;
    h ^= h >> 13;
    h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return h;
}
// Blocked Bloom filter of 2^wlog2 words: a key sets two bits of one word (one read to test).
__device__ __forceinline__ uint32_t bloom_word(uint32_t h, uint32_t wlog2) { return h >> (32 - wlog2); }
__device__ __forceinline__ uint32_t bloom_bits(uint32_t h) { return (1u << (h & 31)) | (1u << ((h >> 5) & 31)); }

// Chunk c of a query's slice holds candidate rows: its mask flag is set and every zone
// predicate can hold (check_mmmeta, core/metadata.cuh:539-547).
__device__ __forceinline__ bool slice_keeps(const SliceDev &s, uint64_t c) {
    bool pass = !s.chunk_mask || s.chunk_mask[c];
    for (uint32_t z = 0; z < SLICE_MAX_ZONES; ++z)
        if (pass && s.zone[z].zmin) pass = s.zone[z].lo <= s.zone[z].zmax[c] && s.zone[z].hi >= s.zone[z].zmin[c];
    return pass;
}

struct StarCol {
    ColView c;
    int window;   // LDS window index when some page of the column in range needs decoding, else -1
    int anchors;  // without a window: LDS prefix-sum region for its delta-bitpack pages, else -1
};

// How one 32-bit column is read at a tile row.
enum : uint32_t { ACC_DIRECT = 0, ACC_WINDOW = 1, ACC_AFFINE = 2, ACC_RLE = 3, ACC_FOR = 4, ACC_BAD = 5, ACC_DBPA = 6 };
struct ColAccess {
    const uint8_t *base;  // DIRECT: payload in HBM; WINDOW: LDS window; RLE: page start; FOR, DBPA: fields
    int64_t a, b;         // AFFINE: v[r] = a + r * b; FOR: a = base; DBPA: a = v[0], b = min delta
    uint32_t mode;
    uint32_t nruns;       // RLE: runs; FOR, DBPA: bits per field
    uint32_t *anc;        // DBPA: LDS prefix sums of the fields, one per 32 rows
    uint32_t words;       // DBPA: packed words
};

// Delta-bitpack field j (row j + 1's delta minus the minimum) of a page's packed words.
__device__ __forceinline__ uint32_t dbp_field(const uint32_t *pk, uint32_t words, uint32_t bw, uint32_t j) {
    const uint64_t bit = (uint64_t)j * bw;
    const uint32_t w = (uint32_t)(bit >> 5);
    const uint64_t lo = pk[w], hi = w + 1 < words ? pk[w + 1] : 0;
    return (uint32_t)(((lo | (hi << 32)) >> (bit & 31)) & (bw == 32 ? 0xffffffffull : ((1ull << bw) - 1)));
}

// Sets up the access to column sc for page p (rows rows); wwin: its LDS window or null; anc: its
// LDS prefix-sum region or null. A delta-bitpack page that is not point-readable is read
// through its window when it has one, else through prefix sums the tile computes (ACC_DBPA).
__device__ inline ColAccess col_access(const StarCol &sc, uint64_t p, uint32_t rows, uint8_t *wwin,
                                       uint32_t *anc = nullptr) {
    ColAccess x{};
    const uint8_t codec = sc.c.page_codec[p];
    const uint8_t *src = sc.c.comp + sc.c.page_off[p];
    const uint32_t csize = sc.c.page_csize[p];
    if (is_direct(codec)) {
        const uint32_t hdr = direct_hdr(codec, rows * 4);
        x.base = src + hdr;
        x.mode = (csize == hdr + rows * 4 && !(reinterpret_cast<uintptr_t>(x.base) & 15)) ? ACC_DIRECT : ACC_BAD;
    } else if ((codec & PAGE_CODEC_MASK) == golap_fmt::DBP && (codec & golap_fmt::PAGE_POINT)) {
        if (csize < 32 || src[20] != 0 || src[21] != 4 || ld_u32_unaligned(src + 16) != rows) {
            x.mode = ACC_BAD;
        } else {
            x.a = (int64_t)ld_u64_unaligned(src);
            x.b = (int64_t)ld_u64_unaligned(src + 8);
            x.mode = ACC_AFFINE;
        }
    } else if ((codec & PAGE_CODEC_MASK) == golap_fmt::FOR) {
        if (for_check(src, csize, 4) != (int64_t)rows) {
            x.mode = ACC_BAD;
        } else {
            x.base = src + 32;
            x.a = (int64_t)ld_u64_unaligned(src);
            x.nruns = src[20];
            x.mode = ACC_FOR;
        }
    } else if ((codec & PAGE_CODEC_MASK) == golap_fmt::RLE) {
        const uint32_t nruns = ld_u32_unaligned(src), n = ld_u32_unaligned(src + 4);
        const bool ok = csize >= 16 && src[8] == 4 && n == rows && nruns >= 1 &&
                        16ull + 4ull * nruns + 4ull * nruns <= csize;
        x.base = src;
        x.nruns = nruns;
        x.mode = ok ? ACC_RLE : ACC_BAD;
    } else if ((codec & PAGE_CODEC_MASK) == golap_fmt::DBP && !wwin && anc) {
        const uint32_t bw = src[20];
        const uint32_t words = rows ? (uint32_t)(((uint64_t)(rows - 1) * bw + 31) / 32) : 0;
        if (csize < 32 || bw == 0 || bw > 32 || src[21] != 4 || ld_u32_unaligned(src + 16) != rows ||
            32ull + 4ull * words > csize || (reinterpret_cast<uintptr_t>(src) & 3)) {
            x.mode = ACC_BAD;
        } else {
            x.base = src + 32;
            x.a = (int64_t)ld_u64_unaligned(src);
            x.b = (int64_t)ld_u64_unaligned(src + 8);
            x.nruns = bw;
            x.words = words;
            x.anc = anc;
            x.mode = ACC_DBPA;
        }
    } else {
        x.base = wwin;
        x.mode = wwin ? ACC_WINDOW : ACC_BAD;
    }
    return x;
}

// ACC_DBPA set-up by one wave: anc[g] = sum of the fields of rows below 32g (mod 2^32), for the
// ceil(rows / 32) groups of the page.
__device__ inline void dbp_anchors_wave(const ColAccess &x, uint32_t rows) {
    const uint32_t lane = threadIdx.x % WAVE;
    const uint32_t *pk = reinterpret_cast<const uint32_t *>(x.base);
    const uint32_t bw = x.nruns, nf = rows ? rows - 1 : 0, groups = (rows + 31) / 32;
    uint32_t carry = 0;
    for (uint32_t g0 = 0; g0 < groups; g0 += WAVE) {
        const uint32_t g = g0 + lane;
        uint32_t s = 0;
        if (g < groups && 32 * g < nf) {
            const uint32_t j1 = min(nf, 32 * g + 32);
            if (bw == 1) {
                const uint32_t k = j1 - 32 * g;
                s = __popc(pk[g] & (k == 32 ? 0xffffffffu : ((1u << k) - 1)));
            } else {
                for (uint32_t j = 32 * g; j < j1; ++j) s += dbp_field(pk, x.words, bw, j);
            }
        }
        uint32_t incl = s;
        for (int d = 1; d < WAVE; d <<= 1) {
            const uint32_t t = __shfl_up(incl, d, WAVE);
            if (lane >= (uint32_t)d) incl += t;
        }
        if (g < groups) x.anc[g] = carry + incl - s;
        carry += __shfl(incl, WAVE - 1, WAVE);
    }
}

// DBPA = false: a kernel that never sets up ACC_DBPA (no prefix-sum regions) leaves its code out.
template <bool DBPA = true>
__device__ __forceinline__ int32_t col_value(const ColAccess &x, uint32_t r) {
    int32_t v;
    switch (DBPA ? x.mode : (x.mode == ACC_DBPA ? ACC_BAD : x.mode)) {
    case ACC_AFFINE: return (int32_t)(uint32_t)(uint64_t)(x.a + (int64_t)r * x.b);
    case ACC_FOR: return (int32_t)(uint32_t)for_value(x.base, (uint64_t)x.a, x.nruns, r);
    case ACC_DBPA: {
        if (!DBPA) return 0;
        // v[r] = v[0] + r * min_delta + (fields of rows 1..r): the group's prefix plus the rest
        const uint32_t g = r >> 5, j0 = g << 5;
        const uint32_t *pk = reinterpret_cast<const uint32_t *>(x.base);
        uint32_t s = x.anc[g];
        if (x.nruns == 1) {
            if (r > j0) s += __popc(pk[g] & ((1u << (r - j0)) - 1));
        } else {
            for (uint32_t j = j0; j < r; ++j) s += dbp_field(pk, x.words, x.nruns, j);
        }
        return (int32_t)((uint32_t)(uint64_t)x.a + r * (uint32_t)(uint64_t)x.b + s);
    }
    case ACC_RLE: {
        // first run whose exclusive end is above r
        const uint8_t *ends = x.base + 16 + 4ull * x.nruns;
        uint32_t lo = 0, hi = x.nruns - 1;
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (ld_u32_unaligned(ends + 4ull * mid) > r) hi = mid;
            else lo = mid + 1;
        }
        return (int32_t)ld_u32_unaligned(x.base + 16 + 4ull * lo);
    }
    default:
        __builtin_memcpy(&v, x.base + 4 * (size_t)r, 4);
        return v;
    }
}

// Driver filter: dense bits over [fkmin, fkmin + fkrange) when fhash_log2 == 0, else a blocked
// Bloom filter of 2^fhash_log2 words.
using DriverFilter = DriverFilterView;
__device__ __forceinline__ bool filter_test(const DriverFilter &df, int32_t key) {
    if (df.fhash_log2 == 0) {
        uint32_t bit = (uint32_t)key - (uint32_t)df.fkmin;
        const bool in = bit < df.fkrange;
        bit = in ? bit : 0u;
        return in && ((df.f[bit >> 5] >> (bit & 31)) & 1u);
    }
    const uint32_t h = fmix32((uint32_t)key), m = bloom_bits(h);
    return (df.f[bloom_word(h, df.fhash_log2)] & m) == m;
}

// Exact probe of an open-addressing table (8-byte slots, key ^ 0x80000000 in the low half).
__device__ __forceinline__ bool probe_exact(const uint64_t *slots, uint32_t log2cap, int32_t key,
                                            uint64_t first, uint32_t h, uint32_t &payload) {
    const uint32_t mask = (1u << log2cap) - 1, sk = (uint32_t)key ^ 0x80000000u;
    uint64_t s = first;
    for (uint32_t t = 0; s != 0 && (uint32_t)s != sk && t < mask; ++t) {
        h = (h + 1) & mask;
        s = slots[h];
    }
    payload = (uint32_t)(s >> 32);
    return s != 0 && (uint32_t)s == sk;
}

// LDS-private hashed group table (keys, counts, int64 sums) with spill to global atomics.
struct GroupTable {
    uint32_t *gkey;
    uint32_t *gcnt;
    unsigned long long *gsum;
    uint32_t slots;  // power of two, 0 = no LDS table
};
__device__ __forceinline__ bool group_add_lds(const GroupTable &T, uint32_t key, int64_t v) {
    if (!T.slots) return false;
    const uint32_t mask = T.slots - 1;
    uint32_t h = (hash32(key) >> 7) & mask;
    for (int i = 0; i < GROUP_PROBES; ++i) {
        uint32_t k = T.gkey[h];
        if (k == GROUP_EMPTY) {
            const uint32_t old = atomicCAS(&T.gkey[h], GROUP_EMPTY, key);
            k = old == GROUP_EMPTY ? key : old;
        }
        if (k == key) {
            atomicAdd(&T.gsum[h], (unsigned long long)v);
            atomicAdd(&T.gcnt[h], 1u);
            return true;
        }
        h = (h + 1) & mask;
    }
    return false;  // spill
}
__device__ __forceinline__ void group_add_global(int64_t *g_sum, uint64_t *g_cnt, uint64_t gidx, int64_t v) {
    atomicAdd(reinterpret_cast<unsigned long long *>(&g_sum[gidx]), (unsigned long long)v);
    atomicAdd(reinterpret_cast<unsigned long long *>(&g_cnt[gidx]), 1ull);
}

}  // namespace stardev
}  // namespace golap_impl
