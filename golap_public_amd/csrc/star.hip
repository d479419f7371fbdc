// Fused star join + group-by over paged fact columns (gfx950).
//
// Replaces the reference's per-chunk pipeline_q2 / pipeline_q3 kernels
// (usecases/ssb/queries/02.cuh:102-193, 03.cuh:118-170) that probe chained HashMaps and
// aggregate with HashAggregate::add (core/dev_structs.cuh:209-244) after nvcomp decompressed
// each chunk. One workgroup handles one page-row tile:
//   1. the driving dimension's foreign keys are tested against its membership filter pinned in
//      LDS (exact bitmap for a dense key range, one-hash bit filter otherwise);
//   2. passing rows are compacted into an LDS candidate list (wave prefix sum + one LDS atomic
//      per wave);
//   3. only for tiles with candidates, coded pages of the other columns are decoded into LDS;
//      in-place pages are read at the candidate rows only;
//   4. candidates probe every dimension's exact table (open addressing, 8-byte slots), form a
//      dense group index from the dimensions' payload attributes and add SUM/COUNT into an
//      LDS-private hashed group table that spills to global int64 atomics when a slot run is
//      full; the table is merged into the global dense result once per workgroup.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "common.hpp"
#include "device_codecs.hpp"

namespace golap_impl {
using namespace golap_dev;
using golap_fmt::PAGE_CODEC_MASK;
using golap_fmt::PAGE_LITERAL_ONLY;

namespace {
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
// Blocked Bloom filter of 2^wlog2 words: a key sets two bits of one word (one LDS read to test).
__device__ __forceinline__ uint32_t bloom_word(uint32_t h, uint32_t wlog2) { return h >> (32 - wlog2); }
__device__ __forceinline__ uint32_t bloom_bits(uint32_t h) { return (1u << (h & 31)) | (1u << ((h >> 5) & 31)); }

}  // namespace

struct StarCol {
    ColView c;
    int window;  // LDS window index when some page of the column in range needs decoding, else -1
};

struct StarArgs {
    StarCol fk[STAR_MAX_DIMS];
    StarCol val;
    uint32_t val_signed;
    StarDimProbe dim[STAR_MAX_DIMS];
    uint32_t n_dims;
    uint32_t driver;
    // driver filter (LDS): dense bits over [fkmin, fkmin + fkrange) when fhash_log2 == 0,
    // else a blocked Bloom filter of 2^fhash_log2 words (bloom_word / bloom_bits of fmix32(key))
    const uint32_t *filter;
    uint32_t filter_words;
    int32_t fkmin;
    uint32_t fkrange;
    uint32_t fhash_log2;
    uint32_t page_rows;
    uint64_t n_rows;
    uint64_t page_begin, page_end;
    const uint16_t *chunk_mask;
    uint32_t pages_per_chunk;
    uint32_t n_windows;
    uint64_t n_groups;
    uint32_t lds_gslots;  // LDS group table slots (power of two), 0 = global atomics only
    int64_t *g_sum;
    uint64_t *g_cnt;
    uint64_t *rows_matched;
    uint32_t *err;
};

struct StarLds {
    uint32_t *filter;
    uint16_t *cand;
    uint8_t *win;
    uint32_t *gkey;
    uint32_t *gcnt;
    unsigned long long *gsum;
};

// How one 32-bit column is read at a tile row.
enum : uint32_t { ACC_DIRECT = 0, ACC_WINDOW = 1, ACC_AFFINE = 2, ACC_RLE = 3, ACC_BAD = 4 };
struct ColAccess {
    const uint8_t *base;  // DIRECT: payload in HBM; WINDOW: LDS window; RLE: page start
    int64_t a, b;         // AFFINE: v[r] = a + r * b
    uint32_t mode;
    uint32_t nruns;       // RLE
};

// Sets up the access to column sc for page p (rows rows); wwin: its LDS window or null.
__device__ inline ColAccess col_access(const StarCol &sc, uint64_t p, uint32_t rows, uint8_t *wwin) {
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
    } else if ((codec & PAGE_CODEC_MASK) == golap_fmt::RLE) {
        const uint32_t nruns = ld_u32_unaligned(src), n = ld_u32_unaligned(src + 4);
        const bool ok = csize >= 16 && src[8] == 4 && n == rows && nruns >= 1 &&
                        16ull + 4ull * nruns + 4ull * nruns <= csize;
        x.base = src;
        x.nruns = nruns;
        x.mode = ok ? ACC_RLE : ACC_BAD;
    } else {
        x.base = wwin;
        x.mode = wwin ? ACC_WINDOW : ACC_BAD;
    }
    return x;
}

__device__ __forceinline__ int32_t col_value(const ColAccess &x, uint32_t r) {
    int32_t v;
    switch (x.mode) {
    case ACC_AFFINE: return (int32_t)(uint32_t)(uint64_t)(x.a + (int64_t)r * x.b);
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

__device__ __forceinline__ bool star_filter(const StarArgs &a, const uint32_t *f, int32_t key) {
    if (a.fhash_log2 == 0) {
        uint32_t bit = (uint32_t)key - (uint32_t)a.fkmin;
        const bool in = bit < a.fkrange;
        bit = in ? bit : 0u;
        return in && ((f[bit >> 5] >> (bit & 31)) & 1u);
    }
    const uint32_t h = fmix32((uint32_t)key), m = bloom_bits(h);
    return (f[bloom_word(h, a.fhash_log2)] & m) == m;
}

__device__ __forceinline__ void group_add(const StarArgs &a, const StarLds &L, uint32_t gidx, int64_t v) {
    if (a.lds_gslots) {
        const uint32_t mask = a.lds_gslots - 1;
        uint32_t h = (hash32(gidx) >> 7) & mask;
        for (int i = 0; i < GROUP_PROBES; ++i) {
            uint32_t k = L.gkey[h];
            if (k == GROUP_EMPTY) {
                const uint32_t old = atomicCAS(&L.gkey[h], GROUP_EMPTY, gidx);
                k = old == GROUP_EMPTY ? gidx : old;
            }
            if (k == gidx) {
                atomicAdd(&L.gsum[h], (unsigned long long)v);
                atomicAdd(&L.gcnt[h], 1u);
                return;
            }
            h = (h + 1) & mask;
        }
    }
    // spill (or no LDS table): straight into the global dense result
    atomicAdd(reinterpret_cast<unsigned long long *>(&a.g_sum[gidx]), (unsigned long long)v);
    atomicAdd(reinterpret_cast<unsigned long long *>(&a.g_cnt[gidx]), 1ull);
}

// PER = page_rows / BLOCK rows per thread (4, 8, 16 or 32), taken as groups of four.
template <int PER>
__global__ __launch_bounds__(BLOCK) void k_star(StarArgs a) {
    extern __shared__ __align__(16) uint8_t lds[];
    __shared__ uint32_t cand_n, tile_bad;
    __shared__ ColAccess acc[STAR_MAX_DIMS + 1];
    constexpr int G = PER / 4;
    const uint32_t P = a.page_rows;
    StarLds L;
    size_t off = 0;
    L.filter = reinterpret_cast<uint32_t *>(lds);
    off += ((size_t)a.filter_words * 4 + 15) & ~(size_t)15;
    L.cand = reinterpret_cast<uint16_t *>(lds + off);
    off += ((size_t)P * 2 + 15) & ~(size_t)15;
    L.win = lds + off;
    off += (size_t)a.n_windows * P * 4;
    L.gkey = reinterpret_cast<uint32_t *>(lds + off);
    off += (size_t)a.lds_gslots * 4;
    L.gcnt = reinterpret_cast<uint32_t *>(lds + off);
    off += (size_t)a.lds_gslots * 4;
    L.gsum = reinterpret_cast<unsigned long long *>(lds + off);

    for (uint32_t i = threadIdx.x; i < a.filter_words; i += BLOCK) L.filter[i] = a.filter[i];
    for (uint32_t i = threadIdx.x; i < a.lds_gslots; i += BLOCK) {
        L.gkey[i] = GROUP_EMPTY;
        L.gcnt[i] = 0;
        L.gsum[i] = 0;
    }
    __syncthreads();

    const uint32_t wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    const uint32_t ncols = a.n_dims + 1;  // fk[0..n_dims) then val
    uint64_t matched = 0;

    // XCD-aware tile order (workgroup i runs on XCD i mod 8; each XCD gets a contiguous range)
    uint64_t first = a.page_begin + blockIdx.x, end = a.page_end, stride = gridDim.x;
    if (gridDim.x % XCDS == 0) {
        const uint64_t n = a.page_end - a.page_begin;
        const uint32_t x = blockIdx.x % XCDS;
        first = a.page_begin + n * x / XCDS + blockIdx.x / XCDS;
        end = a.page_begin + n * (x + 1) / XCDS;
        stride = gridDim.x / XCDS;
    }
    int4 kv[G];  // driver keys of the current tile
    for (uint64_t p = first; p < end; p += stride) {
        if (a.chunk_mask && a.chunk_mask[p / a.pages_per_chunk] == 0) {
                continue;
        }
        const uint32_t rows = (uint32_t)min<uint64_t>(P, a.n_rows - p * P);
        if (threadIdx.x == 0) {
            cand_n = 0;
            tile_bad = 0;
        }
        if (threadIdx.x < ncols) {
            const StarCol &sc = threadIdx.x < a.n_dims ? a.fk[threadIdx.x] : a.val;
            acc[threadIdx.x] = col_access(sc, p, rows, sc.window >= 0 ? L.win + (size_t)sc.window * P * 4 : nullptr);
        }
        __syncthreads();
        bool bad = false;
        for (uint32_t c = 0; c < ncols; ++c) bad |= acc[c].mode == ACC_BAD;
        if (bad) {  // uniform: every thread read the same entries
            if (threadIdx.x == 0) raise_err(a.err, DEVERR_CORRUPT);
                __syncthreads();
            continue;
        }
        const StarCol &dsc = a.fk[a.driver];
        const ColAccess dacc = acc[a.driver];
        if (dacc.mode == ACC_WINDOW) {
            if (wave == 0) {
                const int64_t got = page_decode_wave(dsc.c.page_codec[p], dsc.c.comp + dsc.c.page_off[p],
                                                     dsc.c.page_csize[p], const_cast<uint8_t *>(dacc.base),
                                                     rows * 4, 4);
                if (got != (int64_t)rows * 4 && lane == 0) {
                    tile_bad = 1;
                    raise_err(a.err, DEVERR_CORRUPT);
                }
            }
            __syncthreads();
        }
        // filter + compaction: thread t tests rows 4(t + k*BLOCK) + j; a wave prefix sum over the
        // per-thread counts gives every thread its slots in the candidate list
        uint32_t passbits = 0;
        const bool vec = rows == (uint32_t)(PER * BLOCK) && dacc.mode <= ACC_WINDOW;
        if (!tile_bad) {
            if (vec) {
#pragma unroll
                for (int k = 0; k < G; ++k)
                    kv[k] = *reinterpret_cast<const int4 *>(dacc.base + 16 * ((size_t)threadIdx.x + (size_t)k * BLOCK));
#pragma unroll
                for (int k = 0; k < G; ++k) {
                    passbits |= (uint32_t)star_filter(a, L.filter, kv[k].x) << (4 * k);
                    passbits |= (uint32_t)star_filter(a, L.filter, kv[k].y) << (4 * k + 1);
                    passbits |= (uint32_t)star_filter(a, L.filter, kv[k].z) << (4 * k + 2);
                    passbits |= (uint32_t)star_filter(a, L.filter, kv[k].w) << (4 * k + 3);
                }
            } else {
                for (int k = 0; k < G; ++k)
                    for (int j = 0; j < 4; ++j) {
                        const uint32_t r = 4 * (threadIdx.x + k * BLOCK) + j;
                        if (r < rows && star_filter(a, L.filter, col_value(dacc, r))) passbits |= 1u << (4 * k + j);
                    }
            }
        }
        const uint32_t mine = __popc(passbits);
        uint32_t incl = mine;
#pragma unroll
        for (int d = 1; d < WAVE; d <<= 1) {
            const uint32_t t = __shfl_up(incl, d, WAVE);
            if (lane >= (uint32_t)d) incl += t;
        }
        const uint32_t wtotal = __shfl(incl, WAVE - 1, WAVE);
        uint32_t wbase = 0;
        if (lane == WAVE - 1 && wtotal) wbase = atomicAdd(&cand_n, wtotal);
        wbase = __shfl(wbase, WAVE - 1, WAVE);
        uint32_t slot = wbase + incl - mine;
        for (uint32_t bits = passbits; bits; bits &= bits - 1) {
            const uint32_t b = __ffs(bits) - 1;
            L.cand[slot++] = (uint16_t)(4 * (threadIdx.x + (b >> 2) * BLOCK) + (b & 3));
        }
        __syncthreads();
        const uint32_t nc = cand_n;
        if (nc == 0 || tile_bad) {
            __syncthreads();
            continue;
        }
        // other columns that need a decode: wave w decodes columns w, w+4, ...
        for (uint32_t c = wave; c < ncols; c += BLOCK / WAVE) {
            if (c == a.driver || acc[c].mode != ACC_WINDOW) continue;
            const StarCol &sc = c < a.n_dims ? a.fk[c] : a.val;
            const int64_t got = page_decode_wave(sc.c.page_codec[p], sc.c.comp + sc.c.page_off[p],
                                                 sc.c.page_csize[p], const_cast<uint8_t *>(acc[c].base),
                                                 rows * 4, 4);
            if (got != (int64_t)rows * 4 && lane == 0) {
                tile_bad = 1;
                raise_err(a.err, DEVERR_CORRUPT);
            }
        }
        __syncthreads();
        if (!tile_bad) {
            for (uint32_t i = threadIdx.x; i < nc; i += BLOCK) {
                const uint32_t r = L.cand[i];
                // exact probe of the driver first (its small table stays in L2): filter false
                // positives stop here, before any other column is read
                const StarDimProbe &dd = a.dim[a.driver];
                {
                    const int32_t dk = col_value(acc[a.driver], r);
                    const uint32_t mask = (1u << dd.log2cap) - 1, sk = (uint32_t)dk ^ 0x80000000u;
                    uint32_t hh = hash32((uint32_t)dk) >> (32 - dd.log2cap);
                    uint64_t s = dd.slots[hh];
                    for (uint32_t t = 0; s != 0 && (uint32_t)s != sk && t < mask; ++t) {
                        hh = (hh + 1) & mask;
                        s = dd.slots[hh];
                    }
                    if (s == 0 || (uint32_t)s != sk) continue;
                }
                // then every column value of the row and the first slot of every table: two
                // memory round trips for the row instead of two per dimension
                int32_t key[STAR_MAX_DIMS];
                for (uint32_t d = 0; d < a.n_dims; ++d) key[d] = col_value(acc[d], r);
                const int32_t raw = col_value(acc[a.n_dims], r);
                uint32_t h[STAR_MAX_DIMS];
                uint64_t slot[STAR_MAX_DIMS];
                for (uint32_t d = 0; d < a.n_dims; ++d) {
                    h[d] = hash32((uint32_t)key[d]) >> (32 - a.dim[d].log2cap);
                    slot[d] = a.dim[d].slots[h[d]];
                }
                uint64_t gidx = 0;
                bool ok = true;
                for (uint32_t d = 0; d < a.n_dims && ok; ++d) {
                    const uint32_t mask = (1u << a.dim[d].log2cap) - 1, sk = (uint32_t)key[d] ^ 0x80000000u;
                    uint64_t s = slot[d];
                    uint32_t hh = h[d];
                    for (uint32_t t = 0; s != 0 && (uint32_t)s != sk && t < mask; ++t) {
                        hh = (hh + 1) & mask;
                        s = a.dim[d].slots[hh];
                    }
                    ok = s != 0 && (uint32_t)s == sk;
                    if (ok && a.dim[d].group_card) {
                        const uint32_t gv = (uint32_t)(s >> 32) - (uint32_t)a.dim[d].group_lo;
                        if (gv >= a.dim[d].group_card) {
                            raise_err(a.err, DEVERR_KEYRANGE);
                            ok = false;
                        }
                        gidx += (uint64_t)gv * a.dim[d].group_stride;
                    }
                }
                if (!ok) continue;
                const int64_t v = a.val_signed ? (int64_t)raw : (int64_t)(uint32_t)raw;
                group_add(a, L, (uint32_t)gidx, v);
                ++matched;
            }
        }
        __syncthreads();  // candidate list, windows and access entries are reused by the next tile
    }
    // merge the LDS-private table into the global dense result
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < a.lds_gslots; i += BLOCK) {
        const uint32_t k = L.gkey[i];
        if (k == GROUP_EMPTY) continue;
        atomicAdd(reinterpret_cast<unsigned long long *>(&a.g_sum[k]), L.gsum[i]);
        atomicAdd(reinterpret_cast<unsigned long long *>(&a.g_cnt[k]), (unsigned long long)L.gcnt[i]);
    }
    uint64_t m = matched;
#pragma unroll
    for (int d = WAVE / 2; d > 0; d >>= 1) m += __shfl_xor(m, d, WAVE);
    if (lane == 0 && m) atomicAdd(reinterpret_cast<unsigned long long *>(a.rows_matched), (unsigned long long)m);
}

size_t star_lds_bytes(uint32_t filter_words, uint32_t page_rows, uint32_t n_windows, uint32_t lds_gslots) {
    return (((size_t)filter_words * 4 + 15) & ~(size_t)15) + (((size_t)page_rows * 2 + 15) & ~(size_t)15) +
           (size_t)n_windows * page_rows * 4 + (size_t)lds_gslots * 16;
}

template <typename K>
static hipError_t launch_star_t(golap_ctx *ctx, K kernel, const StarArgs &a, size_t lds, const StarLaunch &l) {
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    int occ = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kernel, BLOCK, lds);
    if (e != hipSuccess) return e;
    occ = std::max(occ, 1);
    const uint64_t tiles = l.page_end - l.page_begin;
    uint64_t grid = std::max<uint64_t>(1, std::min<uint64_t>(tiles, (uint64_t)ctx->cus * occ));
    if (grid >= XCDS) grid -= grid % XCDS;
    hipLaunchKernelGGL(kernel, dim3((uint32_t)grid), dim3(BLOCK), lds, ctx->stream, a);
    return hipGetLastError();
}

hipError_t launch_star(golap_ctx *ctx, const StarLaunch &l) {
    StarArgs a{};
    for (uint32_t d = 0; d < l.n_dims; ++d) {
        a.fk[d] = StarCol{l.fk[d], l.fk_window[d]};
        a.dim[d] = l.dim[d];
    }
    a.val = StarCol{l.val, l.val_window};
    a.val_signed = l.val_signed;
    a.n_dims = l.n_dims;
    a.driver = l.driver;
    a.filter = l.filter;
    a.filter_words = l.filter_words;
    a.fkmin = l.fkmin;
    a.fkrange = l.fkrange;
    a.fhash_log2 = l.fhash_log2;
    a.page_rows = l.page_rows;
    a.n_rows = l.n_rows;
    a.page_begin = l.page_begin;
    a.page_end = l.page_end;
    a.chunk_mask = l.chunk_mask;
    a.pages_per_chunk = l.pages_per_chunk;
    a.n_windows = l.n_windows;
    a.n_groups = l.n_groups;
    a.lds_gslots = l.lds_gslots;
    a.g_sum = l.g_sum;
    a.g_cnt = l.g_cnt;
    a.rows_matched = l.rows_matched;
    a.err = l.err;
    const size_t lds = star_lds_bytes(l.filter_words, l.page_rows, l.n_windows, l.lds_gslots);
    if (lds > (size_t)ctx->lds_max) return hipErrorInvalidConfiguration;
    switch (l.page_rows) {
    case 1024: return launch_star_t(ctx, k_star<4>, a, lds, l);
    case 2048: return launch_star_t(ctx, k_star<8>, a, lds, l);
    case 4096: return launch_star_t(ctx, k_star<16>, a, lds, l);
    case 8192: return launch_star_t(ctx, k_star<32>, a, lds, l);
    default: return hipErrorInvalidValue;
    }
}

// group_by_agg_2col / _4col (core/group.cuh:14-36) over plain device columns: one thread per
// row forms the dense group index of up to four int key columns and adds the value into the
// LDS-private table (spilling to global atomics), merged once per workgroup.
struct GroupByArgs {
    const void *key[4];
    uint32_t key_type[4];
    int32_t key_lo[4];
    uint32_t key_card[4];
    uint64_t key_stride[4];
    uint32_t n_keys;
    const void *val;
    uint32_t val_type;
    uint64_t n;
    uint32_t lds_gslots;
    int64_t *g_sum;
    uint64_t *g_cnt;
    uint32_t *err;
};

__device__ __forceinline__ int64_t ld_typed(const void *col, uint32_t type, uint64_t row) {
    switch (type) {
    case GOLAP_I8: return static_cast<const int8_t *>(col)[row];
    case GOLAP_U8: return static_cast<const uint8_t *>(col)[row];
    case GOLAP_I16: return static_cast<const int16_t *>(col)[row];
    case GOLAP_U16: return static_cast<const uint16_t *>(col)[row];
    case GOLAP_I32: return static_cast<const int32_t *>(col)[row];
    case GOLAP_U32: return static_cast<const uint32_t *>(col)[row];
    default: return static_cast<const int64_t *>(col)[row];
    }
}

__global__ __launch_bounds__(BLOCK) void k_group_by(GroupByArgs g) {
    extern __shared__ __align__(16) uint8_t lds[];
    StarArgs a{};
    a.lds_gslots = g.lds_gslots;
    a.g_sum = g.g_sum;
    a.g_cnt = g.g_cnt;
    StarLds L{};
    L.gkey = reinterpret_cast<uint32_t *>(lds);
    L.gcnt = L.gkey + g.lds_gslots;
    L.gsum = reinterpret_cast<unsigned long long *>(lds + (size_t)g.lds_gslots * 8);
    for (uint32_t i = threadIdx.x; i < g.lds_gslots; i += BLOCK) {
        L.gkey[i] = GROUP_EMPTY;
        L.gcnt[i] = 0;
        L.gsum[i] = 0;
    }
    __syncthreads();
    for (uint64_t r = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; r < g.n; r += (uint64_t)gridDim.x * BLOCK) {
        uint64_t gidx = 0;
        bool ok = true;
        for (uint32_t k = 0; k < g.n_keys; ++k) {
            const uint64_t v = (uint64_t)(ld_typed(g.key[k], g.key_type[k], r) - (int64_t)g.key_lo[k]);
            if (v >= g.key_card[k]) ok = false;
            gidx += v * g.key_stride[k];
        }
        if (!ok) {
            raise_err(g.err, DEVERR_KEYRANGE);
            continue;
        }
        group_add(a, L, (uint32_t)gidx, ld_typed(g.val, g.val_type, r));
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < g.lds_gslots; i += BLOCK) {
        const uint32_t k = L.gkey[i];
        if (k == GROUP_EMPTY) continue;
        atomicAdd(reinterpret_cast<unsigned long long *>(&g.g_sum[k]), L.gsum[i]);
        atomicAdd(reinterpret_cast<unsigned long long *>(&g.g_cnt[k]), (unsigned long long)L.gcnt[i]);
    }
}

hipError_t launch_group_by(golap_ctx *ctx, const GroupByLaunch &l) {
    GroupByArgs g{};
    for (uint32_t k = 0; k < l.n_keys; ++k) {
        g.key[k] = l.key[k];
        g.key_type[k] = l.key_type[k];
        g.key_lo[k] = l.key_lo[k];
        g.key_card[k] = l.key_card[k];
        g.key_stride[k] = l.key_stride[k];
    }
    g.n_keys = l.n_keys;
    g.val = l.val;
    g.val_type = l.val_type;
    g.n = l.n;
    g.lds_gslots = l.lds_gslots;
    g.g_sum = l.g_sum;
    g.g_cnt = l.g_cnt;
    g.err = l.err;
    const size_t lds = (size_t)l.lds_gslots * 16;
    int occ = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_group_by, BLOCK, lds);
    if (e != hipSuccess) return e;
    const uint64_t grid = std::max<uint64_t>(1, std::min<uint64_t>((l.n + BLOCK - 1) / BLOCK,
                                                                   (uint64_t)ctx->cus * std::max(occ, 1)));
    hipLaunchKernelGGL(k_group_by, dim3((uint32_t)grid), dim3(BLOCK), lds, ctx->stream, g);
    return hipGetLastError();
}

// Blocked Bloom filter (2^wlog2 words) over the keys of a table: the driver filter when the key
// range is too wide for the dense bitmap.
__global__ void k_hash_filter(const uint64_t *slots, uint64_t n, uint32_t *filter, uint32_t wlog2) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t s = slots[i];
        if (s == 0) continue;
        const uint32_t h = fmix32((uint32_t)s ^ 0x80000000u);
        atomicOr(&filter[bloom_word(h, wlog2)], bloom_bits(h));
    }
}

hipError_t launch_hash_filter(golap_ctx *ctx, const uint64_t *slots, uint32_t log2cap, uint32_t *filter,
                              uint32_t words_log2) {
    const uint64_t n = 1ull << log2cap;
    const uint32_t grid = (uint32_t)std::min<uint64_t>((n + BLOCK - 1) / BLOCK, 4096);
    hipLaunchKernelGGL(k_hash_filter, dim3(std::max(grid, 1u)), dim3(BLOCK), 0, ctx->stream, slots, n, filter,
                       words_log2);
    return hipGetLastError();
}

}  // namespace golap_impl
