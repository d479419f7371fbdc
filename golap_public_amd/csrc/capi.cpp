// C ABI implementation (include/golap_hip.h). Host orchestration of the reference's plan
// functions (usecases/ssb/queries/01.hpp:53-449) with HIP streams and events.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <limits>
#include <memory>

#include "common.hpp"
#include "format.hpp"

using namespace golap_impl;
using Lock = std::lock_guard<std::recursive_mutex>;

namespace golap_impl {

bool page_direct_host(uint8_t codec) {
    return (codec & golap_fmt::PAGE_CODEC_MASK) == golap_fmt::RAW ||
           (codec & golap_fmt::PAGE_LITERAL_ONLY);
}

int dev_alloc(golap_ctx *ctx, uint64_t bytes, void **p) {
    *p = nullptr;
    if (!bytes) bytes = 16;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return fail(ctx, GOLAP_ENOMEM, std::string("hipMalloc failed: ") + hipGetErrorString(e));
    }
    std::lock_guard<std::mutex> g(ctx->alloc_mu);
    ctx->allocated += bytes;
    ctx->allocs[*p] = bytes;
    return GOLAP_OK;
}

void dev_release(golap_ctx *ctx, void *p) {
    if (!p) return;
    {
        std::lock_guard<std::mutex> g(ctx->alloc_mu);
        auto it = ctx->allocs.find(p);
        if (it != ctx->allocs.end()) {
            ctx->allocated -= it->second;
            ctx->allocs.erase(it);
        }
    }
    (void)hipFree(p);
}

int set_device(golap_ctx *ctx) {
    GOLAP_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    return GOLAP_OK;
}

// Reads and clears the device error word; maps it to a status.
int check_device_errors(golap_ctx *ctx, const char *what) {
    uint32_t flags = 0;
    GOLAP_HIP_CHECK(ctx, hipMemcpyAsync(&flags, ctx->d_scratch, 4, hipMemcpyDeviceToHost, ctx->stream));
    GOLAP_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return error_flags_status(ctx, flags, what);
}

int error_flags_status(golap_ctx *ctx, uint32_t flags, const char *what) {
    if (!flags) return GOLAP_OK;
    GOLAP_HIP_CHECK(ctx, hipMemsetAsync(ctx->d_scratch, 0, 4, ctx->stream));
    GOLAP_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    if (flags & DEVERR_CORRUPT) return fail(ctx, GOLAP_ECORRUPT, std::string(what) + ": corrupt page");
    if (flags & DEVERR_CAPACITY) return fail(ctx, GOLAP_ECAPACITY, std::string(what) + ": table full");
    return fail(ctx, GOLAP_EINVAL, std::string(what) + ": value outside its declared range");
}

bool valid_type(uint32_t t) { return t >= GOLAP_I8 && t <= GOLAP_I64; }

}  // namespace golap_impl

extern "C" {

int golap_api_version(void) { return GOLAP_API_VERSION; }

const char *golap_status_string(int s) {
    switch (s) {
    case GOLAP_OK: return "ok";
    case GOLAP_EINVAL: return "invalid argument";
    case GOLAP_ENOMEM: return "out of memory";
    case GOLAP_EHIP: return "HIP runtime error";
    case GOLAP_ECORRUPT: return "corrupt compressed page";
    case GOLAP_ECAPACITY: return "capacity exceeded";
    case GOLAP_EBUDGET: return "HBM budget exceeded";
    case GOLAP_ENODEV: return "no HIP device";
    default: return "unknown status";
    }
}

int golap_ctx_create(int device, golap_ctx **out) {
    if (!out) return GOLAP_EINVAL;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
        (void)hipGetLastError();
        return GOLAP_ENODEV;
    }
    if (device < 0 || device >= n) return GOLAP_EINVAL;
    std::unique_ptr<golap_ctx> c(new golap_ctx());
    c->device = device;
    if (hipSetDevice(device) != hipSuccess) return GOLAP_EHIP;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return GOLAP_EHIP;
    c->cus = prop.multiProcessorCount;
    int lds = 0;
    if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess && lds > 0)
        c->lds_max = lds;
    // the marketing name can be empty on some drivers: then the architecture name
    strncpy(c->name, prop.name[0] ? prop.name : prop.gcnArchName, sizeof(c->name) - 1);
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return GOLAP_EHIP;
    if (hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) return GOLAP_EHIP;
    void *s = nullptr;
    if (dev_alloc(c.get(), golap_ctx::SCRATCH_BYTES, &s) != GOLAP_OK) return GOLAP_ENOMEM;
    c->d_scratch = static_cast<uint8_t *>(s);
    if (hipMemset(c->d_scratch, 0, golap_ctx::SCRATCH_BYTES) != hipSuccess) return GOLAP_EHIP;
    fixed_scratch_init(c.get());
    *out = c.release();
    return GOLAP_OK;
}

int golap_ctx_destroy(golap_ctx *ctx) {
    if (!ctx) return GOLAP_EINVAL;
    {
        Lock l(ctx->mu);
        (void)hipSetDevice(ctx->device);
        (void)hipStreamSynchronize(ctx->stream);
        for (auto &kv : ctx->allocs) (void)hipFree(kv.first);
        ctx->allocs.clear();
        if (ctx->h_stage) (void)hipHostFree(ctx->h_stage);
        (void)hipEventDestroy(ctx->ev0);
        (void)hipEventDestroy(ctx->ev1);
        (void)hipStreamDestroy(ctx->stream);
    }
    delete ctx;
    return GOLAP_OK;
}

// The message is copied into a per-thread buffer, so it stays valid while other threads of the
// same context report errors.
const char *golap_last_error(golap_ctx *ctx) {
    if (!ctx) return "null context";
    static thread_local std::string copy;
    copy = last_error(ctx);
    return copy.c_str();
}

// Also reports device errors raised by work enqueued on the context's stream without a wait
// (golap_decompress_chunks_async / golap_hashagg_add with a null stream).
int golap_sync(golap_ctx *ctx) {
    if (!ctx) return GOLAP_EINVAL;
    Lock l(ctx->mu);
    if (int r = set_device(ctx)) return r;
    return check_device_errors(ctx, "sync");
}

int golap_device_info(golap_ctx *ctx, int *cus, char *name, int name_len) {
    if (!ctx) return GOLAP_EINVAL;
    if (cus) *cus = ctx->cus;
    if (name && name_len > 0) {
        strncpy(name, ctx->name, (size_t)name_len - 1);
        name[name_len - 1] = 0;
    }
    return GOLAP_OK;
}

// ------------------------------------------------------------------------------ memory
int golap_dev_alloc(golap_ctx *ctx, uint64_t bytes, void **dev_ptr) {
    if (!ctx || !dev_ptr) return GOLAP_EINVAL;
    Lock l(ctx->mu);
    if (int r = set_device(ctx)) return r;
    return dev_alloc(ctx, bytes, dev_ptr);
}
int golap_dev_free(golap_ctx *ctx, void *dev_ptr) {
    if (!ctx) return GOLAP_EINVAL;
    Lock l(ctx->mu);
    if (!dev_ptr) return GOLAP_OK;
    bool known;
    {
        std::lock_guard<std::mutex> g(ctx->alloc_mu);
        known = ctx->allocs.count(dev_ptr) != 0;
    }
    if (!known) return fail(ctx, GOLAP_EINVAL, "dev_free: unknown pointer");
    (void)hipStreamSynchronize(ctx->stream);
    dev_release(ctx, dev_ptr);
    return GOLAP_OK;
}
int golap_h2d(golap_ctx *ctx, void *dst, const void *src, uint64_t bytes) {
    if (!ctx || (bytes && (!dst || !src))) return GOLAP_EINVAL;
    Lock l(ctx->mu);
    if (!bytes) return GOLAP_OK;
    if (int r = set_device(ctx)) return r;
    GOLAP_HIP_CHECK(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    GOLAP_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return GOLAP_OK;
}
int golap_d2h(golap_ctx *ctx, void *dst, const void *src, uint64_t bytes) {
    if (!ctx || (bytes && (!dst || !src))) return GOLAP_EINVAL;
    Lock l(ctx->mu);
    if (!bytes) return GOLAP_OK;
    if (int r = set_device(ctx)) return r;
    GOLAP_HIP_CHECK(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    GOLAP_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return GOLAP_OK;
}
int golap_memset(golap_ctx *ctx, void *dst, int value, uint64_t bytes) {
    if (!ctx || (bytes && !dst)) return GOLAP_EINVAL;
    Lock l(ctx->mu);
    if (!bytes) return GOLAP_OK;
    if (int r = set_device(ctx)) return r;
    GOLAP_HIP_CHECK(ctx, hipMemsetAsync(dst, value, bytes, ctx->stream));
    return GOLAP_OK;
}
int golap_d2d(golap_ctx *ctx, void *dst, const void *src, uint64_t bytes) {
    if (!ctx || (bytes && (!dst || !src))) return GOLAP_EINVAL;
    Lock l(ctx->mu);
    if (!bytes) return GOLAP_OK;
    if (int r = set_device(ctx)) return r;
    GOLAP_HIP_CHECK(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    GOLAP_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return GOLAP_OK;
}
uint64_t golap_device_allocated(golap_ctx *ctx) {
    if (!ctx) return 0;
    std::lock_guard<std::mutex> g(ctx->alloc_mu);
    return ctx->allocated;
}

int golap_read_probe(golap_ctx *ctx, const void *dev_buf, uint64_t bytes, uint32_t blocks_per_cu, double *ms) {
    if (!ctx || !dev_buf || !ms || !bytes) return GOLAP_EINVAL;
    Lock l(ctx->mu);
    if (int r = set_device(ctx)) return r;
    GOLAP_HIP_CHECK(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    const hipError_t e = launch_read_probe(ctx, dev_buf, bytes, blocks_per_cu,
                                           reinterpret_cast<uint32_t *>(ctx->d_scratch + 3072), ctx->stream);
    if (e == hipErrorInvalidValue) return fail(ctx, GOLAP_EINVAL, "read_probe: bytes and buffer must be 16-byte aligned");
    GOLAP_HIP_CHECK(ctx, e);
    GOLAP_HIP_CHECK(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    GOLAP_HIP_CHECK(ctx, hipEventSynchronize(ctx->ev1));
    float t = 0;
    GOLAP_HIP_CHECK(ctx, hipEventElapsedTime(&t, ctx->ev0, ctx->ev1));
    *ms = t;
    return GOLAP_OK;
}

// ------------------------------------------------------------------------------ write path
void golap_encode_opts_default(golap_encode_opts *o) {
    if (!o) return;
    o->codec = GOLAP_CODEC_LZ4;
    o->page_rows = 4096;
    o->chunk_rows = 262144;
    o->lz4_min_match = 4;
    o->threads = 0;
    o->lz4_unit = 1;
}

int golap_encode(const void *host_values, uint64_t n_rows, uint32_t type,
                 const golap_encode_opts *opts, golap_host_column **out) {
    if (!out || !valid_type(type)) return GOLAP_EINVAL;
    *out = nullptr;
    golap_encode_opts o;
    golap_encode_opts_default(&o);
    if (opts) o = *opts;
    std::unique_ptr<golap_host_column> c(new golap_host_column());
    std::string err;
    int r = golap_host::encode_column(host_values, n_rows, type, o, *c, err);
    if (r != GOLAP_OK) return r;
    *out = c.release();
    return GOLAP_OK;
}

int golap_host_column_assemble(uint32_t type, uint64_t n_rows, uint32_t page_rows,
                               uint32_t chunk_rows, const uint8_t *const *pages,
                               const uint32_t *csize, const uint8_t *codec,
                               golap_host_column **out) {
    if (!out || !valid_type(type) || !page_rows || page_rows % 64 || !chunk_rows ||
        chunk_rows % page_rows)
        return GOLAP_EINVAL;
    *out = nullptr;
    std::unique_ptr<golap_host_column> c(new golap_host_column());
    c->type = type;
    c->elem = golap_host::type_size(type);
    c->page_rows = page_rows;
    c->chunk_rows = chunk_rows;
    c->n_rows = n_rows;
    c->n_pages = (n_rows + page_rows - 1) / page_rows;
    c->n_chunks = (n_rows + chunk_rows - 1) / chunk_rows;
    if (c->n_pages && (!pages || !csize || !codec)) return GOLAP_EINVAL;
    c->page_off.resize(c->n_pages);
    c->page_csize.assign(csize, csize + c->n_pages);
    c->page_codec.resize(c->n_pages);
    std::vector<const uint8_t *> ptrs(c->n_pages);
    for (uint64_t p = 0; p < c->n_pages; ++p) {
        uint8_t cd = codec[p] & golap_fmt::PAGE_CODEC_MASK;
        if (cd > golap_fmt::LZ4I) return GOLAP_EINVAL;
        if (csize[p] && !pages[p]) return GOLAP_EINVAL;
        uint64_t rows = std::min<uint64_t>(page_rows, n_rows - p * page_rows);
        if (cd == golap_fmt::LZ4 && golap_host::lz4_is_literal_only(pages[p], csize[p], rows * c->elem))
            cd |= golap_fmt::PAGE_LITERAL_ONLY;
        if (cd == golap_fmt::RLE || cd == golap_fmt::FOR || (cd == golap_fmt::DBP && csize[p] >= 32 && pages[p][20] == 0))
            cd |= golap_fmt::PAGE_POINT;
        c->page_codec[p] = cd;
        ptrs[p] = pages[p];
    }
    golap_host::layout_pages(*c, ptrs);
    // no statistics are known: every chunk may qualify
    c->chunk_min.assign(c->n_chunks, std::numeric_limits<int64_t>::min());
    c->chunk_max.assign(c->n_chunks, std::numeric_limits<int64_t>::max());
    *out = c.release();
    return GOLAP_OK;
}

int golap_host_column_free(golap_host_column *col) {
    delete col;
    return GOLAP_OK;
}

int golap_host_column_info(const golap_host_column *c, golap_column_desc *o) {
    if (!c || !o) return GOLAP_EINVAL;
    o->n_rows = c->n_rows;
    o->n_pages = c->n_pages;
    o->n_chunks = c->n_chunks;
    o->comp_bytes = c->comp.size();
    o->payload_bytes = c->payload_bytes;
    o->uncomp_bytes = c->n_rows * c->elem;
    o->type = c->type;
    o->elem = c->elem;
    o->page_rows = c->page_rows;
    o->chunk_rows = c->chunk_rows;
    o->comp = c->comp.data();
    o->page_off = c->page_off.data();
    o->page_csize = c->page_csize.data();
    o->page_codec = c->page_codec.data();
    o->chunk_min = c->chunk_min.data();
    o->chunk_max = c->chunk_max.data();
    return GOLAP_OK;
}

int golap_column_upload(golap_ctx *ctx, const golap_host_column *h, golap_column **out) {
    if (!ctx || !h || !out) return GOLAP_EINVAL;
    *out = nullptr;
    Lock l(ctx->mu);
    if (int r = set_device(ctx)) return r;
    return golap_impl::create_device_column(ctx, h, h->comp.size(), h->comp.data(), out);
}

}  // extern "C"

int golap_impl::create_device_column(golap_ctx *ctx, const golap_host_column *h, uint64_t comp_bytes,
                                     const uint8_t *comp, golap_column **out) {
    std::unique_ptr<golap_column> c(new golap_column());
    c->ctx = ctx;
    c->type = h->type;
    c->elem = h->elem;
    c->page_rows = h->page_rows;
    c->chunk_rows = h->chunk_rows;
    c->n_rows = h->n_rows;
    c->n_pages = h->n_pages;
    c->n_chunks = h->n_chunks;
    c->comp_bytes = comp_bytes;
    c->payload_bytes = h->payload_bytes;
    c->page_off = h->page_off;
    c->page_csize = h->page_csize;
    c->page_codec = h->page_codec;
    c->chunk_min = h->chunk_min;
    c->chunk_max = h->chunk_max;
    void *p[6];
    // the page data is followed by COMP_TAIL zero bytes: kernels read packed fields as dword
    // pairs, whose second dword may lie past the last page
    const uint64_t sz[6] = {comp_bytes + COMP_TAIL, h->n_pages * 8, h->n_pages * 4, h->n_pages,
                            h->n_chunks * 8, h->n_chunks * 8};
    const void *src[6] = {comp, h->page_off.data(), h->page_csize.data(),
                          h->page_codec.data(), h->chunk_min.data(), h->chunk_max.data()};
    for (int i = 0; i < 6; ++i) {
        if (int r = dev_alloc(ctx, sz[i], &p[i])) {
            for (int j = 0; j < i; ++j) dev_release(ctx, p[j]);
            return r;
        }
        const uint64_t n = i == 0 ? sz[i] - COMP_TAIL : sz[i];
        hipError_t e = i == 0 ? hipMemsetAsync(static_cast<uint8_t *>(p[i]) + n, 0, COMP_TAIL, ctx->stream) : hipSuccess;
        if (e == hipSuccess && n && src[i])  // no page data: the caller fills d_comp
            e = hipMemcpyAsync(p[i], src[i], n, hipMemcpyHostToDevice, ctx->stream);
        if (e != hipSuccess) {
            for (int j = 0; j <= i; ++j) dev_release(ctx, p[j]);
            return fail(ctx, GOLAP_EHIP, std::string("upload: ") + hipGetErrorString(e));
        }
    }
    GOLAP_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    c->d_comp = static_cast<uint8_t *>(p[0]);
    c->d_page_off = static_cast<uint64_t *>(p[1]);
    c->d_page_csize = static_cast<uint32_t *>(p[2]);
    c->d_page_codec = static_cast<uint8_t *>(p[3]);
    c->d_chunk_min = static_cast<int64_t *>(p[4]);
    c->d_chunk_max = static_cast<int64_t *>(p[5]);
    register_extra(c.get());
    if (comp) {
        c->extra.lz4i_max_nseq = lz4i_max_nseq(*h, comp, &c->extra.lz4i_coop_csize);
        c->extra.lz4i_coop_known = true;
    }
    *out = c.release();
    return GOLAP_OK;
}

// Most sequences of an LZ4I page of a host column whose page data is comp.
// coop_csize (optional): the largest of those pages with more sequences than the cooperative
// decode hands to the in-order one.
uint32_t golap_impl::lz4i_max_nseq(const golap_host_column &h, const uint8_t *comp, uint32_t *coop_csize) {
    uint32_t m = 0, cs = 0;
    for (uint64_t p = 0; p < h.n_pages; ++p)
        if ((h.page_codec[p] & golap_fmt::PAGE_CODEC_MASK) == golap_fmt::LZ4I && h.page_csize[p] >= 4) {
            const uint32_t n = (uint32_t)comp[h.page_off[p]] | ((uint32_t)comp[h.page_off[p] + 1] << 8);
            m = std::max(m, n);
            if (n > golap_fmt::LZ4I_SERIAL_MAX) cs = std::max(cs, h.page_csize[p]);
        }
    if (coop_csize) *coop_csize = cs;
    return m;
}

uint32_t golap_impl::lz4i_stage_bytes(const ColumnExtra &ex) {
    if (!ex.has_lz4i) return 0;
    return ((ex.lz4i_coop_known ? ex.lz4i_coop_csize : ex.lz4i_max_csize) + 15) & ~15u;
}

// The planning tables of a column from its host catalog (page kinds, payload per chunk).
void golap_impl::register_extra(golap_column *c) {
    ColumnExtra ex;
    ex.nondirect_prefix.resize(c->n_pages + 1, 0);
    ex.window_prefix.resize(c->n_pages + 1, 0);
    ex.dbp_prefix.resize(c->n_pages + 1, 0);
    for (uint64_t i = 0; i < c->n_pages; ++i) {
        const bool direct = page_direct_host(c->page_codec[i]);
        ex.nondirect_prefix[i + 1] = ex.nondirect_prefix[i] + (direct ? 0 : 1);
        const bool point = (c->page_codec[i] & golap_fmt::PAGE_POINT) != 0;
        ex.window_prefix[i + 1] = ex.window_prefix[i] + (direct || point ? 0 : 1);
        const bool dbp = (c->page_codec[i] & golap_fmt::PAGE_CODEC_MASK) == golap_fmt::DBP && !point;
        if ((c->page_codec[i] & golap_fmt::PAGE_CODEC_MASK) == golap_fmt::LZ4I) {
            ex.has_lz4i = true;
            ex.lz4i_max_csize = std::max(ex.lz4i_max_csize, c->page_csize[i]);
        }
        ex.dbp_prefix[i + 1] = ex.dbp_prefix[i] + (dbp ? 1 : 0);
    }
    ex.chunk_payload.resize(c->n_chunks + 1, 0);
    const uint64_t ppc = c->chunk_rows / c->page_rows;
    for (uint64_t ch = 0; ch < c->n_chunks; ++ch) {
        uint64_t s = 0;
        for (uint64_t i = ch * ppc; i < std::min<uint64_t>(c->n_pages, (ch + 1) * ppc); ++i) s += c->page_csize[i];
        ex.chunk_payload[ch + 1] = ex.chunk_payload[ch] + s;
    }
    c->extra = std::move(ex);
}

int golap_impl::set_column_window(golap_ctx *ctx, golap_column *c, const golap_host_column &w, uint64_t comp_bytes,
                                  hipStream_t stream) {
    c->type = w.type;
    c->elem = w.elem;
    c->page_rows = w.page_rows;
    c->chunk_rows = w.chunk_rows;
    c->n_rows = w.n_rows;
    c->n_pages = w.n_pages;
    c->n_chunks = w.n_chunks;
    c->comp_bytes = comp_bytes;
    c->payload_bytes = w.payload_bytes;
    c->page_off = w.page_off;
    c->page_csize = w.page_csize;
    c->page_codec = w.page_codec;
    c->chunk_min = w.chunk_min;
    c->chunk_max = w.chunk_max;
    // Copied from the column's own host catalog, which stays until the column's next window: no
    // wait here. The caller orders every reader after these copies (an event on the stream) and
    // sets the next window only once the scans of this one have returned.
    hipError_t e = hipMemcpyAsync(c->d_page_off, c->page_off.data(), w.n_pages * 8, hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) e = hipMemcpyAsync(c->d_page_csize, c->page_csize.data(), w.n_pages * 4, hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) e = hipMemcpyAsync(c->d_page_codec, c->page_codec.data(), w.n_pages, hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) e = hipMemcpyAsync(c->d_chunk_min, c->chunk_min.data(), w.n_chunks * 8, hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) e = hipMemcpyAsync(c->d_chunk_max, c->chunk_max.data(), w.n_chunks * 8, hipMemcpyHostToDevice, stream);
    if (e != hipSuccess) return fail(ctx, GOLAP_EHIP, std::string("column window: ") + hipGetErrorString(e));
    register_extra(c);
    return GOLAP_OK;
}

extern "C" {

int golap_column_free(golap_ctx *ctx, golap_column *c) {
    if (!ctx || !c) return GOLAP_EINVAL;
    Lock l(ctx->mu);
    (void)hipStreamSynchronize(ctx->stream);
    dev_release(ctx, c->d_comp);
    dev_release(ctx, c->d_page_off);
    dev_release(ctx, c->d_page_csize);
    dev_release(ctx, c->d_page_codec);
    dev_release(ctx, c->d_chunk_min);
    dev_release(ctx, c->d_chunk_max);
    delete c;
    return GOLAP_OK;
}

int golap_column_info(const golap_column *c, golap_column_desc *o) {
    if (!c || !o) return GOLAP_EINVAL;
    o->n_rows = c->n_rows;
    o->n_pages = c->n_pages;
    o->n_chunks = c->n_chunks;
    o->comp_bytes = c->comp_bytes;
    o->payload_bytes = c->payload_bytes;
    o->uncomp_bytes = c->n_rows * c->elem;
    o->type = c->type;
    o->elem = c->elem;
    o->page_rows = c->page_rows;
    o->chunk_rows = c->chunk_rows;
    o->comp = c->d_comp;
    o->page_off = c->d_page_off;
    o->page_csize = c->d_page_csize;
    o->page_codec = c->d_page_codec;
    o->chunk_min = c->d_chunk_min;
    o->chunk_max = c->d_chunk_max;
    return GOLAP_OK;
}

int golap_prepare_compressed_device(golap_ctx *ctx, const void *host_values, uint64_t n_rows,
                                    uint32_t type, const golap_encode_opts *opts,
                                    golap_column **out, double *comp_ms) {
    if (!ctx || !out) return GOLAP_EINVAL;
    auto t0 = std::chrono::steady_clock::now();
    golap_host_column *h = nullptr;
    golap_encode_opts o;
    golap_encode_opts_default(&o);
    if (opts) o = *opts;
    std::unique_ptr<golap_host_column> hc(new golap_host_column());
    std::string err;
    int r = golap_host::encode_column(host_values, n_rows, type, o, *hc, err);
    if (r != GOLAP_OK) return fail(ctx, r, err);
    h = hc.get();
    auto t1 = std::chrono::steady_clock::now();
    if (comp_ms) *comp_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    return golap_column_upload(ctx, h, out);
}

// ------------------------------------------------------------------------------ decompression
int golap_decompress_chunks(golap_ctx *ctx, const golap_column *c, uint64_t cb, uint64_t ce,
                            void *dev_out, double *kernel_ms) {
    if (!ctx || !c || !dev_out || cb > ce || ce > c->n_chunks) return GOLAP_EINVAL;
    Lock l(ctx->mu);
    if (int r = set_device(ctx)) return r;
    if (kernel_ms) *kernel_ms = 0;
    if (cb == ce) return GOLAP_OK;
    const uint64_t ppc = c->chunk_rows / c->page_rows;
    const uint64_t row0 = cb * c->chunk_rows, row_end = ce * c->chunk_rows;
    if (row_end > c->n_rows) {  // zero-padded tail of the last chunk
        uint64_t pad0 = (c->n_rows - row0) * c->elem;
        GOLAP_HIP_CHECK(ctx, hipMemsetAsync(static_cast<uint8_t *>(dev_out) + pad0, 0,
                                            (row_end - c->n_rows) * c->elem, ctx->stream));
    }
    GOLAP_HIP_CHECK(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    GOLAP_HIP_CHECK(ctx, launch_decode_pages(ctx, c, cb * ppc, std::min<uint64_t>(c->n_pages, ce * ppc),
                                             static_cast<uint8_t *>(dev_out), row0));
    GOLAP_HIP_CHECK(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    if (int r = check_device_errors(ctx, "decompress")) return r;
    if (kernel_ms) {
        float ms = 0;
        GOLAP_HIP_CHECK(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        *kernel_ms = ms;
    }
    return GOLAP_OK;
}

// ------------------------------------------------------------------------------ hash join
// Largest key range given a membership bitmap (2^18 bits = 32 KiB, pinned in LDS by the scans).
constexpr uint64_t HASHMAP_DENSE_MAX_KEYS = 1ull << 18;

static uint32_t log2_ceil(uint64_t v) {
    uint32_t b = 0;
    while ((1ull << b) < v) ++b;
    return b;
}

int golap_hashmap_create(golap_ctx *ctx, uint64_t ht_size, const void *dev_keys, uint32_t key_type,
                         golap_hashmap **out) {
    if (!ctx || !out || !dev_keys || !valid_type(key_type) || ht_size > (1ull << 31))
        return GOLAP_EINVAL;
    *out = nullptr;
    Lock l(ctx->mu);
    if (int r = set_device(ctx)) return r;
    std::unique_ptr<golap_hashmap> m(new golap_hashmap());
    m->ctx = ctx;
    m->ht_size = ht_size;
    m->d_keys = dev_keys;
    m->key_type = key_type;
    m->log2cap = std::max<uint32_t>(6, log2_ceil(std::max<uint64_t>(1, ht_size) * 2));  // load <= 1/2
    void *p = nullptr;
    if (int r = dev_alloc(ctx, 8ull << m->log2cap, &p)) return r;
    m->d_slots = static_cast<uint64_t *>(p);
    GOLAP_HIP_CHECK(ctx, launch_hash_clear(ctx, m->d_slots, 1ull << m->log2cap));
    *out = m.release();
    return GOLAP_OK;
}

int golap_hashmap_build(golap_ctx *ctx, golap_hashmap *m, uint64_t n_rows, const golap_pred *preds,
                        uint32_t n_preds, const int32_t *dev_payload) {
    if (!ctx || !m || m->ctx != ctx || n_preds > MAX_PREDS || (n_preds && !preds)) return GOLAP_EINVAL;
    Lock l(ctx->mu);
    if (int r = set_device(ctx)) return r;
    PredView pv[MAX_PREDS];
    for (uint32_t i = 0; i < n_preds; ++i) {
        if (!preds[i].dev_col || !valid_type(preds[i].type)) return fail(ctx, GOLAP_EINVAL, "build: bad predicate");
        pv[i] = PredView{preds[i].dev_col, preds[i].type, preds[i].lo, preds[i].hi};
    }
    // a rebuild starts from an empty table
    GOLAP_HIP_CHECK(ctx, launch_hash_clear(ctx, m->d_slots, 1ull << m->log2cap));
    // scratch [64,72): inserted count, [72,80): min and max inserted key
    uint64_t *d_filled = reinterpret_cast<uint64_t *>(ctx->d_scratch + 64);
    int32_t *d_kminmax = reinterpret_cast<int32_t *>(ctx->d_scratch + 72);
    static const int32_t init_minmax[2] = {INT32_MAX, INT32_MIN};
    GOLAP_HIP_CHECK(ctx, hipMemsetAsync(d_filled, 0, 8, ctx->stream));
    GOLAP_HIP_CHECK(ctx, hipMemcpyAsync(d_kminmax, init_minmax, 8, hipMemcpyHostToDevice, ctx->stream));
    GOLAP_HIP_CHECK(ctx, launch_hash_build(ctx, m, n_rows, pv, n_preds, dev_payload, d_filled, d_kminmax));
    uint64_t back[2] = {0, 0};
    GOLAP_HIP_CHECK(ctx, hipMemcpyAsync(back, d_filled, 16, hipMemcpyDeviceToHost, ctx->stream));
    if (int r = check_device_errors(ctx, "hashmap_build")) return r;
    m->filled = back[0];
    memcpy(&m->kmin, &back[1], 4);
    memcpy(&m->kmax, reinterpret_cast<uint8_t *>(&back[1]) + 4, 4);
    if (m->filled > m->ht_size)
        return fail(ctx, GOLAP_ECAPACITY, "hashmap_build: more qualifying rows than ht_size");
    // dense form: membership bits over the inserted key range when it is small. Buffers are
    // kept across rebuilds and grown only when a larger one is needed.
    m->krange = 0;
    m->use_bitmap = false;
    if (m->filled && (int64_t)m->kmax - m->kmin + 1 <= (int64_t)HASHMAP_DENSE_MAX_KEYS) {
        m->krange = (uint32_t)((int64_t)m->kmax - m->kmin + 1);
        const uint64_t words = (m->krange + 31) / 32;
        if (words > m->bitmap_words) {
            dev_release(ctx, m->d_bitmap);
            m->d_bitmap = nullptr;
            m->bitmap_words = 0;
            void *p = nullptr;
            if (int r = dev_alloc(ctx, words * 4, &p)) return r;
            m->d_bitmap = static_cast<uint32_t *>(p);
            m->bitmap_words = words;
        }
        GOLAP_HIP_CHECK(ctx, hipMemsetAsync(m->d_bitmap, 0, words * 4, ctx->stream));
        // no sync: the range is exact by construction; a raised flag surfaces at the next check
        GOLAP_HIP_CHECK(ctx, launch_hash_bitmap(ctx, m));
        m->use_bitmap = true;
    }
    // compact copy (load <= 1/2): the table the scans probe
    m->small_log2 = 0;
    const uint32_t sl = std::max<uint32_t>(6, log2_ceil(std::max<uint64_t>(1, m->filled) * 2));
    if (sl < m->log2cap) {
        if (sl > m->small_cap_log2) {
            dev_release(ctx, m->d_small);
            m->d_small = nullptr;
            m->small_cap_log2 = 0;
            void *p = nullptr;
            if (int r = dev_alloc(ctx, 8ull << sl, &p)) return r;
            m->d_small = static_cast<uint64_t *>(p);
            m->small_cap_log2 = sl;
        }
        m->small_log2 = sl;
        GOLAP_HIP_CHECK(ctx, launch_hash_clear(ctx, m->d_small, 1ull << sl));
        // no sync: sized for twice the inserted keys, it cannot fill up
        GOLAP_HIP_CHECK(ctx, launch_hash_compact(ctx, m));
    }
    m->build_id++;
    m->built = true;
    return GOLAP_OK;
}

int golap_hashmap_probe(golap_ctx *ctx, const golap_hashmap *m, const void *keys, uint32_t key_type,
                        uint64_t n, uint64_t *dev_out_rid) {
    if (!ctx || !m || m->ctx != ctx || (n && (!keys || !dev_out_rid)) || !valid_type(key_type))
        return GOLAP_EINVAL;
    Lock l(ctx->mu);
    if (int r = set_device(ctx)) return r;
    GOLAP_HIP_CHECK(ctx, launch_hash_probe(ctx, m, keys, key_type, n, dev_out_rid, nullptr));
    GOLAP_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return GOLAP_OK;
}

int golap_hash_join_count(golap_ctx *ctx, const golap_hashmap *m, const void *keys, uint32_t key_type,
                          uint64_t n, uint64_t *hits) {
    if (!ctx || !m || m->ctx != ctx || !hits || (n && !keys) || !valid_type(key_type)) return GOLAP_EINVAL;
    Lock l(ctx->mu);
    if (int r = set_device(ctx)) return r;
    uint64_t *d_hits = reinterpret_cast<uint64_t *>(ctx->d_scratch + 64);
    GOLAP_HIP_CHECK(ctx, hipMemsetAsync(d_hits, 0, 8, ctx->stream));
    GOLAP_HIP_CHECK(ctx, launch_hash_probe(ctx, m, keys, key_type, n, nullptr, d_hits));
    GOLAP_HIP_CHECK(ctx, hipMemcpyAsync(hits, d_hits, 8, hipMemcpyDeviceToHost, ctx->stream));
    GOLAP_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return GOLAP_OK;
}

int golap_hashmap_filled(golap_ctx *ctx, const golap_hashmap *m, uint64_t *filled) {
    if (!ctx || !m || !filled) return GOLAP_EINVAL;
    *filled = m->filled;
    return GOLAP_OK;
}

int golap_hashmap_free(golap_ctx *ctx, golap_hashmap *m) {
    if (!ctx || !m || m->ctx != ctx) return GOLAP_EINVAL;
    Lock l(ctx->mu);
    (void)hipStreamSynchronize(ctx->stream);
    dev_release(ctx, m->d_slots);
    dev_release(ctx, m->d_small);
    dev_release(ctx, m->d_bitmap);
    dev_release(ctx, m->d_hfilter);
    delete m;
    return GOLAP_OK;
}

// ------------------------------------------------------------------------------ zone maps
int golap_minmax_check(golap_ctx *ctx, const golap_column *c, int64_t lo, int64_t hi, uint16_t *flags) {
    if (!ctx || !c || !flags) return GOLAP_EINVAL;
    Lock l(ctx->mu);
    if (int r = set_device(ctx)) return r;
    GOLAP_HIP_CHECK(ctx, launch_minmax_check(ctx, c, lo, hi, flags));
    return GOLAP_OK;
}

int golap_hashmap_keys(golap_ctx *ctx, const golap_hashmap *m, int32_t *dev_keys, uint64_t cap, uint64_t *n) {
    if (!ctx || !m || !n || m->ctx != ctx || (cap && !dev_keys)) return GOLAP_EINVAL;
    Lock l(ctx->mu);
    if (int r = set_device(ctx)) return r;
    if (!m->built) return fail(ctx, GOLAP_EINVAL, "hashmap_keys: the map is not built");
    return collect_sorted_keys(ctx, m, dev_keys, cap, n);
}

int golap_minmax_check_points(golap_ctx *ctx, const golap_column *c, const int32_t *dev_keys, uint64_t n,
                              uint16_t *flags) {
    if (!ctx || !c || !flags || (n && !dev_keys)) return GOLAP_EINVAL;
    Lock l(ctx->mu);
    if (int r = set_device(ctx)) return r;
    GOLAP_HIP_CHECK(ctx, launch_check_points(ctx, c, dev_keys, n, flags));
    return GOLAP_OK;
}

int golap_combine_and(golap_ctx *ctx, uint16_t *a, const uint16_t *b, uint64_t n) {
    if (!ctx || (n && (!a || !b))) return GOLAP_EINVAL;
    Lock l(ctx->mu);
    if (int r = set_device(ctx)) return r;
    GOLAP_HIP_CHECK(ctx, launch_combine_and(ctx, a, b, n));
    return GOLAP_OK;
}

// ------------------------------------------------------------------------------ Q1 plan
// Inclusive range [lo, hi] intersected with the value range of a column type, as lo + span in
// unsigned arithmetic; false if nothing can pass.
static bool range_span(int64_t lo, int64_t hi, int64_t tmin, int64_t tmax, uint32_t &ulo, uint32_t &span) {
    lo = std::max(lo, tmin);
    hi = std::min(hi, tmax);
    if (lo > hi) return false;
    ulo = (uint32_t)lo;
    span = (uint32_t)(hi - lo);
    return true;
}

int golap_q1_run(golap_ctx *ctx, const golap_q1_plan *plan, golap_q1_result *res) {
    if (!ctx || !plan || !res) return GOLAP_EINVAL;
    const golap_column *cols[4] = {plan->date, plan->disc, plan->qty, plan->price};
    for (auto *c : cols)
        if (!c || c->ctx != ctx) return fail(ctx, GOLAP_EINVAL, "q1: missing column");
    const golap_hashmap *m = plan->date_map;
    if (!m || m->ctx != ctx || !m->built) return fail(ctx, GOLAP_EINVAL, "q1: date map not built");
    const golap_column *d = plan->date;
    for (auto *c : cols)
        if (c->n_rows != d->n_rows || c->page_rows != d->page_rows || c->chunk_rows != d->chunk_rows)
            return fail(ctx, GOLAP_EINVAL, "q1: columns are not row-aligned");
    auto is32 = [](const golap_column *c) { return c->type == GOLAP_I32 || c->type == GOLAP_U32; };
    if (!is32(plan->date) || !is32(plan->price) || plan->disc->type != GOLAP_U8 || plan->qty->type != GOLAP_U8)
        return fail(ctx, GOLAP_EINVAL, "q1: expected date/price 32-bit, discount/quantity uint8");
    for (const golap_column *c : {plan->date, plan->price})
        if (c->type == GOLAP_U32)
            for (int64_t mx : c->chunk_max)
                if (mx > INT32_MAX) return fail(ctx, GOLAP_EINVAL, "q1: uint32 values above INT32_MAX");
    if (plan->chunk_begin > plan->chunk_end || plan->chunk_end > d->n_chunks)
        return fail(ctx, GOLAP_EINVAL, "q1: bad chunk range");
    Lock l(ctx->mu);
    if (int r = set_device(ctx)) return r;
    memset(res, 0, sizeof(*res));

    const uint64_t ppc = d->chunk_rows / d->page_rows;
    Q1Launch a{};
    a.date = view(plan->date);
    a.disc = view(plan->disc);
    a.qty = view(plan->qty);
    a.price = view(plan->price);
    a.page_rows = d->page_rows;
    a.n_rows = d->n_rows;
    a.page_begin = plan->chunk_begin * ppc;
    a.page_end = std::min<uint64_t>(d->n_pages, plan->chunk_end * ppc);
    a.chunk_mask = plan->dev_chunk_mask;
    a.pages_per_chunk = (uint32_t)ppc;
    const int64_t pmin = plan->price->type == GOLAP_U32 ? 0 : INT32_MIN;
    bool any = range_span(plan->disc_lo, plan->disc_hi, 0, 255, a.disc_lo, a.disc_span);
    any = range_span(plan->qty_lo, plan->qty_hi, 0, 255, a.qty_lo, a.qty_span) && any;
    any = range_span(plan->price_lo, plan->price_hi, pmin, INT32_MAX, a.price_lo, a.price_span) && any;
    any = any && m->filled > 0;
    if (m->use_bitmap && (m->krange + 31) / 32 * 4 <= 32768) {
        a.probe = 2;
        a.bitmap = m->d_bitmap;
        a.kmin = m->kmin;
        a.krange = m->krange;
    } else if (m->small_log2 && m->small_log2 <= 12) {
        a.probe = 1;
        a.slots = m->d_small;
        a.log2cap = m->small_log2;
    } else {
        a.probe = 0;
        a.slots = m->d_slots;
        a.log2cap = m->log2cap;
    }
    a.err = reinterpret_cast<uint32_t *>(ctx->d_scratch);
    a.sum = reinterpret_cast<int64_t *>(ctx->d_scratch + 8);
    a.count = reinterpret_cast<uint64_t *>(ctx->d_scratch + 16);
    GOLAP_HIP_CHECK(ctx, hipMemsetAsync(ctx->d_scratch + 8, 0, 16, ctx->stream));

    bool any_general = false, any_direct = true;
    uint32_t max_nseq = 0;
    bool nseq_known = true;
    for (auto *c : cols) {
        auto &ex = c->extra;
        const uint64_t nd = ex.nondirect_prefix[a.page_end] - ex.nondirect_prefix[a.page_begin];
        if (nd) any_general = true;
        if (nd == a.page_end - a.page_begin) any_direct = false;  // no tile has all four in place
        max_nseq = std::max(max_nseq, ex.lz4i_max_nseq);
        nseq_known = nseq_known && (ex.lz4i_max_nseq || !ex.has_lz4i);
        a.stage_cap = std::max(a.stage_cap, lz4i_stage_bytes(ex));
    }
    // one window serves every column of a tile: sized for the widest (4-byte) column
    lz4i_window(d->page_rows * 4, nseq_known ? std::max(max_nseq, 1u) : 0, a.win_cap, a.tok_cap);
    a.max_nseq = nseq_known ? max_nseq : 0;
    GOLAP_HIP_CHECK(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    if (any && a.page_end > a.page_begin) {  // an empty predicate range or map matches nothing
        if (any_direct) GOLAP_HIP_CHECK(ctx, launch_q1(ctx, a, true));
        if (any_general) {
            // LZ4I pages: one workgroup per tile, decoding them cooperatively; other coded pages
            // (and page shapes that kernel does not take): one wave per tile
            hipError_t e = a.stage_cap ? launch_q1_wg(ctx, a) : hipErrorNotSupported;
            if (e == hipErrorNotSupported) e = launch_q1(ctx, a, false);
            GOLAP_HIP_CHECK(ctx, e);
        }
    }
    GOLAP_HIP_CHECK(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    uint64_t host[3] = {0, 0, 0};
    GOLAP_HIP_CHECK(ctx, hipMemcpyAsync(host, ctx->d_scratch, 24, hipMemcpyDeviceToHost, ctx->stream));
    std::vector<uint16_t> mask;
    if (plan->dev_chunk_mask && plan->chunk_end > plan->chunk_begin) {
        mask.resize(plan->chunk_end);
        GOLAP_HIP_CHECK(ctx, hipMemcpyAsync(mask.data() + plan->chunk_begin,
                                            plan->dev_chunk_mask + plan->chunk_begin,
                                            (plan->chunk_end - plan->chunk_begin) * 2,
                                            hipMemcpyDeviceToHost, ctx->stream));
    }
    GOLAP_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    float ms = 0;
    GOLAP_HIP_CHECK(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    if (host[0]) return error_flags_status(ctx, (uint32_t)host[0], "q1");
    res->sum = (int64_t)host[1];
    res->matches = host[2];
    res->kernel_ms = ms;
    {
        for (uint64_t ch = plan->chunk_begin; ch < plan->chunk_end; ++ch) {
            if (!mask.empty() && !mask[ch]) {
                res->chunks_pruned++;
                continue;
            }
            res->chunks_run++;
            for (auto *c : cols) {
                auto &ex = c->extra;
                res->comp_bytes += ex.chunk_payload[ch + 1] - ex.chunk_payload[ch];
            }
        }
    }
    return GOLAP_OK;
}

}  // extern "C"
