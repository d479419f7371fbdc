// C ABI: caller streams and events (the Op protocol's cudaStream_t / cudaEvent_t, core/core.hpp:94-124,
// 192-307), the asynchronous chunk decode on a caller stream, and the device-resident group-by
// accumulator (HashAggregate, core/dev_structs.cuh:157-253). See include/golap_hip.h.
#include <cstring>
#include <string>
#include <vector>

#include "common.hpp"
#include "star_host.hpp"

using namespace golap_impl;
using namespace golap_impl::starhost;
using Lock = std::lock_guard<std::recursive_mutex>;

namespace {
hipStream_t stream_of(golap_ctx *ctx, golap_stream *s) { return s ? s->stream : ctx->stream; }
uint32_t *err_of(golap_ctx *ctx, golap_stream *s) {
    return s ? s->d_err : reinterpret_cast<uint32_t *>(ctx->d_scratch);
}
}  // namespace

int golap_stream_create(golap_ctx *ctx, golap_stream **out) {
    if (!ctx || !out) return GOLAP_EINVAL;
    Lock l(ctx->mu);
    if (int r = set_device(ctx)) return r;
    auto s = new golap_stream;
    s->ctx = ctx;
    void *p = nullptr;
    if (int r = dev_alloc(ctx, 4, &p)) {
        delete s;
        return r;
    }
    s->d_err = static_cast<uint32_t *>(p);
    hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMemsetAsync(s->d_err, 0, 4, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) {
        if (s->stream) (void)hipStreamDestroy(s->stream);
        dev_release(ctx, s->d_err);
        delete s;
        return fail(ctx, GOLAP_EHIP, std::string("stream_create: ") + hipGetErrorString(e));
    }
    *out = s;
    return GOLAP_OK;
}

int golap_stream_destroy(golap_stream *s) {
    if (!s) return GOLAP_EINVAL;
    golap_ctx *ctx = s->ctx;
    {
        Lock l(ctx->mu);
        (void)hipSetDevice(ctx->device);
        (void)hipStreamSynchronize(s->stream);
        (void)hipStreamDestroy(s->stream);
        dev_release(ctx, s->d_err);
    }
    delete s;
    return GOLAP_OK;
}

int golap_stream_sync(golap_stream *s) {
    if (!s) return GOLAP_EINVAL;
    golap_ctx *ctx = s->ctx;
    if (int r = set_device(ctx)) return r;
    uint32_t flags = 0;
    GOLAP_HIP_CHECK(ctx, hipMemcpyAsync(&flags, s->d_err, 4, hipMemcpyDeviceToHost, s->stream));
    GOLAP_HIP_CHECK(ctx, hipStreamSynchronize(s->stream));
    if (!flags) return GOLAP_OK;
    GOLAP_HIP_CHECK(ctx, hipMemsetAsync(s->d_err, 0, 4, s->stream));
    GOLAP_HIP_CHECK(ctx, hipStreamSynchronize(s->stream));
    return status_of_flags(ctx, flags, "stream");
}

int golap_event_create(golap_ctx *ctx, golap_event **out) {
    if (!ctx || !out) return GOLAP_EINVAL;
    Lock l(ctx->mu);
    if (int r = set_device(ctx)) return r;
    auto e = new golap_event;
    e->ctx = ctx;
    hipError_t h = hipEventCreateWithFlags(&e->event, hipEventDisableTiming);
    if (h != hipSuccess) {
        delete e;
        return fail(ctx, GOLAP_EHIP, std::string("event_create: ") + hipGetErrorString(h));
    }
    *out = e;
    return GOLAP_OK;
}

int golap_event_destroy(golap_event *e) {
    if (!e) return GOLAP_EINVAL;
    (void)hipSetDevice(e->ctx->device);
    (void)hipEventSynchronize(e->event);
    (void)hipEventDestroy(e->event);
    delete e;
    return GOLAP_OK;
}

int golap_event_record(golap_event *e, golap_stream *s) {
    if (!e || (s && s->ctx != e->ctx)) return GOLAP_EINVAL;
    golap_ctx *ctx = e->ctx;
    if (int r = set_device(ctx)) return r;
    GOLAP_HIP_CHECK(ctx, hipEventRecord(e->event, stream_of(ctx, s)));
    return GOLAP_OK;
}

int golap_stream_wait_event(golap_ctx *ctx, golap_stream *s, golap_event *e) {
    if (!ctx || !e || e->ctx != ctx || (s && s->ctx != ctx)) return GOLAP_EINVAL;
    if (int r = set_device(ctx)) return r;
    GOLAP_HIP_CHECK(ctx, hipStreamWaitEvent(stream_of(ctx, s), e->event, 0));
    return GOLAP_OK;
}

int golap_event_sync(golap_event *e) {
    if (!e) return GOLAP_EINVAL;
    if (int r = set_device(e->ctx)) return r;
    GOLAP_HIP_CHECK(e->ctx, hipEventSynchronize(e->event));
    return GOLAP_OK;
}

int golap_stream_copy(golap_ctx *ctx, golap_stream *s, void *dst, const void *src, uint64_t bytes) {
    if (!ctx || (s && s->ctx != ctx) || (bytes && (!dst || !src))) return GOLAP_EINVAL;
    if (!bytes) return GOLAP_OK;
    if (int r = set_device(ctx)) return r;
    GOLAP_HIP_CHECK(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, stream_of(ctx, s)));
    return GOLAP_OK;
}

int golap_decompress_chunks_async(golap_ctx *ctx, const golap_column *c, uint64_t cb, uint64_t ce, void *dev_out,
                                  golap_stream *s) {
    if (!ctx || !c || !dev_out || cb > ce || ce > c->n_chunks || c->ctx != ctx || (s && s->ctx != ctx))
        return GOLAP_EINVAL;
    Lock l(ctx->mu);  // held while enqueueing only
    if (int r = set_device(ctx)) return r;
    if (cb == ce) return GOLAP_OK;
    const hipStream_t st = stream_of(ctx, s);
    const uint64_t ppc = c->chunk_rows / c->page_rows;
    const uint64_t row0 = cb * c->chunk_rows, row_end = ce * c->chunk_rows;
    if (row_end > c->n_rows) {  // zero-padded tail of the last chunk
        const uint64_t pad0 = (c->n_rows - row0) * c->elem;
        GOLAP_HIP_CHECK(ctx, hipMemsetAsync(static_cast<uint8_t *>(dev_out) + pad0, 0,
                                            (row_end - c->n_rows) * c->elem, st));
    }
    GOLAP_HIP_CHECK(ctx, launch_decode_pages(ctx, c, cb * ppc, std::min<uint64_t>(c->n_pages, ce * ppc),
                                             static_cast<uint8_t *>(dev_out), row0, st, err_of(ctx, s)));
    return GOLAP_OK;
}

// ------------------------------------------------------------------------------ HashAggregate
int golap_hashagg_create(golap_ctx *ctx, const golap_group_key *keys, uint32_t n_keys, uint32_t agg_mode,
                         golap_hashagg **out) {
    if (!ctx || !keys || !out || n_keys < 1 || n_keys > 4 || agg_mode > GOLAP_AGG_GLOBAL) return GOLAP_EINVAL;
    Lock l(ctx->mu);
    if (int r = set_device(ctx)) return r;
    auto h = new golap_hashagg;
    h->ctx = ctx;
    h->n_keys = n_keys;
    h->agg_mode = agg_mode;
    h->domain = 1;
    for (uint32_t k = 0; k < n_keys; ++k) {
        if (!is_int_type(keys[k].type) || !keys[k].card) {
            delete h;
            return fail(ctx, GOLAP_EINVAL, "hashagg: bad key");
        }
        h->lo.push_back(keys[k].lo);
        h->card.push_back(keys[k].card);
        h->type.push_back(keys[k].type);
        h->domain *= keys[k].card;
        if (h->domain >= MAX_GROUPS) {
            delete h;
            return fail(ctx, GOLAP_EINVAL, "hashagg: group domain too large");
        }
    }
    h->stride.resize(n_keys);
    for (int k = (int)n_keys - 1, s = 1; k >= 0; --k) {
        h->stride[k] = (uint64_t)s;
        s *= (int)h->card[k];
    }
    void *p = nullptr;
    if (int r = dev_alloc(ctx, h->domain * 16, &p)) {
        delete h;
        return r;
    }
    h->d_sum = static_cast<int64_t *>(p);
    h->d_cnt = reinterpret_cast<uint64_t *>(static_cast<uint8_t *>(p) + h->domain * 8);
    GOLAP_HIP_CHECK(ctx, hipMemsetAsync(p, 0, h->domain * 16, ctx->stream));
    GOLAP_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    *out = h;
    return GOLAP_OK;
}

int golap_hashagg_add(golap_hashagg *h, const void *const *dev_keys, const void *dev_val, uint32_t val_type,
                      uint64_t n_rows, golap_stream *s) {
    if (!h || (n_rows && (!dev_keys || !dev_val)) || !is_int_type(val_type) || (s && s->ctx != h->ctx))
        return GOLAP_EINVAL;
    golap_ctx *ctx = h->ctx;
    Lock l(ctx->mu);  // held while enqueueing only
    if (int r = set_device(ctx)) return r;
    if (!n_rows) return GOLAP_OK;
    GroupByLaunch g{};
    for (uint32_t k = 0; k < h->n_keys; ++k) {
        if (!dev_keys[k]) return fail(ctx, GOLAP_EINVAL, "hashagg: missing key column");
        g.key[k] = dev_keys[k];
        g.key_type[k] = h->type[k];
        g.key_lo[k] = h->lo[k];
        g.key_card[k] = h->card[k];
        g.key_stride[k] = h->stride[k];
    }
    g.n_keys = h->n_keys;
    g.val = dev_val;
    g.val_type = val_type;
    g.n = n_rows;
    g.lds_gslots = choose_lds_slots(h->agg_mode, 0, h->domain);
    g.g_sum = h->d_sum;
    g.g_cnt = h->d_cnt;
    g.err = err_of(ctx, s);
    g.stream = stream_of(ctx, s);
    GOLAP_HIP_CHECK(ctx, launch_group_by(ctx, g));
    return GOLAP_OK;
}

int golap_hashagg_results(golap_hashagg *h, uint64_t *n_groups, int32_t *group_keys, int64_t *sums,
                          uint64_t *counts, uint64_t max_groups) {
    if (!h || !n_groups) return GOLAP_EINVAL;
    golap_ctx *ctx = h->ctx;
    Lock l(ctx->mu);
    if (int r = set_device(ctx)) return r;
    if (int r = check_device_errors(ctx, "hashagg")) return r;
    return emit_groups(ctx, h->domain, h->d_sum, h->d_cnt, h->lo, h->card, h->stride, n_groups, nullptr, group_keys,
                       sums, counts, max_groups);
}

int golap_hashagg_clear(golap_hashagg *h) {
    if (!h) return GOLAP_EINVAL;
    golap_ctx *ctx = h->ctx;
    Lock l(ctx->mu);
    if (int r = set_device(ctx)) return r;
    GOLAP_HIP_CHECK(ctx, hipMemsetAsync(h->d_sum, 0, h->domain * 16, ctx->stream));
    GOLAP_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return GOLAP_OK;
}

int golap_hashagg_destroy(golap_hashagg *h) {
    if (!h) return GOLAP_EINVAL;
    golap_ctx *ctx = h->ctx;
    {
        Lock l(ctx->mu);
        (void)hipSetDevice(ctx->device);
        (void)hipStreamSynchronize(ctx->stream);
        dev_release(ctx, h->d_sum);
    }
    delete h;
    return GOLAP_OK;
}
