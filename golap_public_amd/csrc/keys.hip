// Join-key pruning (gfx950): the keys a dimension map holds, sorted on the device, and a
// zone-map check of a foreign-key column against them.
//
// The reference collects the keys of a built HashMap with collect_hashmap_keys
// (core/dev_structs.cuh:136-150: one thread per head slot walks its chain and appends each key
// with a global atomic) and tests every chunk's [min, max] against every key with
// check_mmmeta_multiple_pts (core/metadata.cuh:550-559), O(keys) per chunk; 01.hpp:246-270
// then ANDs the flags into the plan's chunk mask. Here the keys are appended from the map's
// open-addressing slots with one atomic per wave, radix-sorted, and each chunk binary-searches
// the first key >= its minimum: O(log keys) per chunk, the same flags.
#include <hip/hip_runtime.h>
#include <hipcub/device/device_radix_sort.hpp>

#include <algorithm>

#include "common.hpp"
#include "device_codecs.hpp"

namespace golap_impl {
namespace {

using golap_dev::WAVE;
constexpr int BLOCK = 256;

// One thread per slot: occupied slots append their key (wave ballot, one atomic per wave);
// writes past cap are dropped (the count still tells the caller the size it needs).
__global__ __launch_bounds__(BLOCK) void k_collect_keys(const uint64_t *slots, uint64_t cap_slots, int32_t *keys,
                                                        uint64_t cap, unsigned long long *count) {
    const uint32_t lane = threadIdx.x % WAVE;
    for (uint64_t base = (uint64_t)blockIdx.x * BLOCK; base < cap_slots; base += (uint64_t)gridDim.x * BLOCK) {
        const uint64_t i = base + threadIdx.x;
        const uint64_t s = i < cap_slots ? slots[i] : 0;
        const bool occ = s != 0;
        const uint64_t ball = __ballot(occ);
        unsigned long long first = 0;
        if (lane == 0 && ball) first = atomicAdd(count, (unsigned long long)__popcll(ball));
        first = __shfl(first, 0, WAVE);
        if (occ) {
            const uint64_t at = first + (uint64_t)__popcll(lane ? (ball << (64 - lane)) : 0ull);
            if (at < cap) keys[at] = (int32_t)((uint32_t)s ^ 0x80000000u);
        }
    }
}

// check_mmmeta_multiple_pts: chunk c is kept iff some key k of the ascending array satisfies
// min[c] <= k <= max[c] (MinMaxMeta::check_pred with low = high = k, metadata.cuh:203-211).
__global__ __launch_bounds__(BLOCK) void k_check_points(const int64_t *mins, const int64_t *maxs, uint64_t n_chunks,
                                                        const int32_t *keys, uint64_t n_keys, uint16_t *flags) {
    for (uint64_t c = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; c < n_chunks; c += (uint64_t)gridDim.x * BLOCK) {
        const int64_t lo = mins[c], hi = maxs[c];
        uint64_t a = 0, b = n_keys;  // first key >= lo
        while (a < b) {
            const uint64_t m = (a + b) / 2;
            if ((int64_t)keys[m] < lo) a = m + 1;
            else b = m;
        }
        flags[c] = (a < n_keys && (int64_t)keys[a] <= hi) ? 1 : 0;
    }
}

}  // namespace

int collect_sorted_keys(golap_ctx *ctx, const golap_hashmap *m, int32_t *keys, uint64_t cap, uint64_t *n) {
    *n = 0;
    const bool small = m->small_log2 != 0;
    const uint64_t *slots = small ? m->d_small : m->d_slots;
    const uint64_t cap_slots = 1ull << (small ? m->small_log2 : m->log2cap);
    auto *count = reinterpret_cast<unsigned long long *>(ctx->d_scratch + 3080);
    GOLAP_HIP_CHECK(ctx, hipMemsetAsync(count, 0, 8, ctx->stream));
    // unsorted keys first, in a buffer of their own
    void *tmp_keys = nullptr;
    if (int r = dev_alloc(ctx, std::max<uint64_t>(cap, 1) * 4, &tmp_keys)) return r;
    const uint32_t grid = (uint32_t)std::min<uint64_t>((cap_slots + BLOCK - 1) / BLOCK, (uint64_t)ctx->cus * 8);
    hipLaunchKernelGGL(k_collect_keys, dim3(std::max(grid, 1u)), dim3(BLOCK), 0, ctx->stream, slots, cap_slots,
                       static_cast<int32_t *>(tmp_keys), cap, count);
    hipError_t e = hipGetLastError();
    unsigned long long got = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&got, count, 8, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        dev_release(ctx, tmp_keys);
        return fail(ctx, GOLAP_EHIP, std::string("hashmap_keys: ") + hipGetErrorString(e));
    }
    *n = got;
    if (got > cap) {
        dev_release(ctx, tmp_keys);
        return fail(ctx, GOLAP_ECAPACITY, "hashmap_keys: more keys than the output holds");
    }
    if (got) {
        size_t tbytes = 0;
        e = hipcub::DeviceRadixSort::SortKeys(nullptr, tbytes, static_cast<const int32_t *>(tmp_keys), keys, (int)got, 0,
                                              32, ctx->stream);
        void *tstore = nullptr;
        if (e == hipSuccess) {
            if (int r = dev_alloc(ctx, std::max<size_t>(tbytes, 16), &tstore)) {
                dev_release(ctx, tmp_keys);
                return r;
            }
            e = hipcub::DeviceRadixSort::SortKeys(tstore, tbytes, static_cast<const int32_t *>(tmp_keys), keys,
                                                  (int)got, 0, 32, ctx->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
            dev_release(ctx, tstore);
        }
        if (e != hipSuccess) {
            dev_release(ctx, tmp_keys);
            return fail(ctx, GOLAP_EHIP, std::string("hashmap_keys sort: ") + hipGetErrorString(e));
        }
    }
    dev_release(ctx, tmp_keys);
    return GOLAP_OK;
}

hipError_t launch_check_points(golap_ctx *ctx, const golap_column *col, const int32_t *keys, uint64_t n_keys,
                               uint16_t *flags) {
    if (!col->n_chunks) return hipSuccess;
    const uint32_t grid = (uint32_t)std::min<uint64_t>((col->n_chunks + BLOCK - 1) / BLOCK, 1024);
    hipLaunchKernelGGL(k_check_points, dim3(grid), dim3(BLOCK), 0, ctx->stream, col->d_chunk_min, col->d_chunk_max,
                       col->n_chunks, keys, n_keys, flags);
    return hipGetLastError();
}

}  // namespace golap_impl
