// C ABI: the allocation-free scan entry point (golap_scan_run_noalloc, include/golap_hip.h).
//
// golap_scan_run takes the device memory a call needs beyond its inputs and outputs from the
// context: the error word and work-list counters in the context scratch, the work items and the
// dense group result in buffers grown on demand, and a pinned staging page for the plan
// descriptor. Callers that cannot allocate (graph capture, tight serving loops, harnesses that
// forbid hipMalloc) get the same scan over one file-scope device array instead. Each device that
// loads the library holds its own copy of the array until the process ends; its address on a
// device is looked up once, when a context is created on it, and only that address is handed to
// kernels and copies (the array's name is a host-side stand-in, not a device address).
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "common.hpp"

namespace {

static __device__ __align__(256) uint8_t scratch[GOLAP_NOALLOC_SCRATCH_BYTES];

// Calls on one device share its array: one lock per device index serialises them (each call
// ends with its stream synchronised, so the array is idle when the lock is released).
constexpr int MAX_DEVICES = 64;
std::mutex g_dev_mu[MAX_DEVICES];

}  // namespace

void golap_impl::fixed_scratch_init(golap_ctx *ctx) {
    ctx->d_fixed = nullptr;
    ctx->fixed_bytes = 0;
    if (ctx->device < 0 || ctx->device >= MAX_DEVICES) {
        ctx->fixed_err = hipErrorInvalidDevice;
        return;
    }
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    ctx->fixed_err = e;
    if (e != hipSuccess || !p) {
        (void)hipGetLastError();
        if (e == hipSuccess) ctx->fixed_err = hipErrorInvalidSymbol;
        return;
    }
    ctx->d_fixed = static_cast<uint8_t *>(p);
    ctx->fixed_bytes = GOLAP_NOALLOC_SCRATCH_BYTES;
}

extern "C" int golap_scan_run_noalloc(golap_ctx *ctx, const golap_scan_plan *plan, golap_star_result *res,
                                      int32_t *group_keys, int64_t *sums, uint64_t *counts, uint64_t max_groups) {
    if (!ctx || !plan || !res) return GOLAP_EINVAL;
    if (!ctx->d_fixed)
        return golap_impl::fail(ctx, GOLAP_EHIP, std::string("scan_run_noalloc: device array unavailable: ") +
                                                     hipGetErrorString(ctx->fixed_err));
    std::lock_guard<std::mutex> l(g_dev_mu[ctx->device]);
    const golap_impl::ScanFixedScratch f{ctx->d_fixed, ctx->fixed_bytes};
    return golap_impl::scan_run_impl(ctx, plan, res, group_keys, sums, counts, max_groups, &f);
}
