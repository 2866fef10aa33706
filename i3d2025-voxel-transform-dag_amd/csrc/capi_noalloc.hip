// Allocation-free form of the device-resident matcher (tdag_match_level_device_noalloc).
//
// The matcher's device scratch (sort keys, rank-ordered masks and transposed layouts, best[],
// child table, Bloom filter, post-processing arrays, counters) lives in one file-scope device
// array of TDAG_NOALLOC_SCRATCH_BYTES per GPU. Its device address is looked up once per device,
// when the first context on that device is created, and every call uses that address. The call
// itself allocates nothing: no hipMalloc / hipFree, no host heap, no library routine that
// allocates. Every kernel it launches clears or seeds the parts of the array it reads (the array
// keeps the previous call's contents), so consecutive calls of any size are independent; calls
// that overlap on one GPU share the array and race.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "context.hpp"
#include "matcher.hpp"
#include "tdag.h"

namespace {

static __device__ __align__(256) unsigned char scratch[TDAG_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_scratchPtr[kMaxDevices] = {}; // device address of `scratch` on each device (null: not looked up)
hipError_t g_lookupError[kMaxDevices] = {};

// matched (u64) and unique (low 32 bits of the second word) out of the matcher's counters.
__global__ void kCopyCounts(const unsigned long long* counters, uint64_t* out)
{
    if (threadIdx.x == 0) {
        out[0] = counters[0];
        out[1] = counters[1] & 0xFFFFFFFFull;
    }
}

} // namespace

namespace tdag {

// Looks up the device address of `scratch` on `device`, which must be the current device.
// Called once per device at context creation; later calls return the cached result.
hipError_t noallocScratchInit(int device)
{
    if (device < 0 || device >= kMaxDevices)
        return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_mu);
    if (g_scratchPtr[device] || g_lookupError[device] != hipSuccess)
        return g_lookupError[device];
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
        g_lookupError[device] = e != hipSuccess ? e : hipErrorInvalidSymbol;
        return g_lookupError[device];
    }
    g_scratchPtr[device] = ptr;
    return hipSuccess;
}

} // namespace tdag

using namespace tdag::capi;

extern "C" tdag_status tdag_match_level_device_noalloc(tdag_context* ctx, uint32_t level, const uint64_t* d_masks, uint64_t n, const tdag_config* cfg,
    uint64_t* d_encoded, uint64_t* d_mapping, uint32_t* d_unique, uint64_t* d_counts, void* stream)
{
    if (tdag_status s = checkArgs(ctx, level, n, cfg); s != TDAG_OK)
        return s;
    if (n > 0 && !d_masks)
        return fail(TDAG_EINVAL, "d_masks is null");
    const int32_t v = ctx->variant[level];
    const tdag::MatchPlan plan = tdag::planMatch(level, n, v >= 0 ? (uint32_t)v : tdag::defaultVariant(level), ctx->firstBatch[level], ctx->batchGrowth[level],
        ctx->bloomBits[level], ctx->childFrames[level]);
    if (plan.scratchBytes > (size_t)TDAG_NOALLOC_SCRATCH_BYTES)
        return fail(TDAG_EINVAL, "level " + std::to_string(level) + " with " + std::to_string(n) + " items needs " + std::to_string(plan.scratchBytes) +
                " bytes of scratch; the allocation-free matcher has " + std::to_string((uint64_t)TDAG_NOALLOC_SCRATCH_BYTES));
    void* ptr = nullptr;
    hipError_t lookup = hipErrorInvalidDevice;
    {
        std::lock_guard<std::mutex> lock(g_mu);
        if (ctx->device >= 0 && ctx->device < kMaxDevices) {
            ptr = g_scratchPtr[ctx->device];
            lookup = g_lookupError[ctx->device];
        }
    }
    if (!ptr)
        return fail(TDAG_EDEVICE, "the allocation-free scratch of device " + std::to_string(ctx->device) + " could not be set up at context creation (" +
                hipGetErrorName(lookup) + ")");
    int current = -1;
    CAPI_HIP(hipGetDevice(&current), "hipGetDevice");
    if (current != ctx->device)
        CAPI_HIP(hipSetDevice(ctx->device), "hipSetDevice");
    hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
    const uint32_t cfgBits = cfgBitsOf(cfg) | (ctx->counters ? tdag::kCfgCount : 0u);
    CAPI_HIP(tdag::enqueueMatch(plan, d_masks, cfgBits, ptr, st, d_encoded, d_mapping, d_unique, nullptr, nullptr), "matcher launch");
    if (d_counts) {
        hipLaunchKernelGGL(kCopyCounts, dim3(1), dim3(64), 0, st, (const unsigned long long*)((char*)ptr + plan.offCounters), d_counts);
        CAPI_HIP(hipGetLastError(), "counts copy");
    }
    return TDAG_OK;
}
