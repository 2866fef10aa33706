// C-ABI of the matching path (include/tdag.h): contexts, error reporting, matcher calls.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <algorithm>
#include <cstddef>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "batched.hpp"
#include "context.hpp"
#include "errors.hpp"
#include "learned.hpp"
#include "matcher.hpp"
#include "tdag.h"


namespace {
thread_local std::string g_lastError;
} // namespace

namespace tdag {
namespace capi {

tdag_status fail(tdag_status s, const std::string& msg)
{
    g_lastError = msg;
    return s;
}

tdag_status failHip(hipError_t e, const char* what)
{
    g_lastError = std::string(what) + ": " + hipGetErrorName(e) + " (" + hipGetErrorString(e) + ")";
    return e == hipErrorOutOfMemory ? TDAG_ENOMEM : TDAG_EDEVICE;
}

tdag_status ensureBuffer(void*& buf, size_t& have, size_t need, const char* what)
{
    if (have >= need && buf)
        return TDAG_OK;
    if (buf) {
        (void)hipFree(buf);
        buf = nullptr;
        have = 0;
    }
    const size_t bytes = std::max<size_t>(need, 4096);
    CAPI_HIP(hipMalloc(&buf, bytes), what);
    have = bytes;
    return TDAG_OK;
}

uint32_t cfgBitsOf(const tdag_config* cfg)
{
    return (cfg->symmetry ? tdag::kCfgSymmetry : 0u) | (cfg->axis_permutation ? tdag::kCfgAxis : 0u) | (cfg->translation ? tdag::kCfgTranslation : 0u);
}

tdag_status checkArgs(tdag_context* ctx, uint32_t level, uint64_t n, const tdag_config* cfg)
{
    if (!ctx)
        return fail(TDAG_EINVAL, "context is null");
    if (!cfg)
        return fail(TDAG_EINVAL, "config is null");
    if (level < 2 || level > 5)
        return fail(TDAG_EINVAL, "level " + std::to_string(level) + " is outside the supported range [2, 5]");
    // Level 2 stores ranks in 29 bits of a table value, levels 3..5 in 26 bits of a table entry.
    const uint32_t rankBits = level == 2 ? 28 : 26;
    if (n >= (1ull << rankBits))
        return fail(TDAG_EINVAL, "item count " + std::to_string(n) + " exceeds the supported maximum 2^" + std::to_string(rankBits) + " - 1 at level " +
            std::to_string(level));
    return TDAG_OK;
}

} // namespace capi
} // namespace tdag

using namespace tdag::capi;

// A batched matcher call: per-level plans, its own device scratch, the uploaded BatchDesc and,
// when requested, the captured hipGraph of the whole call.
struct tdag_batch {
    tdag_context* ctx = nullptr;
    tdag::BatchPlan plan {};
    std::vector<tdag_level_job> jobs;
    void* scratch = nullptr;
    uint64_t* pinned = nullptr; // counters of every level, then the barrier error word
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint64_t spinTicks = 0;
    uint32_t cfgBits = 0;
    uint32_t grab = 1;
    uint32_t splitEvery = tdag::kBatchSplitEvery;
    // Concurrent mode: every level's own tuned matcher (the per-level plan and kernels) on its own
    // stream, forked from and joined back into the context's stream (graph branches when captured).
    bool concurrent = false;
    tdag::MatchPlan cplans[tdag::kBatchMaxLevels] {};
    size_t cOff[tdag::kBatchMaxLevels] {};
    hipStream_t aux[tdag::kBatchMaxLevels] {};
    hipEvent_t evFork = nullptr;
    hipEvent_t evJoin[tdag::kBatchMaxLevels] {};
    // Learned level-2 job (tdag_batch_create_learned): the proposal matcher on its own branch.
    tdag_mlp* mlp = nullptr;
    uint32_t k = 0;
    int learned = -1; // job index, -1: none
    tdag::MatchPlan lplan {};
    size_t lOff = 0, lWorkOff = 0; // its scratch and proposal workspace inside `scratch`
};

namespace tdag {
tdag_status setError(tdag_status status, const std::string& message)
{
    g_lastError = message;
    return status;
}
} // namespace tdag

extern "C" {

size_t tdag_last_error(char* buf, size_t cap)
{
    if (buf && cap > 0) {
        const size_t k = std::min(cap - 1, g_lastError.size());
        std::memcpy(buf, g_lastError.data(), k);
        buf[k] = '\0';
    }
    return g_lastError.size();
}

const char* tdag_version(void) { return "tdag-mi355x 0.1 gfx950"; }

tdag_status tdag_context_create(int device, tdag_context** out_ctx)
{
    if (!out_ctx)
        return fail(TDAG_EINVAL, "out_ctx is null");
    *out_ctx = nullptr;
    int count = 0;
    CAPI_HIP(hipGetDeviceCount(&count), "hipGetDeviceCount");
    if (device < 0 || device >= count)
        return fail(TDAG_EDEVICE, "HIP device " + std::to_string(device) + " not present (" + std::to_string(count) + " devices)");
    CAPI_HIP(hipSetDevice(device), "hipSetDevice");
    tdag_context* ctx = new tdag_context();
    ctx->device = device;
    hipError_t e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e == hipSuccess)
        e = hipHostMalloc((void**)&ctx->pinned, 64, hipHostMallocDefault);
    for (int i = 0; i < 4 && e == hipSuccess; ++i)
        e = hipEventCreate(&ctx->ev[i]);
    // Set-up of the allocation-free matcher's scratch on this device (once per device). A
    // failure here only disables tdag_match_level_device_noalloc, which then reports it.
    if (e == hipSuccess)
        (void)tdag::noallocScratchInit(device);
    if (e != hipSuccess) {
        tdag_context_destroy(ctx);
        return failHip(e, "context creation");
    }
    *out_ctx = ctx;
    return TDAG_OK;
}

tdag_status tdag_context_destroy(tdag_context* ctx)
{
    if (!ctx)
        return TDAG_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream)
        (void)hipStreamSynchronize(ctx->stream);
    for (int i = 0; i < 4; ++i)
        if (ctx->ev[i])
            (void)hipEventDestroy(ctx->ev[i]);
    if (ctx->scratch)
        (void)hipFree(ctx->scratch);
    if (ctx->io)
        (void)hipFree(ctx->io);
    if (ctx->svdagScratch)
        (void)hipFree(ctx->svdagScratch);
    if (ctx->pinned)
        (void)hipHostFree(ctx->pinned);
    if (ctx->stream)
        (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return TDAG_OK;
}

tdag_status tdag_match_level_device(tdag_context* ctx, uint32_t level, const uint64_t* d_masks, uint64_t n, const tdag_config* cfg,
    uint64_t* d_encoded, uint64_t* d_mapping, uint32_t* d_unique, uint64_t* out_num_unique, void* stream, tdag_match_stats* stats)
{
    if (tdag_status s = checkArgs(ctx, level, n, cfg); s != TDAG_OK)
        return s;
    if (n > 0 && !d_masks)
        return fail(TDAG_EINVAL, "d_masks is null");
    CAPI_HIP(hipSetDevice(ctx->device), "hipSetDevice");
    hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
    const int32_t v = ctx->variant[level];
    const tdag::MatchPlan plan = tdag::planMatch(level, n, v >= 0 ? (uint32_t)v : tdag::defaultVariant(level), ctx->firstBatch[level], ctx->batchGrowth[level],
        ctx->bloomBits[level], ctx->childFrames[level]);
    if (tdag_status s = ensureBuffer(ctx->scratch, ctx->scratchBytes, plan.scratchBytes, "matcher scratch allocation"); s != TDAG_OK)
        return s;
    CAPI_HIP(hipEventRecord(ctx->ev[0], st), "hipEventRecord");
    const uint32_t cfgBits = cfgBitsOf(cfg) | (ctx->counters ? tdag::kCfgCount : 0u);
    CAPI_HIP(tdag::enqueueMatch(plan, d_masks, cfgBits, ctx->scratch, st, d_encoded, d_mapping, d_unique, ctx->ev[1], ctx->ev[2]), "matcher launch");
    CAPI_HIP(hipEventRecord(ctx->ev[3], st), "hipEventRecord");
    const bool needSync = out_num_unique || stats;
    if (needSync) {
        CAPI_HIP(hipMemcpyAsync(ctx->pinned, (char*)ctx->scratch + plan.offCounters, 8 * tdag::kNumCounters, hipMemcpyDeviceToHost, st), "counter readback");
        CAPI_HIP(hipStreamSynchronize(st), "hipStreamSynchronize");
        tdag_match_stats s {};
        s.num_items = n;
        s.num_matched = ctx->pinned[0];
        s.num_unique = n > 0 ? (uint32_t)(ctx->pinned[1] & 0xFFFFFFFFull) : 0;
        s.table_entries = 0;
        s.probes = ctx->pinned[2];
        s.filter_pass = ctx->pinned[3];
        s.key_hits = ctx->pinned[4];
        s.candidates = ctx->pinned[5];
        s.verifications = ctx->pinned[6];
        s.verified = ctx->pinned[7];
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[3]) == hipSuccess)
            s.ms_total = ms;
        if (hipEventElapsedTime(&ms, ctx->ev[1], ctx->ev[2]) == hipSuccess)
            s.ms_search = ms;
        ctx->last = s;
        if (out_num_unique)
            *out_num_unique = s.num_unique;
        if (stats)
            *stats = s;
    }
    return TDAG_OK;
}

} // extern "C"

namespace {

// Child flags by rank: rankFlags[r] = flags[order[r]] (order = rank -> original index).
__global__ void kRankFlags(const uint32_t* __restrict__ order, uint32_t n, const uint32_t* __restrict__ flags, uint32_t* __restrict__ rankFlags)
{
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n)
        rankFlags[r] = flags[order[r]] != 0u ? 1u : 0u;
}

} // namespace

extern "C" {

tdag_status tdag_match_level_subset_device(tdag_context* ctx, uint32_t level, const uint64_t* d_masks, uint64_t n, const tdag_config* cfg,
    const uint32_t* d_child_flags, uint64_t* d_encoded, tdag_match_stats* stats)
{
    if (tdag_status s = checkArgs(ctx, level, n, cfg); s != TDAG_OK)
        return s;
    if (n > 0 && (!d_masks || !d_child_flags))
        return fail(TDAG_EINVAL, "d_masks / d_child_flags is null");
    CAPI_HIP(hipSetDevice(ctx->device), "hipSetDevice");
    hipStream_t st = ctx->stream;
    const int32_t v = ctx->variant[level];
    const tdag::MatchPlan plan = tdag::planMatch(level, n, v >= 0 ? (uint32_t)v : tdag::defaultVariant(level), ctx->firstBatch[level], ctx->batchGrowth[level],
        ctx->bloomBits[level], ctx->childFrames[level]);
    const size_t flagsOff = (plan.scratchBytes + 255) & ~size_t(255);
    if (tdag_status s = ensureBuffer(ctx->scratch, ctx->scratchBytes, flagsOff + std::max<size_t>(256, n * 4), "matcher scratch allocation"); s != TDAG_OK)
        return s;
    uint32_t* rankFlags = (uint32_t*)((char*)ctx->scratch + flagsOff);
    const uint32_t cfgBits = cfgBitsOf(cfg) | (ctx->counters ? tdag::kCfgCount : 0u);
    CAPI_HIP(hipEventRecord(ctx->ev[0], st), "hipEventRecord");
    CAPI_HIP(tdag::enqueuePrepareOrder(plan, d_masks, ctx->scratch, st), "matcher launch");
    if (n > 0) {
        hipLaunchKernelGGL(kRankFlags, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, (const uint32_t*)((char*)ctx->scratch + plan.offOrder), (uint32_t)n,
            d_child_flags, rankFlags);
        CAPI_HIP(hipGetLastError(), "child flags");
        CAPI_HIP(tdag::enqueueBuildTable(plan, ctx->scratch, st, rankFlags), "matcher launch");
    }
    CAPI_HIP(hipEventRecord(ctx->ev[1], st), "hipEventRecord");
    if (n > 0)
        CAPI_HIP(tdag::enqueueSearch(plan, cfgBits, ctx->scratch, st), "matcher launch");
    CAPI_HIP(hipEventRecord(ctx->ev[2], st), "hipEventRecord");
    CAPI_HIP(tdag::enqueueFinish(plan, ctx->scratch, st, d_encoded, nullptr, nullptr), "matcher launch");
    CAPI_HIP(hipEventRecord(ctx->ev[3], st), "hipEventRecord");
    CAPI_HIP(hipMemcpyAsync(ctx->pinned, (char*)ctx->scratch + plan.offCounters, 16, hipMemcpyDeviceToHost, st), "counter readback");
    CAPI_HIP(hipStreamSynchronize(st), "hipStreamSynchronize");
    if (stats) {
        tdag_match_stats s {};
        s.num_items = n;
        s.num_matched = ctx->pinned[0];
        s.num_unique = n > 0 ? (uint32_t)(ctx->pinned[1] & 0xFFFFFFFFull) : 0;
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[3]) == hipSuccess)
            s.ms_total = ms;
        if (hipEventElapsedTime(&ms, ctx->ev[1], ctx->ev[2]) == hipSuccess)
            s.ms_search = ms;
        *stats = s;
    }
    return TDAG_OK;
}

tdag_status tdag_match_level(tdag_context* ctx, uint32_t level, const uint64_t* masks, uint64_t n, const tdag_config* cfg,
    uint64_t* out_encoded, uint64_t* out_mapping, uint32_t* out_unique, uint64_t* out_num_unique, tdag_match_stats* stats)
{
    if (tdag_status s = checkArgs(ctx, level, n, cfg); s != TDAG_OK)
        return s;
    if (n > 0 && (!masks || !out_mapping || !out_unique))
        return fail(TDAG_EINVAL, "masks, out_mapping and out_unique are required");
    CAPI_HIP(hipSetDevice(ctx->device), "hipSetDevice");
    const tdag::MatchPlan plan = tdag::planMatch(level, n, tdag::defaultVariant(level));
    const size_t maskBytes = (size_t)n * plan.W * 8;
    const size_t a = (maskBytes + 255) / 256 * 256;
    const size_t b = ((size_t)n * 8 + 255) / 256 * 256;
    const size_t c = ((size_t)n * 4 + 255) / 256 * 256;
    if (tdag_status s = ensureBuffer(ctx->io, ctx->ioBytes, a + 2 * b + c, "io buffer allocation"); s != TDAG_OK)
        return s;
    uint64_t* dMasks = (uint64_t*)ctx->io;
    uint64_t* dEncoded = (uint64_t*)((char*)ctx->io + a);
    uint64_t* dMapping = (uint64_t*)((char*)ctx->io + a + b);
    uint32_t* dUnique = (uint32_t*)((char*)ctx->io + a + 2 * b);
    if (n > 0)
        CAPI_HIP(hipMemcpyAsync(dMasks, masks, maskBytes, hipMemcpyHostToDevice, ctx->stream), "mask upload");
    uint64_t numUnique = 0;
    tdag_match_stats st {};
    if (tdag_status s = tdag_match_level_device(ctx, level, dMasks, n, cfg, out_encoded ? dEncoded : nullptr, dMapping, dUnique, &numUnique, ctx->stream, &st); s != TDAG_OK)
        return s;
    if (n > 0) {
        if (out_encoded)
            CAPI_HIP(hipMemcpyAsync(out_encoded, dEncoded, (size_t)n * 8, hipMemcpyDeviceToHost, ctx->stream), "encoded readback");
        CAPI_HIP(hipMemcpyAsync(out_mapping, dMapping, (size_t)n * 8, hipMemcpyDeviceToHost, ctx->stream), "mapping readback");
        CAPI_HIP(hipMemcpyAsync(out_unique, dUnique, (size_t)numUnique * 4, hipMemcpyDeviceToHost, ctx->stream), "unique readback");
        CAPI_HIP(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    }
    if (out_num_unique)
        *out_num_unique = numUnique;
    if (stats)
        *stats = st;
    return TDAG_OK;
}

tdag_status tdag_context_set_counters(tdag_context* ctx, int enable)
{
    if (!ctx)
        return fail(TDAG_EINVAL, "context is null");
    ctx->counters = enable != 0;
    return TDAG_OK;
}

tdag_status tdag_context_set_variant(tdag_context* ctx, uint32_t level, int32_t variant)
{
    if (!ctx || level < 2 || level > 5)
        return fail(TDAG_EINVAL, "bad context or level");
    if (variant >= 0 && !tdag::variantValid(level, (uint32_t)variant))
        return fail(TDAG_EINVAL, "variant " + std::to_string(variant) + " is not valid at level " + std::to_string(level));
    ctx->variant[level] = variant < 0 ? -1 : variant;
    return TDAG_OK;
}

tdag_status tdag_context_get_variant(const tdag_context* ctx, uint32_t level, int32_t* variant)
{
    if (!ctx || !variant || level < 2 || level > 5)
        return fail(TDAG_EINVAL, "bad argument");
    *variant = ctx->variant[level] >= 0 ? ctx->variant[level] : (int32_t)tdag::defaultVariant(level);
    return TDAG_OK;
}

namespace {

// Average search time (ms) of `plan` over `repeats` runs, after `warm` untimed runs.
tdag_status timePlan(tdag_context* ctx, const tdag::MatchPlan& plan, const uint64_t* d_masks, uint32_t cfgBits, uint32_t warm, uint32_t repeats, float* avg)
{
    if (tdag_status s = ensureBuffer(ctx->scratch, ctx->scratchBytes, plan.scratchBytes, "matcher scratch allocation"); s != TDAG_OK)
        return s;
    // The whole call is timed (preparation, search, post-processing): the child-table frame
    // count and the filter size change the table build as well as the search.
    float total = 0.f;
    for (uint32_t rep = 0; rep < warm + repeats; ++rep) {
        CAPI_HIP(hipEventRecord(ctx->ev[1], ctx->stream), "hipEventRecord");
        CAPI_HIP(tdag::enqueueMatch(plan, d_masks, cfgBits, ctx->scratch, ctx->stream, nullptr, nullptr, nullptr, nullptr, nullptr), "matcher launch");
        CAPI_HIP(hipEventRecord(ctx->ev[2], ctx->stream), "hipEventRecord");
        CAPI_HIP(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
        float ms = 0.f;
        CAPI_HIP(hipEventElapsedTime(&ms, ctx->ev[1], ctx->ev[2]), "hipEventElapsedTime");
        if (rep >= warm)
            total += ms;
    }
    *avg = total / (float)repeats;
    return TDAG_OK;
}

} // namespace

tdag_status tdag_autotune_level(tdag_context* ctx, uint32_t level, const uint64_t* d_masks, uint64_t n, const tdag_config* cfg, uint32_t repeats,
    float* out_ms, int32_t* out_best, uint32_t* out_first_batch)
{
    if (tdag_status s = checkArgs(ctx, level, n, cfg); s != TDAG_OK)
        return s;
    if (n > 0 && !d_masks)
        return fail(TDAG_EINVAL, "d_masks is null");
    CAPI_HIP(hipSetDevice(ctx->device), "hipSetDevice");
    repeats = std::max<uint32_t>(1, repeats);
    const uint32_t cfgBits = cfgBitsOf(cfg);
    // out_ms holds TDAG_VARIANT_SLOTS per-variant times followed by the tuned time.
    constexpr uint32_t kVariantSlots = TDAG_VARIANT_SLOTS;
    static_assert(tdag::kNumVariants <= kVariantSlots, "out_ms has TDAG_VARIANT_SLOTS variant slots");
    if (out_ms)
        for (uint32_t v = 0; v <= kVariantSlots; ++v)
            out_ms[v] = -1.f;
    // 1. Every valid variant at the default schedule, for both child-table frame counts
    //    (levels 3..5: the parent runs 48 combos against 8 entries per child, or 8 combos against
    //    48 entries per child).
    std::vector<std::pair<float, uint32_t>> byTime;
    uint32_t bestFrames = level == 2 ? 1u : tdag::defaultChildFrames(level);
    {
        const uint32_t frameCands[2] = { 1, 6 };
        float bestOfFrames = -1.f;
        for (uint32_t fi = 0; fi < (level == 2 ? 1u : 2u); ++fi) {
            std::vector<std::pair<float, uint32_t>> times;
            for (uint32_t v = 0; v < tdag::kNumVariants; ++v) {
                if (!tdag::variantValid(level, v))
                    continue;
                float avg = 0.f;
                if (tdag_status s = timePlan(ctx, tdag::planMatch(level, n, v, 0, 0, 0, frameCands[fi]), d_masks, cfgBits, 1, repeats, &avg); s != TDAG_OK)
                    return s;
                times.emplace_back(avg, v);
            }
            std::sort(times.begin(), times.end());
            if (!times.empty() && (bestOfFrames < 0.f || times[0].first < bestOfFrames)) {
                bestOfFrames = times[0].first;
                bestFrames = level == 2 ? 1u : frameCands[fi];
                byTime = times;
            }
        }
        if (out_ms)
            for (const auto& t : byTime)
                out_ms[t.second] = t.first;
    }
    int32_t best = byTime.empty() ? 0 : (int32_t)byTime[0].second;
    uint32_t bestFirst = 0;
    float bestMs = byTime.empty() ? 0.f : byTime[0].first;
    if (n > 0 && !byTime.empty()) {
        // 2. The first search batch (parents searched before the first table pruning), jointly
        //    with the three fastest variants: the best schedule differs between variants.
        const uint32_t cands[4] = { (uint32_t)std::max<uint64_t>(64, n / 256), (uint32_t)std::max<uint64_t>(64, n / 64),
            (uint32_t)std::max<uint64_t>(64, n / 16), (uint32_t)n };
        bool first = true;
        for (size_t vi = 0; vi < std::min<size_t>(3, byTime.size()); ++vi) {
            for (uint32_t ci = 0; ci < 4; ++ci) {
                float avg = 0.f;
                if (tdag_status s = timePlan(ctx, tdag::planMatch(level, n, byTime[vi].second, cands[ci], 2, 0, bestFrames), d_masks, cfgBits, 0, repeats, &avg);
                    s != TDAG_OK)
                    return s;
                if (first || avg < bestMs) {
                    first = false;
                    bestMs = avg;
                    best = (int32_t)byTime[vi].second;
                    bestFirst = cands[ci];
                }
            }
        }
        // 3. The Bloom filter size (bits per table entry): small filters stay in L1.
        const uint32_t bitsCands[4] = { 4, 8, 16, 32 };
        const uint32_t defaultBits = tdag::defaultFilterBits(level);
        uint32_t bestBits = defaultBits;
        for (uint32_t bi = 0; bi < 4; ++bi) {
            if (bitsCands[bi] == defaultBits)
                continue;
            float avg = 0.f;
            if (tdag_status s = timePlan(ctx, tdag::planMatch(level, n, (uint32_t)best, bestFirst, 2, bitsCands[bi], bestFrames), d_masks, cfgBits, 0, repeats, &avg);
                s != TDAG_OK)
                return s;
            if (avg < bestMs) {
                bestMs = avg;
                bestBits = bitsCands[bi];
            }
        }
        // 4. The batch growth (each batch this many times the previous one), jointly with the
        //    first batch around the one found: fewer, larger batches prune the table less often
        //    (a prune sweeps the whole table) but leave more settled children in it for longer.
        uint32_t bestGrowth = 2;
        const uint32_t growthCands[5] = { 3, 4, 6, 8, 12 };
        const uint32_t firstAround[3] = { std::max<uint32_t>(1u, bestFirst / 2), bestFirst, (uint32_t)std::min<uint64_t>(n, 2ull * bestFirst) };
        for (uint32_t gi = 0; gi < 5; ++gi) {
            for (uint32_t fi = 0; fi < 3; ++fi) {
                float avg = 0.f;
                if (tdag_status s = timePlan(ctx, tdag::planMatch(level, n, (uint32_t)best, firstAround[fi], growthCands[gi], bestBits, bestFrames), d_masks, cfgBits, 0,
                        repeats, &avg);
                    s != TDAG_OK)
                    return s;
                if (avg < bestMs) {
                    bestMs = avg;
                    bestGrowth = growthCands[gi];
                    bestFirst = firstAround[fi];
                }
            }
        }
        ctx->firstBatch[level] = bestFirst;
        ctx->batchGrowth[level] = bestGrowth;
        ctx->bloomBits[level] = bestBits;
        if (out_ms)
            out_ms[kVariantSlots] = bestMs;
    }
    ctx->childFrames[level] = bestFrames;
    ctx->variant[level] = best;
    if (out_best)
        *out_best = best;
    if (out_first_batch)
        *out_first_batch = bestFirst;
    return TDAG_OK;
}

tdag_status tdag_context_set_child_frames(tdag_context* ctx, uint32_t level, uint32_t frames)
{
    if (!ctx || level < 3 || level > 5 || (frames != 0 && frames != 1 && frames != 6))
        return fail(TDAG_EINVAL, "bad context, level (3..5) or frame count (0, 1 or 6)");
    ctx->childFrames[level] = frames;
    return TDAG_OK;
}

tdag_status tdag_context_get_child_frames(const tdag_context* ctx, uint32_t level, uint32_t* frames)
{
    if (!ctx || !frames || level < 3 || level > 5)
        return fail(TDAG_EINVAL, "bad argument");
    *frames = ctx->childFrames[level] ? ctx->childFrames[level] : tdag::defaultChildFrames(level);
    return TDAG_OK;
}

tdag_status tdag_context_set_filter_bits(tdag_context* ctx, uint32_t level, uint32_t bits_per_entry)
{
    if (!ctx || level < 2 || level > 5 || bits_per_entry > 64)
        return fail(TDAG_EINVAL, "bad context, level or bits per entry");
    ctx->bloomBits[level] = bits_per_entry;
    return TDAG_OK;
}

tdag_status tdag_context_get_filter_bits(const tdag_context* ctx, uint32_t level, uint32_t* bits_per_entry)
{
    if (!ctx || !bits_per_entry || level < 2 || level > 5)
        return fail(TDAG_EINVAL, "bad argument");
    *bits_per_entry = ctx->bloomBits[level] ? ctx->bloomBits[level] : tdag::defaultFilterBits(level);
    return TDAG_OK;
}

tdag_status tdag_context_set_batching(tdag_context* ctx, uint32_t level, uint32_t first_batch, uint32_t growth)
{
    if (!ctx || level < 2 || level > 5)
        return fail(TDAG_EINVAL, "bad context or level");
    if (growth == 1)
        return fail(TDAG_EINVAL, "batch growth must be 0 (default) or >= 2");
    ctx->firstBatch[level] = first_batch;
    ctx->batchGrowth[level] = growth;
    return TDAG_OK;
}

tdag_status tdag_context_get_batching(const tdag_context* ctx, uint32_t level, uint32_t* first_batch, uint32_t* growth)
{
    if (!ctx || level < 2 || level > 5)
        return fail(TDAG_EINVAL, "bad context or level");
    if (first_batch)
        *first_batch = ctx->firstBatch[level];
    if (growth)
        *growth = ctx->batchGrowth[level];
    return TDAG_OK;
}

tdag_status tdag_last_stats(const tdag_context* ctx, tdag_match_stats* out)
{
    if (!ctx || !out)
        return fail(TDAG_EINVAL, "null argument");
    *out = ctx->last;
    return TDAG_OK;
}


} // extern "C"

namespace {

tdag_status uploadBatchDesc(tdag_batch* b)
{
    tdag::BatchDesc desc;
    tdag::fillBatchDesc(b->plan, b->scratch, b->cfgBits, b->grab, &desc, b->splitEvery);
    CAPI_HIP(hipMemcpy((char*)b->scratch + b->plan.descOff, &desc, sizeof(desc), hipMemcpyHostToDevice), "batch descriptor upload");
    return TDAG_OK;
}

tdag_status enqueueConcurrentCall(tdag_batch* b)
{
    const uint32_t k = (uint32_t)b->jobs.size();
    hipStream_t main = b->ctx->stream;
    CAPI_HIP(hipEventRecord(b->evFork, main), "fork event");
    for (uint32_t i = 0; i < k; ++i) {
        const tdag_level_job& j = b->jobs[i];
        CAPI_HIP(hipStreamWaitEvent(b->aux[i], b->evFork, 0), "fork wait");
        if ((int)i == b->learned)
            CAPI_HIP(tdag::learn::enqueueLearned2(b->cplans[i], &b->mlp->mlp, b->k, j.d_masks, b->cfgBits & ~tdag::kCfgCount, (char*)b->scratch + b->cOff[i],
                         (char*)b->scratch + b->lWorkOff, b->aux[i], j.d_encoded, j.d_mapping, j.d_unique, nullptr),
                "learned matcher launch");
        else
            CAPI_HIP(tdag::enqueueMatch(b->cplans[i], j.d_masks, b->cfgBits, (char*)b->scratch + b->cOff[i], b->aux[i], j.d_encoded, j.d_mapping, j.d_unique, nullptr,
                         nullptr),
                "concurrent matcher launch");
        CAPI_HIP(hipEventRecord(b->evJoin[i], b->aux[i]), "join event");
        CAPI_HIP(hipStreamWaitEvent(main, b->evJoin[i], 0), "join wait");
    }
    return TDAG_OK;
}

tdag_status enqueueBatchCall(tdag_batch* b)
{
    if (b->concurrent)
        return enqueueConcurrentCall(b);
    hipStream_t main = b->ctx->stream;
    if (b->learned >= 0) {
        // The learned level-2 matcher runs on the call's stream ahead of the persistent kernel,
        // never beside it: the persistent grid is sized to the blocks the GPU holds at once, and
        // its grid barriers wait for all of them, so no other kernel may take CUs from it.
        const tdag_level_job& j = b->jobs[b->learned];
        CAPI_HIP(tdag::learn::enqueueLearned2(b->lplan, &b->mlp->mlp, b->k, j.d_masks, b->cfgBits & ~tdag::kCfgCount, (char*)b->scratch + b->lOff,
                     (char*)b->scratch + b->lWorkOff, main, j.d_encoded, j.d_mapping, j.d_unique, nullptr),
            "learned matcher launch");
    }
    const uint32_t k = b->plan.count;
    if (k > 0) {
        // planBatch's job indices refer to the jobs without the learned one; map them back.
        std::vector<const uint64_t*> masks(b->jobs.size());
        std::vector<uint64_t*> enc(b->jobs.size()), map(b->jobs.size());
        std::vector<uint32_t*> uniq(b->jobs.size());
        uint32_t c = 0;
        for (uint32_t i = 0; i < b->jobs.size(); ++i) {
            if ((int)i == b->learned)
                continue;
            masks[c] = b->jobs[i].d_masks;
            enc[c] = b->jobs[i].d_encoded;
            map[c] = b->jobs[i].d_mapping;
            uniq[c] = b->jobs[i].d_unique;
            ++c;
        }
        CAPI_HIP(tdag::enqueueBatch(b->plan, masks.data(), b->scratch, main, enc.data(), map.data(), uniq.data(), nullptr, nullptr, b->spinTicks),
            "batched matcher launch");
    }
    return TDAG_OK;
}

// Caller's job index of the persistent plan's entry i (the learned job is not in the plan).
uint32_t persistentJob(const tdag_batch* b, uint32_t i)
{
    uint32_t j = b->plan.job[i];
    if (b->learned >= 0 && j >= (uint32_t)b->learned)
        ++j;
    return j;
}

} // namespace

extern "C" {

static tdag_status createBatch(tdag_context* ctx, const tdag_level_job* jobs, uint32_t count, const tdag_config* cfg, uint32_t flags, tdag_mlp* mlp, uint32_t kProp,
    tdag_batch** out)
{
    if (!ctx || !jobs || !cfg || !out)
        return fail(TDAG_EINVAL, "null argument");
    if (count == 0 || count > tdag::kBatchMaxLevels)
        return fail(TDAG_EINVAL, "a batch holds 1 to 4 levels");
    int learned = -1;
    if (mlp) {
        if (mlp->ctx != ctx)
            return fail(TDAG_EINVAL, "the network belongs to another context");
        if (kProp == 0 || kProp > tdag::learn::kMaxProposals)
            return fail(TDAG_EINVAL, "k must be in [1, 64]");
        for (uint32_t i = 0; i < count; ++i)
            if (jobs[i].level == 2)
                learned = (int)i;
        if (learned < 0)
            return fail(TDAG_EINVAL, "the learned proposal needs a level-2 job");
    }
    uint32_t levels[tdag::kBatchMaxLevels], first[tdag::kBatchMaxLevels], bits[tdag::kBatchMaxLevels], frames[tdag::kBatchMaxLevels];
    uint64_t ns[tdag::kBatchMaxLevels];
    for (uint32_t i = 0; i < count; ++i) {
        if (tdag_status s = checkArgs(ctx, jobs[i].level, jobs[i].n, cfg); s != TDAG_OK)
            return s;
        if (jobs[i].n > 0 && !jobs[i].d_masks)
            return fail(TDAG_EINVAL, "d_masks is null");
        for (uint32_t j = 0; j < i; ++j)
            if (jobs[j].level == jobs[i].level)
                return fail(TDAG_EINVAL, "a batch holds at most one job per level");
        levels[i] = jobs[i].level;
        ns[i] = jobs[i].n;
        first[i] = ctx->firstBatch[jobs[i].level];
        bits[i] = ctx->bloomBits[jobs[i].level];
        frames[i] = ctx->childFrames[jobs[i].level];
    }
    CAPI_HIP(hipSetDevice(ctx->device), "hipSetDevice");
    auto* b = new tdag_batch();
    b->ctx = ctx;
    b->jobs.assign(jobs, jobs + count);
    b->concurrent = (flags & TDAG_BATCH_CONCURRENT) != 0;
    b->mlp = mlp;
    b->k = kProp;
    b->learned = learned;
    auto cleanup = [&](tdag_status s) {
        tdag_batch_destroy(b);
        return s;
    };
    // The learned job's plan: the one tdag_match_level_learned_device uses.
    const auto learnedPlan = [&](uint64_t n) { return tdag::planMatch(2, n, 0, ctx->firstBatch[2], ctx->batchGrowth[2], ctx->bloomBits[2]); };
    if (learned >= 0)
        if (hipError_t e = tdag::learn::mlpPrepareEmbed(mlp->mlp, jobs[learned].n); e != hipSuccess)
            return cleanup(failHip(e, "network workspace"));
    if (b->concurrent) {
        // The same plans tdag_match_level_device would use for each level (tuned variant and
        // schedule), each with its own scratch, stream and join event.
        size_t off = 0;
        for (uint32_t i = 0; i < count; ++i) {
            const uint32_t l = jobs[i].level;
            const int32_t v = ctx->variant[l];
            b->cplans[i] = (int)i == learned ? learnedPlan(jobs[i].n)
                                             : tdag::planMatch(l, jobs[i].n, v >= 0 ? (uint32_t)v : tdag::defaultVariant(l), ctx->firstBatch[l], ctx->batchGrowth[l],
                                                   ctx->bloomBits[l], ctx->childFrames[l]);
            b->cOff[i] = off;
            off += (b->cplans[i].scratchBytes + 255) & ~size_t(255);
        }
        if (learned >= 0) {
            b->lWorkOff = off;
            off += (tdag::learn::learnedWorkBytes(jobs[learned].n) + 255) & ~size_t(255);
        }
        if (hipError_t e = hipMalloc(&b->scratch, std::max<size_t>(off, 256)); e != hipSuccess)
            return cleanup(failHip(e, "concurrent matcher scratch allocation"));
        for (uint32_t i = 0; i < count; ++i) {
            if (hipError_t e = hipStreamCreateWithFlags(&b->aux[i], hipStreamNonBlocking); e != hipSuccess) // default priority
                return cleanup(failHip(e, "batch stream creation"));
            if (hipError_t e = hipEventCreateWithFlags(&b->evJoin[i], hipEventDisableTiming); e != hipSuccess)
                return cleanup(failHip(e, "batch event creation"));
        }
        if (hipError_t e = hipEventCreateWithFlags(&b->evFork, hipEventDisableTiming); e != hipSuccess)
            return cleanup(failHip(e, "batch event creation"));
    } else {
        // The persistent kernel takes every job but the learned one.
        uint32_t pl[tdag::kBatchMaxLevels], pf[tdag::kBatchMaxLevels], pb[tdag::kBatchMaxLevels], pr[tdag::kBatchMaxLevels];
        uint64_t pn[tdag::kBatchMaxLevels];
        uint32_t c = 0;
        for (uint32_t i = 0; i < count; ++i) {
            if ((int)i == learned)
                continue;
            pl[c] = levels[i];
            pn[c] = ns[i];
            pf[c] = first[i];
            pb[c] = bits[i];
            pr[c] = frames[i];
            ++c;
        }
        size_t total = 0;
        if (c > 0) {
            if (hipError_t e = tdag::planBatch(c, pl, pn, pf, pb, pr, ctx->device, &b->plan); e != hipSuccess)
                return cleanup(failHip(e, "batched matcher plan"));
            total = (b->plan.scratchBytes + 255) & ~size_t(255);
        }
        if (learned >= 0) {
            b->lplan = learnedPlan(jobs[learned].n);
            b->lOff = total;
            total += (b->lplan.scratchBytes + 255) & ~size_t(255);
            b->lWorkOff = total;
            total += (tdag::learn::learnedWorkBytes(jobs[learned].n) + 255) & ~size_t(255);
        }
        if (hipError_t e = hipMalloc(&b->scratch, std::max<size_t>(total, 256)); e != hipSuccess)
            return cleanup(failHip(e, "batched matcher scratch allocation"));
    }
    if (hipError_t e = hipHostMalloc((void**)&b->pinned, 8 * (tdag::kNumCounters * count + 3)); e != hipSuccess)
        return cleanup(failHip(e, "pinned readback allocation"));
    int rateKHz = 0;
    if (hipError_t e = hipDeviceGetAttribute(&rateKHz, hipDeviceAttributeWallClockRate, ctx->device); e != hipSuccess)
        return cleanup(failHip(e, "wall clock rate"));
    b->spinTicks = (uint64_t)std::max(rateKHz, 1) * 1000ull * 10ull; // 10 s
    b->cfgBits = cfgBitsOf(cfg) | (ctx->counters ? tdag::kCfgCount : 0u);
    b->grab = 4;
    if (!b->concurrent && b->plan.count > 0)
        if (tdag_status s = uploadBatchDesc(b); s != TDAG_OK)
            return cleanup(s);
    if (flags & TDAG_BATCH_GRAPH) {
        if (hipError_t e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal); e != hipSuccess)
            return cleanup(failHip(e, "hipStreamBeginCapture"));
        const tdag_status s = enqueueBatchCall(b);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
        b->graph = g;
        if (s != TDAG_OK)
            return cleanup(s);
        if (e != hipSuccess)
            return cleanup(failHip(e, "hipStreamEndCapture"));
        if (hipError_t e2 = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0); e2 != hipSuccess)
            return cleanup(failHip(e2, "hipGraphInstantiate"));
    }
    *out = b;
    return TDAG_OK;
}

tdag_status tdag_batch_create(tdag_context* ctx, const tdag_level_job* jobs, uint32_t count, const tdag_config* cfg, int use_graph, tdag_batch** out)
{
    return createBatch(ctx, jobs, count, cfg, (uint32_t)use_graph, nullptr, 0, out);
}

tdag_status tdag_batch_create_learned(tdag_context* ctx, const tdag_level_job* jobs, uint32_t count, const tdag_config* cfg, uint32_t flags, tdag_mlp* mlp, uint32_t k,
    tdag_batch** out)
{
    if (!mlp)
        return fail(TDAG_EINVAL, "mlp is null");
    return createBatch(ctx, jobs, count, cfg, flags, mlp, k, out);
}

tdag_status tdag_batch_run(tdag_batch* b, tdag_match_stats* out_stats, float* out_ms)
{
    if (!b)
        return fail(TDAG_EINVAL, "batch is null");
    tdag_context* ctx = b->ctx;
    CAPI_HIP(hipSetDevice(ctx->device), "hipSetDevice");
    CAPI_HIP(hipEventRecord(ctx->ev[0], ctx->stream), "hipEventRecord");
    if (b->exec) {
        CAPI_HIP(hipGraphLaunch(b->exec, ctx->stream), "hipGraphLaunch");
    } else if (tdag_status s = enqueueBatchCall(b); s != TDAG_OK) {
        return s;
    }
    CAPI_HIP(hipEventRecord(ctx->ev[3], ctx->stream), "hipEventRecord");
    const uint32_t k = (uint32_t)b->jobs.size();
    const uint32_t numLaunched = b->concurrent ? k : b->plan.count;
    for (uint32_t i = 0; i < numLaunched; ++i) {
        const tdag::LevelView v = b->concurrent ? tdag::levelView(b->cplans[i], (char*)b->scratch + b->cOff[i])
                                                : tdag::levelView(b->plan.plans[i], (char*)b->scratch + b->plan.scratchOff[i]);
        const uint32_t job = b->concurrent ? i : persistentJob(b, i);
        CAPI_HIP(hipMemcpyAsync(b->pinned + tdag::kNumCounters * job, v.counters, 8 * tdag::kNumCounters, hipMemcpyDeviceToHost, ctx->stream), "counter readback");
    }
    if (b->learned >= 0) {
        const tdag::MatchPlan& lp = b->concurrent ? b->cplans[b->learned] : b->lplan;
        char* base = (char*)b->scratch + (b->concurrent ? b->cOff[b->learned] : b->lOff);
        if (!b->concurrent) {
            const tdag::LevelView v = tdag::levelView(lp, base);
            CAPI_HIP(hipMemcpyAsync(b->pinned + tdag::kNumCounters * b->learned, v.counters, 8 * tdag::kNumCounters, hipMemcpyDeviceToHost, ctx->stream),
                "counter readback");
        }
        CAPI_HIP(hipMemcpyAsync(b->pinned + tdag::kNumCounters * k + 1, tdag::learn::learnedCounters(lp, (char*)b->scratch + b->lWorkOff), 16,
                     hipMemcpyDeviceToHost, ctx->stream),
            "learned counter readback");
    }
    if (b->concurrent || b->plan.count == 0)
        b->pinned[tdag::kNumCounters * k] = 0; // no grid barrier in this mode
    else
        CAPI_HIP(hipMemcpyAsync(b->pinned + tdag::kNumCounters * k, (char*)b->scratch + b->plan.syncOff + offsetof(tdag::BatchSync, error), 4, hipMemcpyDeviceToHost,
                     ctx->stream),
            "barrier status readback");
    CAPI_HIP(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    if ((uint32_t)b->pinned[tdag::kNumCounters * k] != 0)
        return fail(TDAG_EINTERNAL, "batched matcher: a grid barrier timed out (not all blocks were resident); results are invalid");
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[3]);
    if (out_ms)
        *out_ms = ms;
    if (out_stats)
        for (uint32_t j = 0; j < k; ++j) {
            const uint64_t* c = b->pinned + tdag::kNumCounters * j;
            tdag_match_stats s {};
            s.num_items = b->jobs[j].n;
            s.num_matched = c[0];
            s.num_unique = b->jobs[j].n > 0 ? (uint32_t)(c[1] & 0xFFFFFFFFull) : 0;
            s.probes = c[2];
            s.filter_pass = c[3];
            s.key_hits = c[4];
            s.candidates = c[5];
            s.verifications = c[6];
            s.verified = c[7];
            s.ms_total = ms;
            s.ms_search = ms;
            out_stats[j] = s;
        }
    return TDAG_OK;
}

tdag_status tdag_batch_autotune(tdag_batch* b, uint32_t repeats, uint32_t* out_grab)
{
    if (!b)
        return fail(TDAG_EINVAL, "batch is null");
    if (b->concurrent || b->plan.count == 0) { // no work queue to tune: each level runs its own tuned kernels
        if (out_grab)
            *out_grab = 0;
        return TDAG_OK;
    }
    repeats = std::max<uint32_t>(1, repeats);
    // Mean time of `repeats` runs after a warm-up with the batch's current tuning.
    auto timeRuns = [&](float* outMs) -> tdag_status {
        if (tdag_status s = uploadBatchDesc(b); s != TDAG_OK)
            return s;
        float total = 0.f;
        for (uint32_t rep = 0; rep <= repeats; ++rep) {
            float ms = 0.f;
            if (tdag_status s = tdag_batch_run(b, nullptr, &ms); s != TDAG_OK)
                return s;
            if (rep > 0)
                total += ms;
        }
        *outMs = total / repeats;
        return TDAG_OK;
    };
    // Tiles per pop first, then the share of blocks that start on the lighter levels' queue.
    const uint32_t grabs[5] = { 1, 2, 4, 8, 16 };
    uint32_t bestGrab = b->grab;
    float bestMs = -1.f;
    for (uint32_t c : grabs) {
        b->grab = c;
        float ms = 0.f;
        if (tdag_status s = timeRuns(&ms); s != TDAG_OK)
            return s;
        if (bestMs < 0.f || ms < bestMs) {
            bestMs = ms;
            bestGrab = c;
        }
    }
    b->grab = bestGrab;
    const uint32_t splits[5] = { 0, 2, 3, 4, 6 };
    uint32_t bestSplit = b->splitEvery;
    bestMs = -1.f;
    for (uint32_t c : splits) {
        b->splitEvery = c;
        float ms = 0.f;
        if (tdag_status s = timeRuns(&ms); s != TDAG_OK)
            return s;
        if (bestMs < 0.f || ms < bestMs) {
            bestMs = ms;
            bestSplit = c;
        }
    }
    b->splitEvery = bestSplit;
    if (tdag_status s = uploadBatchDesc(b); s != TDAG_OK)
        return s;
    if (out_grab)
        *out_grab = bestGrab;
    return TDAG_OK;
}

tdag_status tdag_batch_set_tuning(tdag_batch* b, uint32_t grab, uint32_t split_every)
{
    if (!b)
        return fail(TDAG_EINVAL, "batch is null");
    if (grab == 0 || grab > 64)
        return fail(TDAG_EINVAL, "tiles per pop must be 1..64");
    if (split_every == 1 || split_every > 64)
        return fail(TDAG_EINVAL, "split_every must be 0 (off) or 2..64");
    b->grab = grab;
    b->splitEvery = split_every;
    if (b->concurrent || b->plan.count == 0)
        return TDAG_OK;
    return uploadBatchDesc(b);
}

tdag_status tdag_batch_get_tuning(const tdag_batch* b, uint32_t* out_grab, uint32_t* out_split_every)
{
    if (!b)
        return fail(TDAG_EINVAL, "batch is null");
    if (out_grab)
        *out_grab = b->grab;
    if (out_split_every)
        *out_split_every = b->splitEvery;
    return TDAG_OK;
}

tdag_status tdag_batch_learned_stats(const tdag_batch* b, uint64_t* out_proposal_hits, uint64_t* out_fallback_queries)
{
    if (!b)
        return fail(TDAG_EINVAL, "batch is null");
    if (b->learned < 0)
        return fail(TDAG_EINVAL, "the batch has no learned job");
    const uint32_t k = (uint32_t)b->jobs.size();
    if (out_proposal_hits)
        *out_proposal_hits = b->pinned[tdag::kNumCounters * k + 1];
    if (out_fallback_queries)
        *out_fallback_queries = b->pinned[tdag::kNumCounters * k + 2];
    return TDAG_OK;
}

tdag_status tdag_batch_destroy(tdag_batch* b)
{
    if (!b)
        return TDAG_OK;
    if (b->exec)
        (void)hipGraphExecDestroy(b->exec);
    if (b->graph)
        (void)hipGraphDestroy(b->graph);
    if (b->scratch)
        (void)hipFree(b->scratch);
    if (b->pinned)
        (void)hipHostFree(b->pinned);
    for (auto st : b->aux)
        if (st)
            (void)hipStreamDestroy(st);
    for (auto ev : b->evJoin)
        if (ev)
            (void)hipEventDestroy(ev);
    if (b->evFork)
        (void)hipEventDestroy(b->evFork);
    delete b;
    return TDAG_OK;
}

tdag_status tdag_match_levels_batched_device(tdag_context* ctx, const tdag_level_job* jobs, uint32_t count, const tdag_config* cfg, tdag_match_stats* out_stats)
{
    tdag_batch* b = nullptr;
    if (tdag_status s = tdag_batch_create(ctx, jobs, count, cfg, 0, &b); s != TDAG_OK)
        return s;
    const tdag_status s = tdag_batch_run(b, out_stats, nullptr);
    tdag_batch_destroy(b);
    return s;
}

} // extern "C"
