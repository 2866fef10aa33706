// The translation stage of constructTransformDAG (transform_dag.cpp:165-233) with every step on the
// device: for each level from maxTranslationLevel down to 2, the nodes referenced by the surviving
// parents are marked and compacted in ascending index order, flattened into the matcher's input,
// matched, the unique ones emitted as the level's output nodes, and the parents' child pointers
// rewritten through the matcher's mapping. Level L's kernels follow level L+1's on one stream, so
// the levels are gated on the device; only counts cross to the host, once per level, on the first
// (eager) run. Once the per-level sizes of a build are known, the whole stage is one fixed
// sequence of launches and is captured as a hipGraph that replays the build.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "device_build.hpp"
#include "primitives.hpp"

namespace tdag {

namespace {

constexpr uint32_t kThreads = 256;
constexpr uint64_t kPtrMask = (1ull << 37) - 1;
constexpr uint64_t kSentinel = 0x82082400FFFFFFFFull;

inline uint32_t blocksFor(uint64_t n) { return (uint32_t)std::max<uint64_t>(1, (n + kThreads - 1) / kThreads); }
inline size_t al256(size_t v) { return (v + 255) & ~size_t(255); }

// referenced[child] = 1 for every child pointer of the np parents (transform_dag.cpp:177-185). A
// pointer outside the level sets *bad.
__global__ __launch_bounds__(kThreads) void kMarkReferenced(const uint64_t* __restrict__ parents, uint32_t np, uint32_t numItems, uint32_t* __restrict__ referenced,
    uint32_t* __restrict__ bad)
{
    const uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
    if (i >= (uint64_t)np * 8)
        return;
    const uint64_t c = parents[i];
    if (c == kSentinel)
        return;
    const uint64_t ptr = c & kPtrMask;
    if (ptr >= numItems) {
        atomicOr(bad, 1u);
        return;
    }
    referenced[ptr] = 1u;
}

// indices[pos[i]] = i for every referenced i: the referenced items in ascending order
// (transform_dag.cpp:186-192).
__global__ __launch_bounds__(kThreads) void kCompact(const uint32_t* __restrict__ referenced, const uint32_t* __restrict__ pos, uint32_t numItems,
    uint32_t* __restrict__ indices)
{
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i < numItems && referenced[i])
        indices[pos[i]] = i;
}

// The level's output (transform_dag.cpp:208-221): unique item ui is input node
// indices[unique[ui]], as a TransformNode whose children are identity pointers to the input level
// below (transform_dag.cpp:123-135), or the leaf mask at level 2. nu is read from the device.
__global__ __launch_bounds__(kThreads) void kEmit(DeviceSvdag d, uint32_t level, const uint32_t* __restrict__ indices, const uint32_t* __restrict__ unique,
    const uint32_t* __restrict__ nuDev, uint32_t n, uint64_t* __restrict__ out)
{
    const uint32_t nu = *nuDev;
    if (level == 2) {
        const uint32_t ui = blockIdx.x * kThreads + threadIdx.x;
        if (ui < nu && ui < n)
            out[ui] = d.leaves[indices[unique[ui]]];
        return;
    }
    const uint64_t t = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
    const uint32_t ui = (uint32_t)(t >> 3), c = (uint32_t)(t & 7u);
    if (ui >= nu || ui >= n)
        return;
    const uint32_t child = d.nodes[level][(size_t)indices[unique[ui]] * 8 + c];
    out[t] = child == 0xFFFFFFFFu ? kSentinel : (0x8208240000000000ull | child);
}

// Parent pointers rewritten through the matcher's mapping (transform_dag.cpp:224-233): child ptr p
// (an input index) becomes mapping[pos[p]], pos[p] being p's place among the referenced items.
// A pointer outside the level (already reported by kMarkReferenced) is left as it is.
__global__ __launch_bounds__(kThreads) void kRewrite(uint64_t* __restrict__ parents, uint32_t np, uint32_t numItems, const uint32_t* __restrict__ pos,
    const uint64_t* __restrict__ mapping)
{
    const uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
    if (i >= (uint64_t)np * 8)
        return;
    const uint64_t c = parents[i];
    if (c != kSentinel && (c & kPtrMask) < numItems)
        parents[i] = mapping[pos[c & kPtrMask]];
}

// Level counters out of the matcher's: items, unique, matched.
__global__ void kLevelCounts(const uint32_t* __restrict__ total, const unsigned long long* __restrict__ matcher, uint32_t* __restrict__ out)
{
    if (threadIdx.x == 0) {
        out[0] = total ? *total : 0u;
        out[1] = matcher ? (uint32_t)(matcher[1] & 0xFFFFFFFFull) : 0u;
        out[2] = matcher ? (uint32_t)matcher[0] : 0u;
    }
}

} // namespace

#define DB_TRY(expr)                   \
    do {                               \
        const hipError_t e_ = (expr);  \
        if (e_ != hipSuccess)          \
            return e_;                 \
    } while (0)

hipError_t DeviceTranslation::init(const DeviceSvdag& dag, const Svdag& host, const std::vector<std::array<uint64_t, 8>>& mid, uint32_t maxLevel,
    const LevelTuning* tuning)
{
    release();
    d = dag;
    maxT = maxLevel;
    numMid = (uint32_t)mid.size();
    for (uint32_t l = 0; l < 6; ++l)
        tune[l] = tuning ? tuning[l] : LevelTuning {};
    // Arena: sized with each level's item count as the bound of its referenced items.
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t o = off;
        off += al256(std::max<size_t>(bytes, 8));
        return o;
    };
    uint64_t maxItems = 1, maxMaskWords = 1;
    size_t maxScratch = 256;
    for (uint32_t level = maxT; level >= 2; --level) {
        const uint64_t items = level == 2 ? host.leaves.size() : host.nodes[level].size();
        numItems[level] = (uint32_t)items;
        const uint32_t R = 1u << level;
        const uint64_t W = std::max<uint64_t>(1, (uint64_t)R * R * R / 64);
        maxItems = std::max(maxItems, items);
        maxMaskWords = std::max(maxMaskWords, items * W);
        const LevelTuning& t = tune[level];
        const MatchPlan p = planMatch(level, items, t.variant >= 0 ? (uint32_t)t.variant : defaultVariant(level), t.firstBatch, t.batchGrowth, t.bloomBits,
            t.childFrames);
        maxScratch = std::max(maxScratch, p.scratchBytes);
        maxScratch = std::max(maxScratch, flattenTempBytes(level, items));
        outOff[level] = take(items * (level == 2 ? 8 : 64));
        if (level == 2)
            break;
    }
    midOff = take((size_t)numMid * 64);
    refOff = take((maxItems + 1) * 4);
    posOff = take((maxItems + 1) * 4);
    totalOff = take(64);
    idxOff = take(maxItems * 4);
    maskOff = take(maxMaskWords * 8);
    mapOff = take(maxItems * 8);
    uniqOff = take(maxItems * 4);
    scanOff = take(scanScratchBytes(maxItems));
    scratchOff = take(maxScratch);
    scratchCap = maxScratch;
    countsOff = take(6 * 4 * 4);
    badOff = take(64);
    arenaBytes = off;
    DB_TRY(hipMalloc(&arena, arenaBytes));
    DB_TRY(hipMemcpy(base() + midOff, mid.data(), (size_t)numMid * 64, hipMemcpyHostToDevice));
    midHost = mid;
    return hipSuccess;
}

void DeviceTranslation::release()
{
    if (graphExec)
        (void)hipGraphExecDestroy(graphExec);
    if (graph)
        (void)hipGraphDestroy(graph);
    if (arena)
        (void)hipFree(arena);
    graphExec = nullptr;
    graph = nullptr;
    arena = nullptr;
    known = false;
}

MatchPlan DeviceTranslation::planFor(uint32_t level, uint64_t count) const
{
    const LevelTuning& t = tune[level];
    return planMatch(level, count, t.variant >= 0 ? (uint32_t)t.variant : defaultVariant(level), t.firstBatch, t.batchGrowth, t.bloomBits, t.childFrames);
}

// One level: the np parents -> referenced items (flags, positions, the ascending list) -> n.
hipError_t DeviceTranslation::enqueueMark(uint32_t level, const uint64_t* parents, uint32_t np, hipStream_t s)
{
    char* b = base();
    uint32_t* ref = (uint32_t*)(b + refOff);
    uint32_t* pos = (uint32_t*)(b + posOff);
    uint32_t* total = (uint32_t*)(b + totalOff);
    const uint32_t items = numItems[level];
    DB_TRY(fillU32(ref, (size_t)items + 1, 0u, s));
    hipLaunchKernelGGL(kMarkReferenced, dim3(blocksFor((uint64_t)np * 8)), dim3(kThreads), 0, s, parents, np, items, ref, (uint32_t*)(b + badOff));
    DB_TRY(hipGetLastError());
    DB_TRY(exclusiveScanU32(ref, pos, items, total, b + scanOff, s));
    hipLaunchKernelGGL(kCompact, dim3(blocksFor(items)), dim3(kThreads), 0, s, ref, pos, items, (uint32_t*)(b + idxOff));
    return hipGetLastError();
}

// With n known: flatten, match, emit, rewrite the parents, record the counts.
hipError_t DeviceTranslation::enqueueMatchLevel(uint32_t level, uint32_t count, uint64_t* parents, uint32_t np, uint32_t cfgBits, hipStream_t s)
{
    char* b = base();
    uint32_t* counts = (uint32_t*)(b + countsOff) + 4 * level;
    const uint32_t* total = (const uint32_t*)(b + totalOff);
    if (count == 0) {
        hipLaunchKernelGGL(kLevelCounts, dim3(1), dim3(64), 0, s, total, nullptr, counts);
        return hipGetLastError();
    }
    const MatchPlan p = planFor(level, count);
    if (count > numItems[level] || p.scratchBytes > scratchCap || flattenTempBytes(level, count) > scratchCap)
        return hipErrorInvalidValue;
    const uint32_t* idx = (const uint32_t*)(b + idxOff);
    uint64_t* masks = (uint64_t*)(b + maskOff);
    DB_TRY(flattenDevice(d, level, idx, count, masks, (uint64_t*)(b + scratchOff), s, (uint32_t*)(b + badOff)));
    uint64_t* mapping = (uint64_t*)(b + mapOff);
    uint32_t* uniq = (uint32_t*)(b + uniqOff);
    void* scratch = b + scratchOff;
    DB_TRY(enqueueMatch(p, masks, cfgBits, scratch, s, nullptr, mapping, uniq, nullptr, nullptr));
    const unsigned long long* mc = (const unsigned long long*)((char*)scratch + p.offCounters);
    hipLaunchKernelGGL(kLevelCounts, dim3(1), dim3(64), 0, s, total, mc, counts);
    uint64_t* out = (uint64_t*)(b + outOff[level]);
    const uint64_t emitThreads = level == 2 ? count : (uint64_t)count * 8;
    hipLaunchKernelGGL(kEmit, dim3(blocksFor(emitThreads)), dim3(kThreads), 0, s, d, level, idx, uniq, counts + 1, count, out);
    hipLaunchKernelGGL(kRewrite, dim3(blocksFor((uint64_t)np * 8)), dim3(kThreads), 0, s, parents, np, numItems[level], (const uint32_t*)(b + posOff), mapping);
    return hipGetLastError();
}

uint64_t* DeviceTranslation::parentsOf(uint32_t level)
{
    return level == maxT ? (uint64_t*)(base() + midOff) : (uint64_t*)(base() + outOff[level + 1]);
}

hipError_t DeviceTranslation::runEager(uint32_t cfgBits, hipStream_t s)
{
    char* b = base();
    DB_TRY(fillU32((uint32_t*)(b + badOff), 16, 0u, s));
    DB_TRY(fillU32((uint32_t*)(b + countsOff), 24, 0u, s));
    DB_TRY(hipMemcpyAsync(b + midOff, midHost.data(), (size_t)numMid * 64, hipMemcpyHostToDevice, s));
    for (uint32_t level = maxT; level >= 2; --level) {
        const uint32_t np = level == maxT ? numMid : nu[level + 1];
        DB_TRY(enqueueMark(level, parentsOf(level), np, s));
        uint32_t total = 0;
        DB_TRY(hipMemcpyAsync(&total, b + totalOff, 4, hipMemcpyDeviceToHost, s));
        DB_TRY(hipStreamSynchronize(s));
        n[level] = total;
        DB_TRY(enqueueMatchLevel(level, total, parentsOf(level), np, cfgBits, s));
        uint32_t c[4] = { 0, 0, 0, 0 };
        DB_TRY(hipMemcpyAsync(c, (uint32_t*)(b + countsOff) + 4 * level, 16, hipMemcpyDeviceToHost, s));
        DB_TRY(hipStreamSynchronize(s));
        nu[level] = c[1];
        matched[level] = c[2];
        if (level == 2)
            break;
    }
    uint32_t bad = 0;
    DB_TRY(hipMemcpyAsync(&bad, b + badOff, 4, hipMemcpyDeviceToHost, s));
    DB_TRY(hipStreamSynchronize(s));
    if (bad)
        return hipErrorInvalidValue;
    known = true;
    return hipSuccess;
}

hipError_t DeviceTranslation::enqueueKnown(uint32_t cfgBits, hipStream_t s)
{
    if (!known)
        return hipErrorInvalidValue;
    char* b = base();
    // The mid level's parents are restored from a device copy kept beside the arena's working copy
    // (the rewrites of the previous run changed the working copy; a device-to-device copy, so the
    // captured graph holds no copy from pageable host memory).
    if (numMid)
        DB_TRY(hipMemcpyAsync(b + midOff, midDev, (size_t)numMid * 64, hipMemcpyDeviceToDevice, s));
    for (uint32_t level = maxT; level >= 2; --level) {
        const uint32_t np = level == maxT ? numMid : nu[level + 1];
        DB_TRY(enqueueMark(level, parentsOf(level), np, s));
        DB_TRY(enqueueMatchLevel(level, n[level], parentsOf(level), np, cfgBits, s));
        if (level == 2)
            break;
    }
    return hipSuccess;
}

hipError_t DeviceTranslation::capture(uint32_t cfgBits, hipStream_t s)
{
    if (!known)
        return hipErrorInvalidValue;
    if (!midDev) {
        DB_TRY(hipMalloc((void**)&midDev, std::max<size_t>(64, (size_t)numMid * 64)));
        DB_TRY(hipMemcpy(midDev, midHost.data(), (size_t)numMid * 64, hipMemcpyHostToDevice));
    }
    DB_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    const hipError_t e = enqueueKnown(cfgBits, s);
    hipGraph_t g = nullptr;
    const hipError_t e2 = hipStreamEndCapture(s, &g);
    graph = g;
    DB_TRY(e);
    DB_TRY(e2);
    return hipGraphInstantiate(&graphExec, graph, nullptr, nullptr, 0);
}

hipError_t DeviceTranslation::replay(hipStream_t s)
{
    if (!graphExec)
        return hipErrorInvalidValue;
    return hipGraphLaunch(graphExec, s);
}

hipError_t DeviceTranslation::readResult(TDag& t, hipStream_t s) const
{
    const char* b = (const char*)arena;
    // Levels maxT + 1 (rewritten mid parents), maxT..3 (rewritten outputs), 2 (leaves).
    t.nodes[maxT + 1].resize(numMid);
    if (numMid)
        DB_TRY(hipMemcpyAsync(t.nodes[maxT + 1].data(), b + midOff, (size_t)numMid * 64, hipMemcpyDeviceToHost, s));
    for (uint32_t level = maxT; level >= 2; --level) {
        if (level == 2) {
            t.leaves.resize(nu[2]);
            if (nu[2])
                DB_TRY(hipMemcpyAsync(t.leaves.data(), b + outOff[2], (size_t)nu[2] * 8, hipMemcpyDeviceToHost, s));
            break;
        }
        t.nodes[level].resize(nu[level]);
        if (nu[level])
            DB_TRY(hipMemcpyAsync(t.nodes[level].data(), b + outOff[level], (size_t)nu[level] * 64, hipMemcpyDeviceToHost, s));
    }
    return hipStreamSynchronize(s);
}

DeviceTranslation::~DeviceTranslation()
{
    release();
    if (midDev)
        (void)hipFree(midDev);
}

} // namespace tdag
