// Batched form of the exhaustive matcher on gfx950: the parent searches of several levels run in
// one persistent kernel fed by a device work queue.
//
// The grid is sized so that every block is resident (blocks per CU from the occupancy API times
// the CU count) and loops: it pops `grab` consecutive tile indices from the current phase's queue
// (one atomic per block), runs the tiles, and pops again until the phase is drained. Phase p holds batch p of every
// level's rank-batched search (ranks are popcount-descending, so a batch is a popcount bucket of
// parents; bucket widths come from the per-level schedule the autotuner picks). Between phases a
// grid barrier separates the search from the table pruning and filter rebuild of every level,
// done by all blocks together. Tiles are ordered heaviest level first inside a phase.
//
// Tile shapes (four waves per block):
//   level 5: one family of (direction) combos of one parent on 8 of the 32 planes (8 tiles per
//            parent when the table holds the children in all six permutation frames, else 24)
//   level 4: one family on all 16 planes                                       (2 or 6 per parent)
//   level 3: four wave items, each a family sharing passes 1 and 2 on all 8 planes (2 or 6 per parent)
//   level 2: four parents, one per wave (each with its work list in its LDS slice)
// Levels 3..5 run the shape pre-filter of the Bloom lookups (the children's extent triples in
// LDS) wherever the per-level kernels do: when the parent searches permutation code 0 alone.
// Results are bit-identical to the per-level launches: the search computes the same exhaustive
// minimum, whatever the schedule.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "batched.hpp"
#include "common.hpp"
#include "primitives.hpp"
#include "search_device.hpp"

namespace tdag {

namespace {

using namespace dev;

constexpr int kBatchThreads = 256;
constexpr int kBatchWaves = kBatchThreads / 64;

template <int L>
constexpr size_t rowsBytes()
{
    return Geo<L>::Rows * sizeof(typename Geo<L>::Row); // the layout of the unit's first-pass axis
}

constexpr size_t alignUp16(size_t v) { return (v + 15) / 16 * 16; }

// LDS of a block tile of level 4 or 5 (byte offsets): the rows of the family's first-pass axis,
// the four waves' planes, then the shape pre-filter's pre tables, n1 masks and extent bitmap.
template <int L>
struct BlockTileLds {
    static constexpr size_t planes = alignUp16(rowsBytes<L>());
    static constexpr size_t pre = planes + (size_t)kBatchWaves * Geo<L>::P * Geo<L>::Slab * 4;
    static constexpr size_t n1 = alignUp16(pre + (size_t)shapePreBytes<L>());
    static constexpr size_t bits = n1 + (size_t)shapeN1Words<L>(kBatchWaves) * 4;
    static constexpr size_t bytes = bits + (size_t)shapeLdsWords<L>() * 4;
};

// LDS slice of one wave of a level-3 wave item: rows, two plane buffers, then the shape
// pre-filter's tables for that wave alone.
struct Wave3Lds {
    static constexpr size_t planes0 = alignUp16(rowsBytes<3>());
    static constexpr size_t planes1 = planes0 + (size_t)Geo<3>::P * Geo<3>::Slab * 4;
    static constexpr size_t pre = planes1 + (size_t)Geo<3>::P * Geo<3>::Slab * 4;
    static constexpr size_t n1 = alignUp16(pre + (size_t)shapePreBytes<3>());
    static constexpr size_t bits = n1 + (size_t)shapeN1Words<3>(1) * 4;
    static constexpr size_t bytes = alignUp16(bits + (size_t)shapeLdsWords<3>() * 4);
};
constexpr size_t kL3WaveBytes = Wave3Lds::bytes;

constexpr size_t kL3Bytes = kBatchWaves * kL3WaveBytes;
constexpr size_t kBatchLds = std::max(BlockTileLds<5>::bytes, std::max(BlockTileLds<4>::bytes, kL3Bytes));
static_assert(kBatchLds <= kLdsBytes, "batched tile LDS");

constexpr uint32_t kL5Chunks = Geo<5>::MaxWaves / kBatchWaves; // plane chunks of a level-5 tile
constexpr uint32_t kNoTile = 0xFFFFFFFFu; // a queue pop that found both queues drained
static_assert(Geo<4>::MaxWaves == kBatchWaves, "a level-4 tile covers all planes");
static_assert(Geo<3>::MaxWaves == 1, "a level-3 wave item covers all planes");

// Search configuration bits of level view v (adds kCfgChildFrames where its table uses them).
__device__ __forceinline__ uint32_t levelCfgBits(const LevelView& v, uint32_t cfgBits)
{
    return v.childFrames == 6 ? (cfgBits | kCfgChildFrames) : (cfgBits & ~kCfgChildFrames);
}

uint32_t tilesFor(const MatchPlan& p, uint32_t parents)
{
    const uint32_t cfg = p.childFrames == 6 ? kCfgChildFrames : 0u;
    switch (p.level) {
    case 5:
        return parents * searchUnits(kUnitFamily, cfg) * kL5Chunks;
    case 4:
        return parents * searchUnits(kUnitFamily, cfg);
    case 3:
        return (parents * searchUnits(kUnitShared, cfg) + kBatchWaves - 1) / kBatchWaves;
    default:
        return (parents + kBatchWaves - 1) / kBatchWaves;
    }
}

// Whether a level's search runs the shape pre-filter: the parent searches permutation code 0
// alone (the table holds the children in all six frames, or permutations are off).
__device__ __forceinline__ bool shapeOn(const LevelView& v, uint32_t cfgBits)
{
    return v.shape != nullptr && !((cfgBits & kCfgAxis) != 0 && (cfgBits & kCfgChildFrames) == 0);
}

// Block tile of levels 4 and 5: all four waves search family `unit` of parent r on plane chunk
// `chunk`.
template <int L>
__device__ __forceinline__ void blockTile(const LevelView& v, uint32_t r, uint32_t unit, uint32_t chunk, unsigned char* s_mem, uint32_t* s_occ, uint32_t* qMeta, uint32_t* qHash,
    uint32_t cfgBits)
{
    using G = Geo<L>;
    using Lds = BlockTileLds<L>;
    if (!(cfgBits & kCfgAxis) && (unit >> 1) != 0)
        return; // only permutation code 0 (p0 == 0) is allowed
    auto* s_rows = reinterpret_cast<typename G::Row*>(s_mem);
    uint32_t* planes = reinterpret_cast<uint32_t*>(s_mem + Lds::planes);
    Occupancy occ;
    if (!loadRowsOcc<L, kBatchThreads>(v.sortedMasks, v.n, r, unit >> 1, s_rows, s_occ, occ))
        return;
    ProbeCounts cnt;
    const int wave = (int)(threadX() >> 6);
    if (shapeOn(v, cfgBits)) {
        auto* s_pre = reinterpret_cast<typename G::Row*>(s_mem + Lds::pre);
        uint32_t* s_bits = reinterpret_cast<uint32_t*>(s_mem + Lds::bits);
        buildShapeTables<L>(s_rows, v.shape, s_bits, s_pre, (int)threadX(), kBatchThreads);
        __syncthreads();
        ShapeLds sh;
        sh.pre = s_pre;
        sh.n1 = reinterpret_cast<uint32_t*>(s_mem + Lds::n1);
        sh.bits = s_bits;
        searchUnit<L, kBatchWaves, kUnitFamily, true>(r, unit, chunk, wave, s_rows, occ, planes, planes, qMeta, qHash, v.sortedMasks, v.n, v.tabKeys, v.logCap, v.bloom,
            v.logBloom, v.best, cfgBits, cnt, sh);
    } else {
        searchUnit<L, kBatchWaves, kUnitFamily>(r, unit, chunk, wave, s_rows, occ, planes, planes, qMeta, qHash, v.sortedMasks, v.n, v.tabKeys, v.logCap, v.bloom, v.logBloom,
            v.best, cfgBits, cnt);
    }
    flushCounts(cnt, v.counters, cfgBits);
}

// Wave item of level 3: the calling wave alone searches family `unit` of parent r.
__device__ __forceinline__ void waveItem3(const LevelView& v, uint32_t r, uint32_t unit, unsigned char* s_mem, uint32_t* qMeta, uint32_t* qHash, uint32_t cfgBits)
{
    using G = Geo<3>;
    if (!(cfgBits & kCfgAxis) && (unit >> 1) != 0)
        return;
    unsigned char* base = s_mem + (threadX() >> 6) * kL3WaveBytes;
    auto* s_rows = reinterpret_cast<G::Row*>(base);
    uint32_t* planes0 = reinterpret_cast<uint32_t*>(base + Wave3Lds::planes0);
    uint32_t* planes1 = reinterpret_cast<uint32_t*>(base + Wave3Lds::planes1);
    Occupancy occ;
    if (!loadRowsWave<3>(v.sortedMasks, v.n, r, unit >> 1, s_rows, occ))
        return;
    ProbeCounts cnt;
    if (shapeOn(v, cfgBits)) {
        auto* s_pre = reinterpret_cast<G::Row*>(base + Wave3Lds::pre);
        uint32_t* s_bits = reinterpret_cast<uint32_t*>(base + Wave3Lds::bits);
        buildShapeTables<3>(s_rows, v.shape, s_bits, s_pre, (int)(threadX() & 63u), 64);
        waveSync();
        ShapeLds sh;
        sh.pre = s_pre;
        sh.n1 = reinterpret_cast<uint32_t*>(base + Wave3Lds::n1);
        sh.bits = s_bits;
        searchUnit<3, 1, kUnitShared, true>(r, unit, 0, 0, s_rows, occ, planes0, planes1, qMeta, qHash, v.sortedMasks, v.n, v.tabKeys, v.logCap, v.bloom, v.logBloom, v.best,
            cfgBits, cnt, sh);
    } else {
        searchUnit<3, 1, kUnitShared>(r, unit, 0, 0, s_rows, occ, planes0, planes1, qMeta, qHash, v.sortedMasks, v.n, v.tabKeys, v.logCap, v.bloom, v.logBloom, v.best, cfgBits,
            cnt);
    }
    flushCounts(cnt, v.counters, cfgBits);
}

// Wave item of level 2: the calling wave searches parent r with its work list in its LDS slice.
static_assert(kWalkQueue2 * (8 + 2) <= kL3WaveBytes, "level-2 work list fits a wave's LDS slice");
__device__ __forceinline__ void waveItem2(const LevelView& v, uint32_t r, unsigned char* s_mem, uint32_t cfgBits)
{
    unsigned char* base = s_mem + (threadX() >> 6) * kL3WaveBytes;
    uint64_t* qKey = reinterpret_cast<uint64_t*>(base);
    uint16_t* qMeta = reinterpret_cast<uint16_t*>(base + kWalkQueue2 * 8);
    ProbeCounts cnt;
    search2ParentQueued(r, v.sortedMasks, v.tabKeys, v.tabVals, v.logCap, v.bloom, v.logBloom, v.best, cfgBits, qKey, qMeta, cnt);
    flushCounts(cnt, v.counters, cfgBits);
}

// Grid-wide barrier of the persistent kernel (all blocks are resident). Thread 0 of each block
// arrives with an agent-scope release and, once the generation has moved, takes one agent-scope
// acquire, so table, filter and best[] writes made before the barrier are visible after it on
// every XCD. The wait itself polls with relaxed loads: an agent-scope acquire invalidates the
// caches of the waiting CU's XCD, and blocks polling that way would keep evicting the tables
// the blocks still searching on that XCD read. A wait longer than `spinTicks` wall-clock ticks
// sets sync->error and gives up (the results are then flagged as invalid by the host) instead
// of hanging; once the flag is set no block waits again.
__device__ __forceinline__ void gridBarrier(BatchSync* sync, uint64_t spinTicks)
{
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t gen = __hip_atomic_load(&sync->generation, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __threadfence();
        const uint32_t arrived = __hip_atomic_fetch_add(&sync->arrive, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) + 1u;
        if (arrived == gridDim.x) {
            __hip_atomic_store(&sync->arrive, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_fetch_add(&sync->generation, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            const uint64_t t0 = (uint64_t)wall_clock64();
            while (__hip_atomic_load(&sync->generation, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gen) {
                __builtin_amdgcn_s_sleep(4);
                if (__hip_atomic_load(&sync->error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u)
                    break; // another block already gave up: do not wait again
                if ((uint64_t)wall_clock64() - t0 > spinTicks) {
                    __hip_atomic_store(&sync->error, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    break;
                }
            }
        }
        __threadfence();
    }
    __syncthreads();
}

// Three blocks per CU: a budget of 168 VGPRs holds every tile path without spills (a fourth
// block would fit the LDS, but the 128-VGPR budget it leaves spills and measured no faster).
__global__ __launch_bounds__(kBatchThreads, 3) void kBatched(const BatchDesc* __restrict__ desc, BatchSync* __restrict__ sync, uint64_t spinTicks)
{
    __shared__ __align__(16) unsigned char s_mem[kBatchLds];
    __shared__ uint32_t s_tile;
    __shared__ uint32_t s_occ[3 * kBatchWaves];
    __shared__ uint32_t s_qMeta[kBatchWaves][kProbeQueue];
    __shared__ uint32_t s_qHash[kBatchWaves][kProbeQueue];
    const uint32_t numLevels = desc->numLevels;
    const uint32_t numPhases = desc->numPhases;
    const uint32_t cfgBits = desc->cfgBits;
    const uint32_t wave = threadIdx.x >> 6;
    const uint64_t gtid = (uint64_t)blockIdx.x * kBatchThreads + threadIdx.x;
    const uint64_t gstride = (uint64_t)gridDim.x * kBatchThreads;
    for (uint32_t p = 0; p < numPhases; ++p) {
        const uint32_t total = desc->totalTiles[p];
        const uint32_t split = desc->splitTile[p];
        const uint32_t grab = desc->grab;
        // Queue A holds tiles [0, split) (the heaviest level), queue B tiles [split, total). A
        // block takes from its preferred queue while it has tiles, then from the other one; the
        // phase ends for the block when both are drained.
        const bool preferB = split < total && desc->splitEvery != 0u && blockIdx.x % desc->splitEvery == desc->splitEvery - 1u;
        bool doneA = split == 0u, doneB = split >= total;
        for (;;) {
            if (threadIdx.x == 0) {
                uint32_t t = kNoTile;
                for (int attempt = 0; attempt < 2 && t == kNoTile; ++attempt) {
                    if ((attempt == 0) == preferB) {
                        if (!doneB) {
                            const uint32_t x = split + atomicAdd(&sync->queueB[p], grab);
                            if (x < total)
                                t = x;
                            else
                                doneB = true;
                        }
                    } else if (!doneA) {
                        const uint32_t x = atomicAdd(&sync->queue[p], grab);
                        if (x < split)
                            t = x;
                        else
                            doneA = true;
                    }
                }
                s_tile = t;
            }
            __syncthreads();
            // Uniform over the block: in a scalar register, the tile's level view, parent and
            // unit stay scalar too (their pointers in SGPRs, the descriptor read with scalar loads).
            const uint32_t t0 = __builtin_amdgcn_readfirstlane(s_tile);
            __syncthreads();
            if (t0 == kNoTile)
                break;
            const uint32_t tEnd = min(t0 < split ? split : total, t0 + grab);
            for (uint32_t t = t0; t < tEnd; ++t) {
                if (t != t0)
                    __syncthreads(); // the previous tile's waves are done with the LDS
                uint32_t li = 0;
                while (li + 1 < numLevels && t >= desc->ph[p][li + 1].tileBase)
                    ++li;
                const BatchPhaseLevel& pl = desc->ph[p][li];
                const LevelView& v = desc->lv[li];
                const uint32_t k = t - pl.tileBase;
                const uint32_t lvCfg = levelCfgBits(v, cfgBits);
                switch (v.level) {
                case 5: {
                    const uint32_t per = searchUnits(kUnitFamily, lvCfg) * kL5Chunks;
                    const uint32_t m = k % per;
                    blockTile<5>(v, pl.r0 + k / per, m / kL5Chunks, m % kL5Chunks, s_mem, s_occ, s_qMeta[wave], s_qHash[wave], lvCfg);
                    break;
                }
                case 4: {
                    const uint32_t per = searchUnits(kUnitFamily, lvCfg);
                    blockTile<4>(v, pl.r0 + k / per, k % per, 0, s_mem, s_occ, s_qMeta[wave], s_qHash[wave], lvCfg);
                    break;
                }
                case 3: {
                    const uint32_t per = searchUnits(kUnitShared, lvCfg);
                    const uint32_t item = k * kBatchWaves + wave;
                    if (item < (pl.r1 - pl.r0) * per)
                        waveItem3(v, pl.r0 + item / per, item % per, s_mem, s_qMeta[wave], s_qHash[wave], lvCfg);
                    break;
                }
                default: {
                    const uint32_t r = pl.r0 + k * kBatchWaves + wave;
                    if (r < pl.r1)
                        waveItem2(v, r, s_mem, cfgBits);
                    break;
                }
                }
            }
        }
        if (p + 1 == numPhases)
            break;
        // Between phases: prune every level that has further batches (rank-batched search).
        gridBarrier(sync, spinTicks);
        for (uint32_t li = 0; li < numLevels; ++li)
            if (desc->ph[p][li].prune) {
                const LevelView& v = desc->lv[li];
                const uint64_t words = 1ull << v.logBloom;
                for (uint64_t i = gtid; i < words; i += gstride)
                    v.bloom[i] = 0u;
            }
        gridBarrier(sync, spinTicks);
        for (uint32_t li = 0; li < numLevels; ++li)
            if (desc->ph[p][li].prune) {
                const LevelView& v = desc->lv[li];
                const uint64_t cap = 1ull << v.logCap;
                if (v.level == 2)
                    pruneSplit(v.tabKeys, v.tabVals, cap, v.best, desc->ph[p][li].pruneKey, v.bloom, v.logBloom, gtid, gstride);
                else
                    prunePacked(v.tabKeys, cap, v.best, desc->ph[p][li].pruneKey, v.bloom, v.logBloom, gtid, gstride);
            }
        gridBarrier(sync, spinTicks);
    }
}

inline size_t alignUp(size_t v, size_t a) { return (v + a - 1) / a * a; }

} // namespace

hipError_t planBatch(uint32_t count, const uint32_t* levels, const uint64_t* ns, const uint32_t* firstBatch, const uint32_t* bloomBits, const uint32_t* childFrames,
    int device, BatchPlan* out)
{
    if (count == 0 || count > kBatchMaxLevels)
        return hipErrorInvalidValue;
    BatchPlan b;
    b.count = count;
    // Heaviest level first: its tiles lead every phase's queue.
    std::vector<uint32_t> idx(count);
    for (uint32_t i = 0; i < count; ++i)
        idx[i] = i;
    std::sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t c) { return levels[a] > levels[c]; });
    for (uint32_t i = 0; i + 1 < count; ++i)
        if (levels[idx[i]] == levels[idx[i + 1]])
            return hipErrorInvalidValue; // one job per level
    size_t off = 0;
    for (uint32_t i = 0; i < count; ++i) {
        const uint32_t j = idx[i];
        if (levels[j] < 2 || levels[j] > 5 || ns[j] >= (levels[j] == 2 ? (1ull << 28) : (1ull << 26)))
            return hipErrorInvalidValue;
        b.job[i] = j;
        b.plans[i] = planMatch(levels[j], ns[j], 0, firstBatch ? firstBatch[j] : 0, 2, bloomBits ? bloomBits[j] : 0, childFrames ? childFrames[j] : 0);
        b.scratchOff[i] = off;
        off += alignUp(b.plans[i].scratchBytes, 256);
    }
    b.descOff = off;
    off += alignUp(sizeof(BatchDesc), 256);
    b.syncOff = off;
    off += alignUp(sizeof(BatchSync), 256);
    b.scratchBytes = off;
    uint32_t phases = 0;
    for (uint32_t i = 0; i < count; ++i)
        phases = std::max<uint32_t>(phases, (uint32_t)batchEnds(b.plans[i]).size());
    if (phases > kBatchMaxPhases)
        return hipErrorInvalidValue;
    b.numPhases = std::max<uint32_t>(phases, 1u);
    int blocksPerCu = 0, cus = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocksPerCu, kBatched, kBatchThreads, 0);
    if (e != hipSuccess)
        return e;
    e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
    if (e != hipSuccess)
        return e;
    if (blocksPerCu < 1 || cus < 1)
        return hipErrorInvalidConfiguration;
    b.gridBlocks = (uint32_t)(blocksPerCu * cus);
    *out = b;
    return hipSuccess;
}

void fillBatchDesc(const BatchPlan& b, void* d_scratch, uint32_t cfgBits, uint32_t grab, BatchDesc* desc, uint32_t splitEvery)
{
    *desc = BatchDesc {};
    desc->grab = std::max<uint32_t>(1u, grab);
    desc->splitEvery = splitEvery;
    desc->numLevels = b.count;
    desc->numPhases = b.numPhases;
    desc->cfgBits = cfgBits;
    std::vector<std::vector<uint32_t>> ends(b.count);
    for (uint32_t i = 0; i < b.count; ++i) {
        desc->lv[i] = levelView(b.plans[i], (char*)d_scratch + b.scratchOff[i]);
        ends[i] = batchEnds(b.plans[i]);
    }
    for (uint32_t p = 0; p < b.numPhases; ++p) {
        uint32_t base = 0;
        for (uint32_t i = 0; i < b.count; ++i) {
            BatchPhaseLevel& pl = desc->ph[p][i];
            pl.tileBase = base;
            if (p < ends[i].size()) {
                pl.r0 = p == 0 ? 0u : ends[i][p - 1];
                pl.r1 = ends[i][p];
                pl.tiles = tilesFor(b.plans[i], pl.r1 - pl.r0);
                pl.prune = pl.r1 < (uint32_t)b.plans[i].n ? 1u : 0u;
                pl.pruneKey = pruneKey(b.plans[i].level, pl.r1);
            }
            base += pl.tiles;
        }
        desc->totalTiles[p] = base;
        desc->splitTile[p] = b.count > 1 ? desc->ph[p][1].tileBase : base;
    }
}

hipError_t enqueueBatch(const BatchPlan& b, const uint64_t* const* d_masks, void* d_scratch, hipStream_t stream, uint64_t* const* d_encoded,
    uint64_t* const* d_mapping, uint32_t* const* d_unique, hipEvent_t evSearch0, hipEvent_t evSearch1, uint64_t spinTicks)
{
    char* base = (char*)d_scratch;
    for (uint32_t i = 0; i < b.count; ++i) {
        const hipError_t e = enqueuePrepare(b.plans[i], d_masks[b.job[i]], base + b.scratchOff[i], stream);
        if (e != hipSuccess)
            return e;
    }
    hipError_t e = fillU32((uint32_t*)(base + b.syncOff), sizeof(BatchSync) / 4, 0u, stream);
    if (e != hipSuccess)
        return e;
    if (evSearch0 && (e = hipEventRecord(evSearch0, stream)) != hipSuccess)
        return e;
    bool any = false;
    for (uint32_t i = 0; i < b.count; ++i)
        any = any || b.plans[i].n > 0;
    if (any) {
        hipLaunchKernelGGL(kBatched, dim3(b.gridBlocks), dim3(kBatchThreads), 0, stream, (const BatchDesc*)(base + b.descOff), (BatchSync*)(base + b.syncOff), spinTicks);
        if ((e = hipGetLastError()) != hipSuccess)
            return e;
    }
    if (evSearch1 && (e = hipEventRecord(evSearch1, stream)) != hipSuccess)
        return e;
    for (uint32_t i = 0; i < b.count; ++i) {
        const uint32_t j = b.job[i];
        e = enqueueFinish(b.plans[i], base + b.scratchOff[i], stream, d_encoded ? d_encoded[j] : nullptr, d_mapping ? d_mapping[j] : nullptr,
            d_unique ? d_unique[j] : nullptr);
        if (e != hipSuccess)
            return e;
    }
    return hipSuccess;
}

} // namespace tdag
