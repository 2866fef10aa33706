// The translation stage of constructTransformDAG with every step on the device, and its hipGraph
// (see device_build.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <vector>

#include "matcher.hpp"
#include "pipeline.hpp"

namespace tdag {

// Matcher schedule of one level (the context's tuning; negative variant / zeros: defaults).
struct LevelTuning {
    int32_t variant = -1;
    uint32_t firstBatch = 0, batchGrowth = 0, bloomBits = 0, childFrames = 0;
};

class DeviceTranslation {
public:
    ~DeviceTranslation();
    // Plans the stage for an SVDAG (device copy `dag`, host copy `host`) whose level maxT + 1
    // parents (TransformNodes after the hierarchical pass) are `mid`; allocates the arena, sized
    // with each level's node count as the bound of its referenced nodes. tuning: 6 entries by level.
    hipError_t init(const DeviceSvdag& dag, const Svdag& host, const std::vector<std::array<uint64_t, 8>>& mid, uint32_t maxLevel, const LevelTuning* tuning);
    // Runs the stage on `s`, reading each level's referenced count back before planning its
    // matcher; afterwards the per-level sizes are known. hipErrorInvalidValue: a child pointer or
    // flattened item outside its level.
    hipError_t runEager(uint32_t cfgBits, hipStream_t s);
    // The stage with the known sizes, enqueue only (no host synchronisation): capturable.
    hipError_t enqueueKnown(uint32_t cfgBits, hipStream_t s);
    // Captures enqueueKnown into a hipGraph (after runEager), and replays it.
    hipError_t capture(uint32_t cfgBits, hipStream_t s);
    hipError_t replay(hipStream_t s);
    // Copies the translated levels (maxT + 1 down to the leaves) into t; synchronises s.
    hipError_t readResult(TDag& t, hipStream_t s) const;
    void release();

    uint32_t maxT = 0;
    uint32_t n[6] = {}, nu[6] = {}, matched[6] = {}; // per level: referenced items, unique, matched

private:
    char* base() { return (char*)arena; }
    MatchPlan planFor(uint32_t level, uint64_t count) const;
    uint64_t* parentsOf(uint32_t level);
    hipError_t enqueueMark(uint32_t level, const uint64_t* parents, uint32_t np, hipStream_t s);
    hipError_t enqueueMatchLevel(uint32_t level, uint32_t count, uint64_t* parents, uint32_t np, uint32_t cfgBits, hipStream_t s);

    DeviceSvdag d {};
    LevelTuning tune[6] {};
    uint32_t numItems[6] = {};
    uint32_t numMid = 0;
    std::vector<std::array<uint64_t, 8>> midHost;
    uint64_t* midDev = nullptr; // pristine device copy of the mid parents (graph replays restore from it)
    void* arena = nullptr;
    size_t arenaBytes = 0;
    size_t outOff[6] = {}, midOff = 0, refOff = 0, posOff = 0, totalOff = 0, idxOff = 0, maskOff = 0, mapOff = 0, uniqOff = 0, scanOff = 0, scratchOff = 0,
           countsOff = 0, badOff = 0, scratchCap = 0;
    bool known = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graphExec = nullptr;
};

} // namespace tdag
