// Batched form of the exhaustive matcher: several levels (their inputs already known) are
// matched together, with every level's parent search running in ONE persistent kernel that
// pulls work tiles from a device queue.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "matcher.hpp"

namespace tdag {

constexpr uint32_t kBatchMaxLevels = 4;
constexpr uint32_t kBatchMaxPhases = 32;
// Default share of blocks that start on the lighter levels' queue (one in kBatchSplitEvery).
constexpr uint32_t kBatchSplitEvery = 3;

// Work of one level in one phase: parents [r0, r1) as tiles [tileBase, tileBase + tiles) of the
// phase's queue; after the phase, children whose best key is below pruneKey leave the table
// (prune != 0).
struct BatchPhaseLevel {
    uint32_t r0, r1, tileBase, tiles;
    uint64_t pruneKey;
    uint32_t prune, pad;
};

// Device-resident description of one batched launch.
struct BatchDesc {
    uint32_t numLevels, numPhases, cfgBits;
    uint32_t grab; // consecutive tiles a block takes per queue pop (tuned)
    // Every splitEvery-th block (0: none) takes its tiles from the phase's second queue first:
    // tiles [splitTile, totalTiles) of the lighter levels, so those run beside the heaviest
    // level's tiles [0, splitTile) through the whole phase instead of after them.
    uint32_t splitEvery, pad;
    LevelView lv[kBatchMaxLevels];
    uint32_t totalTiles[kBatchMaxPhases];
    uint32_t splitTile[kBatchMaxPhases];
    BatchPhaseLevel ph[kBatchMaxPhases][kBatchMaxLevels];
};

// Queue heads, grid-barrier state and the error flag of one batched launch (reset by the
// enqueue, in-graph, with a fill kernel).
struct BatchSync {
    uint32_t queue[kBatchMaxPhases]; // heads of the heaviest level's tiles
    uint32_t queueB[kBatchMaxPhases]; // heads of the other levels' tiles, counted from splitTile
    uint32_t arrive, generation, error, pad;
};

struct BatchPlan {
    uint32_t count = 0;
    uint32_t job[kBatchMaxLevels]; // caller's job index of plans[i] (plans are ordered by level, descending)
    MatchPlan plans[kBatchMaxLevels];
    size_t scratchOff[kBatchMaxLevels];
    size_t descOff = 0, syncOff = 0, scratchBytes = 0;
    uint32_t numPhases = 0;
    uint32_t gridBlocks = 0; // persistent grid (all blocks resident)
};

// Plans a batched call over `count` levels (distinct, 2..5) with per-level schedules taken from
// firstBatch[i] / bloomBits[i] / childFrames[i] (0 = defaults; each array may be null). The phase
// p of the persistent kernel runs batch p of every level: ranks are popcount-descending, so each
// batch is a popcount bucket of parents.
hipError_t planBatch(uint32_t count, const uint32_t* levels, const uint64_t* ns, const uint32_t* firstBatch, const uint32_t* bloomBits, const uint32_t* childFrames,
    int device, BatchPlan* out);

// Writes the BatchDesc for scratch base `d_scratch` into host memory `desc` (splitEvery: see
// BatchDesc).
void fillBatchDesc(const BatchPlan& plan, void* d_scratch, uint32_t cfgBits, uint32_t grab, BatchDesc* desc, uint32_t splitEvery = kBatchSplitEvery);

// Enqueues prepare (all levels), the persistent search and finish (all levels) on `stream`. The
// BatchDesc must already be at d_scratch + plan.descOff. Array arguments are indexed by the
// caller's job index. Capturable into a hipGraph.
hipError_t enqueueBatch(const BatchPlan& plan, const uint64_t* const* d_masks, void* d_scratch, hipStream_t stream, uint64_t* const* d_encoded,
    uint64_t* const* d_mapping, uint32_t* const* d_unique, hipEvent_t evSearch0, hipEvent_t evSearch1, uint64_t spinTicks);

} // namespace tdag
