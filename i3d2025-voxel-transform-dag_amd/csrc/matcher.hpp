// Host-side plan and launcher of the exhaustive matcher (one level per call).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <algorithm>
#include <vector>

namespace tdag {

// TransformDAGConfig flags as one word (transform_dag.h:157-160).
constexpr uint32_t kCfgSymmetry = 1u;
constexpr uint32_t kCfgAxis = 2u;
constexpr uint32_t kCfgTranslation = 4u;
// Not a TransformDAGConfig flag: collect per-stage probe counters (diagnostics).
constexpr uint32_t kCfgCount = 8u;
// Not a TransformDAGConfig flag: the level-3..5 child table holds every child under all six
// axis permutations (48 entries per child), so the parent search runs the identity permutation
// only (8 scan directions instead of 48 combos).
constexpr uint32_t kCfgChildFrames = 16u;
// Device counters of one call: matched, unique (u32), probes, filter passes, key hits,
// candidates, verifications, verified.
constexpr int kNumCounters = 8;

// Search-kernel variants for levels 3..5 (the autotuner's tile shapes). variant % 3: a block
// runs one (permutation, direction) combo (0), a family of 8 combos (1), or a family of 8 combos
// that shares passes 1 and 2 between the two pass-3 directions (2). (variant / 3) % 3: log2 of
// the split of the R planes of a unit over blocks (0, 1, 2). variant / 9: block order (0: parent
// fastest, 1: all blocks of a parent together on one XCD). Variants whose blocks exceed the LDS
// are invalid.
constexpr uint32_t kNumVariants = 18;

bool variantValid(uint32_t level, uint32_t variant);
uint32_t defaultVariant(uint32_t level);
// Bloom filter bits per table entry used when a plan asks for 0.
uint32_t defaultFilterBits(uint32_t level);

// Device scratch layout of one matcher call.
struct MatchPlan {
    uint32_t level, R, R3, W, logCap, logBloom, variant;
    uint32_t firstBatch, batchGrowth; // rank-batched search schedule
    uint32_t bloomBitsPerEntry; // Bloom filter size per table entry (bits)
    uint32_t childFrames; // levels 3..5: child-table permutation frames, 1 or 6 (kCfgChildFrames)
    uint64_t n;
    size_t offKeys, offOrder, offSortScratch, offSortedMasks, offBest;
    size_t offTabKeys, offTabVals, offTickets, offBloom, offShape, offParentTp, offIsUnique, offNewIdx, offScanScratch, offCounters;
    size_t scratchBytes;
};

// firstBatch = 0 / batchGrowth = 0 select the defaults (max(64, n/64) parents, doubling).
// bloomBitsPerEntry = 0 selects defaultFilterBits(level); childFrames = 0 selects
// defaultChildFrames(level) (levels 3..5; 1 or 6).
MatchPlan planMatch(uint32_t level, uint64_t n, uint32_t variant, uint32_t firstBatch = 0, uint32_t batchGrowth = 0, uint32_t bloomBitsPerEntry = 0,
    uint32_t childFrames = 0);
uint32_t defaultChildFrames(uint32_t level);
// Configuration bits of the search for `plan` (adds kCfgChildFrames where the plan uses it).
uint32_t searchCfgBits(const MatchPlan& plan, uint32_t cfgBits);

// Enqueues the whole matcher on `stream`. Outputs may be null. The device counters at
// scratch + offCounters hold {num_matched (u64), num_unique (u32)} when the stream completes.
// Equivalent to enqueuePrepare, enqueueSearch and enqueueFinish in order (with the events
// around the search).
hipError_t enqueueMatch(const MatchPlan& plan, const uint64_t* d_masks, uint32_t cfgBits, void* d_scratch, hipStream_t stream,
    uint64_t* d_encoded, uint64_t* d_mapping, uint32_t* d_unique, hipEvent_t evSearch0, hipEvent_t evSearch1);

// Stages of one matcher call. Prepare: rank order, sorted masks, best[] seeds, child table and
// filter. Search: the rank-batched parent search with table pruning between batches. Finish:
// parents, unique flags and the mapping.
hipError_t enqueuePrepare(const MatchPlan& plan, const uint64_t* d_masks, void* d_scratch, hipStream_t stream);
// Prepare in two parts: rank order, sorted masks and best[] seeds; then the child table. With
// childFlags (indexed by rank) only ranks whose flag is non-zero enter the table.
hipError_t enqueuePrepareOrder(const MatchPlan& plan, const uint64_t* d_masks, void* d_scratch, hipStream_t stream);
hipError_t enqueueBuildTable(const MatchPlan& plan, void* d_scratch, hipStream_t stream, const uint32_t* childFlags);
hipError_t enqueueSearch(const MatchPlan& plan, uint32_t cfgBits, void* d_scratch, hipStream_t stream);
hipError_t enqueueFinish(const MatchPlan& plan, void* d_scratch, hipStream_t stream, uint64_t* d_encoded, uint64_t* d_mapping, uint32_t* d_unique);

// Device pointers of one matcher call's search state inside its scratch.
struct LevelView {
    uint32_t level, n;
    uint64_t* sortedMasks; // rank-ordered masks; levels >= 3 followed by the two transposed layouts
    uint64_t* best;
    uint64_t* tabKeys; // level 2: keys; levels >= 3: packed (hash << 32 | rank << 6 | perm << 3 | corner)
    uint32_t* tabVals; // level 2 only
    uint32_t* tickets; // levels >= 3: entries started per home slot while the table is built
    uint32_t* bloom;
    uint32_t* shape; // levels >= 3: the children's extent triples (spans e = extent - 1), bit shapeBit(e0) of word shapeWord(e1, e2)
    unsigned long long* counters;
    uint32_t logCap, logBloom;
    uint32_t childFrames;
};
LevelView levelView(const MatchPlan& plan, void* d_scratch);

// End ranks (exclusive) of the rank-batched search's batches.
std::vector<uint32_t> batchEnds(const MatchPlan& plan);
// Smallest key a parent of rank >= r1 can produce at `level`: children whose best key is below
// it are final once parents [0, r1) have been searched.
uint64_t pruneKey(uint32_t level, uint32_t r1);

} // namespace tdag
