// Host-side data model and pipeline stages around the matcher (C++; compiled into libtdag.so).
//
//   Svdag  mirrors voxcom::EditStructure<void, uint32_t> after toDAG (structure.h:159-211)
//   TDag   mirrors voxcom::EditStructure<void, TransformPointer> (transform_dag.h:161-162)
#pragma once
#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <string>
#include <vector>

#include "tdag.h"

namespace tdag {

constexpr uint32_t kEmptyChild = 0xFFFFFFFFu;

struct Svdag {
    uint32_t resolution = 0;
    uint32_t rootLevel = 0;
    int32_t structureType = 1; // voxcom::StructureType::DAG
    // nodes[level] holds 8 child indices per node (kEmptyChild for none); levels 0..2 empty.
    std::vector<std::vector<std::array<uint32_t, 8>>> nodes;
    // Leaves (level 2): 4x4x4 masks, bit = morton(x, y, z) (structure.h:116-125).
    std::vector<uint64_t> leaves;
};

struct TDag {
    uint32_t resolution = 0;
    uint32_t rootLevel = 0;
    int32_t structureType = 0; // StructureType::Tree, as the reference's default-constructed output
    std::vector<std::vector<std::array<uint64_t, 8>>> nodes; // TransformPointer bit patterns
    std::vector<uint64_t> leaves;
};

// The .tdag16 compact encoding (TransformDAG16, transform_dag_encoding.h:138-153): 16-bit node
// words per level, leaves, per-level pointer tables and Huffman code words (LSB first).
struct HuffCode {
    uint64_t code;
    uint32_t bits;
};
struct Tdag16 {
    tdag_config cfg {};
    bool tables = false, huffman = false;
    uint32_t resolution = 0, rootLevel = 0;
    std::vector<std::vector<uint16_t>> levels;
    std::vector<uint64_t> subgrids;
    std::vector<std::vector<uint64_t>> ptrTables;
    std::vector<std::vector<HuffCode>> codes; // empty when the Huffman code is off
};
std::vector<uint32_t> huffmanCodeLengths(const std::vector<uint64_t>& freq, uint32_t maxLen);
std::vector<HuffCode> huffmanCodeWords(const std::vector<uint32_t>& len, bool lsbFirst);
Tdag16 encodeTdag16(const TDag& in, const tdag_config& cfg, bool tables, bool huffman);
std::vector<uint8_t> serializeTdag16(const Tdag16& d);
bool parseTdag16(const uint8_t* bytes, size_t len, Tdag16& out, std::string& err);
tdag_tdag16* newTdag16Handle(Tdag16&& e);

// Per-level record of one constructTransformDAG run.
struct LevelRecord {
    uint32_t level = 0;
    uint64_t items = 0; // matcher input size (referenced nodes)
    uint64_t unique = 0;
    uint64_t matched = 0; // items whose parent is another item
    float msMatch = 0.f; // device time of the matcher call
    bool learned = false; // matched with the learned proposal
    uint64_t proposalHits = 0, fallback = 0;
    std::vector<uint64_t> masks; // the matcher's input (kept when requested)
};

bool generateSceneDense(uint32_t kind, uint32_t R, uint64_t seed, uint8_t* bits, int threads);

// EditStructure(VoxelGrid) + toDAG, built bottom-up in Morton order.
Svdag svdagFromDense(const uint8_t* bits, uint32_t R, int threads);

// The same SVDAG built on the device from a dense grid in device memory (R^3 / 8 bytes), work
// enqueued on `stream` in a workspace (scratch, scratchBytes) that is grown as needed and kept
// by the caller; returns after the result has been copied into `out`.
hipError_t svdagFromDenseDevice(const uint8_t* d_bits, uint32_t R, hipStream_t stream, void*& scratch, size_t& scratchBytes, Svdag& out);

// createLargeSubGrid for level-`level` nodes (level 2 = leaves): linear masks, W words each.
void flattenNodes(const Svdag& s, uint32_t level, const uint32_t* nodes, uint64_t n, uint64_t* out, int threads);

// Device copy of an SVDAG's leaves and node levels 3..maxLevel (maxLevel <= 5) for flattening.
struct DeviceSvdag {
    const uint64_t* leaves = nullptr;
    const uint32_t* nodes[6] = { nullptr, nullptr, nullptr, nullptr, nullptr, nullptr }; // 8 children per node
    uint32_t count[6] = { 0, 0, 0, 0, 0, 0 }; // leaves at [2], nodes at [3..5]
    uint32_t maxLevel = 0;
};
hipError_t uploadSvdag(const Svdag& s, uint32_t maxLevel, DeviceSvdag& d);
void freeDeviceSvdag(DeviceSvdag& d);

// createLargeSubGrid on the device for the level-`level` nodes d_items[0..n) into d_out (n * W
// words); d_temp holds flattenTempBytes(level, n) bytes. Enqueued on `stream`. An item index
// outside the level reads as empty and sets *d_badItem (when given).
size_t flattenTempBytes(uint32_t level, uint64_t n);
hipError_t flattenDevice(const DeviceSvdag& d, uint32_t level, const uint32_t* d_items, uint64_t n, uint64_t* d_out, uint64_t* d_temp, hipStream_t stream,
    uint32_t* d_badItem = nullptr);

// constructTransformDAGHierarchical with midLevel; fills outMidInv (transform_dag_hierarchical.cpp:304-523).
// With staticOffsets, groups are keyed as constructStaticTransformDAGHierarchical keys them
// (transform_dag_hierarchical.cpp:531-743: children by output word offset), which changes the
// output order of levels >= 4 but not the set of kept nodes; the result stays in node-index form.
TDag hierarchicalDedup(const Svdag& in, uint32_t midLevel, std::vector<uint32_t>& outMidInv, const tdag_config& cfg, int threads,
    bool staticOffsets = false);

// Transform tables of the hierarchical pass (state = mirror bits | permutation code << 3, the
// transform-ID numbering) and its 256-entry canonical 2x2x2 leaf table, for the device form.
struct HierTables {
    uint8_t compose[48][48]; // child state after applying transform id to state
    uint8_t childPos[48][8]; // where child i lands under transform id
    uint8_t required[48][48]; // invariance ID needed for state l == state r
    uint8_t inverseId[48]; // pointer state stored on a hit with transform id
    uint8_t ids[48]; // the allowed transform IDs, ascending
    uint32_t numIds;
    uint32_t leafPtr[256]; // canonical 2x2x2 table: representative byte, invariance, state
    uint64_t leafInv[256];
    uint8_t leafState[256];
};
void hierarchicalTables(const tdag_config& cfg, HierTables& out);

// hierarchicalDedup computed on the device (same TDag and outMidInv, bit for bit) with work on
// `stream`, in a workspace (work, workBytes) that is grown as needed and kept by the caller. A
// failed internal check (an index outside its table, an inconsistent class) returns
// hipErrorInvalidValue with a description in *failure.
hipError_t hierarchicalDedupDevice(const Svdag& in, uint32_t midLevel, std::vector<uint32_t>& outMidInv, const tdag_config& cfg, bool staticOffsets,
    hipStream_t stream, void*& work, size_t& workBytes, TDag& out, std::string* failure = nullptr);

// .tdag byte stream (voxelizer/src/main.cpp:147-150).
std::vector<uint8_t> serializeTdag(const TDag& t, const tdag_config& cfg);

// StaticStructure packing (structure.h:299-320): a node is its child-mask word followed by one
// word per present child in child-index order. Word offset of every node of an 8-slot level
// (absent children: 0xFFFFFFFF in an SVDAG level, the sentinel TransformPointer in a TDag level).
std::vector<uint64_t> packedNodeOffsets(const std::vector<std::array<uint32_t, 8>>& nodes);
std::vector<uint64_t> packedNodeOffsets(const std::vector<std::array<uint64_t, 8>>& nodes);

// StaticStructure<void, uint32_t>(EditStructure) words of one level (structure.h:766-798).
std::vector<uint32_t> staticSvdagWords(const Svdag& s, uint32_t level);

// Reads a StaticStructure<void, uint32_t> into the node-index form; false with a message when a
// level does not parse or a child word is not a node offset (leaf index) of the level below.
bool svdagFromStatic(uint32_t resolution, uint32_t rootLevel, const uint32_t* const* levelWords, const uint64_t* levelWordCounts, const uint64_t* leaves,
    uint64_t numLeaves, Svdag& out, std::string& err);

// StaticStructure<void, TransformPointer> words of one level of a node-index TDag (children by
// word offset, leaf index at level 3), and its writeTo byte stream (structure.h:888-895).
std::vector<uint64_t> staticTdagWords(const TDag& t, uint32_t level);
std::vector<uint8_t> serializeStaticTdag(const TDag& t);

} // namespace tdag
