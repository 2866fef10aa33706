// Calls the C++ drop-ins of include/tdag.hpp the way the reference's code calls its own entry
// points (transform_dag.cpp:201 and voxelizer/src/main.cpp:146), on stand-in types with the
// reference's member names and layouts, and writes the results for tests/test_gpu_dropin.py.
//
//   dropin_check construct <in> <out> <sym> <axis> <translation> <maxT>
//       in:  u32 resolution, u32 rootLevel, u64 leaf count, leaf masks, then for level 3..root:
//            u64 node count and 8 u32 children per node (an EditStructure<void, uint32_t>)
//       out: u32 rootLevel, i32 structureType, then for level 2..root: u64 words and the words
//            (level 2: leaf masks; above: 8 TransformPointer bit patterns per node)
//   dropin_check match <level> <in> <out> <sym> <axis> <translation> <maxT>
//       in:  u64 n, n * W u64 LargeSubGrid words;  out: n mapping words, u64 unique count, unique u32
#include <array>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

#include "tdag.hpp"

// Stand-ins with the layouts of voxcom's types (structure.h:45-51, 117-125, 127-171;
// transform_dag.h:19-33, 157-160; large_sub_grid.h:80-97).
namespace standin {
enum class StructureType { Tree, DAG };
template <typename Child>
struct EditNode {
    std::array<Child, 8> children;
};
struct EditSubGridVoid {
    uint64_t bitmask = 0;
};
template <typename Child>
struct EditStructure {
    std::vector<std::vector<EditNode<Child>>> nodesPerLevel;
    std::vector<EditSubGridVoid> subGrids;
    unsigned resolution = 0;
    uint32_t rootLevel = 0;
    StructureType structureType = StructureType::Tree;
};
struct TransformPointer {
    uint64_t ptr : 37 = 0;
    uint64_t sx : 1 = 0, sy : 1 = 0, sz : 1 = 0;
    uint64_t p0 : 2 = 0, p1 : 2 = 1, p2 : 2 = 2;
    uint64_t tx : 6 = 0, ty : 6 = 0, tz : 6 = 0;
};
static_assert(sizeof(TransformPointer) == 8);
template <uint32_t Resolution>
struct LargeSubGrid {
    uint64_t storage[Resolution * Resolution * Resolution / 64];
};
struct TransformDAGConfig {
    bool symmetry = true, axisPermutation = true, translation = true;
    uint32_t maxTranslationLevel = 4;
};
} // namespace standin

template <typename T>
static void rd(std::ifstream& f, T* p, size_t n)
{
    f.read(reinterpret_cast<char*>(p), (std::streamsize)(n * sizeof(T)));
    if (!f)
        throw std::runtime_error("short input file");
}
template <typename T>
static void wr(std::ofstream& f, const T* p, size_t n)
{
    f.write(reinterpret_cast<const char*>(p), (std::streamsize)(n * sizeof(T)));
}

static int construct(const char* inPath, const char* outPath, const standin::TransformDAGConfig& config)
{
    std::ifstream f(inPath, std::ios::binary);
    standin::EditStructure<uint32_t> in;
    uint32_t resolution = 0, root = 0;
    uint64_t n = 0;
    rd(f, &resolution, 1);
    rd(f, &root, 1);
    rd(f, &n, 1);
    in.resolution = resolution;
    in.rootLevel = root;
    in.structureType = standin::StructureType::DAG;
    in.subGrids.resize(n);
    rd(f, in.subGrids.data(), n);
    in.nodesPerLevel.resize(root + 1);
    for (uint32_t level = 3; level <= root; ++level) {
        rd(f, &n, 1);
        in.nodesPerLevel[level].resize(n);
        rd(f, in.nodesPerLevel[level].data(), n);
    }
    // The reference's call: constructTransformDAG(inDag, config) -> EditStructure<void, TransformPointer>.
    const auto out = tdag_cpp::constructTransformDAG<standin::EditStructure<standin::TransformPointer>>(in, config);
    std::ofstream o(outPath, std::ios::binary);
    const int32_t type = (int32_t)out.structureType;
    wr(o, &out.rootLevel, 1);
    wr(o, &type, 1);
    for (uint32_t level = 2; level <= out.rootLevel; ++level) {
        if (level == 2) {
            n = out.subGrids.size();
            wr(o, &n, 1);
            wr(o, out.subGrids.data(), n);
        } else {
            n = out.nodesPerLevel[level].size() * 8;
            wr(o, &n, 1);
            wr(o, out.nodesPerLevel[level].data(), n / 8);
        }
    }
    return 0;
}

template <uint32_t Level>
static int match(const char* inPath, const char* outPath, const standin::TransformDAGConfig& config)
{
    constexpr uint32_t R = 1u << Level;
    std::ifstream f(inPath, std::ios::binary);
    uint64_t n = 0;
    rd(f, &n, 1);
    std::vector<standin::LargeSubGrid<R>> flattenedGrids(n);
    rd(f, flattenedGrids.data(), n);
    std::vector<uint32_t> indices(n);
    for (uint64_t i = 0; i < n; ++i)
        indices[i] = (uint32_t)(1000 + i); // payloads distinct from positions
    std::vector<standin::TransformPointer> nodeMapping;
    // The reference's call (transform_dag.cpp:201).
    const auto remainingIndices = tdag_cpp::findDuplicatesAmongFlattenedItems_gpu<Level, uint32_t>(indices, flattenedGrids, nodeMapping, config);
    std::ofstream o(outPath, std::ios::binary);
    wr(o, nodeMapping.data(), nodeMapping.size());
    const uint64_t nu = remainingIndices.size();
    wr(o, &nu, 1);
    wr(o, remainingIndices.data(), nu);
    return 0;
}

int main(int argc, char** argv)
{
    try {
        const std::string mode = argc > 1 ? argv[1] : "";
        const int base = mode == "construct" ? 4 : 5;
        if ((mode != "construct" && mode != "match") || argc != base + 4) {
            std::fprintf(stderr, "usage: dropin_check construct <in> <out> s a t maxT | match <level> <in> <out> s a t maxT\n");
            return 2;
        }
        standin::TransformDAGConfig config;
        config.symmetry = std::atoi(argv[base]) != 0;
        config.axisPermutation = std::atoi(argv[base + 1]) != 0;
        config.translation = std::atoi(argv[base + 2]) != 0;
        config.maxTranslationLevel = (uint32_t)std::atoi(argv[base + 3]);
        if (mode == "construct")
            return construct(argv[2], argv[3], config);
        switch (std::atoi(argv[2])) {
        case 2:
            return match<2>(argv[3], argv[4], config);
        case 3:
            return match<3>(argv[3], argv[4], config);
        case 4:
            return match<4>(argv[3], argv[4], config);
        case 5:
            return match<5>(argv[3], argv[4], config);
        default:
            return 2;
        }
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
}
