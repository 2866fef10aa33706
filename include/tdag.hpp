// tdag.hpp — C++ drop-ins over the C-ABI in tdag.h, mirroring the reference's entry points:
//
//   findDuplicatesAmongFlattenedItems_gpu<Level, Item>   transform_dag.h:310-311
//   constructTransformDAG(structure, config)              transform_dag.h:161-162
//   constructStaticTransformDAG(structure, config)        transform_dag.h:163
//   constructTransformDAG16(structure, config, encoding)  svdag_encoding/src/transform_dag_encoding.cpp:241-478
//
// Header-only; link against libtdag.so. Errors are rethrown as std::runtime_error, which is
// what the reference's assert_always does (error_handling.h:17-45).
#pragma once

#include <cstdint>
#include <cstring>
#include <memory>
#include <span>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "tdag.h"

namespace tdag_cpp {

inline void check(tdag_status s, const char* what)
{
    if (s != TDAG_OK) {
        char buf[1024];
        tdag_last_error(buf, sizeof(buf));
        throw std::runtime_error(std::string(what) + ": " + buf);
    }
}

// Owns a tdag_context (one HIP stream, cached device scratch). Not thread-safe.
class Context {
public:
    explicit Context(int device = 0) { check(tdag_context_create(device, &m_ctx), "tdag_context_create"); }
    ~Context() { tdag_context_destroy(m_ctx); }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    tdag_context* get() const { return m_ctx; }

    static Context& instance()
    {
        static Context ctx(0);
        return ctx;
    }

private:
    tdag_context* m_ctx = nullptr;
};

inline tdag_config toC(bool symmetry, bool axisPermutation, bool translation, uint32_t maxTranslationLevel)
{
    tdag_config c {};
    c.symmetry = symmetry;
    c.axis_permutation = axisPermutation;
    c.translation = translation;
    c.max_translation_level = maxTranslationLevel;
    return c;
}

namespace detail {

// TransformDAGConfig (transform_dag.h:157-160) or tdag_config: 3 bools, padding, uint32.
template <typename Config>
tdag_config toConfig(const Config& config)
{
    static_assert(sizeof(Config) == sizeof(tdag_config) && std::is_trivially_copyable_v<Config>,
        "config must have the byte layout of voxcom::TransformDAGConfig");
    tdag_config c;
    std::memcpy(&c, &config, sizeof(c));
    return c;
}

} // namespace detail

// findDuplicatesAmongFlattenedItems_gpu<Level, Item> (transform_dag.h:310-311), typed like the
// reference's declaration and called the same way (transform_dag.cpp:201):
//   findDuplicatesAmongFlattenedItems_gpu<Level, uint32_t>(indices, flattenedGrids, nodeMapping, config)
//   items                one payload per grid (node index, leaf, ...)
//   grids                a contiguous range (std::vector, std::span) of voxcom::LargeSubGrid<2^Level>
//                        (large_sub_grid.h:80-97: W uint64 words each), or of n * W uint64 words
//   outPrevLevelMapping  std::vector of voxcom::TransformPointer (or uint64_t): resized to n, filled
//                        with the TransformPointer bit patterns (transform_dag.h:24-33)
//   config               voxcom::TransformDAGConfig (or tdag_config)
// Returns the unique items in original order, as the reference does.
template <size_t Level, typename Item, typename Grids, typename Pointer, typename Config>
std::vector<Item> findDuplicatesAmongFlattenedItems_gpu(std::span<const Item> items, const Grids& grids, std::vector<Pointer>& outPrevLevelMapping,
    const Config& config, Context& ctx = Context::instance())
{
    static_assert(Level >= 2 && Level <= 5, "levels 2..5 are supported");
    using Grid = std::remove_cv_t<std::remove_reference_t<decltype(*std::data(grids))>>;
    static_assert(std::is_trivially_copyable_v<Grid> && sizeof(Grid) % 8 == 0, "grids must be LargeSubGrid<R> or uint64 words");
    static_assert(sizeof(Pointer) == 8 && std::is_trivially_copyable_v<Pointer>, "the mapping holds 64-bit TransformPointers");
    constexpr size_t R = size_t(1) << Level;
    constexpr size_t W = R * R * R / 64 > 0 ? R * R * R / 64 : 1;
    static_assert(sizeof(Grid) == 8 || sizeof(Grid) == W * 8, "a grid must be W = R^3/64 uint64 words (LargeSubGrid<R>)");
    const uint64_t* words = reinterpret_cast<const uint64_t*>(std::data(grids));
    const size_t numWords = std::size(grids) * (sizeof(Grid) / 8);
    if (numWords != items.size() * W)
        throw std::runtime_error("findDuplicatesAmongFlattenedItems_gpu: the grids do not hold W words per item");
    const uint64_t n = items.size();
    const tdag_config cfg = detail::toConfig(config);
    std::vector<uint64_t> mapping(n);
    std::vector<uint32_t> unique(n);
    uint64_t numUnique = 0;
    check(tdag_match_level(ctx.get(), (uint32_t)Level, words, n, &cfg, nullptr, mapping.data(), unique.data(), &numUnique, nullptr), "tdag_match_level");
    outPrevLevelMapping.resize(n);
    if (n)
        std::memcpy(static_cast<void*>(outPrevLevelMapping.data()), mapping.data(), n * 8);
    std::vector<Item> out;
    out.reserve(numUnique);
    for (uint64_t k = 0; k < numUnique; ++k)
        out.push_back(items[unique[k]]);
    return out;
}

// RAII owners of the pipeline handles.
struct SvdagDeleter {
    void operator()(tdag_svdag* p) const { tdag_svdag_free(p); }
};
struct TdagDeleter {
    void operator()(tdag_tdag* p) const { tdag_tdag_free(p); }
};
using Svdag = std::unique_ptr<tdag_svdag, SvdagDeleter>;
using TDag = std::unique_ptr<tdag_tdag, TdagDeleter>;

inline Svdag svdagFromDense(const uint8_t* bits, uint32_t resolution)
{
    tdag_svdag* s = nullptr;
    check(tdag_svdag_from_dense(bits, resolution, &s), "tdag_svdag_from_dense");
    return Svdag(s);
}

// EditStructure(VoxelGrid) + toDAG built on the device from a dense grid already in device
// memory (resolution^3 / 8 bytes, bit i = x + R*y + R*R*z); the same SVDAG as svdagFromDense.
inline Svdag svdagFromDenseDevice(const uint8_t* d_bits, uint32_t resolution, Context& ctx = Context::instance(), float* out_ms = nullptr)
{
    tdag_svdag* s = nullptr;
    check(tdag_svdag_from_dense_device(ctx.get(), d_bits, resolution, &s, out_ms), "tdag_svdag_from_dense_device");
    return Svdag(s);
}

// FlatGridGenerator / createLargeSubGrid on the device: d_nodes (n level-`level` node indices)
// -> d_out (n * W words, the LargeSubGrid layout tdag_match_level_device reads).
inline void flattenDevice(const tdag_svdag* svdag, uint32_t level, const uint32_t* d_nodes, uint64_t n, uint64_t* d_out, Context& ctx = Context::instance())
{
    check(tdag_svdag_flatten_device(ctx.get(), svdag, level, d_nodes, n, d_out), "tdag_svdag_flatten_device");
}

// constructTransformDAG (transform_dag.h:161-162) on the GPU matcher.
inline TDag constructTransformDAG(const tdag_svdag* inDag, const tdag_config& config, Context& ctx = Context::instance())
{
    tdag_tdag* t = nullptr;
    check(tdag_construct_transform_dag(ctx.get(), inDag, &config, 0, &t), "tdag_construct_transform_dag");
    return TDag(t);
}

// An EditStructure<void, uint32_t> / EditStructureOOC<void, uint32_t> (structure.h:133-171) as
// the C-ABI view: any type with the reference's members nodesPerLevel (per level a contiguous
// range of EditNode<uint32_t>: 8 uint32 children), subGrids (EditSubGrid<void>: one uint64),
// resolution, rootLevel and structureType. Validated and copied by the library.
template <typename InStructure>
Svdag svdagFromStructure(const InStructure& inDag)
{
    const uint32_t root = (uint32_t)inDag.rootLevel;
    std::vector<const uint32_t*> nodes(root + 1, nullptr);
    std::vector<uint64_t> counts(root + 1, 0);
    for (uint32_t level = 3; level <= root && level < std::size(inDag.nodesPerLevel); ++level) {
        const auto& lv = inDag.nodesPerLevel[level];
        using Node = std::remove_cv_t<std::remove_reference_t<decltype(*std::data(lv))>>;
        static_assert(sizeof(Node) == 32 && std::is_trivially_copyable_v<Node>, "EditNode<uint32_t>: 8 uint32 children");
        nodes[level] = reinterpret_cast<const uint32_t*>(std::data(lv));
        counts[level] = std::size(lv);
    }
    using Sub = std::remove_cv_t<std::remove_reference_t<decltype(*std::data(inDag.subGrids))>>;
    static_assert(sizeof(Sub) == 8 && std::is_trivially_copyable_v<Sub>, "EditSubGrid<void>: one uint64 bitmask");
    const tdag_svdag_view view { (uint32_t)inDag.resolution, root, nodes.data(), counts.data(), reinterpret_cast<const uint64_t*>(std::data(inDag.subGrids)),
        (uint64_t)std::size(inDag.subGrids), (int32_t)inDag.structureType, 0u };
    tdag_svdag* s = nullptr;
    check(tdag_svdag_from_view(&view, &s), "tdag_svdag_from_view");
    return Svdag(s);
}

// constructTransformDAG (transform_dag.h:161-162) typed like the reference:
//   EditStructure<void, TransformPointer> constructTransformDAG(const Structure<void, uint32_t>&, const TransformDAGConfig&)
// OutStructure is the reference's EditStructure<void, TransformPointer> (or any type with its
// members: nodesPerLevel as std::vector<std::vector<EditNode<TransformPointer>>>, subGrids as a
// std::vector of EditSubGrid<void>, resolution, rootLevel, structureType); it comes back with the
// reference's output: every level filled, structureType Tree (transform_dag_hierarchical.cpp:307).
// A voxcom build binds its own template to this one:
//   template <template <typename, typename> typename Structure>
//   EditStructure<void, TransformPointer> constructTransformDAG(const Structure<void, uint32_t>& in, const TransformDAGConfig& c)
//   { return tdag_cpp::constructTransformDAG<EditStructure<void, TransformPointer>>(in, c); }
template <typename OutStructure, typename InStructure, typename Config>
OutStructure constructTransformDAG(const InStructure& inDag, const Config& config, Context& ctx = Context::instance())
{
    const Svdag in = svdagFromStructure(inDag);
    const tdag_config cfg = detail::toConfig(config);
    tdag_tdag* raw = nullptr;
    check(tdag_construct_transform_dag(ctx.get(), in.get(), &cfg, 0, &raw), "tdag_construct_transform_dag");
    const TDag t(raw);
    OutStructure out {};
    uint32_t root = 0;
    check(tdag_tdag_info(t.get(), &root, nullptr, 0), "tdag_tdag_info");
    out.resolution = inDag.resolution;
    out.rootLevel = root;
    out.structureType = static_cast<decltype(out.structureType)>(0);
    out.nodesPerLevel.resize(root + 1);
    for (uint32_t level = 2; level <= root; ++level) {
        uint64_t words = 0;
        check(tdag_tdag_copy_level(t.get(), level, nullptr, 0, &words), "tdag_tdag_copy_level");
        if (level == 2) {
            static_assert(sizeof(*std::data(out.subGrids)) == 8, "EditSubGrid<void>: one uint64 bitmask");
            out.subGrids.resize(words);
            check(tdag_tdag_copy_level(t.get(), level, reinterpret_cast<uint64_t*>(out.subGrids.data()), words, &words), "tdag_tdag_copy_level");
        } else {
            auto& lv = out.nodesPerLevel[level];
            static_assert(sizeof(*std::data(lv)) == 64, "EditNode<TransformPointer>: 8 64-bit TransformPointers");
            lv.resize(words / 8);
            check(tdag_tdag_copy_level(t.get(), level, reinterpret_cast<uint64_t*>(lv.data()), words, &words), "tdag_tdag_copy_level");
        }
    }
    return out;
}

// The .tdag byte stream the voxelizer writes (voxelizer/src/main.cpp:144-151).
inline std::vector<uint8_t> tdagBytes(const tdag_tdag* t, const tdag_config& config)
{
    uint64_t size = 0;
    check(tdag_tdag_serialize(t, &config, nullptr, 0, &size), "tdag_tdag_serialize");
    std::vector<uint8_t> out(size);
    check(tdag_tdag_serialize(t, &config, out.data(), out.size(), &size), "tdag_tdag_serialize");
    return out;
}

// The fields of voxcom::StaticStructure (structure.h:299-320) that the static driver reads and
// writes: per level the packed words (child mask, then one word per present child), the 4x4x4
// leaf masks (EditSubGrid<void> is one uint64), resolution and root level.
template <typename Word>
struct StaticStructure {
    std::vector<std::vector<Word>> nodesPerLevel;
    std::vector<uint64_t> subGrids;
    uint32_t resolution = 0;
    uint32_t rootLevel = 0;
    int32_t structureType = 0;
};

// StaticStructure(const EditStructure&) (structure.h:766-798) of an SVDAG built by the library.
inline StaticStructure<uint32_t> toStaticStructure(const tdag_svdag* svdag)
{
    StaticStructure<uint32_t> out;
    check(tdag_svdag_info(svdag, &out.resolution, &out.rootLevel, nullptr, 0), "tdag_svdag_info");
    out.structureType = 1;
    out.nodesPerLevel.resize(out.rootLevel + 1);
    uint64_t n = 0;
    check(tdag_svdag_static_words(svdag, 2, nullptr, 0, &n), "tdag_svdag_static_words");
    out.subGrids.resize(n);
    check(tdag_svdag_static_words(svdag, 2, out.subGrids.data(), n, &n), "tdag_svdag_static_words");
    for (uint32_t level = 3; level <= out.rootLevel; ++level) {
        check(tdag_svdag_static_words(svdag, level, nullptr, 0, &n), "tdag_svdag_static_words");
        out.nodesPerLevel[level].resize(n);
        check(tdag_svdag_static_words(svdag, level, out.nodesPerLevel[level].data(), n, &n), "tdag_svdag_static_words");
    }
    return out;
}

// constructStaticTransformDAG (transform_dag.h:163) on the GPU matcher.
inline StaticStructure<uint64_t> constructStaticTransformDAG(const StaticStructure<uint32_t>& inDag, const tdag_config& config, Context& ctx = Context::instance())
{
    std::vector<const uint32_t*> words(inDag.rootLevel + 1, nullptr);
    std::vector<uint64_t> counts(inDag.rootLevel + 1, 0);
    for (uint32_t level = 3; level <= inDag.rootLevel && level < inDag.nodesPerLevel.size(); ++level) {
        words[level] = inDag.nodesPerLevel[level].data();
        counts[level] = inDag.nodesPerLevel[level].size();
    }
    const tdag_static_svdag_view view { inDag.resolution, inDag.rootLevel, words.data(), counts.data(), inDag.subGrids.data(), inDag.subGrids.size() };
    tdag_tdag* raw = nullptr;
    check(tdag_construct_static_transform_dag(ctx.get(), &view, &config, 0, &raw), "tdag_construct_static_transform_dag");
    TDag t(raw);
    StaticStructure<uint64_t> out;
    out.resolution = inDag.resolution;
    out.rootLevel = inDag.rootLevel;
    out.nodesPerLevel.resize(inDag.rootLevel + 1);
    for (uint32_t level = 2; level <= inDag.rootLevel; ++level) {
        uint64_t n = 0;
        check(tdag_tdag_static_words(t.get(), level, nullptr, 0, &n), "tdag_tdag_static_words");
        std::vector<uint64_t>& dst = level == 2 ? out.subGrids : out.nodesPerLevel[level];
        dst.resize(n);
        check(tdag_tdag_static_words(t.get(), level, dst.data(), n, &n), "tdag_tdag_static_words");
    }
    return out;
}

struct Tdag16Deleter {
    void operator()(tdag_tdag16* h) const { tdag_tdag16_free(h); }
};
using TDag16 = std::unique_ptr<tdag_tdag16, Tdag16Deleter>;

// constructTransformDAG16 (transform_dag_encoding.cpp:241-478) of a constructTransformDAG result;
// TransformEncodingConfig{pointerTables, huffmanCode} (transform_dag_encoding.h:118-120).
inline TDag16 constructTransformDAG16(const tdag_tdag* t, const tdag_config& config, bool pointerTables, bool huffmanCode)
{
    tdag_tdag16* h = nullptr;
    check(tdag_tdag16_encode(t, &config, (pointerTables ? TDAG16_POINTER_TABLES : 0u) | (huffmanCode ? TDAG16_HUFFMAN : 0u), &h), "tdag_tdag16_encode");
    return TDag16(h);
}

// TransformDAG16::writeTo (transform_dag_encoding.cpp:1020-1027).
inline std::vector<uint8_t> tdag16Bytes(const tdag_tdag16* h)
{
    uint64_t n = 0;
    check(tdag_tdag16_serialize(h, nullptr, 0, &n), "tdag_tdag16_serialize");
    std::vector<uint8_t> out(n);
    check(tdag_tdag16_serialize(h, out.data(), n, &n), "tdag_tdag16_serialize");
    return out;
}

// TransformDAG16::readFrom (transform_dag_encoding.cpp:1029-1036).
inline TDag16 tdag16FromBytes(const std::vector<uint8_t>& bytes)
{
    tdag_tdag16* h = nullptr;
    check(tdag_tdag16_from_bytes(bytes.data(), bytes.size(), &h), "tdag_tdag16_from_bytes");
    return TDag16(h);
}

// TranslateTransformBaseDAG::get (transform_dag_encoding.cpp:897-942) for every voxel, into a dense
// grid in device memory (resolution^3 / 32 words, bit i = x + R*y + R*R*z).
inline float decodeTdag16Device(const tdag_tdag16* h, uint32_t* d_words, uint64_t numWords, Context& ctx = Context::instance())
{
    float ms = 0.f;
    check(tdag_tdag16_decode_device(ctx.get(), h, d_words, numWords, &ms), "tdag_tdag16_decode_device");
    return ms;
}

// First-hit ray casting through the encoded DAG on the device (the renderer's use of the format,
// renderer/src/dags/transform_dag/transform_dag.cu:5-41): d_rays n x 8 floats (origin, tmin,
// direction, tmax), d_hits n x 4 int32 (voxel x y z or -1, steps), d_t n floats. Returns the
// device time in milliseconds.
inline float traceTdag16Device(const tdag_tdag16* h, const float* d_rays, uint64_t n, int32_t* d_hits, float* d_t, Context& ctx = Context::instance())
{
    float ms = 0.f;
    check(tdag_tdag16_trace_device(ctx.get(), h, d_rays, n, d_hits, d_t, &ms), "tdag_tdag16_trace_device");
    return ms;
}

} // namespace tdag_cpp
