// SVDAG producer, flattening and .tdag serialisation (host code).
//
// The reference builds the octree by a post-order DFS with children in Morton order and drops
// empty subtrees (structure.h:658-701), then removes duplicates level by level keeping the
// first occurrence (structure.h:703-763). A post-order DFS in Morton order emits the non-empty
// nodes of every level in ascending Morton code of their position, so this builder computes the
// same indices bottom-up: leaves sorted by Morton code, parents grouped by code >> 3.
#include <algorithm>
#include <cstring>
#include <thread>
#include <unordered_map>

#include "pipeline.hpp"

namespace tdag {

namespace {

inline uint32_t spread3(uint32_t v)
{
    uint32_t x = v & 0x3FFu;
    x = (x | (x << 16)) & 0x030000FFu;
    x = (x | (x << 8)) & 0x0300F00Fu;
    x = (x | (x << 4)) & 0x030C30C3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}
inline uint32_t compact3(uint32_t x)
{
    x &= 0x09249249u;
    x = (x ^ (x >> 2)) & 0x030C30C3u;
    x = (x ^ (x >> 4)) & 0x0300F00Fu;
    x = (x ^ (x >> 8)) & 0x030000FFu;
    x = (x ^ (x >> 16)) & 0x000003FFu;
    return x;
}
inline uint64_t compact3_64(uint64_t x)
{
    x &= 0x1249249249249249ull;
    x = (x ^ (x >> 2)) & 0x10C30C30C30C30C3ull;
    x = (x ^ (x >> 4)) & 0x100F00F00F00F00Full;
    x = (x ^ (x >> 8)) & 0x1F0000FF0000FFull;
    x = (x ^ (x >> 16)) & 0x1F00000000FFFFull;
    x = (x ^ (x >> 32)) & 0x1FFFFFull;
    return x;
}

struct ArrayHash {
    size_t operator()(const std::array<uint32_t, 8>& a) const
    {
        uint64_t h = 0x243F6A8885A308D3ull;
        for (uint32_t v : a)
            h = (h ^ v) * 0x100000001B3ull + (h >> 29);
        return (size_t)h;
    }
};

template <typename F>
void parallelRange(uint64_t n, int threads, F&& f)
{
    threads = std::max(1, threads);
    if (n < 4096 || threads == 1) {
        f(0, n, 0);
        return;
    }
    std::vector<std::thread> pool;
    const uint64_t chunk = (n + threads - 1) / threads;
    for (int t = 0; t < threads; ++t) {
        const uint64_t b = std::min<uint64_t>(n, chunk * t), e = std::min<uint64_t>(n, chunk * (t + 1));
        pool.emplace_back([&, b, e, t]() { f(b, e, t); });
    }
    for (auto& th : pool)
        th.join();
}

} // namespace

Svdag svdagFromDense(const uint8_t* bits, uint32_t R, int threads)
{
    Svdag s;
    s.resolution = R;
    uint32_t levels = 0;
    while ((1u << levels) < R)
        ++levels;
    s.rootLevel = levels;
    s.nodes.resize(levels + 1);
    s.structureType = 0;
    // Row nibble (x = 0..3 of one (y, z) row) -> Morton bit positions {0, 1, 8, 9}.
    uint64_t nibble[16];
    for (uint32_t v = 0; v < 16; ++v) {
        uint64_t m = 0;
        for (uint32_t x = 0; x < 4; ++x)
            if ((v >> x) & 1u)
                m |= 1ull << spread3(x);
        nibble[v] = m;
    }
    // 1. Non-empty leaves in ascending Morton code of their 4^3 block.
    const uint32_t B = R / 4;
    const uint64_t numBlocks = (uint64_t)B * B * B;
    std::vector<std::vector<std::pair<uint64_t, uint64_t>>> parts(std::max(1, threads));
    parallelRange(numBlocks, threads, [&](uint64_t b, uint64_t e, int t) {
        auto& out = parts[t];
        for (uint64_t code = b; code < e; ++code) {
            const uint64_t bx = compact3_64(code), by = compact3_64(code >> 1), bz = compact3_64(code >> 2);
            uint64_t mask = 0;
            for (uint32_t z = 0; z < 4; ++z)
                for (uint32_t y = 0; y < 4; ++y) {
                    const uint64_t i = 4 * bx + (uint64_t)R * (4 * by + y) + (uint64_t)R * R * (4 * bz + z);
                    const uint32_t nib = (bits[i >> 3] >> (i & 7)) & 0xFu;
                    if (nib)
                        mask |= nibble[nib] << ((spread3(y) << 1) | (spread3(z) << 2));
                }
            if (mask)
                out.emplace_back(code, mask);
        }
    });
    std::vector<uint64_t> codes;
    for (auto& p : parts)
        for (auto& cm : p) {
            codes.push_back(cm.first);
            s.leaves.push_back(cm.second);
        }
    parts.clear();
    // 2. Inner levels: a node's code is its children's code >> 3.
    for (uint32_t level = 3; level <= levels; ++level) {
        std::vector<uint64_t> parentCodes;
        auto& nodes = s.nodes[level];
        for (size_t i = 0; i < codes.size(); ++i) {
            const uint64_t pc = codes[i] >> 3;
            if (parentCodes.empty() || parentCodes.back() != pc) {
                parentCodes.push_back(pc);
                std::array<uint32_t, 8> n;
                n.fill(kEmptyChild);
                nodes.push_back(n);
            }
            nodes.back()[codes[i] & 7u] = (uint32_t)i;
        }
        codes.swap(parentCodes);
    }
    // 3. toDAG: keep the first occurrence per level, bottom-up.
    std::vector<uint32_t> remap;
    {
        std::unordered_map<uint64_t, uint32_t> lut;
        lut.reserve(s.leaves.size());
        std::vector<uint64_t> out;
        remap.resize(s.leaves.size());
        for (size_t i = 0; i < s.leaves.size(); ++i) {
            auto ins = lut.emplace(s.leaves[i], (uint32_t)out.size());
            if (ins.second)
                out.push_back(s.leaves[i]);
            remap[i] = ins.first->second;
        }
        s.leaves.swap(out);
    }
    for (uint32_t level = 3; level <= levels; ++level) {
        auto& nodes = s.nodes[level];
        for (auto& n : nodes)
            for (uint32_t& c : n)
                if (c != kEmptyChild)
                    c = remap[c];
        std::unordered_map<std::array<uint32_t, 8>, uint32_t, ArrayHash> lut;
        lut.reserve(nodes.size());
        std::vector<std::array<uint32_t, 8>> out;
        std::vector<uint32_t> next(nodes.size());
        for (size_t i = 0; i < nodes.size(); ++i) {
            auto ins = lut.emplace(nodes[i], (uint32_t)out.size());
            if (ins.second)
                out.push_back(nodes[i]);
            next[i] = ins.first->second;
        }
        nodes.swap(out);
        remap.swap(next);
    }
    s.structureType = 1;
    return s;
}

void flattenNodes(const Svdag& s, uint32_t level, const uint32_t* nodeIdx, uint64_t n, uint64_t* out, int threads)
{
    const uint32_t R = 1u << level;
    const uint64_t W = std::max<uint64_t>(1, (uint64_t)R * R * R / 64);
    parallelRange(n, threads, [&](uint64_t b, uint64_t e, int) {
        struct Frame {
            uint32_t level, node, x, y, z;
        };
        std::vector<Frame> stack;
        for (uint64_t i = b; i < e; ++i) {
            uint64_t* m = out + i * W;
            std::memset(m, 0, W * 8);
            stack.clear();
            stack.push_back({ level, nodeIdx[i], 0, 0, 0 });
            while (!stack.empty()) {
                const Frame f = stack.back();
                stack.pop_back();
                if (f.level == 2) {
                    uint64_t leaf = s.leaves[f.node];
                    while (leaf) {
                        const uint32_t bit = (uint32_t)__builtin_ctzll(leaf);
                        leaf &= leaf - 1;
                        const uint32_t x = f.x + compact3(bit), y = f.y + compact3(bit >> 1), z = f.z + compact3(bit >> 2);
                        const uint64_t li = (uint64_t)x + (uint64_t)R * y + (uint64_t)R * R * z;
                        m[li >> 6] |= 1ull << (li & 63);
                    }
                    continue;
                }
                const auto& node = s.nodes[f.level][f.node];
                const uint32_t half = 1u << (f.level - 1);
                for (uint32_t c = 0; c < 8; ++c)
                    if (node[c] != kEmptyChild)
                        stack.push_back({ f.level - 1, node[c], f.x + (c & 1u) * half, f.y + ((c >> 1) & 1u) * half, f.z + ((c >> 2) & 1u) * half });
            }
        }
    });
}

std::vector<uint8_t> serializeTdag(const TDag& t, const tdag_config& cfg)
{
    // BinaryWriter::write (binary_writer.h:48-78): vectors are a uint64 count followed by raw
    // items; EditStructure::writeTo writes nodesPerLevel, subGrids, resolution, rootLevel,
    // structureType (structure.h:493-501); the voxelizer then writes the 8-byte config.
    size_t bytes = 8 + 8 + t.leaves.size() * 8 + 4 + 4 + 4 + 8;
    for (const auto& lvl : t.nodes)
        bytes += 8 + lvl.size() * 64;
    std::vector<uint8_t> b(bytes);
    uint8_t* p = b.data();
    auto put = [&](const void* src, size_t n) {
        std::memcpy(p, src, n);
        p += n;
    };
    const uint64_t nl = t.nodes.size();
    put(&nl, 8);
    for (const auto& lvl : t.nodes) {
        const uint64_t n = lvl.size();
        put(&n, 8);
        if (n)
            put(lvl.data(), n * 64);
    }
    const uint64_t ns = t.leaves.size();
    put(&ns, 8);
    if (ns)
        put(t.leaves.data(), ns * 8);
    put(&t.resolution, 4);
    put(&t.rootLevel, 4);
    put(&t.structureType, 4);
    const uint8_t c4[4] = { cfg.symmetry, cfg.axis_permutation, cfg.translation, 0 };
    put(c4, 4);
    put(&cfg.max_translation_level, 4);
    return b;
}

} // namespace tdag
