// StaticStructure form of the pipeline (host code): the packed node layout of
// voxcom::StaticStructure (structure.h:299-320, 766-798, 859-874, 888-895), read into and written
// from the node-index form the rest of the pipeline works on.
//
// A packed level holds, per node, its child-mask word followed by one word per present child in
// child-index order; a child word is the child's word offset in the level below, or a leaf index
// at level 3. Node i of a level sits at the offset of node i-1 plus 1 + popcount(mask of i-1).
#include <algorithm>
#include <cstring>
#include <string>

#include "pipeline.hpp"

namespace tdag {

namespace {

constexpr uint64_t kSentinelTp = 0x82082400FFFFFFFFull;
constexpr uint64_t kPtrMask = (1ull << 37) - 1;

template <typename Node, typename IsChild>
std::vector<uint64_t> offsetsOf(const std::vector<Node>& nodes, IsChild isChild)
{
    std::vector<uint64_t> out(nodes.size());
    uint64_t at = 0;
    for (size_t i = 0; i < nodes.size(); ++i) {
        out[i] = at;
        uint32_t n = 0;
        for (auto c : nodes[i])
            n += isChild(c) ? 1u : 0u;
        at += 1 + n;
    }
    return out;
}

} // namespace

std::vector<uint64_t> packedNodeOffsets(const std::vector<std::array<uint32_t, 8>>& nodes)
{
    return offsetsOf(nodes, [](uint32_t c) { return c != kEmptyChild; });
}

std::vector<uint64_t> packedNodeOffsets(const std::vector<std::array<uint64_t, 8>>& nodes)
{
    return offsetsOf(nodes, [](uint64_t c) { return c != kSentinelTp; });
}

std::vector<uint32_t> staticSvdagWords(const Svdag& s, uint32_t level)
{
    std::vector<uint32_t> out;
    if (level < 3 || level > s.rootLevel)
        return out;
    const std::vector<uint64_t> below = level > 3 ? packedNodeOffsets(s.nodes[level - 1]) : std::vector<uint64_t> {};
    for (const auto& node : s.nodes[level]) {
        const size_t at = out.size();
        out.push_back(0);
        for (uint32_t c = 0; c < 8; ++c)
            if (node[c] != kEmptyChild) {
                out.push_back(level > 3 ? (uint32_t)below[node[c]] : node[c]);
                out[at] |= 1u << c;
            }
    }
    return out;
}

bool svdagFromStatic(uint32_t resolution, uint32_t rootLevel, const uint32_t* const* levelWords, const uint64_t* levelWordCounts, const uint64_t* leaves,
    uint64_t numLeaves, Svdag& out, std::string& err)
{
    out = Svdag {};
    out.resolution = resolution;
    out.rootLevel = rootLevel;
    out.structureType = 1;
    out.nodes.resize(rootLevel + 1);
    out.leaves.assign(leaves, leaves + numLeaves);
    std::vector<uint64_t> belowStarts; // node offsets of the level below (empty at level 3)
    for (uint32_t level = 3; level <= rootLevel; ++level) {
        const uint32_t* w = levelWords[level];
        const uint64_t nw = levelWordCounts[level];
        if (nw && !w) {
            err = "level " + std::to_string(level) + " has words but a null pointer";
            return false;
        }
        std::vector<uint64_t> starts;
        auto& nodes = out.nodes[level];
        for (uint64_t at = 0; at < nw;) {
            const uint32_t mask = w[at];
            if (mask > 0xFFu) {
                err = "level " + std::to_string(level) + " word " + std::to_string(at) + ": child mask " + std::to_string(mask) + " has bits above 8";
                return false;
            }
            const uint32_t nc = (uint32_t)__builtin_popcount(mask);
            if (at + 1 + nc > nw) {
                err = "level " + std::to_string(level) + ": the node at word " + std::to_string(at) + " runs past the level's end";
                return false;
            }
            std::array<uint32_t, 8> node;
            node.fill(kEmptyChild);
            uint32_t o = 0;
            for (uint32_t c = 0; c < 8; ++c) {
                if (!((mask >> c) & 1u))
                    continue;
                const uint32_t child = w[at + 1 + o++];
                if (level == 3) {
                    if (child >= numLeaves) {
                        err = "level 3 word " + std::to_string(at) + ": leaf index " + std::to_string(child) + " out of range";
                        return false;
                    }
                    node[c] = child;
                } else {
                    const auto it = std::lower_bound(belowStarts.begin(), belowStarts.end(), (uint64_t)child);
                    if (it == belowStarts.end() || *it != child) {
                        err = "level " + std::to_string(level) + " word " + std::to_string(at) + ": child word " + std::to_string(child) +
                            " is not a node offset of level " + std::to_string(level - 1);
                        return false;
                    }
                    node[c] = (uint32_t)(it - belowStarts.begin());
                }
            }
            starts.push_back(at);
            nodes.push_back(node);
            at += 1 + nc;
        }
        belowStarts = std::move(starts);
    }
    return true;
}

std::vector<uint64_t> staticTdagWords(const TDag& t, uint32_t level)
{
    std::vector<uint64_t> out;
    if (level == 2) {
        out = t.leaves;
        return out;
    }
    if (level < 3 || level >= t.nodes.size())
        return out;
    const std::vector<uint64_t> below = level > 3 ? packedNodeOffsets(t.nodes[level - 1]) : std::vector<uint64_t> {};
    for (const auto& node : t.nodes[level]) {
        const size_t at = out.size();
        out.push_back(0);
        for (uint32_t c = 0; c < 8; ++c)
            if (node[c] != kSentinelTp) {
                const uint64_t ptr = node[c] & kPtrMask;
                out.push_back((node[c] & ~kPtrMask) | (level > 3 ? below[ptr] : ptr));
                out[at] |= 1u << c;
            }
    }
    return out;
}

std::vector<uint8_t> serializeStaticTdag(const TDag& t)
{
    std::vector<uint8_t> b;
    auto put = [&](const void* p, size_t n) {
        const uint8_t* c = (const uint8_t*)p;
        b.insert(b.end(), c, c + n);
    };
    // nodesPerLevel: u64 level count, then per level u64 word count + the words.
    const uint64_t nl = t.nodes.size();
    put(&nl, 8);
    for (uint32_t level = 0; level < nl; ++level) {
        const std::vector<uint64_t> w = level >= 3 ? staticTdagWords(t, level) : std::vector<uint64_t> {};
        const uint64_t n = w.size();
        put(&n, 8);
        put(w.data(), n * 8);
    }
    const uint64_t ns = t.leaves.size();
    put(&ns, 8);
    put(t.leaves.data(), ns * 8);
    put(&t.resolution, 4);
    put(&t.rootLevel, 4);
    put(&t.structureType, 4);
    return b;
}

} // namespace tdag
