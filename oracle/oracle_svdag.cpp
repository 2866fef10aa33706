// TEST INFRASTRUCTURE ONLY: CPU restatement of the reference's pipeline around the matcher.
//
//   SVO build from a dense grid       structure.h:423-429 (ctor), 658-701 (recursion)
//   toDAG first-occurrence dedup      structure.h:703-763
//   structure.get / createLargeSubGrid structure.h:334-355, large_sub_grid.h:127-141
//   hierarchical sym+perm dedup       transform_dag_hierarchical.cpp:23-523
//   constructTransformDAG driver      transform_dag.cpp:118-240
//   StaticStructure conversion        structure.h:766-798, 859-874, 888-895
//   static hierarchical dedup         transform_dag_hierarchical.cpp:531-743
//   constructStaticTransformDAG       transform_dag.cpp:242-390
//   .tdag serialisation               voxelizer/src/main.cpp:144-151, structure.h:493-501,
//                                     binary_writer.h:48-78
// The matcher it calls is the oracle restatement in oracle_match.cpp.
#include <algorithm>
#include <array>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <numeric>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "oracle.h"

namespace orc2 {

static const uint32_t kEmpty = 0xFFFFFFFFu;

static inline uint32_t morton3(uint32_t x, uint32_t y, uint32_t z)
{
    uint32_t m = 0;
    for (int b = 0; b < 10; ++b)
        m |= (((x >> b) & 1u) << (3 * b)) | (((y >> b) & 1u) << (3 * b + 1)) | (((z >> b) & 1u) << (3 * b + 2));
    return m;
}
static inline void demorton3(uint32_t m, uint32_t& x, uint32_t& y, uint32_t& z)
{
    x = y = z = 0;
    for (int b = 0; b < 10; ++b) {
        x |= ((m >> (3 * b)) & 1u) << b;
        y |= ((m >> (3 * b + 1)) & 1u) << b;
        z |= ((m >> (3 * b + 2)) & 1u) << b;
    }
}

// EditStructure<void, uint32_t> (structure.h:159-211).
struct Svdag {
    std::vector<std::vector<std::array<uint32_t, 8>>> nodesPerLevel;
    std::vector<uint64_t> subGrids;
    uint32_t resolution = 0;
    uint32_t rootLevel = 0;
    int32_t structureType = 0; // Tree = 0, DAG = 1
};

// VoxelGrid<void> (voxel_grid.h): index = x + R*y + R*R*z, bit-packed, out of range -> false.
struct Dense {
    const uint8_t* bits;
    uint32_t R;
    bool get(uint32_t x, uint32_t y, uint32_t z) const
    {
        const uint64_t i = (uint64_t)x + (uint64_t)R * y + (uint64_t)R * R * z;
        if (i >= (uint64_t)R * R * R)
            return false;
        return (bits[i >> 3] >> (i & 7)) & 1u;
    }
};

// constructOctreeFromVoxelGridRecurse (structure.h:658-701).
static uint32_t buildRecurse(Svdag& s, const Dense& g, int level, uint32_t x0, uint32_t y0, uint32_t z0)
{
    if (level == 2) {
        uint64_t mask = 0;
        for (uint32_t i = 0; i < 64; ++i) {
            uint32_t x, y, z;
            demorton3(i, x, y, z);
            if (g.get(x0 + x, y0 + y, z0 + z))
                mask |= 1ull << i;
        }
        if (mask == 0)
            return kEmpty;
        s.subGrids.push_back(mask);
        return (uint32_t)s.subGrids.size() - 1;
    }
    const uint32_t half = 1u << (level - 1);
    std::array<uint32_t, 8> node;
    for (uint32_t i = 0; i < 8; ++i) {
        uint32_t dx, dy, dz;
        demorton3(i, dx, dy, dz);
        node[i] = buildRecurse(s, g, level - 1, x0 + dx * half, y0 + dy * half, z0 + dz * half);
    }
    bool allEmpty = true;
    for (uint32_t c : node)
        allEmpty = allEmpty && c == kEmpty;
    if (allEmpty)
        return kEmpty;
    s.nodesPerLevel[level].push_back(node);
    return (uint32_t)s.nodesPerLevel[level].size() - 1;
}

static void toDAG(Svdag& s) // structure.h:703-763
{
    std::vector<uint32_t> prev;
    {
        std::unordered_map<uint64_t, uint32_t> lut;
        std::vector<uint64_t> out;
        for (uint64_t sg : s.subGrids) {
            auto it = lut.find(sg);
            if (it != lut.end()) {
                prev.push_back(it->second);
            } else {
                const uint32_t h = (uint32_t)out.size();
                out.push_back(sg);
                lut[sg] = h;
                prev.push_back(h);
            }
        }
        s.subGrids = std::move(out);
    }
    for (uint32_t level = 3; level <= s.rootLevel; ++level) {
        auto& nodes = s.nodesPerLevel[level];
        for (auto& n : nodes)
            for (auto& c : n)
                if (c != kEmpty)
                    c = prev[c];
        prev.clear();
        std::map<std::array<uint32_t, 8>, uint32_t> lut;
        std::vector<std::array<uint32_t, 8>> out;
        for (const auto& n : nodes) {
            auto it = lut.find(n);
            if (it != lut.end()) {
                prev.push_back(it->second);
            } else {
                const uint32_t h = (uint32_t)out.size();
                out.push_back(n);
                lut[n] = h;
                prev.push_back(h);
            }
        }
        nodes = std::move(out);
    }
    s.structureType = 1;
}

// EditStructure::get(voxel, level, root) (structure.h:334-355).
static bool structureGet(const Svdag& s, uint32_t x, uint32_t y, uint32_t z, uint32_t level, uint32_t root)
{
    uint32_t nodeIdx = root;
    for (uint32_t l = level; l > 2; --l) {
        const auto& node = s.nodesPerLevel[l][nodeIdx];
        const uint32_t cl = l - 1;
        const uint32_t childIdx = morton3((x >> cl) & 1u, (y >> cl) & 1u, (z >> cl) & 1u);
        nodeIdx = node[childIdx];
        if (nodeIdx == kEmpty)
            return false;
    }
    const uint32_t vi = morton3(x & 3u, y & 3u, z & 3u);
    return (s.subGrids[nodeIdx] >> vi) & 1ull;
}

// createLargeSubGrid (large_sub_grid.h:127-141): linear layout x + R*y + R*R*z.
static void flatten(const Svdag& s, uint32_t level, uint32_t node, uint64_t* out)
{
    const uint32_t R = 1u << level;
    const uint32_t W = std::max<uint32_t>(1, R * R * R / 64);
    std::memset(out, 0, W * 8);
    for (uint32_t z = 0; z < R; ++z)
        for (uint32_t y = 0; y < R; ++y)
            for (uint32_t x = 0; x < R; ++x)
                if (structureGet(s, x, y, z, level, node)) {
                    const uint32_t i = x + R * y + R * R * z;
                    out[i >> 6] |= 1ull << (i & 63);
                }
}

// ---------------------------------------------------------------------------------------
// TransformPointer helpers (transform_dag.h) on a small struct.
// ---------------------------------------------------------------------------------------
struct TP {
    uint64_t ptr = 0;
    uint32_t sym[3] = { 0, 0, 0 };
    uint32_t perm[3] = { 0, 1, 2 };
    int t[3] = { 0, 0, 0 };
};

static uint64_t tpBits(const TP& tp)
{
    uint64_t v = tp.ptr & ((1ull << 37) - 1);
    v |= (uint64_t)tp.sym[0] << 37 | (uint64_t)tp.sym[1] << 38 | (uint64_t)tp.sym[2] << 39;
    v |= (uint64_t)tp.perm[0] << 40 | (uint64_t)tp.perm[1] << 42 | (uint64_t)tp.perm[2] << 44;
    v |= (uint64_t)((unsigned)(tp.t[0] + 32) & 63) << 46 | (uint64_t)((unsigned)(tp.t[1] + 32) & 63) << 52 | (uint64_t)((unsigned)(tp.t[2] + 32) & 63) << 58;
    return v;
}
static TP tpFromBits(uint64_t v)
{
    TP o;
    o.ptr = v & ((1ull << 37) - 1);
    for (int a = 0; a < 3; ++a) {
        o.sym[a] = (v >> (37 + a)) & 1;
        o.perm[a] = (v >> (40 + 2 * a)) & 3;
        o.t[a] = (int)((v >> (46 + 6 * a)) & 63) - 32;
    }
    return o;
}
static TP tpCreate(uint64_t ptr)
{
    TP o;
    o.ptr = ptr & ((1ull << 37) - 1);
    return o;
}
static const TP kSentinelTP = tpCreate(0xFFFFFFFFull);
static bool hasTransform(const TP& tp)
{
    const bool t = tp.t[0] || tp.t[1] || tp.t[2];
    const bool s = tp.sym[0] || tp.sym[1] || tp.sym[2];
    const bool p = !(tp.perm[0] == 0 && tp.perm[1] == 1 && tp.perm[2] == 2);
    return t || s || p;
}

static const uint32_t kPerms[6][3] = { { 0, 1, 2 }, { 0, 2, 1 }, { 1, 0, 2 }, { 1, 2, 0 }, { 2, 0, 1 }, { 2, 1, 0 } };
static uint32_t encodePerm(const uint32_t p[3]) { return 2 * p[0] + (p[2] < p[1] ? 1 : 0); }

// applyAxisPermutation(vec, perm) = (vec[p0], vec[p1], vec[p2]) (transform_dag.h:181-189)
static void permVec(const uint32_t in[3], const uint32_t perm[3], uint32_t out[3])
{
    out[0] = in[perm[0]];
    out[1] = in[perm[1]];
    out[2] = in[perm[2]];
}
static void invertPerm(const uint32_t p[3], uint32_t out[3]) // transform_dag.h:247-254
{
    out[p[0]] = 0;
    out[p[1]] = 1;
    out[p[2]] = 2;
}
// applyAxisPermutation(childIndex, permutation) (transform_dag.h:256-273)
static uint32_t permChild(uint32_t ci, const uint32_t p[3])
{
    uint32_t o = 0;
    if (ci & (1u << p[0]))
        o |= 1;
    if (ci & (1u << p[1]))
        o |= 2;
    if (ci & (1u << p[2]))
        o |= 4;
    return o;
}
// applySymmetry(childIndex, symmetry) (transform_dag.h:240-245): mirror the 2x2x2 position.
static uint32_t symChild(uint32_t ci, const uint32_t sym[3]) { return ci ^ (sym[0] | sym[1] << 1 | sym[2] << 2); }

// applyTransformationToPointer (transform_dag_hierarchical.cpp:60-66)
static TP transformPointer(const TP& tp, const uint32_t sym[3], const uint32_t perm[3])
{
    TP o = tp;
    uint32_t ps[3];
    permVec(tp.sym, perm, ps);
    for (int a = 0; a < 3; ++a)
        o.sym[a] = sym[a] ^ ps[a];
    uint32_t pp[3];
    permVec(tp.perm, perm, pp);
    for (int a = 0; a < 3; ++a)
        o.perm[a] = pp[a];
    return o;
}

struct Tagged { // TaggedPointer (transform_dag_hierarchical.cpp:68-123)
    TP tp;
    uint64_t inv = 0;
};
static Tagged taggedSentinel() { return Tagged { kSentinelTP, ~0ull }; }

static bool taggedEq(const Tagged& l, const Tagged& r)
{
    if (l.tp.ptr != r.tp.ptr)
        return false;
    if (l.tp.t[0] != r.tp.t[0] || l.tp.t[1] != r.tp.t[1] || l.tp.t[2] != r.tp.t[2] || l.inv != r.inv)
        throw std::runtime_error("TaggedPointer equality precondition violated");
    uint32_t lInvPerm[3];
    invertPerm(l.tp.perm, lInvPerm);
    uint32_t x[3];
    for (int a = 0; a < 3; ++a)
        x[a] = r.tp.sym[a] ^ l.tp.sym[a];
    uint32_t reqSym[3], reqPerm[3];
    permVec(x, lInvPerm, reqSym);
    permVec(r.tp.perm, lInvPerm, reqPerm);
    const uint32_t id = (reqSym[0] | reqSym[1] << 1 | reqSym[2] << 2) | (encodePerm(reqPerm) << 3);
    return (l.inv >> id) & 1ull;
}

struct TNode {
    Tagged c[8];
};
static bool nodeEq(const TNode& a, const TNode& b)
{
    for (int i = 0; i < 8; ++i)
        if (!taggedEq(a.c[i], b.c[i]))
            return false;
    return true;
}
static TNode nodeTransform(const TNode& n, const uint32_t sym[3], const uint32_t perm[3]) // :239-260
{
    TNode out;
    for (uint32_t i = 0; i < 8; ++i) {
        const uint32_t o = symChild(permChild(i, perm), sym);
        out.c[o] = n.c[i];
        out.c[o].tp = transformPointer(n.c[i].tp, sym, perm);
    }
    return out;
}

static inline void hashCombine(uint64_t& seed, uint64_t v) { seed ^= v + 0x9e3779b9ull + (seed << 6) + (seed >> 2); }

// 2x2x2 masks in Morton bit order (transform_dag_hierarchical.cpp:47-58).
static uint32_t transform2(uint32_t m, const uint32_t sym[3], const uint32_t perm[3])
{
    uint32_t out = 0;
    for (uint32_t i = 0; i < 8; ++i) {
        if (!((m >> i) & 1u))
            continue;
        const uint32_t v[3] = { i & 1u, (i >> 1) & 1u, (i >> 2) & 1u };
        uint32_t pv[3];
        permVec(v, perm, pv);
        const uint32_t o = (pv[0] ^ sym[0]) | (pv[1] ^ sym[1]) << 1 | (pv[2] ^ sym[2]) << 2;
        out |= 1u << o;
    }
    return out;
}

static std::array<Tagged, 256> canonicalLeaves(const orc_config& cfg) // :158-203
{
    std::array<uint64_t, 256> inv {};
    std::array<Tagged, 256> out;
    for (auto& t : out) {
        t.tp = TP {};
        t.tp.ptr = 0xFFFF;
    }
    for (uint32_t c = 0; c < 256; ++c)
        for (uint32_t id = 0; id < 48; ++id) {
            const uint32_t symId = id & 7, permId = id >> 3;
            const uint32_t sym[3] = { symId & 1, (symId >> 1) & 1, (symId >> 2) & 1 };
            const uint32_t* perm = kPerms[permId];
            const uint32_t tr = transform2(c, sym, perm);
            if (!cfg.axisPermutation && permId != 0)
                continue;
            if (!cfg.symmetry && symId != 0)
                continue;
            if (c < out[tr].tp.ptr) {
                TP tp;
                tp.ptr = c;
                for (int a = 0; a < 3; ++a) {
                    tp.sym[a] = sym[a];
                    tp.perm[a] = perm[a];
                }
                out[tr] = Tagged { tp, 0 };
            }
            if (tr == c)
                inv[c] |= 1ull << id;
        }
    for (auto& t : out)
        t.inv = inv[t.tp.ptr];
    return out;
}

struct NodeKeyHash {
    size_t operator()(const std::array<uint64_t, 8>& k) const
    {
        uint64_t seed = 0;
        for (uint64_t v : k)
            hashCombine(seed, v);
        return seed;
    }
};

// processLevel (transform_dag_hierarchical.cpp:359-448). outItems gets the input indices of
// kept nodes in emission order.
static void processLevel(const std::vector<TNode>& nodes, const orc_config& cfg, std::vector<uint32_t>& outItems, std::vector<Tagged>& outMapping)
{
    const size_t n = nodes.size();
    outMapping.assign(n, Tagged {});
    std::vector<std::pair<uint64_t, uint32_t>> keys(n);
    for (uint32_t i = 0; i < n; ++i) {
        std::array<uint64_t, 8> ptrs;
        for (int c = 0; c < 8; ++c)
            ptrs[c] = nodes[i].c[c].tp.ptr;
        std::sort(ptrs.begin(), ptrs.end());
        uint64_t seed = 0;
        for (uint64_t p : ptrs)
            hashCombine(seed, p);
        keys[i] = { seed, i };
    }
    std::sort(keys.begin(), keys.end());
    std::vector<Tagged> found(n);
    std::vector<bool> isFound(n, false);
    std::vector<uint64_t> invMasks(n, 0);
    size_t begin = 0;
    while (begin < n) {
        size_t end = begin;
        while (end < n && keys[end].first == keys[begin].first)
            ++end;
        // Group-local map: ptrs-in-position-order -> list of (node, index, invariance).
        std::unordered_map<std::array<uint64_t, 8>, std::vector<std::pair<uint32_t, uint64_t>>, NodeKeyHash> lut;
        for (size_t j = begin; j < end; ++j) {
            const uint32_t index = keys[j].second;
            const TNode& node = nodes[index];
            uint64_t invMask = 0;
            bool hit = false;
            for (uint32_t id = 0; id < 48 && !hit; ++id) {
                const uint32_t symId = id & 7, permId = id >> 3;
                const uint32_t sym[3] = { symId & 1, (symId >> 1) & 1, (symId >> 2) & 1 };
                const uint32_t* perm = kPerms[permId];
                if (!cfg.axisPermutation && permId != 0)
                    continue;
                if (!cfg.symmetry && symId != 0)
                    continue;
                const TNode tn = nodeTransform(node, sym, perm);
                std::array<uint64_t, 8> k;
                for (int c = 0; c < 8; ++c)
                    k[c] = tn.c[c].tp.ptr;
                auto it = lut.find(k);
                if (it != lut.end()) {
                    for (const auto& cand : it->second) {
                        if (nodeEq(tn, nodes[cand.first])) { // lookup key on the left, as unordered_map::find compares
                            uint32_t invPerm[3];
                            invertPerm(perm, invPerm);
                            TP tp = tpCreate(cand.first);
                            uint32_t s2[3];
                            permVec(sym, invPerm, s2);
                            for (int a = 0; a < 3; ++a) {
                                tp.sym[a] = s2[a];
                                tp.perm[a] = invPerm[a];
                            }
                            found[index] = Tagged { tp, cand.second };
                            isFound[index] = true;
                            invMask = cand.second;
                            hit = true;
                            break;
                        }
                    }
                }
                if (!hit && nodeEq(tn, node))
                    invMask |= 1ull << id;
            }
            if (!hit) {
                std::array<uint64_t, 8> k;
                for (int c = 0; c < 8; ++c)
                    k[c] = node.c[c].tp.ptr;
                lut[k].push_back({ index, invMask });
            }
            invMasks[index] = invMask;
        }
        begin = end;
    }
    for (const auto& kv : keys) {
        const uint32_t index = kv.second;
        if (isFound[index]) {
            Tagged t = found[index];
            t.tp.ptr = outMapping[t.tp.ptr].tp.ptr;
            outMapping[index] = t;
        } else {
            const uint32_t handle = (uint32_t)outItems.size();
            outItems.push_back(index);
            outMapping[index] = Tagged { tpCreate(handle), invMasks[index] };
        }
    }
}

// EditStructure<void, TransformPointer>
struct TDag {
    std::vector<std::vector<std::array<uint64_t, 8>>> nodesPerLevel;
    std::vector<uint64_t> subGrids;
    uint32_t resolution = 0, rootLevel = 0;
    int32_t structureType = 0;
};

// constructTransformDAGHierarchical (transform_dag_hierarchical.cpp:304-523)
static TDag hierarchical(const Svdag& in, uint32_t midLevel, std::vector<uint32_t>& midInv, const orc_config& cfg)
{
    TDag out;
    out.resolution = in.resolution;
    out.rootLevel = in.rootLevel;
    out.structureType = 0; // default-constructed output (:307): StructureType::Tree (structure.h:171)
    auto storeMid = [&](size_t numOut, const std::vector<Tagged>& mapping) {
        midInv.assign(numOut, 0);
        for (uint32_t old = 0; old < mapping.size(); ++old)
            if (!hasTransform(mapping[old].tp))
                midInv[mapping[old].tp.ptr] = old;
    };
    std::vector<Tagged> prev;
    {
        const auto canon = canonicalLeaves(cfg);
        std::vector<TNode> tnodes(in.subGrids.size());
        for (size_t i = 0; i < in.subGrids.size(); ++i)
            for (uint32_t k = 0; k < 8; ++k)
                tnodes[i].c[k] = canon[(in.subGrids[i] >> (8 * k)) & 0xFF];
        std::vector<uint32_t> kept;
        processLevel(tnodes, cfg, kept, prev);
        for (uint32_t k : kept)
            out.subGrids.push_back(in.subGrids[k]);
        if (midLevel == 2)
            storeMid(out.subGrids.size(), prev);
    }
    out.nodesPerLevel.resize(in.nodesPerLevel.size());
    for (uint32_t level = 3; level <= in.rootLevel; ++level) {
        const auto& inNodes = in.nodesPerLevel[level];
        std::vector<TNode> tnodes(inNodes.size());
        for (size_t i = 0; i < inNodes.size(); ++i)
            for (int c = 0; c < 8; ++c)
                tnodes[i].c[c] = inNodes[i][c] == kEmpty ? taggedSentinel() : prev[inNodes[i][c]];
        std::vector<uint32_t> kept;
        std::vector<Tagged> mapping;
        processLevel(tnodes, cfg, kept, mapping);
        auto& outNodes = out.nodesPerLevel[level];
        for (uint32_t k : kept) {
            std::array<uint64_t, 8> n;
            for (int c = 0; c < 8; ++c)
                n[c] = tpBits(tnodes[k].c[c].tp);
            outNodes.push_back(n);
        }
        prev = std::move(mapping);
        if (level == midLevel)
            storeMid(outNodes.size(), prev);
    }
    return out;
}

static std::array<uint64_t, 8> toTransformNode(const std::array<uint32_t, 8>& n) // transform_dag.cpp:123-135
{
    std::array<uint64_t, 8> o;
    for (int c = 0; c < 8; ++c)
        o[c] = n[c] == kEmpty ? tpBits(kSentinelTP) : tpBits(tpCreate(n[c]));
    return o;
}

// Matcher inputs recorded per level by the last orc_construct_tdag call (for tests that
// drive the GPU matcher with exactly the driver's inputs).
static std::vector<std::vector<uint64_t>> g_levelInputs(8);
static std::vector<std::vector<uint64_t>> g_levelEncoded(8); // encodedParents of each level's matcher call

// constructTransformDAG (transform_dag.cpp:138-240), with the oracle matcher.
static TDag construct(const Svdag& in, const orc_config& cfg, int nthreads)
{
    if (in.structureType != 1)
        throw std::runtime_error("input must be a DAG");
    std::vector<uint32_t> translationParents;
    const uint32_t maxT = cfg.maxTranslationLevel;
    TDag out = hierarchical(in, maxT + 1, translationParents, cfg);
    for (auto& v : g_levelInputs)
        v.clear();
    for (auto& v : g_levelEncoded)
        v.clear();
    if (!cfg.translation)
        return out;
    out.subGrids.clear();
    for (uint32_t level = 3; level <= maxT; ++level)
        out.nodesPerLevel[level].clear();
    if (out.nodesPerLevel[maxT + 1].size() != translationParents.size())
        throw std::runtime_error("mid level size mismatch");
    for (size_t i = 0; i < translationParents.size(); ++i)
        out.nodesPerLevel[maxT + 1][i] = toTransformNode(in.nodesPerLevel[maxT + 1][translationParents[i]]);
    for (uint32_t level = maxT; level >= 2; --level) {
        const size_t numItems = level == 2 ? in.subGrids.size() : in.nodesPerLevel[level].size();
        std::vector<bool> referenced(numItems, false);
        auto& parents = out.nodesPerLevel[level + 1];
        for (const auto& node : parents)
            for (uint64_t c : node)
                if (c != tpBits(kSentinelTP))
                    referenced[tpFromBits(c).ptr] = true;
        std::vector<uint32_t> indices;
        for (uint32_t i = 0; i < numItems; ++i)
            if (referenced[i])
                indices.push_back(i);
        const uint32_t R = 1u << level;
        const uint32_t W = std::max<uint32_t>(1, R * R * R / 64);
        std::vector<uint64_t> masks((size_t)indices.size() * W);
        for (size_t i = 0; i < indices.size(); ++i)
            flatten(in, level, indices[i], masks.data() + i * W);
        g_levelInputs[level] = masks;
        std::vector<uint64_t> enc(indices.size()), mapping(indices.size());
        std::vector<uint32_t> uniq(indices.size());
        uint64_t nu = 0;
        orc_match_encoded(level, masks.data(), indices.size(), &cfg, enc.data(), nthreads, 0);
        g_levelEncoded[level] = enc;
        orc_postprocess(level, masks.data(), indices.size(), enc.data(), mapping.data(), uniq.data(), &nu);
        if (level == 2) {
            out.subGrids.resize(nu);
            for (uint64_t k = 0; k < nu; ++k)
                out.subGrids[k] = in.subGrids[indices[uniq[k]]];
        } else {
            out.nodesPerLevel[level].resize(nu);
            for (uint64_t k = 0; k < nu; ++k)
                out.nodesPerLevel[level][k] = toTransformNode(in.nodesPerLevel[level][indices[uniq[k]]]);
        }
        for (auto& node : parents)
            for (uint64_t& c : node) {
                if (c == tpBits(kSentinelTP))
                    continue;
                const uint64_t ptr = tpFromBits(c).ptr;
                const auto it = std::lower_bound(indices.begin(), indices.end(), (uint32_t)ptr);
                c = mapping[it - indices.begin()];
            }
        if (level == 2)
            break;
    }
    return out;
}

// .tdag bytes: writer.write(EditStructure) then writer.write(config).
static std::vector<uint8_t> serialize(const TDag& t, const orc_config& cfg)
{
    std::vector<uint8_t> b;
    auto put = [&](const void* p, size_t n) {
        const uint8_t* c = (const uint8_t*)p;
        b.insert(b.end(), c, c + n);
    };
    const uint64_t nl = t.nodesPerLevel.size();
    put(&nl, 8);
    for (const auto& lvl : t.nodesPerLevel) {
        const uint64_t n = lvl.size();
        put(&n, 8);
        for (const auto& node : lvl)
            put(node.data(), 64);
    }
    const uint64_t ns = t.subGrids.size();
    put(&ns, 8);
    put(t.subGrids.data(), ns * 8);
    put(&t.resolution, 4);
    put(&t.rootLevel, 4);
    put(&t.structureType, 4);
    const uint8_t c4[4] = { cfg.symmetry, cfg.axisPermutation, cfg.translation, 0 };
    put(c4, 4);
    put(&cfg.maxTranslationLevel, 4);
    return b;
}

// ---------------------------------------------------------------------------------------
// StaticStructure path: constructStaticTransformDAG (transform_dag.cpp:242-390).
// ---------------------------------------------------------------------------------------

// StaticStructure<void, Child> (structure.h:299-320): per level one packed array; a node is its
// child-mask word followed by one word per present child, in child-index order. Pointers are word
// offsets of the child in the level below (leaf indices at level 3).
template <typename Word>
struct Static {
    std::vector<std::vector<Word>> nodesPerLevel;
    std::vector<uint64_t> subGrids;
    uint32_t resolution = 0, rootLevel = 0;
    int32_t structureType = 0; // StructureType::Tree unless set (structure.h:311)
};
using SSvdag = Static<uint32_t>; // StaticStructure<void, uint32_t>
using STDag = Static<uint64_t>; // StaticStructure<void, TransformPointer>

static inline uint32_t popc(uint32_t v) { return (uint32_t)__builtin_popcount(v); }
static const uint64_t kPtrMask = (1ull << 37) - 1;

// StaticStructure(const EditStructure&) (structure.h:766-798).
static SSvdag toStatic(const Svdag& in)
{
    SSvdag out;
    out.resolution = in.resolution;
    out.rootLevel = in.rootLevel;
    out.structureType = in.structureType;
    out.subGrids = in.subGrids;
    std::vector<uint32_t> prev(in.subGrids.size());
    std::iota(prev.begin(), prev.end(), 0u);
    out.nodesPerLevel.resize(in.nodesPerLevel.size());
    for (uint32_t level = 3; level <= in.rootLevel; ++level) {
        const auto& nodes = in.nodesPerLevel[level];
        auto& words = out.nodesPerLevel[level];
        std::vector<uint32_t> cur(nodes.size());
        for (size_t i = 0; i < nodes.size(); ++i) {
            const uint32_t at = (uint32_t)words.size();
            cur[i] = at;
            words.push_back(0);
            for (uint32_t c = 0; c < 8; ++c)
                if (nodes[i][c] != kEmpty) {
                    words.push_back(prev[nodes[i][c]]);
                    words[at] |= 1u << c;
                }
        }
        prev = std::move(cur);
    }
    return out;
}

// getLevelNodeStarts (structure.h:859-874).
template <typename Word>
static std::vector<uint32_t> levelStarts(const Static<Word>& s, uint32_t level)
{
    std::vector<uint32_t> out;
    if (level == 2) {
        out.resize(s.subGrids.size());
        std::iota(out.begin(), out.end(), 0u);
        return out;
    }
    const auto& w = s.nodesPerLevel[level];
    for (size_t cur = 0; cur < w.size(); cur += 1 + popc((uint32_t)w[cur]))
        out.push_back((uint32_t)cur);
    return out;
}

// StaticNode::getChildPointerAtIndex (structure.h:236-247); a TransformPointer child converts to
// its ptr field (transform_dag.h:339-342).
template <typename Word>
static uint32_t childAt(const std::vector<Word>& w, uint32_t start, uint32_t childIdx)
{
    const uint32_t mask = (uint32_t)w[start];
    const uint32_t off = popc(mask & ~(0xFFFFFFFFu << childIdx));
    return (uint32_t)((uint64_t)w[start + 1 + off] & kPtrMask);
}

// StaticStructure::get(voxel, level, root) (structure.h:836-855).
template <typename Word>
static bool staticGet(const Static<Word>& s, uint32_t x, uint32_t y, uint32_t z, uint32_t level, uint32_t root)
{
    uint32_t at = root;
    for (uint32_t l = level; l > 2; --l) {
        const auto& w = s.nodesPerLevel[l];
        const uint32_t cl = l - 1;
        const uint32_t childIdx = morton3((x >> cl) & 1u, (y >> cl) & 1u, (z >> cl) & 1u);
        if (!((w[at] >> childIdx) & 1u))
            return false;
        at = childAt(w, at, childIdx);
    }
    return (s.subGrids[at] >> morton3(x & 3u, y & 3u, z & 3u)) & 1ull;
}

// createLargeSubGrid over a StaticStructure (transform_dag.cpp:399-403).
template <typename Word>
static void staticFlatten(const Static<Word>& s, uint32_t level, uint32_t start, uint64_t* out)
{
    const uint32_t R = 1u << level;
    const uint32_t W = std::max<uint32_t>(1, R * R * R / 64);
    std::memset(out, 0, W * 8);
    for (uint32_t z = 0; z < R; ++z)
        for (uint32_t y = 0; y < R; ++y)
            for (uint32_t x = 0; x < R; ++x)
                if (staticGet(s, x, y, z, level, start)) {
                    const uint32_t i = x + R * y + R * R * z;
                    out[i >> 6] |= 1ull << (i & 63);
                }
}

// constructStaticTransformDAGHierarchical (transform_dag_hierarchical.cpp:531-743). The groups
// are keyed by the children's output word offsets, so the output order differs from the
// EditStructure pass while the set of kept nodes is the same.
static STDag staticHierarchical(const SSvdag& in, uint32_t midLevel, std::vector<uint32_t>& midStarts, const orc_config& cfg)
{
    STDag out; // default-constructed: StructureType::Tree (:533)
    out.resolution = in.resolution;
    out.rootLevel = in.rootLevel;
    using Map = std::unordered_map<uint32_t, Tagged>;
    auto storeMid = [&](const Map& m) { // :538-548
        std::vector<std::pair<uint32_t, Tagged>> canon;
        for (const auto& kv : m)
            if (!hasTransform(kv.second.tp))
                canon.push_back(kv);
        std::sort(canon.begin(), canon.end(), [](const auto& a, const auto& b) { return a.second.tp.ptr < b.second.tp.ptr; });
        midStarts.clear();
        for (const auto& kv : canon)
            midStarts.push_back(kv.first);
    };
    Map prev;
    {
        const auto canon = canonicalLeaves(cfg);
        std::vector<TNode> tnodes(in.subGrids.size());
        for (size_t i = 0; i < in.subGrids.size(); ++i)
            for (uint32_t k = 0; k < 8; ++k)
                tnodes[i].c[k] = canon[(in.subGrids[i] >> (8 * k)) & 0xFF];
        std::vector<uint32_t> kept;
        std::vector<Tagged> mapping;
        processLevel(tnodes, cfg, kept, mapping);
        for (uint32_t k : kept)
            out.subGrids.push_back(in.subGrids[k]);
        for (uint32_t i = 0; i < in.subGrids.size(); ++i)
            prev[i] = mapping[i];
        if (midLevel == 2)
            storeMid(prev);
    }
    out.nodesPerLevel.resize(in.nodesPerLevel.size());
    for (uint32_t level = 3; level <= in.rootLevel; ++level) {
        const auto& w = in.nodesPerLevel[level];
        const std::vector<uint32_t> starts = levelStarts(in, level);
        std::vector<TNode> tnodes(starts.size());
        for (size_t j = 0; j < starts.size(); ++j)
            for (uint32_t c = 0; c < 8; ++c)
                tnodes[j].c[c] = ((w[starts[j]] >> c) & 1u) ? prev.at(childAt(w, starts[j], c)) : taggedSentinel();
        std::vector<uint32_t> kept;
        std::vector<Tagged> mapping;
        processLevel(tnodes, cfg, kept, mapping);
        auto& ow = out.nodesPerLevel[level];
        std::vector<uint32_t> outStart(kept.size());
        for (size_t h = 0; h < kept.size(); ++h) { // :717-728
            outStart[h] = (uint32_t)ow.size();
            ow.push_back(w[starts[kept[h]]] & 0xFFFFFFFFu);
            for (uint32_t c = 0; c < 8; ++c)
                if (tnodes[kept[h]].c[c].tp.ptr != kSentinelTP.ptr)
                    ow.push_back(tpBits(tnodes[kept[h]].c[c].tp));
        }
        Map next; // :731-737
        for (size_t j = 0; j < starts.size(); ++j) {
            Tagged t = mapping[j];
            t.tp.ptr = outStart[t.tp.ptr];
            next[starts[j]] = t;
        }
        if (level == midLevel)
            storeMid(next);
        prev = std::move(next);
    }
    return out;
}

static uint64_t withPtr(uint64_t tp, uint64_t ptr) { return (tp & ~kPtrMask) | (ptr & kPtrMask); }

// constructStaticTransformDAG (transform_dag.cpp:242-390), with the oracle matcher.
static STDag constructStatic(const SSvdag& in, const orc_config& cfg, int nthreads)
{
    const uint32_t maxT = cfg.maxTranslationLevel;
    std::vector<uint32_t> translationParents;
    STDag out = staticHierarchical(in, maxT + 1, translationParents, cfg);
    for (auto& v : g_levelInputs)
        v.clear();
    if (!cfg.translation)
        return out;
    out.subGrids = in.subGrids; // :261-264
    for (uint32_t level = 3; level <= maxT + 1; ++level) { // :268-293
        const std::vector<uint32_t> starts = level == maxT + 1 ? translationParents : levelStarts(in, level);
        const auto& w = in.nodesPerLevel[level];
        auto& ow = out.nodesPerLevel[level];
        const size_t before = ow.size();
        ow.clear();
        for (uint32_t st : starts) {
            ow.push_back(w[st]);
            for (uint32_t c = 0; c < 8; ++c)
                if ((w[st] >> c) & 1u)
                    ow.push_back(tpBits(tpCreate(childAt(w, st, c))));
        }
        if (level == maxT + 1 && ow.size() != before)
            throw std::runtime_error("mid level size mismatch");
    }
    for (uint32_t level = maxT; level >= 2; --level) { // :298-387
        std::vector<uint32_t> itemStarts = levelStarts(out, level);
        const auto& pw = out.nodesPerLevel[level + 1];
        std::vector<bool> referenced(level == 2 ? out.subGrids.size() : out.nodesPerLevel[level].size(), false);
        for (size_t cur = 0; cur < pw.size(); cur += 1 + popc((uint32_t)pw[cur]))
            for (uint32_t c = 0; c < 8; ++c)
                if ((pw[cur] >> c) & 1u)
                    referenced[childAt(pw, (uint32_t)cur, c)] = true;
        size_t kept = 0;
        for (uint32_t st : itemStarts)
            if (referenced[st])
                itemStarts[kept++] = st;
        itemStarts.resize(kept);
        const uint32_t R = 1u << level;
        const uint32_t W = std::max<uint32_t>(1, R * R * R / 64);
        std::vector<uint64_t> masks((size_t)itemStarts.size() * W);
        for (size_t i = 0; i < itemStarts.size(); ++i)
            staticFlatten(out, level, itemStarts[i], masks.data() + i * W);
        g_levelInputs[level] = masks;
        std::vector<uint64_t> enc(itemStarts.size()), mapping(itemStarts.size());
        std::vector<uint32_t> uniq(itemStarts.size());
        uint64_t nu = 0;
        orc_match_encoded(level, masks.data(), itemStarts.size(), &cfg, enc.data(), nthreads, 0);
        orc_postprocess(level, masks.data(), itemStarts.size(), enc.data(), mapping.data(), uniq.data(), &nu);
        if (level == 2) {
            std::vector<uint64_t> sub(nu);
            for (uint64_t k = 0; k < nu; ++k)
                sub[k] = out.subGrids[itemStarts[uniq[k]]];
            out.subGrids = std::move(sub);
        } else {
            auto& ow = out.nodesPerLevel[level];
            std::vector<uint64_t> nw;
            std::vector<uint32_t> pointerMapping(nu);
            for (uint64_t k = 0; k < nu; ++k) {
                const uint32_t st = itemStarts[uniq[k]];
                pointerMapping[k] = (uint32_t)nw.size();
                for (uint32_t i = 0; i < 1 + popc((uint32_t)ow[st]); ++i)
                    nw.push_back(ow[st + i]);
            }
            ow = std::move(nw);
            for (auto& m : mapping)
                m = withPtr(m, pointerMapping[m & kPtrMask]);
        }
        auto& parents = out.nodesPerLevel[level + 1]; // :373-386
        for (size_t cur = 0; cur < parents.size(); cur += 1 + popc((uint32_t)parents[cur])) {
            const uint32_t nc = popc((uint32_t)parents[cur]);
            for (uint32_t o = 0; o < nc; ++o) {
                const uint32_t old = (uint32_t)(parents[cur + 1 + o] & kPtrMask);
                const auto it = std::lower_bound(itemStarts.begin(), itemStarts.end(), old);
                if (it == itemStarts.end() || *it != old)
                    throw std::runtime_error("parent pointer not among the level's items");
                parents[cur + 1 + o] = mapping[it - itemStarts.begin()];
            }
        }
        if (level == 2)
            break;
    }
    return out;
}

// StaticStructure::writeTo (structure.h:888-895) through BinaryWriter (binary_writer.h:48-78).
static std::vector<uint8_t> serializeStatic(const STDag& t)
{
    std::vector<uint8_t> b;
    auto put = [&](const void* p, size_t n) {
        const uint8_t* c = (const uint8_t*)p;
        b.insert(b.end(), c, c + n);
    };
    const uint64_t nl = t.nodesPerLevel.size();
    put(&nl, 8);
    for (const auto& lvl : t.nodesPerLevel) {
        const uint64_t n = lvl.size();
        put(&n, 8);
        put(lvl.data(), n * 8);
    }
    const uint64_t ns = t.subGrids.size();
    put(&ns, 8);
    put(t.subGrids.data(), ns * 8);
    put(&t.resolution, 4);
    put(&t.rootLevel, 4);
    put(&t.structureType, 4);
    return b;
}

} // namespace orc2

using namespace orc2;

extern "C" {

// Builds EditStructure<void,uint32_t>(VoxelGrid) and runs toDAG. `bits` is the dense grid,
// bit i = x + R*y + R*R*z. Returns an opaque handle through *out.
int orc_svdag_build(const uint8_t* bits, uint32_t resolution, void* out)
{
    if (!bits || !out || resolution < 4 || (resolution & (resolution - 1)))
        return 1;
    Svdag* s = new Svdag();
    s->resolution = resolution;
    uint32_t levels = 0;
    while ((1u << levels) < resolution)
        ++levels;
    s->rootLevel = levels;
    s->nodesPerLevel.resize(levels + 1);
    const Dense g { bits, resolution };
    buildRecurse(*s, g, (int)levels, 0, 0, 0);
    toDAG(*s);
    *(void**)out = s;
    return 0;
}

void orc_svdag_free(void* h) { delete (Svdag*)h; }

uint64_t orc_svdag_level_size(void* h, uint32_t level)
{
    const Svdag* s = (const Svdag*)h;
    if (level == 2)
        return s->subGrids.size();
    if (level >= s->nodesPerLevel.size())
        return 0;
    return s->nodesPerLevel[level].size();
}

void orc_svdag_copy_level(void* h, uint32_t level, uint32_t* out)
{
    const Svdag* s = (const Svdag*)h;
    const auto& nodes = s->nodesPerLevel[level];
    for (size_t i = 0; i < nodes.size(); ++i)
        std::memcpy(out + 8 * i, nodes[i].data(), 32);
}

void orc_svdag_copy_leaves(void* h, uint64_t* out)
{
    const Svdag* s = (const Svdag*)h;
    std::memcpy(out, s->subGrids.data(), s->subGrids.size() * 8);
}

uint32_t orc_svdag_root_level(void* h) { return ((const Svdag*)h)->rootLevel; }

int orc_flatten(void* h, uint32_t level, const uint32_t* nodes, uint64_t n, uint64_t* out)
{
    const Svdag* s = (const Svdag*)h;
    if (level < 2 || level > s->rootLevel)
        return 1;
    const uint32_t R = 1u << level;
    const uint32_t W = std::max<uint32_t>(1, R * R * R / 64);
    for (uint64_t i = 0; i < n; ++i)
        flatten(*s, level, nodes[i], out + i * W);
    return 0;
}

// Runs constructTransformDAG and writes the .tdag bytes to `path`. Returns the byte count, or
// a negative value on error.
int64_t orc_construct_tdag(void* h, const orc_config* cfg, int nthreads, const char* path)
{
    try {
        const TDag t = construct(*(const Svdag*)h, *cfg, nthreads);
        const std::vector<uint8_t> bytes = serialize(t, *cfg);
        FILE* f = std::fopen(path, "wb");
        if (!f)
            return -2;
        std::fwrite(bytes.data(), 1, bytes.size(), f);
        std::fclose(f);
        return (int64_t)bytes.size();
    } catch (const std::exception& e) {
        std::fprintf(stderr, "orc_construct_tdag: %s\n", e.what());
        return -1;
    }
}

// StaticStructure<void, uint32_t>(EditStructure) words of `level` (level 2: the leaf masks as
// uint64). Returns the word count; copies when `out` holds at least that many.
int64_t orc_svdag_static_level(void* h, uint32_t level, void* out, uint64_t cap)
{
    const SSvdag st = toStatic(*(const Svdag*)h);
    if (level == 2) {
        if (out && cap >= st.subGrids.size())
            std::memcpy(out, st.subGrids.data(), st.subGrids.size() * 8);
        return (int64_t)st.subGrids.size();
    }
    if (level >= st.nodesPerLevel.size())
        return -1;
    const auto& w = st.nodesPerLevel[level];
    if (out && cap >= w.size())
        std::memcpy(out, w.data(), w.size() * 4);
    return (int64_t)w.size();
}

// Runs constructStaticTransformDAG on the static form of the SVDAG and writes the
// StaticStructure<void, TransformPointer>::writeTo bytes to `path`. Returns the byte count, or a
// negative value on error.
int64_t orc_construct_static_tdag(void* h, const orc_config* cfg, int nthreads, const char* path)
{
    try {
        const Svdag& in = *(const Svdag*)h;
        if (cfg->maxTranslationLevel + 1 > in.rootLevel)
            return -3;
        const STDag t = constructStatic(toStatic(in), *cfg, nthreads);
        const std::vector<uint8_t> bytes = serializeStatic(t);
        FILE* f = std::fopen(path, "wb");
        if (!f)
            return -2;
        std::fwrite(bytes.data(), 1, bytes.size(), f);
        std::fclose(f);
        return (int64_t)bytes.size();
    } catch (const std::exception& e) {
        std::fprintf(stderr, "orc_construct_static_tdag: %s\n", e.what());
        return -1;
    }
}

// Copies the level-`level` matcher input masks of the last orc_construct_tdag call.
int64_t orc_last_level_inputs(uint32_t level, uint64_t* out, uint64_t cap)
{
    if (level >= g_levelInputs.size())
        return -1;
    const auto& v = g_levelInputs[level];
    if (out && cap >= v.size())
        std::memcpy(out, v.data(), v.size() * 8);
    return (int64_t)v.size();
}

// Copies the encodedParents (minimum keys) of the level-`level` matcher call of the last
// orc_construct_tdag call.
int64_t orc_last_level_encoded(uint32_t level, uint64_t* out, uint64_t cap)
{
    if (level >= g_levelEncoded.size())
        return -1;
    const auto& v = g_levelEncoded[level];
    if (out && cap >= v.size())
        std::memcpy(out, v.data(), v.size() * 8);
    return (int64_t)v.size();
}

} // extern "C"
