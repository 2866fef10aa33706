// Hierarchical symmetry + axis-permutation dedup (host code).
//
// Semantics of constructTransformDAGHierarchical (transform_dag_hierarchical.cpp:304-523):
// every level is reduced to canonical nodes under the 48 mirror/permutation transforms,
// bottom-up. Nodes are grouped by a hash_combine of their sorted child pointers
// (hash.h:10-15 with libstdc++'s identity std::hash<uint64_t>); inside a group, in (hash,
// index) order, a node takes the first transform ID (mirror | perm << 3) under which it equals
// an already kept node, where child equality allows the child's own invariances. Kept nodes
// are numbered in global (hash, index) order.
//
// This implementation keeps every child's (mirror, permutation) as one of 48 "states" and
// precomputes the 48 x 48 state-composition and required-invariance tables, so a node
// transform is 8 table lookups and a child comparison is one lookup and one bit test.
#include <algorithm>
#include <atomic>
#include <cstring>
#include <thread>
#include <unordered_map>

#include "pipeline.hpp"

namespace tdag {

namespace {

const uint32_t kPermTab[6][3] = { { 0, 1, 2 }, { 0, 2, 1 }, { 1, 0, 2 }, { 1, 2, 0 }, { 2, 0, 1 }, { 2, 1, 0 } }; // transform_dag.h:138-145

uint32_t permCodeOf(const uint32_t p[3]) { return 2 * p[0] + (p[2] < p[1] ? 1u : 0u); } // transform_dag.h:326-331

// State = mirror bits | permCode << 3 (the same numbering as the transform IDs).
struct Tables {
    uint8_t compose[48][48]; // compose[id][state]: child state after applying transform id
    uint8_t childPos[48][8]; // childPos[id][i]: where child i lands under transform id
    uint8_t required[48][48]; // required[l][r]: invariance ID needed for l == r
    uint8_t inverseId[48]; // the pointer (mirror, perm) stored on a hit with transform id
    Tables()
    {
        for (uint32_t id = 0; id < 48; ++id) {
            const uint32_t sym = id & 7u;
            const uint32_t* perm = kPermTab[id >> 3];
            for (uint32_t i = 0; i < 8; ++i) {
                // applyAxisPermutation(childIndex, perm) then applySymmetry (transform_dag.h:240-273).
                uint32_t o = 0;
                for (uint32_t a = 0; a < 3; ++a)
                    if (i & (1u << perm[a]))
                        o |= 1u << a;
                childPos[id][i] = (uint8_t)(o ^ sym);
            }
            for (uint32_t st = 0; st < 48; ++st) {
                // applyTransformationToPointer (transform_dag_hierarchical.cpp:60-66).
                const uint32_t cs = st & 7u;
                const uint32_t* cp = kPermTab[st >> 3];
                uint32_t ns = 0, np[3];
                for (uint32_t a = 0; a < 3; ++a) {
                    ns |= (((sym >> a) & 1u) ^ ((cs >> perm[a]) & 1u)) << a;
                    np[a] = cp[perm[a]];
                }
                compose[id][st] = (uint8_t)(ns | permCodeOf(np) << 3);
            }
            // Hit with transform (sym, perm): pointer gets perm^-1 and applyAxisPermutation(sym, perm^-1).
            uint32_t inv[3];
            inv[perm[0]] = 0;
            inv[perm[1]] = 1;
            inv[perm[2]] = 2;
            uint32_t s2 = 0;
            for (uint32_t a = 0; a < 3; ++a)
                s2 |= ((sym >> inv[a]) & 1u) << a;
            inverseId[id] = (uint8_t)(s2 | permCodeOf(inv) << 3);
        }
        for (uint32_t l = 0; l < 48; ++l)
            for (uint32_t r = 0; r < 48; ++r) {
                // TaggedPointer::operator== (transform_dag_hierarchical.cpp:79-122).
                const uint32_t* lp = kPermTab[l >> 3];
                const uint32_t* rp = kPermTab[r >> 3];
                uint32_t linv[3];
                linv[lp[0]] = 0;
                linv[lp[1]] = 1;
                linv[lp[2]] = 2;
                const uint32_t x = (r & 7u) ^ (l & 7u);
                uint32_t rs = 0, rq[3];
                for (uint32_t a = 0; a < 3; ++a) {
                    rs |= ((x >> linv[a]) & 1u) << a;
                    rq[a] = rp[linv[a]];
                }
                required[l][r] = (uint8_t)(rs | permCodeOf(rq) << 3);
            }
    }
};

const Tables& tables()
{
    static const Tables t;
    return t;
}

struct Child {
    uint64_t ptr; // 0xFFFFFFFF for an empty child
    uint64_t inv; // invariance mask of the canonical node `ptr` (all ones for empty)
    uint32_t state;
};
struct Node {
    Child c[8];
};

constexpr uint64_t kSentinelPtr = 0xFFFFFFFFull;

// Child equality of TaggedPointer::operator== (transform_dag_hierarchical.cpp:79-122): same
// pointer, and the pointer's node is invariant under the transform that maps one state onto the
// other.
inline bool childEq(const Child& l, const Child& r, const Tables& T)
{
    if (l.ptr != r.ptr)
        return false;
    return (l.inv >> T.required[l.state][r.state]) & 1ull;
}

inline void hashCombine(uint64_t& seed, uint64_t v) { seed ^= v + 0x9e3779b9ull + (seed << 6) + (seed >> 2); }

// nodeEq(transformNode(n, id), other) without building the transformed node: child i of n lands
// at position childPos[id][i] with its state composed with id; stops at the first mismatch.
inline bool transformedEq(const Node& n, uint32_t id, const Node& other, const Tables& T)
{
    for (uint32_t i = 0; i < 8; ++i) {
        Child c = n.c[i];
        c.state = T.compose[id][c.state];
        if (!childEq(c, other.c[T.childPos[id][i]], T))
            return false;
    }
    return true;
}

// Runs f(begin, end) over [0, n) split into one contiguous range per thread.
template <typename F>
void parallelFor(size_t n, int threads, F&& f)
{
    const size_t t = (size_t)std::max(1, threads);
    if (n < 8192 || t == 1) {
        f(size_t(0), n);
        return;
    }
    std::vector<std::thread> pool;
    const size_t chunk = (n + t - 1) / t;
    for (size_t i = 0; i < t; ++i) {
        const size_t b = std::min(n, chunk * i), e = std::min(n, chunk * (i + 1));
        if (b < e)
            pool.emplace_back([&f, b, e]() { f(b, e); });
    }
    for (auto& th : pool)
        th.join();
}

// Sorts unique (key, index) pairs: sorted chunks, merged pairwise. Pairs are distinct, so the
// result is the single ascending order whatever the split.
void parallelSortPairs(std::vector<std::pair<uint64_t, uint32_t>>& v, int threads)
{
    const size_t n = v.size();
    size_t parts = 1;
    while (parts < (size_t)std::max(1, threads) && parts < 64)
        parts <<= 1;
    if (n < 16384 || parts == 1) {
        std::sort(v.begin(), v.end());
        return;
    }
    std::vector<size_t> bounds(parts + 1);
    for (size_t i = 0; i <= parts; ++i)
        bounds[i] = n * i / parts;
    {
        std::vector<std::thread> pool;
        for (size_t i = 0; i < parts; ++i)
            pool.emplace_back([&, i]() { std::sort(v.begin() + bounds[i], v.begin() + bounds[i + 1]); });
        for (auto& th : pool)
            th.join();
    }
    for (size_t width = 1; width < parts; width <<= 1) {
        std::vector<std::thread> pool;
        for (size_t i = 0; i + width < parts; i += 2 * width) {
            const size_t b = bounds[i], m = bounds[i + width], e = bounds[std::min(parts, i + 2 * width)];
            pool.emplace_back([&v, b, m, e]() { std::inplace_merge(v.begin() + b, v.begin() + m, v.begin() + e); });
        }
        for (auto& th : pool)
            th.join();
    }
}

uint64_t stateToTp(uint64_t ptr, uint32_t state)
{
    const uint32_t* p = kPermTab[state >> 3];
    uint64_t v = ptr & ((1ull << 37) - 1);
    v |= (uint64_t)(state & 7u) << 37;
    v |= (uint64_t)p[0] << 40 | (uint64_t)p[1] << 42 | (uint64_t)p[2] << 44;
    v |= 32ull << 46 | 32ull << 52 | 32ull << 58;
    return v;
}

struct PtrKeyHash {
    size_t operator()(const std::array<uint64_t, 8>& k) const
    {
        uint64_t s = 0;
        for (uint64_t v : k)
            hashCombine(s, v);
        return (size_t)s;
    }
};

struct Mapping {
    uint64_t ptr;
    uint32_t state;
    uint64_t inv;
};

// Reduces one level. `kept` receives the input indices of kept nodes in output order and
// `mapping[i]` the pointer of input node i. Nodes are grouped by a hash of their sorted child
// pointers; with `hashOf`, a child pointer p enters that hash as hashOf[p] (the empty child as
// itself), which is how the StaticStructure pass keys its groups by child word offsets.
void reduceLevel(const std::vector<Node>& nodes, uint32_t allowedIds, const std::vector<uint32_t>& idList, int threads,
    std::vector<uint32_t>& kept, std::vector<Mapping>& mapping, const std::vector<uint64_t>* hashOf = nullptr)
{
    const Tables& T = tables();
    const size_t n = nodes.size();
    std::vector<std::pair<uint64_t, uint32_t>> keys(n);
    parallelFor(n, threads, [&](size_t kb, size_t ke) {
    for (size_t i = kb; i < ke; ++i) {
        uint64_t ptrs[8];
        for (int c = 0; c < 8; ++c) {
            const uint64_t p = nodes[i].c[c].ptr;
            ptrs[c] = (hashOf && p != kSentinelPtr) ? (*hashOf)[p] : p;
        }
        std::sort(ptrs, ptrs + 8);
        uint64_t seed = 0;
        for (uint64_t p : ptrs)
            hashCombine(seed, p);
        keys[i] = { seed, (uint32_t)i };
    }
    });
    parallelSortPairs(keys, threads);
    std::vector<size_t> groupStart;
    for (size_t j = 0; j < n; ++j)
        if (j == 0 || keys[j].first != keys[j - 1].first)
            groupStart.push_back(j);
    groupStart.push_back(n);
    // Per input node: hit (found canonical input index + transform id) or kept (invariance).
    std::vector<int32_t> hitIndex(n, -1);
    std::vector<uint8_t> hitId(n, 0);
    std::vector<uint64_t> invMask(n, 0);
    std::atomic<size_t> nextGroup { 0 };
    auto worker = [&]() {
        std::unordered_map<std::array<uint64_t, 8>, std::vector<uint32_t>, PtrKeyHash> lut;
        for (;;) {
            const size_t g = nextGroup.fetch_add(1);
            if (g + 1 >= groupStart.size())
                break;
            const size_t b = groupStart[g], e = groupStart[g + 1];
            lut.clear();
            for (size_t j = b; j < e; ++j) {
                const uint32_t index = keys[j].second;
                const Node& node = nodes[index];
                uint64_t inv = 0;
                bool hit = false;
                for (uint32_t id : idList) {
                    if (e - b > 1) {
                        std::array<uint64_t, 8> k;
                        for (uint32_t i = 0; i < 8; ++i)
                            k[T.childPos[id][i]] = node.c[i].ptr;
                        auto it = lut.find(k);
                        if (it != lut.end()) {
                            for (uint32_t cand : it->second)
                                if (transformedEq(node, id, nodes[cand], T)) {
                                    hitIndex[index] = (int32_t)cand;
                                    hitId[index] = (uint8_t)id;
                                    inv = invMask[cand];
                                    hit = true;
                                    break;
                                }
                        }
                        if (hit)
                            break;
                    }
                    if (transformedEq(node, id, node, T))
                        inv |= 1ull << id;
                }
                invMask[index] = inv;
                if (!hit && e - b > 1) {
                    std::array<uint64_t, 8> k;
                    for (int c = 0; c < 8; ++c)
                        k[c] = node.c[c].ptr;
                    lut[k].push_back(index);
                }
            }
        }
    };
    if (n < 4096 || threads <= 1) {
        worker(); // small levels: starting threads would cost more than the work
    } else {
        std::vector<std::thread> pool;
        for (int t = 0; t < threads; ++t)
            pool.emplace_back(worker);
        for (auto& th : pool)
            th.join();
    }
    (void)allowedIds;
    mapping.assign(n, Mapping { 0, 0, 0 });
    kept.clear();
    for (const auto& kv : keys) {
        const uint32_t index = kv.second;
        if (hitIndex[index] >= 0) {
            const uint32_t cand = (uint32_t)hitIndex[index];
            mapping[index] = Mapping { mapping[cand].ptr, T.inverseId[hitId[index]], invMask[index] };
        } else {
            mapping[index] = Mapping { (uint64_t)kept.size(), 0u, invMask[index] };
            kept.push_back(index);
        }
    }
}

// 256-entry canonical 2x2x2 table (transform_dag_hierarchical.cpp:158-203): the canonical
// representative is the smallest byte in the orbit, reached by the first transform ID.
void canonical2x2x2(const std::vector<uint32_t>& idList, Child out[256])
{
    uint64_t inv[256] = {};
    uint32_t bestPtr[256];
    uint32_t bestState[256];
    for (int i = 0; i < 256; ++i) {
        bestPtr[i] = 0xFFFF;
        bestState[i] = 0;
    }
    for (uint32_t c = 0; c < 256; ++c)
        for (uint32_t id : idList) {
            const uint32_t sym = id & 7u;
            const uint32_t* perm = kPermTab[id >> 3];
            uint32_t tr = 0;
            for (uint32_t i = 0; i < 8; ++i) {
                if (!((c >> i) & 1u))
                    continue;
                const uint32_t v[3] = { i & 1u, (i >> 1) & 1u, (i >> 2) & 1u };
                const uint32_t o = (v[perm[0]] ^ (sym & 1u)) | (v[perm[1]] ^ ((sym >> 1) & 1u)) << 1 | (v[perm[2]] ^ ((sym >> 2) & 1u)) << 2;
                tr |= 1u << o;
            }
            if (c < bestPtr[tr]) {
                bestPtr[tr] = c;
                bestState[tr] = id;
            }
            if (tr == c)
                inv[c] |= 1ull << id;
        }
    for (int b = 0; b < 256; ++b)
        out[b] = Child { bestPtr[b], inv[bestPtr[b]], bestState[b] };
}

} // namespace

void hierarchicalTables(const tdag_config& cfg, HierTables& out)
{
    const Tables& T = tables();
    std::memcpy(out.compose, T.compose, sizeof(out.compose));
    std::memcpy(out.childPos, T.childPos, sizeof(out.childPos));
    std::memcpy(out.required, T.required, sizeof(out.required));
    std::memcpy(out.inverseId, T.inverseId, sizeof(out.inverseId));
    std::vector<uint32_t> idList;
    out.numIds = 0;
    for (uint32_t id = 0; id < 48; ++id) {
        if (!cfg.axis_permutation && (id >> 3) != 0)
            continue;
        if (!cfg.symmetry && (id & 7u) != 0)
            continue;
        idList.push_back(id);
        out.ids[out.numIds++] = (uint8_t)id;
    }
    Child canon[256];
    canonical2x2x2(idList, canon);
    for (int b = 0; b < 256; ++b) {
        out.leafPtr[b] = (uint32_t)canon[b].ptr;
        out.leafInv[b] = canon[b].inv;
        out.leafState[b] = (uint8_t)canon[b].state;
    }
}

TDag hierarchicalDedup(const Svdag& in, uint32_t midLevel, std::vector<uint32_t>& outMidInv, const tdag_config& cfg, int threads, bool staticOffsets)
{
    std::vector<uint32_t> idList;
    for (uint32_t id = 0; id < 48; ++id) {
        if (!cfg.axis_permutation && (id >> 3) != 0)
            continue;
        if (!cfg.symmetry && (id & 7u) != 0)
            continue;
        idList.push_back(id);
    }
    TDag out;
    out.resolution = in.resolution;
    out.rootLevel = in.rootLevel;
    // The reference's output structure is default-constructed (transform_dag_hierarchical.cpp:307)
    // and no later step sets its type, so a .tdag records StructureType::Tree (structure.h:171).
    out.structureType = 0;
    out.nodes.resize(in.nodes.size());
    auto storeMid = [&](size_t numOut, const std::vector<Mapping>& m) {
        // transform_dag_hierarchical.cpp:311-318: canonical (untransformed) pointers only.
        outMidInv.assign(numOut, 0);
        for (uint32_t old = 0; old < m.size(); ++old)
            if (m[old].state == 0)
                outMidInv[m[old].ptr] = old;
    };
    std::vector<Mapping> prev;
    {
        Child canon[256];
        canonical2x2x2(idList, canon);
        std::vector<Node> leaves(in.leaves.size());
        parallelFor(in.leaves.size(), threads, [&](size_t b, size_t e) {
            for (size_t i = b; i < e; ++i)
                for (uint32_t k = 0; k < 8; ++k)
                    leaves[i].c[k] = canon[(in.leaves[i] >> (8 * k)) & 0xFFu];
        });
        std::vector<uint32_t> kept;
        reduceLevel(leaves, 0, idList, threads, kept, prev);
        out.leaves.resize(kept.size());
        for (size_t k = 0; k < kept.size(); ++k)
            out.leaves[k] = in.leaves[kept[k]];
        if (midLevel == 2)
            storeMid(out.leaves.size(), prev);
    }
    for (uint32_t level = 3; level <= in.rootLevel; ++level) {
        const auto& inNodes = in.nodes[level];
        std::vector<Node> nodes(inNodes.size());
        parallelFor(inNodes.size(), threads, [&](size_t b, size_t e) {
            for (size_t i = b; i < e; ++i)
                for (int c = 0; c < 8; ++c) {
                    const uint32_t ch = inNodes[i][c];
                    nodes[i].c[c] = (ch == kEmptyChild) ? Child { kSentinelPtr, ~0ull, 0u } : Child { prev[ch].ptr, prev[ch].inv, prev[ch].state };
                }
        });
        std::vector<uint32_t> kept;
        std::vector<Mapping> mapping;
        // constructStaticTransformDAGHierarchical (transform_dag_hierarchical.cpp:691-704) points
        // at inner children by their word offset in the output level below (leaf indices stay).
        std::vector<uint64_t> offsets;
        if (staticOffsets && level >= 4)
            offsets = packedNodeOffsets(out.nodes[level - 1]);
        reduceLevel(nodes, 0, idList, threads, kept, mapping, (staticOffsets && level >= 4) ? &offsets : nullptr);
        auto& outNodes = out.nodes[level];
        outNodes.resize(kept.size());
        parallelFor(kept.size(), threads, [&](size_t b, size_t e) {
            for (size_t k = b; k < e; ++k)
                for (int c = 0; c < 8; ++c) {
                    const Child& ch = nodes[kept[k]].c[c];
                    outNodes[k][c] = stateToTp(ch.ptr, ch.state);
                }
        });
        prev.swap(mapping);
        if (level == midLevel)
            storeMid(outNodes.size(), prev);
    }
    return out;
}

} // namespace tdag
