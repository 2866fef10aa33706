// The .tdag16 compact encoding of a Transform-Aware SVDAG, and its decoder on gfx950.
//
// Encoder (host): constructTransformDAG16 (tools/svdag_encoding/src/transform_dag_encoding.cpp:
// 241-478): reference-count ordering with pruning of unreferenced nodes (pointer_encoding.h:28-74),
// per-level length-limited Huffman code of the pointer types (huffman.cpp:93-202), pointer tables
// for frequently used pointers, and nodes of 16-bit words: a child mask of 2-bit size fields
// followed by the child pointers. Byte layout: TransformDAG16::writeTo
// (transform_dag_encoding.cpp:1020-1027, transform_dag_encoding.h:64-70, huffman.cpp:227-230).
//
// Tie rules. The reference leaves some orders to its standard library: pointers of equal reference
// count come out of an unordered_map and a parallel sort, and nodes of equal reference count and
// code words of equal length out of std::sort. Here nodes and code words use std::sort with the
// reference's comparators, and pointers of equal count keep the order of their first use (node
// order, then child order). Any such order decodes to the same volume.
//
// Decoder (device): one thread per 32 voxels of the dense output grid runs the reference's voxel
// lookup (transform_dag_encoding.cpp:897-942) for each: child index through the accumulated
// transform, pointer decoding (777-811, 138-216), the translation bounds test, and the transformed
// 4^3 leaf. All reads are bounds-checked; malformed data raises a flag instead of reading outside
// the buffers.
#include <algorithm>
#include <cstring>
#include <functional>
#include <numeric>
#include <stdexcept>
#include <unordered_map>

#include "common.hpp"
#include "context.hpp"
#include "errors.hpp"
#include "pipeline.hpp"

namespace tdag {

// ---- Huffman ------------------------------------------------------------------------------
std::vector<uint32_t> huffmanCodeLengths(const std::vector<uint64_t>& freq, uint32_t maxLen)
{
    // Package-merge over a node pool. Node order within a list: weight, leaves before packages,
    // leaves by symbol, packages by their members (huffman.cpp:118-132 defines the same order).
    struct Node {
        uint64_t w;
        int32_t sym; // -1 for a package
        int32_t a, b; // members of a package (pool indices)
    };
    std::vector<Node> pool;
    std::function<bool(int32_t, int32_t)> less = [&](int32_t x, int32_t y) -> bool {
        const Node &p = pool[x], &q = pool[y];
        if (p.w != q.w)
            return p.w < q.w;
        if ((p.sym >= 0) != (q.sym >= 0))
            return p.sym >= 0;
        if (p.sym >= 0)
            return p.sym < q.sym;
        if (less(p.a, q.a))
            return true;
        if (less(q.a, p.a))
            return false;
        return less(p.b, q.b);
    };
    const uint32_t n = (uint32_t)freq.size();
    std::vector<uint32_t> len(n, 0);
    if (n < 2 || maxLen == 0)
        return len;
    std::vector<std::vector<int32_t>> lists(maxLen);
    for (uint32_t l = 0; l < maxLen; ++l) {
        auto& cur = lists[l];
        if (l > 0) {
            const auto& prev = lists[l - 1];
            for (size_t i = 0; i + 1 < prev.size(); i += 2) {
                pool.push_back({ pool[prev[i]].w + pool[prev[i + 1]].w, -1, prev[i], prev[i + 1] });
                cur.push_back((int32_t)pool.size() - 1);
            }
        }
        for (uint32_t s = 0; s < n; ++s) {
            pool.push_back({ freq[s], (int32_t)s, -1, -1 });
            cur.push_back((int32_t)pool.size() - 1);
        }
        std::sort(cur.begin(), cur.end(), less);
    }
    size_t take = 2 * (size_t)n - 2;
    for (int l = (int)maxLen - 1; l >= 0; --l) {
        if (take > lists[l].size())
            return {};
        size_t packages = 0;
        for (size_t i = 0; i < take; ++i) {
            const Node& nd = pool[lists[l][i]];
            if (nd.sym >= 0)
                ++len[nd.sym];
            else
                ++packages;
        }
        take = 2 * packages;
    }
    return len;
}

std::vector<HuffCode> huffmanCodeWords(const std::vector<uint32_t>& len, bool lsbFirst)
{
    std::vector<uint32_t> order(len.size());
    std::iota(order.begin(), order.end(), 0u);
    std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return len[a] < len[b]; });
    const uint32_t maxLen = len.empty() ? 0 : *std::max_element(len.begin(), len.end());
    std::vector<HuffCode> out(len.size());
    uint64_t next = 0; // next free code, left-aligned to maxLen bits
    for (uint32_t s : order) {
        uint64_t c = next >> (maxLen - len[s]);
        if (lsbFirst) {
            uint64_t r = 0;
            for (uint32_t i = 0; i < len[s]; ++i)
                r |= ((c >> i) & 1ull) << (len[s] - 1 - i);
            c = r;
        }
        out[s] = HuffCode { c, len[s] };
        next += 1ull << (maxLen - len[s]);
    }
    return out;
}

namespace {

constexpr uint64_t kPtrMask = (1ull << 37) - 1;

struct TpFields {
    uint64_t ptr;
    uint32_t sym, perm;
    int t[3];
    bool transformed() const { return sym || perm || t[0] || t[1] || t[2]; }
};

TpFields tpFields(uint64_t v)
{
    TpFields f;
    f.ptr = v & kPtrMask;
    f.sym = (uint32_t)(v >> 37) & 7u;
    f.perm = permCode((uint32_t)(v >> 40) & 3u, (uint32_t)(v >> 42) & 3u, (uint32_t)(v >> 44) & 3u);
    for (int a = 0; a < 3; ++a)
        f.t[a] = (int)((v >> (46 + 6 * a)) & 63u) - 32;
    return f;
}

// encodePointerType (transform_dag_encoding.cpp:122-136).
uint32_t pointerType(const TpFields& f, uint32_t childLevel, const tdag_config& c)
{
    uint32_t out = 0;
    if (c.axis_permutation)
        out |= f.perm;
    if (c.symmetry)
        out = (out << 3) | f.sym;
    if (c.translation && childLevel <= c.max_translation_level)
        out = (out << 3) | (f.t[0] != 0 ? 1u : 0u) | (f.t[1] != 0 ? 2u : 0u) | (f.t[2] != 0 ? 4u : 0u);
    return out;
}

uint32_t bitWidth(uint64_t v) { return v ? 64u - (uint32_t)__builtin_clzll(v) : 0u; }
uint32_t divUp(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

// Reference-count order of one level's items, pruning items no parent references
// (pointer_encoding.h:28-74). Parents' child pointers are rewritten to the new indices.
template <typename Item>
void sortByReferenceCount(std::vector<Item>& items, std::vector<std::array<uint64_t, 8>>& parents)
{
    const uint64_t sentinel = tpSentinel();
    std::vector<uint32_t> refs(items.size(), 0);
    std::vector<char> used(items.size(), 0);
    for (const auto& p : parents)
        for (uint64_t c : p)
            if (c != sentinel) {
                const uint64_t i = c & kPtrMask;
                ++refs[i];
                used[i] = 1;
            }
    std::vector<uint32_t> order(items.size());
    std::iota(order.begin(), order.end(), 0u);
    std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return refs[a] > refs[b]; });
    std::vector<Item> sorted;
    sorted.reserve(items.size());
    std::vector<uint32_t> map(items.size(), 0);
    for (uint32_t i : order) {
        if (!used[i])
            continue;
        map[i] = (uint32_t)sorted.size();
        sorted.push_back(items[i]);
    }
    for (auto& p : parents)
        for (uint64_t& c : p)
            if (c != sentinel)
                c = (c & ~kPtrMask) | map[c & kPtrMask];
    items.swap(sorted);
}

void pushWords(std::vector<uint16_t>& out, uint64_t v, uint32_t words)
{
    for (uint32_t i = 0; i < words; ++i)
        out.push_back((uint16_t)(v >> (16 * i)));
}

} // namespace

Tdag16 encodeTdag16(const TDag& in, const tdag_config& cfg, bool tables, bool huffman)
{
    const uint64_t sentinel = tpSentinel();
    const uint32_t root = in.rootLevel;
    std::vector<std::vector<std::array<uint64_t, 8>>> nodes = in.nodes;
    const size_t numLevels = std::max<size_t>(nodes.size(), root + 1);
    nodes.resize(numLevels);
    std::vector<uint64_t> leaves = in.leaves;
    for (uint32_t level = root - 1; level > 2 && level < root; --level)
        sortByReferenceCount(nodes[level], nodes[level + 1]);
    if (root >= 3)
        sortByReferenceCount(leaves, nodes[3]);

    Tdag16 out;
    out.cfg = cfg;
    out.tables = tables;
    out.huffman = huffman;
    out.resolution = in.resolution;
    out.rootLevel = root;
    out.levels.resize(numLevels);
    out.ptrTables.resize(numLevels);
    if (huffman)
        out.codes.resize(numLevels);
    out.subgrids = std::move(leaves);

    std::vector<uint64_t> prevMap(out.subgrids.size());
    std::iota(prevMap.begin(), prevMap.end(), 0ull);
    uint32_t numTypes = (cfg.symmetry ? 8u : 1u) * (cfg.axis_permutation ? 6u : 1u) * (cfg.translation ? 8u : 1u);
    for (uint32_t level = 3; level <= root; ++level) {
        const uint32_t childLevel = level - 1;
        auto& inNodes = nodes[level];
        // Distinct child pointers with their reference counts, in order of first use.
        std::unordered_map<uint64_t, uint32_t> slot;
        std::vector<std::pair<uint32_t, uint64_t>> counted; // (references, pointer)
        for (auto& n : inNodes)
            for (uint64_t& c : n) {
                if (c == sentinel)
                    continue;
                c = (c & ~kPtrMask) | prevMap[c & kPtrMask];
                auto ins = slot.emplace(c, (uint32_t)counted.size());
                if (ins.second)
                    counted.push_back({ 0u, c });
                ++counted[ins.first->second].first;
            }
        std::stable_sort(counted.begin(), counted.end(), [](const auto& a, const auto& b) { return a.first > b.first; });

        std::vector<HuffCode> codes;
        if (huffman) {
            std::vector<uint64_t> hist(numTypes, 0);
            for (const auto& [refs, p] : counted)
                ++hist[pointerType(tpFields(p), childLevel, cfg)];
            codes = huffmanCodeWords(huffmanCodeLengths(hist, 16), true);
            out.codes[childLevel] = codes;
        }
        const auto encodePointer = [&](uint64_t v) -> uint64_t {
            const TpFields f = tpFields(v);
            const uint32_t tb = childLevel + 1;
            const int bias = 1 << childLevel;
            if (huffman) {
                const HuffCode code = codes[pointerType(f, childLevel, cfg)];
                uint64_t e = f.ptr;
                for (int a = 0; a < 3; ++a)
                    if (f.t[a])
                        e = (e << tb) | (uint64_t)(uint32_t)(f.t[a] + bias);
                return (e << code.bits) | code.code;
            }
            const bool tr = f.transformed();
            uint64_t e = f.ptr;
            if (cfg.axis_permutation && tr)
                e = (e << 3) | f.perm;
            if (cfg.symmetry && tr)
                e = (e << 3) | f.sym;
            if (cfg.translation && childLevel <= cfg.max_translation_level && tr)
                for (int a = 2; a >= 0; --a)
                    e = (e << tb) | (uint64_t)(uint32_t)(f.t[a] + bias);
            if (cfg.axis_permutation || cfg.symmetry || cfg.translation)
                e = (e << 1) | (tr ? 1u : 0u);
            return e;
        };

        // Each distinct pointer is stored inline or through the level's pointer table, whichever
        // costs fewer 16-bit words over all its references (transform_dag_encoding.cpp:384-431).
        std::unordered_map<uint64_t, std::pair<uint64_t, uint32_t>> stored; // pointer -> (bits, words)
        stored.reserve(counted.size());
        auto& table = out.ptrTables[childLevel];
        for (const auto& [refs, p] : counted) {
            const uint64_t e = encodePointer(p);
            if (tables) {
                const uint64_t inl = e << 1;
                const uint32_t inlWords = std::max(2u, divUp(bitWidth(inl), 16));
                uint32_t inlCost = refs * inlWords;
                if (inlWords > 3)
                    inlCost = UINT32_MAX;
                uint64_t tab;
                uint32_t tabWords;
                if (table.size() <= 0xFFFFu) {
                    tab = table.size();
                    tabWords = 1;
                } else {
                    tab = ((uint64_t)table.size() << 1) | 1u;
                    tabWords = divUp(bitWidth(table.size()), 16);
                }
                const uint32_t tabCost = refs * tabWords + 4;
                if (tabCost < inlCost) {
                    table.push_back(e);
                    stored[p] = { tab, tabWords };
                } else {
                    stored[p] = { inl, inlWords };
                }
            } else {
                const uint32_t words = std::max(1u, divUp(bitWidth(e), 16));
                if (words > 3)
                    throw std::runtime_error("a pointer does not fit in 48 bits without pointer tables");
                stored[p] = { e, words };
            }
        }

        auto& words = out.levels[level];
        std::vector<uint64_t> map;
        map.reserve(inNodes.size());
        for (const auto& n : inNodes) {
            map.push_back(words.size());
            const size_t maskAt = words.size();
            words.push_back(0);
            uint16_t mask = 0;
            for (uint32_t c = 0; c < 8; ++c) {
                if (n[c] == sentinel)
                    continue;
                const auto& [bits, w] = stored.find(n[c])->second;
                pushWords(words, bits, w);
                mask |= (uint16_t)(w << (2 * c));
            }
            words[maskAt] = mask;
        }
        prevMap.swap(map);
    }
    return out;
}

std::vector<uint8_t> serializeTdag16(const Tdag16& d)
{
    std::vector<uint8_t> b;
    auto put = [&](const void* p, size_t n) {
        const uint8_t* s = (const uint8_t*)p;
        b.insert(b.end(), s, s + n);
    };
    auto putU64 = [&](uint64_t v) { put(&v, 8); };
    put(&d.cfg, sizeof(tdag_config));
    const uint8_t enc[2] = { (uint8_t)d.tables, (uint8_t)d.huffman };
    put(enc, 2);
    putU64(d.levels.size());
    for (const auto& l : d.levels) {
        putU64(l.size());
        put(l.data(), l.size() * 2);
    }
    putU64(d.subgrids.size());
    put(d.subgrids.data(), d.subgrids.size() * 8);
    put(&d.resolution, 4);
    put(&d.rootLevel, 4);
    putU64(d.ptrTables.size());
    for (const auto& t : d.ptrTables) {
        putU64(t.size());
        put(t.data(), t.size() * 8);
    }
    putU64(d.codes.size());
    for (const auto& cs : d.codes) {
        putU64(cs.size());
        for (const HuffCode& c : cs) {
            const uint32_t pad = 0;
            put(&c.code, 8);
            put(&c.bits, 4);
            put(&pad, 4);
        }
    }
    return b;
}

bool parseTdag16(const uint8_t* p, size_t len, Tdag16& d, std::string& err)
{
    const uint8_t* end = p + len;
    bool ok = true;
    auto take = [&](void* dst, size_t n) {
        if (!ok || (size_t)(end - p) < n) {
            ok = false;
            return;
        }
        std::memcpy(dst, p, n);
        p += n;
    };
    auto u64 = [&]() {
        uint64_t v = 0;
        take(&v, 8);
        return v;
    };
    auto count = [&](size_t itemBytes, uint64_t limit) -> uint64_t {
        const uint64_t n = u64();
        if (ok && (n > limit || n > (uint64_t)(end - p) / itemBytes))
            ok = false;
        return ok ? n : 0;
    };
    take(&d.cfg, sizeof(tdag_config));
    uint8_t enc[2] = { 0, 0 };
    take(enc, 2);
    d.tables = enc[0] != 0;
    d.huffman = enc[1] != 0;
    d.levels.resize(count(1, 32));
    for (auto& l : d.levels) {
        l.resize(count(2, UINT64_MAX));
        take(l.data(), l.size() * 2);
    }
    d.subgrids.resize(count(8, UINT64_MAX));
    take(d.subgrids.data(), d.subgrids.size() * 8);
    take(&d.resolution, 4);
    take(&d.rootLevel, 4);
    d.ptrTables.resize(count(1, 32));
    for (auto& t : d.ptrTables) {
        t.resize(count(8, UINT64_MAX));
        take(t.data(), t.size() * 8);
    }
    d.codes.resize(count(1, 32));
    for (auto& cs : d.codes) {
        cs.resize(count(16, 4096));
        for (auto& c : cs) {
            uint32_t pad;
            take(&c.code, 8);
            take(&c.bits, 4);
            take(&pad, 4);
            // A code word must fit its length: the decode table indexes (suffix << bits) | code
            // over 2^maxLen entries.
            if (ok && (c.bits > 16 || (c.code >> c.bits) != 0))
                ok = false;
        }
    }
    if (!ok) {
        err = "truncated or malformed .tdag16 stream";
        return false;
    }
    if (d.rootLevel < 2 || d.rootLevel > 12 || d.rootLevel >= d.levels.size() || d.resolution != (1u << d.rootLevel)) {
        err = "inconsistent resolution / root level";
        return false;
    }
    if (d.huffman && d.codes.size() < d.rootLevel) {
        err = "missing Huffman code tables";
        return false;
    }
    return true;
}

// ---- device decoder -----------------------------------------------------------------------
namespace {

constexpr int kMaxLevels = 16;

struct Dev16 {
    const uint16_t* words;
    const uint64_t* subgrids;
    const uint64_t* tables;
    const uint32_t* lut; // symbol | code length << 16
    const uint8_t* xf; // child[48][8], combine[48*48], perm[48][3]
    uint32_t* bad;
    uint64_t levelOff[kMaxLevels], levelSize[kMaxLevels];
    uint64_t tableOff[kMaxLevels], tableSize[kMaxLevels];
    uint64_t lutOff[kMaxLevels];
    uint32_t lutMask[kMaxLevels];
    uint32_t hasLut; // bit per level
    uint64_t numSubgrids;
    uint32_t rootLevel, log2R, maxT;
    uint8_t sym, perm, trans, useTables, useHuffman;
};

constexpr uint32_t kXfChild = 0, kXfCombine = 48 * 8, kXfPerm = kXfCombine + 48 * 48, kXfBytes = kXfPerm + 48 * 3;

// The 48 child-order transforms (transform_dag_encoding.cpp:840-892): id = symmetry bits |
// axis-permutation code << 3; child[id][out] is the input child shown at output child `out`.
std::vector<uint8_t> transformTables()
{
    std::vector<uint8_t> x(kXfBytes, 0);
    for (uint32_t id = 0; id < 48; ++id) {
        for (uint32_t a = 0; a < 3; ++a)
            x[kXfPerm + id * 3 + a] = (uint8_t)permAxis(id >> 3, a);
        for (uint32_t in = 0; in < 8; ++in) {
            uint32_t out = 0;
            for (uint32_t a = 0; a < 3; ++a)
                out |= (((in >> permAxis(id >> 3, a)) ^ (id >> a)) & 1u) << a;
            x[kXfChild + id * 8 + out] = (uint8_t)in;
        }
    }
    for (uint32_t l = 0; l < 48; ++l)
        for (uint32_t r = 0; r < 48; ++r) {
            uint8_t c[8];
            for (uint32_t i = 0; i < 8; ++i)
                c[i] = x[kXfChild + r * 8 + x[kXfChild + l * 8 + i]];
            for (uint32_t id = 0; id < 48; ++id)
                if (std::memcmp(c, &x[kXfChild + id * 8], 8) == 0)
                    x[kXfCombine + l * 48 + r] = (uint8_t)id;
        }
    return x;
}

__device__ __forceinline__ void flagBad(const Dev16& d) { atomicOr(d.bad, 1u); }

// Child pointer of stored child `c` of node `node` at `level` (transform_dag_encoding.cpp:777-811,
// 138-216): 0 for an empty child; 1 with the child's index, transform id and translation (in the
// child's own frame) set; 2 for malformed data (flagged).
__device__ __forceinline__ int readChild(const Dev16& d, uint32_t level, uint64_t node, uint32_t c, uint64_t* child, uint32_t* outXid, int t[3])
{
    const uint32_t cl = level - 1;
    const uint64_t size = d.levelSize[level];
    const uint16_t* w = d.words + d.levelOff[level];
    if (node >= size) {
        flagBad(d);
        return 2;
    }
    const uint32_t mask = w[node];
    const uint32_t sz = (mask >> (2 * c)) & 3u;
    if (!sz)
        return 0;
    // Words before child c: the sum of the lower 2-bit size fields.
    const uint32_t below = mask & ((1u << (2 * c)) - 1u);
    const uint64_t at = node + 1 + __popc(below & 0x5555u) + 2 * __popc(below & 0xAAAAu);
    if (at + sz > size) {
        flagBad(d);
        return 2;
    }
    uint64_t e = 0;
    for (uint32_t i = 0; i < sz; ++i)
        e |= (uint64_t)w[at + i] << (16 * i);
    if (d.useTables && (sz == 1 || (e & 1ull))) {
        const uint64_t idx = sz == 1 ? e : e >> 1;
        if (idx >= d.tableSize[cl]) {
            flagBad(d);
            return 2;
        }
        e = d.tables[d.tableOff[cl] + idx];
    } else if (d.useTables) {
        e >>= 1;
    }
    const bool hasT = d.trans && cl <= d.maxT;
    const uint32_t tb = cl + 1;
    const uint64_t tm = (1ull << tb) - 1ull;
    const int bias = 1 << cl;
    t[0] = t[1] = t[2] = 0;
    uint32_t xid = 0;
    if (d.useHuffman) {
        if (!((d.hasLut >> cl) & 1u)) {
            flagBad(d);
            return 2;
        }
        const uint32_t ent = d.lut[d.lutOff[cl] + ((uint32_t)e & d.lutMask[cl])];
        uint32_t type = ent & 0xFFFFu;
        e >>= ent >> 16;
        uint32_t axes = 0;
        if (hasT) {
            axes = type & 7u;
            type >>= 3;
        }
        for (int a = 2; a >= 0; --a)
            if (axes & (1u << a)) {
                t[a] = (int)(e & tm) - bias;
                e >>= tb;
            }
        if (d.sym) {
            xid |= type & 7u;
            type >>= 3;
        }
        if (d.perm)
            xid |= (type & 7u) << 3;
    } else if (d.sym || d.perm || d.trans) {
        const bool tr = e & 1ull;
        e >>= 1;
        if (tr) {
            if (hasT)
                for (int a = 0; a < 3; ++a) {
                    t[a] = (int)(e & tm) - bias;
                    e >>= tb;
                }
            if (d.sym) {
                xid |= (uint32_t)e & 7u;
                e >>= 3;
            }
            if (d.perm) {
                xid |= ((uint32_t)e & 7u) << 3;
                e >>= 3;
            }
        }
    }
    if (xid >= 48) {
        flagBad(d);
        return 2;
    }
    *child = (uint32_t)e;
    *outXid = xid;
    return 1;
}

// The translation of a pointer in the frame of the transform `tid` accumulated above it.
__device__ __forceinline__ int frameTranslation(const uint8_t* xf, uint32_t tid, const int t[3], int a)
{
    const uint32_t src = xf[kXfPerm + tid * 3 + a];
    const int ta = src == 0 ? t[0] : (src == 1 ? t[1] : t[2]);
    return ((tid >> a) & 1u) ? -ta : ta;
}

// Value of one voxel (transform_dag_encoding.cpp:897-942).
__device__ uint32_t voxelValue(const Dev16& d, const uint8_t* xf, int vx, int vy, int vz)
{
    int v[3] = { vx, vy, vz };
    uint32_t tid = 0;
    uint64_t node = 0;
    for (uint32_t level = d.rootLevel; level > 2; --level) {
        const uint32_t cl = level - 1;
        const uint32_t c0 = (((uint32_t)v[0] >> cl) & 1u) | ((((uint32_t)v[1] >> cl) & 1u) << 1) | ((((uint32_t)v[2] >> cl) & 1u) << 2);
        const uint32_t c = xf[kXfChild + tid * 8 + c0];
        uint64_t child = 0;
        uint32_t xid = 0;
        int t[3];
        if (readChild(d, level, node, c, &child, &xid, t) != 1)
            return 0;
        node = child;
        if (t[0] | t[1] | t[2]) {
            // The translation in the frame of the transform accumulated so far; the voxel must stay
            // inside the same child cell.
            for (int a = 0; a < 3; ++a) {
                const uint32_t before = (uint32_t)v[a] >> cl;
                v[a] -= frameTranslation(xf, tid, t, a);
                if (((uint32_t)v[a] >> cl) != before)
                    return 0;
            }
        }
        tid = xf[kXfCombine + tid * 48 + xid];
    }
    if (node >= d.numSubgrids) {
        flagBad(d);
        return 0;
    }
    const uint32_t lo = ((uint32_t)v[0] & 1u) | (((uint32_t)v[1] & 1u) << 1) | (((uint32_t)v[2] & 1u) << 2);
    const uint32_t hi = (((uint32_t)v[0] >> 1) & 1u) | ((((uint32_t)v[1] >> 1) & 1u) << 1) | ((((uint32_t)v[2] >> 1) & 1u) << 2);
    const uint32_t bit = xf[kXfChild + tid * 8 + hi] * 8u + xf[kXfChild + tid * 8 + lo];
    return (uint32_t)(d.subgrids[node] >> bit) & 1u;
}

// One thread per 32-bit word of the dense grid (bit i = x + R*y + R*R*z).
__global__ __launch_bounds__(256) void kDecodeVolume(Dev16 d, uint32_t* out, uint64_t numWords)
{
    __shared__ uint8_t xf[kXfBytes];
    for (uint32_t i = threadIdx.x; i < kXfBytes; i += blockDim.x)
        xf[i] = d.xf[i];
    __syncthreads();
    const uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= numWords)
        return;
    const uint32_t k = d.log2R;
    const uint64_t rmask = (1ull << k) - 1ull;
    uint32_t word = 0;
    for (uint32_t b = 0; b < 32; ++b) {
        const uint64_t i = w * 32 + b;
        const int x = (int)(i & rmask), y = (int)((i >> k) & rmask), z = (int)(i >> (2 * k));
        word |= voxelValue(d, xf, x, y, z) << b;
    }
    out[w] = word;
}

// The same lookup for a whole run of 32 voxels along x (R >= 32, so a run shares y and z). Voxel
// x+1 follows voxel x's walk down to the first level where x, or x minus that level's translation,
// crosses a cell boundary of the level's children; only the levels below are walked again. The
// walk state of each level is kept in registers: the loops over levels are unrolled for the
// root level ROOT, so every state index is a compile-time constant.
template <int ROOT>
__global__ __launch_bounds__(256) void kDecodeRuns(Dev16 d, uint32_t* out, uint64_t numWords)
{
    __shared__ uint8_t xf[kXfBytes];
    for (uint32_t i = threadIdx.x; i < kXfBytes; i += blockDim.x)
        xf[i] = d.xf[i];
    __syncthreads();
    const uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= numWords)
        return;
    constexpr int N = ROOT - 2; // inner levels ROOT..3 at indices 0..N-1; index N is the leaf stage
    const uint64_t rmask = (1ull << ROOT) - 1ull;
    const uint64_t i0 = w * 32;
    const int x0 = (int)(i0 & rmask);
    // Entry state of the level at index li: x offset (translations subtracted so far), y, z, the
    // accumulated transform and the node; txA[li] is the x translation (global frame) applied there.
    int offX[N + 1], vyA[N + 1], vzA[N + 1], txA[N + 1];
    uint32_t tidA[N + 1];
    uint64_t nodeA[N + 1];
    offX[0] = 0;
    vyA[0] = (int)((i0 >> ROOT) & rmask);
    vzA[0] = (int)(i0 >> (2 * ROOT));
    tidA[0] = 0;
    nodeA[0] = 0;
#pragma unroll
    for (int li = 0; li <= N; ++li)
        txA[li] = 0;
    int stop = -1; // level index where the last walk ended with an empty result (-1: reached a leaf)
    uint32_t word = 0;
    for (int b = 0; b < 32; ++b) {
        const int x = x0 + b;
        int from = 0;
        if (b > 0) {
            // First level whose decision differs between voxel x-1 and x; levels below the one the
            // last walk stopped at hold no state for x-1.
            const int limit = stop >= 0 ? stop : N - 1;
            from = -1;
#pragma unroll
            for (int li = 0; li < N; ++li) {
                const int cl = ROOT - li - 1;
                const int e = x - offX[li];
                const bool moved = (e >> cl) != ((e - 1) >> cl) || (((e - txA[li]) >> cl) != ((e - 1 - txA[li]) >> cl));
                if (from < 0 && li <= limit && moved)
                    from = li;
            }
            if (from < 0)
                from = stop >= 0 ? -2 : N; // -2: the same empty result as voxel x-1
        }
        uint32_t bit = 0;
        if (from != -2) {
            stop = -1;
#pragma unroll
            for (int li = 0; li < N; ++li) {
                if (li < from || stop >= 0)
                    continue;
                const int level = ROOT - li;
                const uint32_t cl = (uint32_t)level - 1u;
                int v[3] = { x - offX[li], vyA[li], vzA[li] };
                const uint32_t tid = tidA[li];
                const uint64_t node = nodeA[li];
                txA[li] = 0;
                const uint64_t size = d.levelSize[level];
                const uint16_t* wds = d.words + d.levelOff[level];
                if (node >= size) {
                    flagBad(d);
                    stop = li;
                    continue;
                }
                const uint32_t c0 = (((uint32_t)v[0] >> cl) & 1u) | ((((uint32_t)v[1] >> cl) & 1u) << 1) | ((((uint32_t)v[2] >> cl) & 1u) << 2);
                const uint32_t c = xf[kXfChild + tid * 8 + c0];
                const uint32_t mask = wds[node];
                const uint32_t sz = (mask >> (2 * c)) & 3u;
                if (!sz) {
                    stop = li;
                    continue;
                }
                const uint32_t below = mask & ((1u << (2 * c)) - 1u);
                const uint64_t at = node + 1 + __popc(below & 0x5555u) + 2 * __popc(below & 0xAAAAu);
                if (at + sz > size) {
                    flagBad(d);
                    stop = li;
                    continue;
                }
                uint64_t e = 0;
                for (uint32_t i = 0; i < sz; ++i)
                    e |= (uint64_t)wds[at + i] << (16 * i);
                bool ok = true;
                if (d.useTables && (sz == 1 || (e & 1ull))) {
                    const uint64_t idx = sz == 1 ? e : e >> 1;
                    if (idx >= d.tableSize[cl])
                        ok = false;
                    else
                        e = d.tables[d.tableOff[cl] + idx];
                } else if (d.useTables) {
                    e >>= 1;
                }
                const bool hasT = d.trans && cl <= d.maxT;
                const uint32_t tb = cl + 1;
                const uint64_t tm = (1ull << tb) - 1ull;
                const int bias = 1 << cl;
                int t[3] = { 0, 0, 0 };
                uint32_t xid = 0;
                if (ok && d.useHuffman) {
                    if (!((d.hasLut >> cl) & 1u)) {
                        ok = false;
                    } else {
                        const uint32_t ent = d.lut[d.lutOff[cl] + ((uint32_t)e & d.lutMask[cl])];
                        uint32_t type = ent & 0xFFFFu;
                        e >>= ent >> 16;
                        uint32_t axes = 0;
                        if (hasT) {
                            axes = type & 7u;
                            type >>= 3;
                        }
                        if (axes & 4u) {
                            t[2] = (int)(e & tm) - bias;
                            e >>= tb;
                        }
                        if (axes & 2u) {
                            t[1] = (int)(e & tm) - bias;
                            e >>= tb;
                        }
                        if (axes & 1u) {
                            t[0] = (int)(e & tm) - bias;
                            e >>= tb;
                        }
                        if (d.sym) {
                            xid |= type & 7u;
                            type >>= 3;
                        }
                        if (d.perm)
                            xid |= (type & 7u) << 3;
                    }
                } else if (ok && (d.sym || d.perm || d.trans)) {
                    const bool tr = e & 1ull;
                    e >>= 1;
                    if (tr) {
                        if (hasT) {
                            t[0] = (int)(e & tm) - bias;
                            e >>= tb;
                            t[1] = (int)(e & tm) - bias;
                            e >>= tb;
                            t[2] = (int)(e & tm) - bias;
                            e >>= tb;
                        }
                        if (d.sym) {
                            xid |= (uint32_t)e & 7u;
                            e >>= 3;
                        }
                        if (d.perm) {
                            xid |= ((uint32_t)e & 7u) << 3;
                            e >>= 3;
                        }
                    }
                }
                if (!ok || xid >= 48) {
                    flagBad(d);
                    stop = li;
                    continue;
                }
                // The translation in the global frame: permuted and mirrored by the transform
                // accumulated so far.
                int tg[3];
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    const uint32_t src = xf[kXfPerm + tid * 3 + a];
                    int ta = src == 0 ? t[0] : (src == 1 ? t[1] : t[2]);
                    tg[a] = ((tid >> a) & 1u) ? -ta : ta;
                }
                txA[li] = tg[0];
                bool exits = false;
#pragma unroll
                for (int a = 0; a < 3; ++a)
                    exits |= (((uint32_t)v[a] >> cl) != ((uint32_t)(v[a] - tg[a]) >> cl));
                if (exits) {
                    stop = li;
                    continue;
                }
                offX[li + 1] = offX[li] + tg[0];
                vyA[li + 1] = v[1] - tg[1];
                vzA[li + 1] = v[2] - tg[2];
                tidA[li + 1] = xf[kXfCombine + tid * 48 + xid];
                nodeA[li + 1] = (uint32_t)e;
            }
            if (stop < 0) {
                const uint64_t node = nodeA[N];
                if (node >= d.numSubgrids) {
                    flagBad(d);
                } else {
                    const int v0 = x - offX[N], v1 = vyA[N], v2 = vzA[N];
                    const uint32_t tid = tidA[N];
                    const uint32_t lo = ((uint32_t)v0 & 1u) | (((uint32_t)v1 & 1u) << 1) | (((uint32_t)v2 & 1u) << 2);
                    const uint32_t hi = (((uint32_t)v0 >> 1) & 1u) | ((((uint32_t)v1 >> 1) & 1u) << 1) | ((((uint32_t)v2 >> 1) & 1u) << 2);
                    bit = (uint32_t)(d.subgrids[node] >> (xf[kXfChild + tid * 8 + hi] * 8u + xf[kXfChild + tid * 8 + lo])) & 1u;
                }
            }
        }
        word |= bit << b;
    }
    out[w] = word;
}

// ---- ray casting ------------------------------------------------------------------------------
// First voxel a ray enters, through the same lookup as the decoder (the renderer's use of the
// format: renderer/src/dags/transform_dag/transform_dag.cu:5-41, tracer_impl.h:874-1010). This
// traversal is its own design: every step walks from the root to the voxel the ray is in, the
// way voxelValue does, and stops at the first level where that voxel is empty: an empty child
// cell, or a translated child whose shifted content does not reach the voxel. The ray then jumps
// to the exit of that empty region (the whole child cell, or up to the part of the cell the
// shifted content covers), so empty space is skipped a cell at a time; inside a 4^3 leaf it steps
// voxel by voxel. Coordinates are integer voxels; only the jump times are floating point.

struct RayState {
    float o[3], dir[3], inv[3];
    float t; // current entry time
    int v[3]; // current voxel (world coordinates)
};

// Moves the ray to where it leaves the world box [lo, hi) (it is inside the box now): the new
// voxel is one past the box on every axis whose face is crossed first.
__device__ __forceinline__ void exitBox(RayState& r, const int lo[3], const int hi[3])
{
    float te[3];
    float tx = 3.0e38f;
    for (int a = 0; a < 3; ++a) {
        te[a] = r.dir[a] > 0.f ? ((float)hi[a] - r.o[a]) * r.inv[a] : (r.dir[a] < 0.f ? ((float)lo[a] - r.o[a]) * r.inv[a] : 3.0e38f);
        tx = fminf(tx, te[a]);
    }
    r.t = fmaxf(r.t, tx);
    for (int a = 0; a < 3; ++a) {
        if (te[a] == tx && r.dir[a] != 0.f) {
            r.v[a] = r.dir[a] > 0.f ? hi[a] : lo[a] - 1;
        } else {
            const int f = (int)floorf(r.o[a] + r.t * r.dir[a]);
            r.v[a] = min(max(f, lo[a]), hi[a] - 1);
        }
    }
}

// Moves the ray to where it enters the world box [lo, hi) ahead of it (entry time te > r.t).
__device__ __forceinline__ void enterBox(RayState& r, const int lo[3], const int hi[3], float te, const float ta[3])
{
    r.t = te;
    for (int a = 0; a < 3; ++a) {
        if (ta[a] == te && r.dir[a] != 0.f) {
            r.v[a] = r.dir[a] > 0.f ? lo[a] : hi[a] - 1;
        } else {
            const int f = (int)floorf(r.o[a] + r.t * r.dir[a]);
            r.v[a] = min(max(f, lo[a]), hi[a] - 1);
        }
    }
}

// Parametric interval of the ray inside the world box [lo, hi); per-axis entry times in ta.
__device__ __forceinline__ void boxInterval(const RayState& r, const int lo[3], const int hi[3], float* t0, float* t1, float ta[3])
{
    float a0 = -3.0e38f, a1 = 3.0e38f;
    for (int a = 0; a < 3; ++a) {
        if (r.dir[a] == 0.f) {
            ta[a] = -3.0e38f;
            if (r.o[a] < (float)lo[a] || r.o[a] >= (float)hi[a])
                a1 = -3.0e38f;
            continue;
        }
        float x0 = ((float)lo[a] - r.o[a]) * r.inv[a];
        float x1 = ((float)hi[a] - r.o[a]) * r.inv[a];
        if (x0 > x1) {
            const float tmp = x0;
            x0 = x1;
            x1 = tmp;
        }
        ta[a] = x0;
        a0 = fmaxf(a0, x0);
        a1 = fminf(a1, x1);
    }
    *t0 = a0;
    *t1 = a1;
}

// One thread per ray. rays: 8 floats each (origin xyz, tmin, direction xyz, tmax), voxel units,
// the volume is [0, R)^3. hits: 4 int32 (voxel x, y, z, steps taken) with x = -1 for no hit; ts:
// the entry distance of the hit voxel. A ray that runs out of its step budget (more than 4R + 64
// walks, which a correct traversal never needs) is reported as a miss with steps = -1.
__global__ __launch_bounds__(256) void kTrace(Dev16 d, const float* __restrict__ rays, uint64_t n, int32_t* __restrict__ hits, float* __restrict__ ts)
{
    __shared__ uint8_t xf[kXfBytes];
    for (uint32_t i = threadIdx.x; i < kXfBytes; i += blockDim.x)
        xf[i] = d.xf[i];
    __syncthreads();
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const float* in = rays + i * 8;
    RayState r;
    for (int a = 0; a < 3; ++a) {
        r.o[a] = in[a];
        r.dir[a] = in[4 + a];
        r.inv[a] = r.dir[a] != 0.f ? 1.0f / r.dir[a] : 0.f;
    }
    const int R = 1 << d.log2R;
    int32_t* h = hits + i * 4;
    h[0] = h[1] = h[2] = -1;
    h[3] = 0;
    ts[i] = 0.f;
    const int zero[3] = { 0, 0, 0 }, full[3] = { R, R, R };
    float t0, t1, ta[3];
    boxInterval(r, zero, full, &t0, &t1, ta);
    t0 = fmaxf(t0, in[3]);
    t1 = fminf(t1, in[7]);
    if (!(t0 < t1) || (r.dir[0] == 0.f && r.dir[1] == 0.f && r.dir[2] == 0.f))
        return; // the ray does not cross the volume
    r.t = t0;
    for (int a = 0; a < 3; ++a)
        r.v[a] = min(max((int)floorf(r.o[a] + t0 * r.dir[a]), 0), R - 1);
    const int budget = 4 * R + 64;
    for (int step = 0; step < budget; ++step) {
        if (r.t > t1 || r.v[0] < 0 || r.v[1] < 0 || r.v[2] < 0 || r.v[0] >= R || r.v[1] >= R || r.v[2] >= R) {
            h[3] = step;
            return; // left the volume or the ray's interval: no hit
        }
        // Walk from the root to the voxel, as voxelValue does, tracking the accumulated translation
        // and the world box in which every choice of this walk stays the same: the intersection of
        // the child cells taken and, below each translated pointer, of the part of its cell that
        // the shifted content covers.
        int vv[3] = { r.v[0], r.v[1], r.v[2] }; // virtual coordinates = world - shift
        int shift[3] = { 0, 0, 0 };
        int blo[3] = { 0, 0, 0 }, bhi[3] = { R, R, R }; // world box where this walk applies
        uint32_t tid = 0;
        uint64_t node = 0;
        int stopLevel = -1; // child level of an empty cell (-1: reached the leaf)
        bool clipped = false;
        int tt[3] = { 0, 0, 0 };
        int clo[3] = { 0, 0, 0 }; // world corner of the current child cell (its size is 2^cl)
        uint32_t cellLevel = 0;
        for (uint32_t level = d.rootLevel; level > 2; --level) {
            const uint32_t cl = level - 1;
            cellLevel = cl;
            const uint32_t c0 = (((uint32_t)vv[0] >> cl) & 1u) | ((((uint32_t)vv[1] >> cl) & 1u) << 1) | ((((uint32_t)vv[2] >> cl) & 1u) << 2);
            const uint32_t c = xf[kXfChild + tid * 8 + c0];
            for (int a = 0; a < 3; ++a) { // the child cell, in world coordinates
                clo[a] = ((vv[a] >> cl) << cl) + shift[a];
                blo[a] = max(blo[a], clo[a]);
                bhi[a] = min(bhi[a], clo[a] + (1 << cl));
            }
            uint64_t child = 0;
            uint32_t xid = 0;
            int t[3];
            const int rc = readChild(d, level, node, c, &child, &xid, t);
            if (rc == 2)
                return; // malformed data (flagged)
            if (rc == 0) {
                stopLevel = (int)cl;
                break;
            }
            if (t[0] | t[1] | t[2]) {
                for (int a = 0; a < 3; ++a)
                    tt[a] = frameTranslation(xf, tid, t, a);
                if ((((uint32_t)(vv[0] - tt[0])) >> cl) != ((uint32_t)vv[0] >> cl) || (((uint32_t)(vv[1] - tt[1])) >> cl) != ((uint32_t)vv[1] >> cl) ||
                    (((uint32_t)(vv[2] - tt[2])) >> cl) != ((uint32_t)vv[2] >> cl)) {
                    stopLevel = (int)cl;
                    clipped = true;
                    break;
                }
                for (int a = 0; a < 3; ++a) { // the part of the cell the shifted content covers
                    blo[a] = max(blo[a], clo[a] + tt[a]);
                    bhi[a] = min(bhi[a], clo[a] + (1 << cl) + tt[a]);
                    vv[a] -= tt[a];
                    shift[a] += tt[a];
                }
            }
            tid = xf[kXfCombine + tid * 48 + xid];
            node = child;
        }
        if (stopLevel >= 0) {
            if (!clipped) {
                exitBox(r, blo, bhi); // an empty child cell (within the walk's box)
                continue;
            }
            // Translated child whose content does not reach the voxel: its content covers the
            // part of the box that is also inside the box shifted by the translation. Jump to
            // where the ray enters that part if it is ahead, else leave the box (one voxel at a
            // time when float rounding disagrees with the integer test).
            int olo[3], ohi[3];
            for (int a = 0; a < 3; ++a) {
                olo[a] = max(blo[a], clo[a] + tt[a]);
                ohi[a] = min(bhi[a], clo[a] + (1 << cellLevel) + tt[a]);
            }
            float e0, e1, ea[3];
            boxInterval(r, olo, ohi, &e0, &e1, ea);
            if (olo[0] >= ohi[0] || olo[1] >= ohi[1] || olo[2] >= ohi[2] || e0 > e1 || e1 < r.t) {
                exitBox(r, blo, bhi);
            } else if (e0 > r.t) {
                enterBox(r, olo, ohi, e0, ea);
            } else {
                const int vlo[3] = { r.v[0], r.v[1], r.v[2] }, vhi[3] = { r.v[0] + 1, r.v[1] + 1, r.v[2] + 1 };
                exitBox(r, vlo, vhi);
            }
            continue;
        }
        // Leaf: step voxel by voxel inside its 4^3 box (within the walk's box).
        if (node >= d.numSubgrids) {
            flagBad(d);
            return;
        }
        const uint64_t leaf = d.subgrids[node];
        for (int a = 0; a < 3; ++a) {
            const int lo = ((vv[a] >> 2) << 2) + shift[a];
            blo[a] = max(blo[a], lo);
            bhi[a] = min(bhi[a], lo + 4);
        }
        for (;;) {
            const int w0 = r.v[0] - shift[0], w1 = r.v[1] - shift[1], w2 = r.v[2] - shift[2];
            const uint32_t lo = ((uint32_t)w0 & 1u) | (((uint32_t)w1 & 1u) << 1) | (((uint32_t)w2 & 1u) << 2);
            const uint32_t hi = (((uint32_t)w0 >> 1) & 1u) | ((((uint32_t)w1 >> 1) & 1u) << 1) | ((((uint32_t)w2 >> 1) & 1u) << 2);
            if ((leaf >> (xf[kXfChild + tid * 8 + hi] * 8u + xf[kXfChild + tid * 8 + lo])) & 1ull) {
                h[0] = r.v[0];
                h[1] = r.v[1];
                h[2] = r.v[2];
                h[3] = step;
                ts[i] = r.t;
                return;
            }
            const int vlo[3] = { r.v[0], r.v[1], r.v[2] }, vhi[3] = { r.v[0] + 1, r.v[1] + 1, r.v[2] + 1 };
            exitBox(r, vlo, vhi);
            if (++step >= budget || r.t > t1)
                break;
            if (r.v[0] < blo[0] || r.v[1] < blo[1] || r.v[2] < blo[2] || r.v[0] >= bhi[0] || r.v[1] >= bhi[1] || r.v[2] >= bhi[2])
                break;
        }
    }
    h[3] = -1; // step budget exhausted
}

} // namespace
} // namespace tdag

// Opaque handle of an encoded DAG, with its device copy (built on first decode).
struct tdag_tdag16 {
    tdag::Tdag16 e;
    int devDevice = -1;
    void* dev = nullptr;
    tdag::Dev16 d {};
    ~tdag_tdag16()
    {
        if (dev) {
            int cur = 0;
            const bool restore = hipGetDevice(&cur) == hipSuccess;
            if (hipSetDevice(devDevice) == hipSuccess)
                (void)hipFree(dev);
            if (restore)
                (void)hipSetDevice(cur);
        }
    }
};

namespace tdag {

tdag_tdag16* newTdag16Handle(Tdag16&& e)
{
    auto* h = new tdag_tdag16();
    h->e = std::move(e);
    return h;
}

namespace {

tdag_status err(tdag_status s, const std::string& m) { return setError(s, m); }
tdag_status hipErr(hipError_t e, const char* what)
{
    return setError(e == hipErrorOutOfMemory ? TDAG_ENOMEM : TDAG_EDEVICE, std::string(what) + ": " + hipGetErrorName(e));
}

// One device buffer: xf tables, bad flag, words, subgrids, tables, LUTs (256-byte aligned parts).
tdag_status upload(tdag_context* ctx, tdag_tdag16* h)
{
    const Tdag16& e = h->e;
    if (e.levels.size() > (size_t)kMaxLevels || e.rootLevel >= (uint32_t)kMaxLevels)
        return err(TDAG_EINVAL, "too many levels for the device decoder");
    Dev16 d {};
    auto al = [](uint64_t b) { return (b + 255) & ~255ull; };
    uint64_t off = 0;
    const uint64_t xfAt = off;
    off += al(kXfBytes);
    const uint64_t badAt = off;
    off += 256;
    const uint64_t wordsAt = off;
    uint64_t nw = 0;
    for (size_t l = 0; l < e.levels.size(); ++l) {
        d.levelOff[l] = nw;
        d.levelSize[l] = e.levels[l].size();
        nw += e.levels[l].size();
    }
    off += al(nw * 2 + 8);
    const uint64_t subAt = off;
    d.numSubgrids = e.subgrids.size();
    off += al(e.subgrids.size() * 8 + 8);
    const uint64_t tabAt = off;
    uint64_t nt = 0;
    for (size_t l = 0; l < e.ptrTables.size() && l < (size_t)kMaxLevels; ++l) {
        d.tableOff[l] = nt;
        d.tableSize[l] = e.ptrTables[l].size();
        nt += e.ptrTables[l].size();
    }
    off += al(nt * 8 + 8);
    const uint64_t lutAt = off;
    std::vector<uint32_t> lut;
    for (size_t l = 0; e.huffman && l < e.codes.size() && l < (size_t)kMaxLevels; ++l) {
        const auto& cs = e.codes[l];
        if (cs.empty())
            continue;
        uint32_t maxLen = 0;
        for (const auto& c : cs) {
            if (c.bits > 16 || (c.code >> c.bits) != 0)
                return hipErr(hipErrorInvalidValue, "Huffman code word longer than its length");
            maxLen = std::max(maxLen, c.bits);
        }
        d.hasLut |= 1u << l;
        d.lutOff[l] = lut.size();
        d.lutMask[l] = (1u << maxLen) - 1u;
        const size_t base = lut.size();
        lut.resize(base + (1ull << maxLen), 0u);
        // HuffmanDecodeLUT (huffman.cpp:204-218): every completion of a code word's high bits.
        for (uint32_t s = 0; s < cs.size(); ++s)
            for (uint64_t suf = 0; suf < (1ull << (maxLen - cs[s].bits)); ++suf)
                lut[base + ((suf << cs[s].bits) | cs[s].code)] = s | (cs[s].bits << 16);
    }
    off += al(lut.size() * 4 + 8);

    void* buf = nullptr;
    if (hipError_t x = hipMalloc(&buf, off); x != hipSuccess)
        return hipErr(x, "hipMalloc of the .tdag16 device copy");
    uint8_t* b = (uint8_t*)buf;
    std::vector<uint8_t> host(off, 0);
    const auto xf = transformTables();
    std::memcpy(&host[xfAt], xf.data(), xf.size());
    for (size_t l = 0; l < e.levels.size(); ++l)
        if (!e.levels[l].empty())
            std::memcpy(&host[wordsAt + d.levelOff[l] * 2], e.levels[l].data(), e.levels[l].size() * 2);
    if (!e.subgrids.empty())
        std::memcpy(&host[subAt], e.subgrids.data(), e.subgrids.size() * 8);
    for (size_t l = 0; l < e.ptrTables.size() && l < (size_t)kMaxLevels; ++l)
        if (!e.ptrTables[l].empty())
            std::memcpy(&host[tabAt + d.tableOff[l] * 8], e.ptrTables[l].data(), e.ptrTables[l].size() * 8);
    if (!lut.empty())
        std::memcpy(&host[lutAt], lut.data(), lut.size() * 4);
    if (hipError_t x = hipMemcpy(buf, host.data(), off, hipMemcpyHostToDevice); x != hipSuccess) {
        (void)hipFree(buf);
        return hipErr(x, "upload of the .tdag16 device copy");
    }
    d.xf = b + xfAt;
    d.bad = (uint32_t*)(b + badAt);
    d.words = (const uint16_t*)(b + wordsAt);
    d.subgrids = (const uint64_t*)(b + subAt);
    d.tables = (const uint64_t*)(b + tabAt);
    d.lut = (const uint32_t*)(b + lutAt);
    d.rootLevel = e.rootLevel;
    d.log2R = e.rootLevel;
    d.maxT = e.cfg.max_translation_level;
    d.sym = e.cfg.symmetry;
    d.perm = e.cfg.axis_permutation;
    d.trans = e.cfg.translation;
    d.useTables = e.tables;
    d.useHuffman = e.huffman;
    h->dev = buf;
    h->devDevice = ctx->device;
    h->d = d;
    return TDAG_OK;
}

// The device copy of the encoded DAG on ctx's device (made on first use; ctx's device current).
tdag_status ensureDeviceCopy(tdag_context* ctx, tdag_tdag16* h)
{
    if (hipError_t x = hipSetDevice(ctx->device); x != hipSuccess)
        return hipErr(x, "hipSetDevice");
    if (h->devDevice == ctx->device)
        return TDAG_OK;
    if (h->dev) {
        const int cur = ctx->device;
        if (hipSetDevice(h->devDevice) == hipSuccess)
            (void)hipFree(h->dev);
        (void)hipSetDevice(cur);
        h->dev = nullptr;
        h->devDevice = -1;
    }
    return upload(ctx, h);
}

} // namespace
} // namespace tdag

extern "C" {

tdag_status tdag_huffman_code_lengths(const uint64_t* freq, uint32_t n, uint32_t max_len, uint32_t* out_len)
{
    if ((!freq || !out_len) && n)
        return tdag::err(TDAG_EINVAL, "null argument");
    if (max_len == 0 || max_len > 32)
        return tdag::err(TDAG_EINVAL, "max_len must be in [1, 32]");
    const auto len = tdag::huffmanCodeLengths(std::vector<uint64_t>(freq, freq + n), max_len);
    if (len.size() != n)
        return tdag::err(TDAG_EINVAL, "more symbols than codes of at most max_len bits");
    std::copy(len.begin(), len.end(), out_len);
    return TDAG_OK;
}

tdag_status tdag_huffman_code_words(const uint32_t* len, uint32_t n, int lsb_first, uint64_t* out_code)
{
    if ((!len || !out_code) && n)
        return tdag::err(TDAG_EINVAL, "null argument");
    for (uint32_t i = 0; i < n; ++i)
        if (len[i] > 63)
            return tdag::err(TDAG_EINVAL, "code length above 63");
    const auto codes = tdag::huffmanCodeWords(std::vector<uint32_t>(len, len + n), lsb_first != 0);
    for (uint32_t i = 0; i < n; ++i)
        out_code[i] = codes[i].code;
    return TDAG_OK;
}

tdag_status tdag_tdag16_from_bytes(const uint8_t* bytes, uint64_t len, tdag_tdag16** out)
{
    if (!bytes || !out)
        return tdag::err(TDAG_EINVAL, "null argument");
    tdag::Tdag16 e;
    std::string why;
    if (!tdag::parseTdag16(bytes, len, e, why))
        return tdag::err(TDAG_EINVAL, why);
    *out = tdag::newTdag16Handle(std::move(e));
    return TDAG_OK;
}

tdag_status tdag_tdag16_serialize(const tdag_tdag16* h, uint8_t* buf, uint64_t cap, uint64_t* out_size)
{
    if (!h || !out_size)
        return tdag::err(TDAG_EINVAL, "null argument");
    const std::vector<uint8_t> bytes = tdag::serializeTdag16(h->e);
    *out_size = bytes.size();
    if (buf) {
        if (cap < bytes.size())
            return tdag::err(TDAG_EINVAL, "buffer too small");
        std::memcpy(buf, bytes.data(), bytes.size());
    }
    return TDAG_OK;
}

tdag_status tdag_tdag16_info(const tdag_tdag16* h, uint32_t* resolution, uint32_t* root_level, uint64_t* level_words, uint64_t* table_entries, uint32_t cap)
{
    if (!h)
        return tdag::err(TDAG_EINVAL, "null argument");
    if (resolution)
        *resolution = h->e.resolution;
    if (root_level)
        *root_level = h->e.rootLevel;
    for (uint32_t l = 0; l < cap; ++l) {
        if (level_words)
            level_words[l] = l == 2 ? h->e.subgrids.size() * 4 : (l < h->e.levels.size() ? h->e.levels[l].size() : 0);
        if (table_entries)
            table_entries[l] = l < h->e.ptrTables.size() ? h->e.ptrTables[l].size() : 0;
    }
    return TDAG_OK;
}

tdag_status tdag_tdag16_decode_device(tdag_context* ctx, const tdag_tdag16* hc, uint32_t* d_out_words, uint64_t out_words, float* out_ms)
{
    if (!ctx || !hc || !d_out_words)
        return tdag::err(TDAG_EINVAL, "null argument");
    auto* h = const_cast<tdag_tdag16*>(hc);
    const uint64_t R = h->e.resolution;
    const uint64_t need = std::max<uint64_t>(1, R * R * R / 32);
    if (R < 4 || out_words < need)
        return tdag::err(TDAG_EINVAL, "output buffer smaller than resolution^3 / 32 words");
    if (tdag_status s = tdag::ensureDeviceCopy(ctx, h); s != TDAG_OK)
        return s;
    hipStream_t st = ctx->stream;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (out_ms) {
        (void)hipEventCreate(&e0);
        (void)hipEventCreate(&e1);
        (void)hipEventRecord(e0, st);
    }
    const uint32_t zero = 0;
    hipError_t x = hipMemcpyAsync(h->d.bad, &zero, 4, hipMemcpyHostToDevice, st);
    if (x == hipSuccess) {
        const uint64_t blocks = (need + 255) / 256;
        const dim3 grid((uint32_t)blocks), block(256);
        switch (h->e.rootLevel) {
        case 5: hipLaunchKernelGGL(tdag::kDecodeRuns<5>, grid, block, 0, st, h->d, d_out_words, need); break;
        case 6: hipLaunchKernelGGL(tdag::kDecodeRuns<6>, grid, block, 0, st, h->d, d_out_words, need); break;
        case 7: hipLaunchKernelGGL(tdag::kDecodeRuns<7>, grid, block, 0, st, h->d, d_out_words, need); break;
        case 8: hipLaunchKernelGGL(tdag::kDecodeRuns<8>, grid, block, 0, st, h->d, d_out_words, need); break;
        case 9: hipLaunchKernelGGL(tdag::kDecodeRuns<9>, grid, block, 0, st, h->d, d_out_words, need); break;
        case 10: hipLaunchKernelGGL(tdag::kDecodeRuns<10>, grid, block, 0, st, h->d, d_out_words, need); break;
        case 11: hipLaunchKernelGGL(tdag::kDecodeRuns<11>, grid, block, 0, st, h->d, d_out_words, need); break;
        case 12: hipLaunchKernelGGL(tdag::kDecodeRuns<12>, grid, block, 0, st, h->d, d_out_words, need); break;
        default: // resolution below 32: a word spans several rows
            hipLaunchKernelGGL(tdag::kDecodeVolume, grid, block, 0, st, h->d, d_out_words, need);
        }
        x = hipGetLastError();
    }
    uint32_t bad = 0;
    if (x == hipSuccess && out_ms)
        x = hipEventRecord(e1, st);
    if (x == hipSuccess)
        x = hipMemcpyAsync(&bad, h->d.bad, 4, hipMemcpyDeviceToHost, st);
    if (x == hipSuccess)
        x = hipStreamSynchronize(st);
    if (out_ms) {
        if (x == hipSuccess)
            (void)hipEventElapsedTime(out_ms, e0, e1);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
    if (x != hipSuccess)
        return tdag::hipErr(x, ".tdag16 device decode");
    if (bad)
        return tdag::err(TDAG_EINVAL, "malformed .tdag16 data (an index outside its level, table or leaf array)");
    return TDAG_OK;
}

tdag_status tdag_tdag16_trace_device(tdag_context* ctx, const tdag_tdag16* hc, const float* d_rays, uint64_t n, int32_t* d_hits, float* d_t, float* out_ms)
{
    if (!ctx || !hc || (n && (!d_rays || !d_hits || !d_t)))
        return tdag::err(TDAG_EINVAL, "null argument");
    auto* h = const_cast<tdag_tdag16*>(hc);
    if (h->e.resolution < 4)
        return tdag::err(TDAG_EINVAL, "resolution below 4");
    if (tdag_status s = tdag::ensureDeviceCopy(ctx, h); s != TDAG_OK)
        return s;
    if (n == 0)
        return TDAG_OK;
    hipStream_t st = ctx->stream;
    hipError_t x = hipEventRecord(ctx->ev[0], st);
    const uint32_t zero = 0;
    if (x == hipSuccess)
        x = hipMemcpyAsync(h->d.bad, &zero, 4, hipMemcpyHostToDevice, st);
    if (x == hipSuccess) {
        hipLaunchKernelGGL(tdag::kTrace, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, st, h->d, d_rays, n, d_hits, d_t);
        x = hipGetLastError();
    }
    if (x == hipSuccess)
        x = hipEventRecord(ctx->ev[3], st);
    uint32_t bad = 0;
    if (x == hipSuccess)
        x = hipMemcpyAsync(&bad, h->d.bad, 4, hipMemcpyDeviceToHost, st);
    if (x == hipSuccess)
        x = hipStreamSynchronize(st);
    if (x != hipSuccess)
        return tdag::hipErr(x, ".tdag16 ray casting");
    if (out_ms)
        (void)hipEventElapsedTime(out_ms, ctx->ev[0], ctx->ev[3]);
    if (bad)
        return tdag::err(TDAG_EINVAL, "malformed .tdag16 data (an index outside its level, table or leaf array)");
    return TDAG_OK;
}

tdag_status tdag_tdag16_free(tdag_tdag16* h)
{
    delete h;
    return TDAG_OK;
}

} // extern "C"
