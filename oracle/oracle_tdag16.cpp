// CPU oracle of the .tdag16 compact encoding: its length-limited Huffman code and its decoder.
// TEST INFRASTRUCTURE ONLY: used by tests/, the smoke check and the CPU baseline, never by the
// product path.
//
// Restated from the reference (tools/svdag_encoding/src/):
//   huffman.cpp:113-183    length-limited code lengths by package-merge (Moffat 2019, alg. 3)
//   huffman.cpp:93-111     canonical code words from lengths (Moffat 2019, sec. 2.5)
//   huffman.cpp:192-202    conversion of the code words to LSB-first bit order
//   huffman.cpp:204-218    decode table indexed by the low maxLength bits
//   transform_dag_encoding.cpp:1020-1036, transform_dag_encoding.h:64-78  byte layout
//   transform_dag_encoding.cpp:777-811     child pointer lookup (2-bit size fields, tables)
//   transform_dag_encoding.cpp:138-216     pointer decoding with and without the Huffman prefix
//   transform_dag_encoding.cpp:840-942     voxel lookup: 48-transform composition, translation
//                                          bounds test, transformed 4^3 leaf
// Pinning: the Huffman functions against the reference's own known-answer tests
// (huffman.cpp:246-365); the decoder by decoding encoded scenes back to the dense grids they were
// built from, voxel for voxel.
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <set>
#include <vector>

#include "oracle.h"

namespace {

// ---- package-merge ------------------------------------------------------------------------
// A node is a leaf (symbol) or a package of two nodes. Nodes of one level are ordered by weight,
// then leaves before packages, leaves by symbol, packages by their two members in order.
struct PMNode {
    uint64_t weight;
    int symbol; // >= 0 for a leaf
    std::vector<PMNode> members;
    bool operator<(const PMNode& o) const
    {
        if (weight != o.weight)
            return weight < o.weight;
        const bool leaf = symbol >= 0, oleaf = o.symbol >= 0;
        if (leaf != oleaf)
            return leaf;
        if (leaf)
            return symbol < o.symbol;
        return std::lexicographical_compare(members.begin(), members.end(), o.members.begin(), o.members.end());
    }
};

std::vector<uint32_t> packageMerge(const std::vector<uint64_t>& freq, uint32_t maxLen)
{
    const size_t n = freq.size();
    std::vector<std::set<PMNode>> lists(maxLen);
    auto addLeaves = [&](std::set<PMNode>& s) {
        for (size_t i = 0; i < n; ++i)
            s.insert(PMNode { freq[i], (int)i, {} });
    };
    addLeaves(lists[0]);
    for (uint32_t l = 1; l < maxLen; ++l) {
        auto it = lists[l - 1].begin();
        while (it != lists[l - 1].end()) {
            const PMNode a = *it++;
            if (it == lists[l - 1].end())
                break;
            const PMNode b = *it++;
            lists[l].insert(PMNode { a.weight + b.weight, -1, { a, b } });
        }
        addLeaves(lists[l]);
    }
    std::vector<uint32_t> len(n, 0);
    if (n < 2)
        return len;
    // Walk back from the deepest list: take its 2n-2 smallest nodes, then at each shallower level
    // twice as many nodes as there were packages among those taken one level deeper.
    size_t take = 2 * n - 2;
    for (int l = (int)maxLen - 1; l >= 0; --l) {
        if (take > lists[l].size())
            return {};
        size_t packages = 0, k = 0;
        for (auto it = lists[l].begin(); k < take; ++it, ++k) {
            if (it->symbol >= 0)
                ++len[it->symbol];
            else
                ++packages;
        }
        take = 2 * packages;
    }
    return len;
}

struct Code {
    uint64_t code;
    uint32_t bits;
};

std::vector<Code> codeWords(const std::vector<uint32_t>& len)
{
    std::vector<uint32_t> order(len.size());
    std::iota(order.begin(), order.end(), 0u);
    // Ties between equal lengths follow std::sort, as in the reference.
    std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return len[a] < len[b]; });
    const uint32_t maxLen = len.empty() ? 0 : *std::max_element(len.begin(), len.end());
    std::vector<Code> out(len.size());
    uint64_t acc = 0;
    for (uint32_t s : order) {
        out[s] = Code { acc >> (maxLen - len[s]), len[s] };
        acc += 1ull << (maxLen - len[s]);
    }
    return out;
}

uint64_t reverseBits(uint64_t v, uint32_t bits)
{
    uint64_t r = 0;
    for (uint32_t i = 0; i < bits; ++i)
        r |= ((v >> i) & 1ull) << (bits - 1 - i);
    return r;
}

// ---- .tdag16 reader -----------------------------------------------------------------------
struct Reader {
    const uint8_t* p;
    const uint8_t* end;
    bool ok = true;
    template <typename T>
    T get()
    {
        T v {};
        if ((size_t)(end - p) < sizeof(T)) {
            ok = false;
            return v;
        }
        std::memcpy(&v, p, sizeof(T));
        p += sizeof(T);
        return v;
    }
    template <typename T>
    std::vector<T> vec()
    {
        const uint64_t n = get<uint64_t>();
        std::vector<T> v;
        if (!ok || n > (uint64_t)(end - p) / sizeof(T)) {
            ok = false;
            return v;
        }
        v.resize(n);
        if (n)
            std::memcpy(v.data(), p, n * sizeof(T));
        p += n * sizeof(T);
        return v;
    }
};

struct Lut {
    uint32_t mask = 0;
    std::vector<std::pair<uint32_t, uint32_t>> entries; // (symbol, code length)
};

struct Tdag16 {
    bool sym, perm, trans;
    uint32_t maxT;
    bool tables, huffman;
    std::vector<std::vector<uint16_t>> levels;
    std::vector<uint64_t> subgrids;
    uint32_t resolution, rootLevel;
    std::vector<std::vector<uint64_t>> ptrTables;
    std::vector<Lut> luts;
};

bool parse(const uint8_t* bytes, uint64_t len, Tdag16& t)
{
    Reader r { bytes, bytes + len };
    t.sym = r.get<uint8_t>() != 0;
    t.perm = r.get<uint8_t>() != 0;
    t.trans = r.get<uint8_t>() != 0;
    r.get<uint8_t>();
    t.maxT = r.get<uint32_t>();
    t.tables = r.get<uint8_t>() != 0;
    t.huffman = r.get<uint8_t>() != 0;
    const uint64_t nl = r.get<uint64_t>();
    if (!r.ok || nl > 64)
        return false;
    t.levels.resize(nl);
    for (auto& l : t.levels)
        l = r.vec<uint16_t>();
    t.subgrids = r.vec<uint64_t>();
    t.resolution = r.get<uint32_t>();
    t.rootLevel = r.get<uint32_t>();
    const uint64_t nt = r.get<uint64_t>();
    if (!r.ok || nt > 64)
        return false;
    t.ptrTables.resize(nt);
    for (auto& v : t.ptrTables)
        v = r.vec<uint64_t>();
    const uint64_t nh = r.get<uint64_t>();
    if (!r.ok || nh > 64)
        return false;
    t.luts.resize(nh);
    for (auto& lut : t.luts) {
        const uint64_t nc = r.get<uint64_t>();
        std::vector<Code> codes;
        for (uint64_t i = 0; r.ok && i < nc && i < 4096; ++i) {
            Code c;
            c.code = r.get<uint64_t>();
            c.bits = r.get<uint32_t>();
            r.get<uint32_t>(); // padding of the 16-byte record
            codes.push_back(c);
        }
        if (!r.ok || nc > 4096)
            return false;
        if (codes.empty())
            continue;
        uint32_t maxLen = 0;
        for (const Code& c : codes)
            maxLen = std::max(maxLen, c.bits);
        if (maxLen > 24)
            return false;
        lut.mask = (1u << maxLen) - 1u;
        lut.entries.assign(1ull << maxLen, { 0u, 0u });
        for (uint32_t s = 0; s < codes.size(); ++s)
            for (uint64_t suffix = 0; suffix < (1ull << (maxLen - codes[s].bits)); ++suffix)
                lut.entries[(suffix << codes[s].bits) | codes[s].code] = { s, codes[s].bits };
    }
    return r.ok && t.rootLevel < nl && t.rootLevel >= 2 && t.resolution == (1u << t.rootLevel);
}

// The 48 child-order transforms: id = symmetry bits | axis-permutation code << 3. Entry [out] of
// a transform is the child index read for output child `out`.
struct Transforms {
    uint8_t child[48][8];
    uint8_t combine[48 * 48];
    int perm[48][3];
    int sym[48][3];
};

const Transforms& transforms()
{
    static const Transforms T = [] {
        static const int perms[6][3] = { { 0, 1, 2 }, { 0, 2, 1 }, { 1, 0, 2 }, { 1, 2, 0 }, { 2, 0, 1 }, { 2, 1, 0 } };
        Transforms t {};
        for (int id = 0; id < 48; ++id) {
            const int* p = perms[id >> 3];
            for (int a = 0; a < 3; ++a) {
                t.perm[id][a] = p[a];
                t.sym[id][a] = (id >> a) & 1;
            }
            for (int in = 0; in < 8; ++in) {
                const int pos[3] = { in & 1, (in >> 1) & 1, (in >> 2) & 1 };
                int out = 0;
                for (int a = 0; a < 3; ++a)
                    out |= ((pos[p[a]] ^ t.sym[id][a]) & 1) << a;
                t.child[id][out] = (uint8_t)in;
            }
        }
        for (int l = 0; l < 48; ++l)
            for (int r = 0; r < 48; ++r) {
                uint8_t c[8];
                for (int i = 0; i < 8; ++i)
                    c[i] = t.child[r][t.child[l][i]];
                for (int id = 0; id < 48; ++id)
                    if (std::memcmp(c, t.child[id], 8) == 0)
                        t.combine[l * 48 + r] = (uint8_t)id;
            }
        return t;
    }();
    return T;
}

struct Ptr {
    uint32_t index;
    uint32_t transformID;
    int t[3];
};

bool childPointer(const Tdag16& d, uint32_t level, uint64_t nodeWord, uint32_t childIdx, Ptr& out)
{
    const uint32_t childLevel = level - 1;
    const auto& words = d.levels[level];
    if (nodeWord >= words.size())
        return false;
    const uint16_t mask = words[nodeWord];
    uint64_t at = nodeWord + 1;
    for (uint32_t i = 0; i < childIdx; ++i)
        at += (mask >> (2 * i)) & 3u;
    const uint32_t size = (mask >> (2 * childIdx)) & 3u;
    if (at + size > words.size())
        return false;
    uint64_t enc = 0;
    for (uint32_t i = 0; i < size; ++i)
        enc |= (uint64_t)words[at + i] << (16 * i);
    if (d.tables) {
        if (size == 1 || (enc & 1ull)) {
            const uint64_t idx = size == 1 ? enc : enc >> 1;
            if (childLevel >= d.ptrTables.size() || idx >= d.ptrTables[childLevel].size())
                return false;
            enc = d.ptrTables[childLevel][idx];
        } else {
            enc >>= 1;
        }
    }
    const bool hasT = d.trans && childLevel <= d.maxT;
    const int bias = 1 << childLevel;
    const uint32_t tb = childLevel + 1;
    const uint64_t tm = (1ull << tb) - 1ull;
    out = Ptr { 0, 0, { 0, 0, 0 } };
    if (d.huffman) {
        if (childLevel >= d.luts.size() || d.luts[childLevel].entries.empty())
            return false;
        const auto& lut = d.luts[childLevel];
        auto [type, bits] = lut.entries[(uint32_t)enc & lut.mask];
        enc >>= bits;
        uint32_t axes = 0;
        if (hasT) {
            axes = type & 7u;
            type >>= 3;
        }
        for (int a = 2; a >= 0; --a)
            if (axes & (1u << a)) {
                out.t[a] = (int)(enc & tm) - bias;
                enc >>= tb;
            }
        if (d.sym) {
            out.transformID |= type & 7u;
            type >>= 3;
        }
        if (d.perm)
            out.transformID |= (type & 7u) << 3;
    } else {
        bool transformed = false;
        if (d.sym || d.perm || d.trans) {
            transformed = enc & 1ull;
            enc >>= 1;
        }
        if (transformed) {
            if (hasT)
                for (int a = 0; a < 3; ++a) {
                    out.t[a] = (int)(enc & tm) - bias;
                    enc >>= tb;
                }
            if (d.sym) {
                out.transformID |= enc & 7u;
                enc >>= 3;
            }
            if (d.perm) {
                out.transformID |= (enc & 7u) << 3;
                enc >>= 3;
            }
        }
    }
    if (out.transformID >= 48)
        return false;
    out.index = (uint32_t)enc;
    return true;
}

// -1 on malformed data, else the voxel's value.
int getVoxel(const Tdag16& d, int vx, int vy, int vz)
{
    const Transforms& T = transforms();
    int v[3] = { vx, vy, vz };
    uint32_t tid = 0;
    uint64_t node = 0;
    for (uint32_t level = d.rootLevel; level > 2; --level) {
        const uint32_t cl = level - 1;
        if (node >= d.levels[level].size())
            return -1;
        uint32_t c = 0;
        for (int a = 0; a < 3; ++a)
            c |= (((uint32_t)v[a] >> cl) & 1u) << a;
        c = T.child[tid][c];
        if (!((d.levels[level][node] >> (2 * c)) & 3u))
            return 0;
        Ptr p;
        if (!childPointer(d, level, node, c, p))
            return -1;
        node = p.index;
        if (p.t[0] || p.t[1] || p.t[2]) {
            // The pointer's translation, expressed in the frame of the transform accumulated so far.
            int t[3];
            for (int a = 0; a < 3; ++a) {
                t[a] = p.t[T.perm[tid][a]];
                if (T.sym[tid][a])
                    t[a] = -t[a];
            }
            for (int a = 0; a < 3; ++a) {
                const uint32_t before = (uint32_t)v[a] >> cl;
                v[a] -= t[a];
                if (((uint32_t)v[a] >> cl) != before)
                    return 0;
            }
        }
        tid = T.combine[tid * 48 + p.transformID];
    }
    if (node >= d.subgrids.size())
        return -1;
    uint32_t lo = 0, hi = 0;
    for (int a = 0; a < 3; ++a) {
        lo |= ((uint32_t)v[a] & 1u) << a;
        hi |= (((uint32_t)v[a] >> 1) & 1u) << a;
    }
    const uint32_t bit = T.child[tid][hi] * 8u + T.child[tid][lo];
    return (int)((d.subgrids[node] >> bit) & 1ull);
}

} // namespace

extern "C" {

int orc_huffman_code_lengths(const uint64_t* freq, uint32_t n, uint32_t maxLen, uint32_t* outLen)
{
    if (maxLen == 0 || maxLen > 32)
        return -1;
    const auto len = packageMerge(std::vector<uint64_t>(freq, freq + n), maxLen);
    if (len.size() != n)
        return -1;
    std::copy(len.begin(), len.end(), outLen);
    return 0;
}

int orc_huffman_code_words(const uint32_t* len, uint32_t n, int lsbFirst, uint64_t* outCode)
{
    const auto codes = codeWords(std::vector<uint32_t>(len, len + n));
    for (uint32_t i = 0; i < n; ++i)
        outCode[i] = lsbFirst ? reverseBits(codes[i].code, codes[i].bits) : codes[i].code;
    return 0;
}

// Decodes a .tdag16 stream into a dense grid (bit i = x + R*y + R*R*z of R^3/8 bytes, R =
// resolution). Returns the resolution, or -1 on malformed data.
int64_t orc_tdag16_decode(const uint8_t* bytes, uint64_t len, uint8_t* outBits, uint64_t outCap)
{
    Tdag16 d;
    if (!parse(bytes, len, d))
        return -1;
    const uint64_t R = d.resolution;
    if (!outBits)
        return (int64_t)R;
    if (outCap < R * R * R / 8)
        return -1;
    std::memset(outBits, 0, R * R * R / 8);
    for (uint64_t z = 0; z < R; ++z)
        for (uint64_t y = 0; y < R; ++y)
            for (uint64_t x = 0; x < R; ++x) {
                const int b = getVoxel(d, (int)x, (int)y, (int)z);
                if (b < 0)
                    return -1;
                const uint64_t i = x + R * y + R * R * z;
                outBits[i >> 3] |= (uint8_t)(b << (i & 7));
            }
    return (int64_t)R;
}

} // extern "C"
