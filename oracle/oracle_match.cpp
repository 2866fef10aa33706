// TEST INFRASTRUCTURE ONLY. This is the CPU oracle: a restatement of the reference's
// Transform-Aware SVDAG matcher, used by tests/, the smoke check and bench.py's cpu_baseline
// leg as the checker. Nothing in the product library links or calls it.
//
// It restates, line for line in behaviour (not in text):
//   tools/lib/include/voxcom/voxel/transform_dag.h:19-151, 217-233, 313-397   (TransformPointer codec)
//   tools/lib/include/voxcom/voxel/large_sub_grid.h:10-97                     (dense R^3 bit grids)
//   tools/lib/src/voxel/find_translation.cpp:66-125                           (translate / mirror / permute)
//   tools/lib/src/voxel/find_translation.cpp:202-264, 466-495                 (nested corner hash passes)
//   tools/lib/src/voxel/find_translation.cpp:616-1000                         (findDuplicatesAmongFlattenedItems_cpp2)
// plus an independent brute-force matcher that enumerates the candidate set directly.
//
// Parity pins (see tests/test_oracle_*.py): closed-form codec values (sentinel and identity
// bit patterns, sentinel keys per level), group properties of the 48 transforms, brute force
// vs. hash search on random sets with planted transformed copies, and the exact-equation /
// no-smaller-key invariants of every chosen parent.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <thread>
#include <unordered_map>
#include <vector>

#include "oracle.h"

namespace orc {

// ----------------------------------------------------------------------------------------
// TransformPointer (transform_dag.h:19-54). GCC x86-64 bit-field layout, LSB first:
// ptr:37 | sx sy sz | p0:2 p1:2 p2:2 | tx:6 ty:6 tz:6 (translations biased by 32).
// ----------------------------------------------------------------------------------------
struct TP {
    uint64_t ptr;
    uint32_t sx, sy, sz;
    uint32_t p[3];
    int t[3];
};

static const uint8_t kPerms[6][3] = { { 0, 1, 2 }, { 0, 2, 1 }, { 1, 0, 2 }, { 1, 2, 0 }, { 2, 0, 1 }, { 2, 1, 0 } }; // transform_dag.h:138-145

static uint32_t encodePerm(const uint32_t p[3]) { return 2 * p[0] + (p[2] < p[1] ? 1 : 0); } // transform_dag.h:326-331

static TP tpCreate(uint64_t ptr, uint32_t symBits = 0, uint32_t permCode = 0, int tx = 0, int ty = 0, int tz = 0)
{
    TP o;
    o.ptr = ptr & ((1ull << 37) - 1);
    o.sx = symBits & 1;
    o.sy = (symBits >> 1) & 1;
    o.sz = (symBits >> 2) & 1;
    for (int a = 0; a < 3; ++a)
        o.p[a] = kPerms[permCode][a];
    // setTranslation stores (t + 32) in 6-bit unsigned fields (transform_dag.h:75-85).
    o.t[0] = (int)(((unsigned)(tx + 32)) & 63) - 32;
    o.t[1] = (int)(((unsigned)(ty + 32)) & 63) - 32;
    o.t[2] = (int)(((unsigned)(tz + 32)) & 63) - 32;
    return o;
}

static uint64_t tpBits(const TP& tp)
{
    uint64_t v = tp.ptr;
    v |= (uint64_t)tp.sx << 37;
    v |= (uint64_t)tp.sy << 38;
    v |= (uint64_t)tp.sz << 39;
    v |= (uint64_t)tp.p[0] << 40;
    v |= (uint64_t)tp.p[1] << 42;
    v |= (uint64_t)tp.p[2] << 44;
    v |= (uint64_t)((unsigned)(tp.t[0] + 32) & 63) << 46;
    v |= (uint64_t)((unsigned)(tp.t[1] + 32) & 63) << 52;
    v |= (uint64_t)((unsigned)(tp.t[2] + 32) & 63) << 58;
    return v;
}

static TP tpFromBits(uint64_t v)
{
    TP o;
    o.ptr = v & ((1ull << 37) - 1);
    o.sx = (v >> 37) & 1;
    o.sy = (v >> 38) & 1;
    o.sz = (v >> 39) & 1;
    o.p[0] = (v >> 40) & 3;
    o.p[1] = (v >> 42) & 3;
    o.p[2] = (v >> 44) & 3;
    o.t[0] = (int)((v >> 46) & 63) - 32;
    o.t[1] = (int)((v >> 52) & 63) - 32;
    o.t[2] = (int)((v >> 58) & 63) - 32;
    return o;
}

static uint32_t symBitsOf(const TP& tp) { return tp.sx | (tp.sy << 1) | (tp.sz << 2); } // bvec3ToU64, transform_dag.h:226-232

// encodeFixed64 (transform_dag.h:375-397).
static uint64_t encodeFixed64(const TP& tp, uint32_t level)
{
    uint64_t out = tp.ptr;
    out <<= 3;
    out |= encodePerm(tp.p);
    out <<= 3;
    out |= symBitsOf(tp);
    if (level != 0) {
        const uint32_t b = level + 1;
        const uint64_t ox = (uint64_t)(uint32_t)(tp.t[0] + (1 << level));
        const uint64_t oy = (uint64_t)(uint32_t)(tp.t[1] + (1 << level));
        const uint64_t oz = (uint64_t)(uint32_t)(tp.t[2] + (1 << level));
        out <<= b;
        out |= oz;
        out <<= b;
        out |= oy;
        out <<= b;
        out |= ox;
    }
    return out;
}

// decodeFixed64 (transform_dag.h:349-373).
static TP decodeFixed64(uint64_t e, uint32_t level)
{
    int t[3] = { 0, 0, 0 };
    if (level != 0) {
        const uint32_t b = level + 1;
        const uint64_t m = (1ull << b) - 1ull;
        unsigned u[3];
        u[0] = unsigned(e & m);
        e >>= b;
        u[1] = unsigned(e & m);
        e >>= b;
        u[2] = unsigned(e & m);
        e >>= b;
        for (int a = 0; a < 3; ++a)
            t[a] = (int)u[a] - (1 << level);
    }
    const uint32_t sym = (uint32_t)(e & 7);
    e >>= 3;
    const uint32_t permCode = (uint32_t)(e & 7);
    e >>= 3;
    return tpCreate(e, sym, permCode, t[0], t[1], t[2]);
}

static TP sentinelTP() { return tpCreate(0xFFFFFFFFull); } // transform_dag.h:344-347

// ----------------------------------------------------------------------------------------
// Dense R^3 grids (large_sub_grid.h:80-97): bit i = x + R*y + R*R*z, word i >> 6.
// ----------------------------------------------------------------------------------------
struct Grid {
    uint32_t R;
    std::vector<uint64_t> w;
    explicit Grid(uint32_t R_ = 4)
        : R(R_)
        , w(std::max<uint32_t>(1, R_ * R_ * R_ / 64), 0)
    {
    }
    bool get(uint32_t x, uint32_t y, uint32_t z) const
    {
        const uint32_t i = x + R * y + R * R * z;
        return (w[i >> 6] >> (i & 63)) & 1;
    }
    void set(uint32_t x, uint32_t y, uint32_t z)
    {
        const uint32_t i = x + R * y + R * R * z;
        w[i >> 6] |= 1ull << (i & 63);
    }
    bool operator==(const Grid& o) const { return w == o.w; }
    bool operator!=(const Grid& o) const { return w != o.w; }
    uint32_t popcount() const
    {
        uint32_t c = 0;
        for (uint64_t v : w)
            c += (uint32_t)__builtin_popcountll(v);
        return c;
    }
};

// translateSubGrid (find_translation.cpp:66-84): out[o] = in[o - t] when in range.
static Grid translateGrid(const Grid& in, const int t[3])
{
    const int R = (int)in.R;
    Grid out(in.R);
    for (int z = 0; z < R; ++z)
        for (int y = 0; y < R; ++y)
            for (int x = 0; x < R; ++x) {
                const int ix = x - t[0], iy = y - t[1], iz = z - t[2];
                if (ix < 0 || iy < 0 || iz < 0 || ix >= R || iy >= R || iz >= R)
                    continue;
                if (in.get(ix, iy, iz))
                    out.set(x, y, z);
            }
    return out;
}

// applySymmetryToSubGrid (find_translation.cpp:85-107).
static Grid mirrorGrid(const Grid& in, uint32_t sym)
{
    const uint32_t R = in.R;
    Grid out(R);
    for (uint32_t z = 0; z < R; ++z)
        for (uint32_t y = 0; y < R; ++y)
            for (uint32_t x = 0; x < R; ++x) {
                if (!in.get(x, y, z))
                    continue;
                uint32_t o[3] = { x, y, z };
                for (int a = 0; a < 3; ++a)
                    if ((sym >> a) & 1)
                        o[a] = R - 1 - o[a];
                out.set(o[0], o[1], o[2]);
            }
    return out;
}

// permuteAxisOfSubGrid (find_translation.cpp:109-125): out[(v[p0], v[p1], v[p2])] = in[v].
static Grid permuteGrid(const Grid& in, const uint32_t p[3])
{
    const uint32_t R = in.R;
    Grid out(R);
    for (uint32_t z = 0; z < R; ++z)
        for (uint32_t y = 0; y < R; ++y)
            for (uint32_t x = 0; x < R; ++x) {
                if (!in.get(x, y, z))
                    continue;
                const uint32_t v[3] = { x, y, z };
                out.set(v[p[0]], v[p[1]], v[p[2]]);
            }
    return out;
}

// ----------------------------------------------------------------------------------------
// Nested corner hash (find_translation.cpp:202-264 scalar passes; HASH_FUNCTION = identity,
// hash_combine from voxcom/utility/hash.h:16-20).
// ----------------------------------------------------------------------------------------
struct HashGrid {
    uint32_t R;
    std::vector<uint64_t> v;
    explicit HashGrid(uint32_t R_)
        : R(R_)
        , v((size_t)R_ * R_ * R_, 0)
    {
    }
    uint64_t& at(int x, int y, int z) { return v[(size_t)x + (size_t)R * y + (size_t)R * R * z]; }
};

static inline void combineStep(uint64_t& combined, uint64_t h)
{
    const uint64_t started = ((h | combined) == 0) ? 0 : ~0ull;
    combined ^= h + 0x9e3779b9ull + (combined << 6) + (combined >> 2);
    combined = std::max<uint64_t>(combined, 1);
    combined &= started;
}

// One pass along `axis` (0=x,1=y,2=z). Positive direction scans from R-1 down to 0.
static void hashPass(HashGrid& g, int axis, bool positive)
{
    const int R = (int)g.R;
    const int start = positive ? R - 1 : 0;
    const int end = positive ? -1 : R;
    const int step = positive ? -1 : +1;
    for (int a = 0; a < R; ++a)
        for (int b = 0; b < R; ++b) {
            uint64_t combined = 0;
            for (int i = start; i != end; i += step) {
                int c[3];
                c[axis] = i;
                c[(axis + 1) % 3] = a;
                c[(axis + 2) % 3] = b;
                uint64_t& cell = g.at(c[0], c[1], c[2]);
                combineStep(combined, cell);
                cell = combined;
            }
        }
}

static HashGrid initialHashGrid(const Grid& g)
{
    HashGrid h(g.R);
    for (size_t i = 0; i < h.v.size(); ++i)
        h.v[i] = (g.w[i >> 6] >> (i & 63)) & 1;
    return h;
}

// computePermutedGridHashInPlace (find_translation.cpp:481-495): pass along perm.x with
// direction dx, then perm.y with dy, then perm.z with dz.
static void permutedHash(HashGrid& h, const uint32_t p[3], const bool dir[3])
{
    hashPass(h, (int)p[0], dir[0]);
    hashPass(h, (int)p[1], dir[1]);
    hashPass(h, (int)p[2], dir[2]);
}

static inline uint32_t mortonEncode3(uint32_t x, uint32_t y, uint32_t z)
{
    uint32_t m = 0;
    for (int b = 0; b < 10; ++b) {
        m |= ((x >> b) & 1u) << (3 * b);
        m |= ((y >> b) & 1u) << (3 * b + 1);
        m |= ((z >> b) & 1u) << (3 * b + 2);
    }
    return m;
}
static inline void mortonDecode3(uint32_t m, uint32_t& x, uint32_t& y, uint32_t& z)
{
    x = y = z = 0;
    for (int b = 0; b < 10; ++b) {
        x |= ((m >> (3 * b)) & 1u) << b;
        y |= ((m >> (3 * b + 1)) & 1u) << b;
        z |= ((m >> (3 * b + 2)) & 1u) << b;
    }
}

// The 8 search directions in the reference's call order (find_translation.cpp:712-719, 869-876).
static const bool kDirs[8][3] = {
    { true, true, true }, { true, true, false }, { true, false, true }, { true, false, false },
    { false, true, true }, { false, true, false }, { false, false, true }, { false, false, false }
};

static void atomicMin64(std::atomic<uint64_t>& a, uint64_t value) // find_translation.cpp:543-548
{
    uint64_t expected = a.load(std::memory_order_relaxed);
    while (value < expected && !a.compare_exchange_weak(expected, value, std::memory_order_relaxed, std::memory_order_relaxed))
        ;
}

static bool configAllows(const TP& tp, const orc_config& cfg) // find_translation.cpp:840-845
{
    const bool hasPerm = !(tp.p[0] == 0 && tp.p[1] == 1 && tp.p[2] == 2);
    const bool hasSym = (tp.sx | tp.sy | tp.sz) != 0;
    const bool hasT = tp.t[0] != 0 || tp.t[1] != 0 || tp.t[2] != 0;
    if (!cfg.axisPermutation && hasPerm)
        return false;
    if (!cfg.symmetry && hasSym)
        return false;
    if (!cfg.translation && hasT)
        return false;
    return true;
}

template <typename F>
static void parallelFor(size_t n, int nthreads, F&& f)
{
    if (nthreads <= 1 || n < 64) {
        for (size_t i = 0; i < n; ++i)
            f(i);
        return;
    }
    std::atomic<size_t> next { 0 };
    std::vector<std::thread> pool;
    for (int t = 0; t < nthreads; ++t)
        pool.emplace_back([&]() {
            for (;;) {
                const size_t begin = next.fetch_add(64);
                if (begin >= n)
                    break;
                const size_t end = std::min(n, begin + 64);
                for (size_t i = begin; i < end; ++i)
                    f(i);
            }
        });
    for (auto& th : pool)
        th.join();
}

static std::vector<Grid> loadGrids(uint32_t level, const uint64_t* masks, uint64_t n)
{
    const uint32_t R = 1u << level;
    const uint32_t W = R * R * R / 64;
    std::vector<Grid> grids(n, Grid(R));
    for (uint64_t i = 0; i < n; ++i)
        std::memcpy(grids[i].w.data(), masks + i * W, W * 8);
    return grids;
}

// findDuplicatesAmongFlattenedItems_cpp2 search phase (find_translation.cpp:616-921), returning
// encodedParents indexed by the original item index (key ptr = popcount-sorted rank of the parent).
//
// Timing form (cpu_baseline): with queryStride > 1 only the queries of rank 0, s, 2s, ... are
// searched (the table still holds every item); setupS / searchS receive the wall time of the
// table build (sort, corner hashes, ranges) and of the search loop.
static void matchHash(uint32_t level, const std::vector<Grid>& grids, const orc_config& cfg, uint64_t* outEncoded, int nthreads, uint32_t queryStride = 1,
    double* setupS = nullptr, double* searchS = nullptr, uint64_t* numQueries = nullptr)
{
    const auto tSetup = std::chrono::steady_clock::now();
    queryStride = std::max<uint32_t>(1u, queryStride);
    const uint32_t R = 1u << level;
    const uint32_t R3 = R * R * R;
    const size_t N = grids.size();

    // Stable sort by voxel count, descending (find_translation.cpp:631-643).
    std::vector<uint32_t> byCount(N);
    std::iota(byCount.begin(), byCount.end(), 0u);
    std::vector<uint16_t> counts(N);
    for (size_t i = 0; i < N; ++i)
        counts[i] = (uint16_t)grids[i].popcount();
    std::stable_sort(byCount.begin(), byCount.end(), [&](uint32_t l, uint32_t r) { return counts[l] > counts[r]; });

    const uint64_t encodedSentinel = encodeFixed64(sentinelTP(), level); // :652
    std::vector<std::atomic<uint64_t>> encodedParents(N);
    for (auto& a : encodedParents)
        a.store(encodedSentinel, std::memory_order_relaxed);
    if (N == 0)
        return;

    // Corner hashes of every item seen from each of its 8 corners (:677-719).
    std::vector<std::pair<uint64_t, TP>> hashAndPointers(8 * N);
    for (int d = 0; d < 8; ++d) {
        const bool* dir = kDirs[d];
        const int anchor[3] = { dir[0] ? 0 : (int)R - 1, dir[1] ? 0 : (int)R - 1, dir[2] ? 0 : (int)R - 1 };
        const uint32_t dirIdx = (dir[0] ? 1u : 0u) | (dir[1] ? 2u : 0u) | (dir[2] ? 4u : 0u); // morton_encode32(direction)
        const uint32_t dirSym = dirIdx;
        parallelFor(N, nthreads, [&](size_t item) {
            HashGrid h = initialHashGrid(grids[item]);
            hashPass(h, 0, dir[0]);
            hashPass(h, 1, dir[1]);
            hashPass(h, 2, dir[2]);
            hashAndPointers[dirIdx * N + item] = { h.at(anchor[0], anchor[1], anchor[2]), tpCreate(item, dirSym) };
        });
    }
    // Drop empty (0) hashes, sort by hash, build ranges and the hash -> range LUT (:721-759).
    hashAndPointers.erase(std::remove_if(hashAndPointers.begin(), hashAndPointers.end(),
                              [](const std::pair<uint64_t, TP>& hp) { return hp.first == 0; }),
        hashAndPointers.end());
    std::sort(hashAndPointers.begin(), hashAndPointers.end(),
        [](const std::pair<uint64_t, TP>& l, const std::pair<uint64_t, TP>& r) { return l.first < r.first; });
    std::vector<std::pair<uint32_t, uint32_t>> ranges;
    std::unordered_map<uint64_t, uint32_t> lut;
    std::vector<TP> pointers(hashAndPointers.size());
    if (!hashAndPointers.empty()) {
        uint32_t start = 0;
        uint64_t prev = hashAndPointers[0].first;
        std::vector<uint64_t> rangeHashes;
        for (uint32_t i = 0; i < hashAndPointers.size(); ++i) {
            if (hashAndPointers[i].first != prev) {
                ranges.push_back({ start, i });
                rangeHashes.push_back(prev);
                prev = hashAndPointers[i].first;
                start = i;
            }
        }
        ranges.push_back({ start, (uint32_t)hashAndPointers.size() });
        rangeHashes.push_back(hashAndPointers.back().first);
        for (size_t i = 0; i < hashAndPointers.size(); ++i)
            pointers[i] = hashAndPointers[i].second;
        for (uint32_t i = 0; i < rangeHashes.size(); ++i)
            lut[rangeHashes[i]] = i;
    }
    hashAndPointers.clear();
    const auto tSearch = std::chrono::steady_clock::now();
    uint64_t searched = 0;

    // Search in batches of queries in sorted order (:769-918).
    const size_t maxBatch = std::max<size_t>(N / 100, (size_t)100000);
    for (size_t batchStart = 0; batchStart < N; batchStart += maxBatch) {
        const size_t batchEnd = std::min(batchStart + maxBatch, N);
        // Sampled ranks of this batch: the multiples of queryStride in [batchStart, batchEnd).
        const size_t first = (batchStart + queryStride - 1) / queryStride * queryStride;
        const size_t count = first < batchEnd ? (batchEnd - first + queryStride - 1) / queryStride : 0;
        searched += count;
        parallelFor(count, nthreads, [&](size_t off) {
            const uint32_t itemIdx = (uint32_t)(first + off * queryStride);
            const uint64_t minKey = encodeFixed64(tpCreate(itemIdx, 0, 0, 1 - (int)R, 1 - (int)R, 1 - (int)R), level);
            const Grid& grid = grids[byCount[itemIdx]];
            const HashGrid initial = initialHashGrid(grid);
            for (int d = 0; d < 8; ++d) {
                const bool* dir = kDirs[d];
                const int anchor[3] = { dir[0] ? 0 : (int)R - 1, dir[1] ? 0 : (int)R - 1, dir[2] ? 0 : (int)R - 1 };
                const uint32_t dirBits = (dir[0] ? 1u : 0u) | (dir[1] ? 2u : 0u) | (dir[2] ? 4u : 0u);
                for (uint32_t permId = 0; permId < 6; ++permId) {
                    const uint32_t perm[3] = { kPerms[permId][0], kPerms[permId][1], kPerms[permId][2] };
                    HashGrid h = initial;
                    permutedHash(h, perm, dir);
                    for (uint32_t hv = 0; hv < R3; ++hv) {
                        uint32_t hx, hy, hz;
                        mortonDecode3(hv, hx, hy, hz);
                        const uint64_t hash = h.at(hx, hy, hz);
                        if (hash == 0)
                            continue;
                        auto it = lut.find(hash);
                        if (it == lut.end())
                            continue;
                        const auto [begin, end] = ranges[it->second];
                        const uint32_t hvArr[3] = { hx, hy, hz };
                        const int voxel[3] = { (int)hvArr[perm[0]], (int)hvArr[perm[1]], (int)hvArr[perm[2]] }; // applyAxisPermutation
                        for (uint32_t gi = begin; gi < end; ++gi) {
                            const TP& match = pointers[gi];
                            const uint64_t matchIdx = match.ptr;
                            const uint64_t currentBest = encodedParents[matchIdx].load(std::memory_order_relaxed);
                            if (minKey >= currentBest)
                                continue;
                            const uint32_t sym = dirBits ^ symBitsOf(match);
                            int t[3];
                            for (int a = 0; a < 3; ++a)
                                t[a] = ((sym >> a) & 1) ? voxel[a] - anchor[a] : anchor[a] - voxel[a]; // glm::mix
                            const TP sp = tpCreate(itemIdx, sym, permId, t[0], t[1], t[2]);
                            if (!configAllows(sp, cfg))
                                continue;
                            const uint64_t key = encodeFixed64(sp, level);
                            if (key >= currentBest)
                                continue;
                            // Exact verification (:857-862).
                            const Grid permuted = permuteGrid(grid, sp.p);
                            const Grid mirrored = mirrorGrid(permuted, symBitsOf(sp));
                            const Grid shifted = translateGrid(mirrored, sp.t);
                            if (grids[matchIdx] != shifted)
                                continue;
                            atomicMin64(encodedParents[matchIdx], key);
                        }
                    }
                }
            }
        });
        // Range pruning (:887-915). Optimisation only: drops pointers that can no longer improve.
        const size_t nextStart = batchEnd + 1;
        const uint64_t nextMinKey = encodeFixed64(tpCreate(nextStart, 0, 0, 1 - (int)R, 1 - (int)R, 1 - (int)R), level);
        for (auto& range : ranges) {
            uint32_t newEnd = range.first;
            for (uint32_t i = range.first; i < range.second; ++i) {
                const uint64_t ep = encodedParents[pointers[i].ptr].load(std::memory_order_relaxed);
                if (ep > nextMinKey)
                    pointers[newEnd++] = pointers[i];
            }
            range.second = newEnd;
        }
    }
    const auto tEnd = std::chrono::steady_clock::now();
    for (size_t i = 0; i < N; ++i)
        outEncoded[i] = encodedParents[i].load();
    if (setupS)
        *setupS = std::chrono::duration<double>(tSearch - tSetup).count();
    if (searchS)
        *searchS = std::chrono::duration<double>(tEnd - tSearch).count();
    if (numQueries)
        *numQueries = searched;
}

// Independent brute force over the candidate set: all refs (in sorted rank order), all 48
// (perm, mirror) pairs and all t in [-(R-1), R-1]^3, enumerated in increasing key order, so
// the first exact hit is the minimum key. Empty items have no corner hash in the reference,
// so they keep the sentinel (find_translation.cpp:721-722).
//
// Equality test for c == translate(Q, t), straight from the definition out[o] = Q[o - t]:
//   (a) every voxel o of c has Q[o - t] set with o - t inside the grid (c is contained), and
//   (b) the number of Q voxels v with v + t inside the grid equals popcount(c) (no extras).
// (b) is answered in O(1) with a 3D summed-area table of Q.
struct Transformed {
    Grid q;
    std::vector<uint32_t> sat; // (R+1)^3 prefix counts
};

static void buildSat(Transformed& t)
{
    const int R = (int)t.q.R;
    const int S = R + 1;
    t.sat.assign((size_t)S * S * S, 0);
    for (int z = 1; z <= R; ++z)
        for (int y = 1; y <= R; ++y)
            for (int x = 1; x <= R; ++x) {
                auto at = [&](int a, int b, int c) -> uint32_t& { return t.sat[(size_t)a + (size_t)S * b + (size_t)S * S * c]; };
                at(x, y, z) = (t.q.get(x - 1, y - 1, z - 1) ? 1u : 0u) + at(x - 1, y, z) + at(x, y - 1, z) + at(x, y, z - 1)
                    - at(x - 1, y - 1, z) - at(x - 1, y, z - 1) - at(x, y - 1, z - 1) + at(x - 1, y - 1, z - 1);
            }
}

static uint32_t boxCount(const Transformed& t, const int lo[3], const int hi[3])
{
    const int S = (int)t.q.R + 1;
    auto at = [&](int a, int b, int c) { return (int64_t)t.sat[(size_t)a + (size_t)S * b + (size_t)S * S * c]; };
    const int64_t v = at(hi[0], hi[1], hi[2]) - at(lo[0], hi[1], hi[2]) - at(hi[0], lo[1], hi[2]) - at(hi[0], hi[1], lo[2])
        + at(lo[0], lo[1], hi[2]) + at(lo[0], hi[1], lo[2]) + at(hi[0], lo[1], lo[2]) - at(lo[0], lo[1], lo[2]);
    return (uint32_t)v;
}

static void matchBrute(uint32_t level, const std::vector<Grid>& grids, const orc_config& cfg, uint64_t* outEncoded, int nthreads)
{
    const int R = 1 << level;
    const size_t N = grids.size();
    std::vector<uint32_t> order(N);
    std::iota(order.begin(), order.end(), 0u);
    std::stable_sort(order.begin(), order.end(), [&](uint32_t l, uint32_t r) { return grids[l].popcount() > grids[r].popcount(); });
    const uint64_t sentinelKey = encodeFixed64(sentinelTP(), level);
    std::vector<Transformed> transformed(N * 48, Transformed { Grid(R), {} });
    parallelFor(N, nthreads, [&](size_t rank) {
        const Grid& g = grids[order[rank]];
        for (uint32_t pc = 0; pc < 6; ++pc) {
            const uint32_t p[3] = { kPerms[pc][0], kPerms[pc][1], kPerms[pc][2] };
            const Grid pg = permuteGrid(g, p);
            for (uint32_t s = 0; s < 8; ++s) {
                Transformed& tr = transformed[rank * 48 + pc * 8 + s];
                tr.q = mirrorGrid(pg, s);
                buildSat(tr);
            }
        }
    });
    parallelFor(N, nthreads, [&](size_t child) {
        const Grid& c = grids[child];
        const uint32_t cpop = c.popcount();
        std::vector<int> vox; // occupied voxels of the child, as (x, y, z) triples
        for (int z = 0; z < R; ++z)
            for (int y = 0; y < R; ++y)
                for (int x = 0; x < R; ++x)
                    if (c.get(x, y, z)) {
                        vox.push_back(x);
                        vox.push_back(y);
                        vox.push_back(z);
                    }
        uint64_t best = sentinelKey;
        if (cpop != 0) {
            for (size_t rank = 0; rank < N && best == sentinelKey; ++rank) {
                if (grids[order[rank]].popcount() < cpop)
                    break; // translation never adds voxels
                for (uint32_t pc = 0; pc < 6 && best == sentinelKey; ++pc)
                    for (uint32_t s = 0; s < 8 && best == sentinelKey; ++s) {
                        const Transformed& tr = transformed[rank * 48 + pc * 8 + s];
                        for (int tz = 1 - R; tz <= R - 1 && best == sentinelKey; ++tz)
                            for (int ty = 1 - R; ty <= R - 1 && best == sentinelKey; ++ty)
                                for (int tx = 1 - R; tx <= R - 1 && best == sentinelKey; ++tx) {
                                    const TP tp = tpCreate(rank, s, pc, tx, ty, tz);
                                    if (!configAllows(tp, cfg))
                                        continue;
                                    const int t[3] = { tx, ty, tz };
                                    bool contained = true;
                                    for (size_t k = 0; k < vox.size(); k += 3) {
                                        const int ix = vox[k] - tx, iy = vox[k + 1] - ty, iz = vox[k + 2] - tz;
                                        if (ix < 0 || iy < 0 || iz < 0 || ix >= R || iy >= R || iz >= R || !tr.q.get(ix, iy, iz)) {
                                            contained = false;
                                            break;
                                        }
                                    }
                                    if (!contained)
                                        continue;
                                    int lo[3], hi[3];
                                    for (int a = 0; a < 3; ++a) {
                                        lo[a] = std::max(0, -t[a]);
                                        hi[a] = std::min(R, R - t[a]);
                                    }
                                    if (boxCount(tr, lo, hi) == cpop)
                                        best = encodeFixed64(tp, level);
                                }
                    }
            }
        }
        outEncoded[child] = best;
    });
}

} // namespace orc

using namespace orc;

extern "C" {

uint64_t orc_tp_create(uint64_t ptr, uint32_t symBits, uint32_t permCode, int tx, int ty, int tz)
{
    return tpBits(tpCreate(ptr, symBits, permCode, tx, ty, tz));
}
uint64_t orc_tp_sentinel(void) { return tpBits(sentinelTP()); }
uint64_t orc_encode_fixed64(uint64_t tpBitsIn, uint32_t level) { return encodeFixed64(tpFromBits(tpBitsIn), level); }
uint64_t orc_decode_fixed64(uint64_t key, uint32_t level) { return tpBits(decodeFixed64(key, level)); }

int orc_transform(uint32_t level, const uint64_t* in, uint32_t permCode, uint32_t symBits, int tx, int ty, int tz, uint64_t* out)
{
    if (level < 1 || level > 5 || permCode > 5 || symBits > 7)
        return 1;
    const uint32_t R = 1u << level;
    Grid g(R);
    std::memcpy(g.w.data(), in, g.w.size() * 8);
    const uint32_t p[3] = { kPerms[permCode][0], kPerms[permCode][1], kPerms[permCode][2] };
    const int t[3] = { tx, ty, tz };
    const Grid r = translateGrid(mirrorGrid(permuteGrid(g, p), symBits), t);
    std::memcpy(out, r.w.data(), r.w.size() * 8);
    return 0;
}

int orc_corner_hash(uint32_t level, const uint64_t* in, uint32_t dirBits, uint64_t* outHash)
{
    if (level < 2 || level > 5 || dirBits > 7)
        return 1;
    const uint32_t R = 1u << level;
    Grid g(R);
    std::memcpy(g.w.data(), in, g.w.size() * 8);
    HashGrid h = initialHashGrid(g);
    const bool dir[3] = { (dirBits & 1) != 0, (dirBits & 2) != 0, (dirBits & 4) != 0 };
    hashPass(h, 0, dir[0]);
    hashPass(h, 1, dir[1]);
    hashPass(h, 2, dir[2]);
    *outHash = h.at(dir[0] ? 0 : R - 1, dir[1] ? 0 : R - 1, dir[2] ? 0 : R - 1);
    return 0;
}

int orc_match_encoded(uint32_t level, const uint64_t* masks, uint64_t n, const orc_config* cfg, uint64_t* outEncoded, int nthreads, int brute)
{
    if (level < 2 || level > 5 || cfg == nullptr)
        return 1;
    const std::vector<Grid> grids = loadGrids(level, masks, n);
    if (brute)
        matchBrute(level, grids, *cfg, outEncoded, nthreads);
    else
        matchHash(level, grids, *cfg, outEncoded, nthreads);
    return 0;
}

// The hash search of orc_match_encoded on every item's table but only the queries of rank
// 0, stride, 2*stride, ...; reports the table-build and search wall times and the number of
// queries searched (timing only: the keys of a sampled run are not the full run's).
int orc_match_timed(uint32_t level, const uint64_t* masks, uint64_t n, const orc_config* cfg, int nthreads, uint32_t queryStride, double* outSetupS,
    double* outSearchS, uint64_t* outQueries)
{
    if (level < 2 || level > 5 || cfg == nullptr || queryStride == 0)
        return 1;
    const std::vector<Grid> grids = loadGrids(level, masks, n);
    std::vector<uint64_t> enc(std::max<uint64_t>(n, 1));
    matchHash(level, grids, *cfg, enc.data(), nthreads, queryStride, outSetupS, outSearchS, outQueries);
    return 0;
}

// Post-processing (find_translation.cpp:923-997): decode, map sorted rank -> original index,
// isUnique = the set of parents, unique items in original order, mapping.
int orc_postprocess(uint32_t level, const uint64_t* masks, uint64_t n, const uint64_t* encoded, uint64_t* outMapping, uint32_t* outUnique, uint64_t* outNumUnique)
{
    if (level < 2 || level > 5)
        return 1;
    const std::vector<Grid> grids = loadGrids(level, masks, n);
    std::vector<uint32_t> byCount(n);
    std::iota(byCount.begin(), byCount.end(), 0u);
    std::vector<uint16_t> counts(n);
    for (size_t i = 0; i < n; ++i)
        counts[i] = (uint16_t)grids[i].popcount();
    std::stable_sort(byCount.begin(), byCount.end(), [&](uint32_t l, uint32_t r) { return counts[l] > counts[r]; });
    const uint64_t encodedSentinel = encodeFixed64(sentinelTP(), level);
    const uint64_t sentinelBits = tpBits(sentinelTP());
    std::vector<uint64_t> parents(n);
    for (size_t i = 0; i < n; ++i) {
        if (encoded[i] == encodedSentinel) {
            parents[i] = sentinelBits;
        } else {
            TP tp = decodeFixed64(encoded[i], level);
            tp.ptr = byCount[tp.ptr];
            parents[i] = tpBits(tp);
        }
    }
    std::vector<bool> isUnique(n, false);
    for (size_t i = 0; i < n; ++i) {
        if (parents[i] == sentinelBits)
            isUnique[i] = true;
        else
            isUnique[tpFromBits(parents[i]).ptr] = true;
    }
    for (size_t i = 0; i < n; ++i)
        outMapping[i] = sentinelBits;
    uint64_t numUnique = 0;
    for (size_t i = 0; i < n; ++i)
        if (isUnique[i]) {
            outMapping[i] = tpBits(tpCreate(numUnique));
            outUnique[numUnique++] = (uint32_t)i;
        }
    for (size_t i = 0; i < n; ++i)
        if (!isUnique[i]) {
            TP parent = tpFromBits(parents[i]);
            parent.ptr = tpFromBits(outMapping[parent.ptr]).ptr;
            outMapping[i] = tpBits(parent);
        }
    *outNumUnique = numUnique;
    return 0;
}

} // extern "C"
