// Exhaustive Transform-Aware SVDAG matcher for gfx950.
//
// Problem (reference: find_translation.cpp:616-1000): for every item c of a level, find the
// minimum 64-bit key (parent rank, perm, mirror, translation) over all exact matches
//     c == translate(mirror(permute(P, perm), mirror), t),   t in [-(R-1), R-1]^3,
// where P ranges over all items in stable popcount-descending order. The reference finds the
// candidates with a nested corner hash and verifies each one; its result equals the
// exhaustive minimum. This file computes the same minimum with its own GPU formulation:
//
//  * Rank order: stable LSD radix sort of (R^3 - popcount, index).
//  * Child table: each item's content as seen from each of its 8 corners ("anchored content")
//    goes into an open-addressing multi-map. At level 2 the anchored content is itself a
//    64-bit mask and is the exact key; at levels 3-5 the key is a 32-bit nested sub-box hash.
//  * Search: every parent enumerates 6 permutations x 8 directions x R^3 anchored sub-boxes.
//    A table hit (child c, corner k) yields mirror = d ^ k and the translation the reference
//    derives (find_translation.cpp:835-837); the candidate is verified bitwise and merged into
//    best[c] with a 64-bit atomicMin. The self-match seeds best[c], so only ranks < c probe.
#include <hip/hip_runtime.h>

#include <type_traits>
#include <vector>

#include "common.hpp"
#include "matcher.hpp"
#include "primitives.hpp"
#include "search_device.hpp"

namespace tdag {

namespace {

using namespace dev;

inline size_t ceilDiv(size_t a, size_t b) { return (a + b - 1) / b; }
inline size_t alignUp(size_t v, size_t a) { return (v + a - 1) / a * a; }

// Host side of the XCD-grouped grid decoding (xcdBlock): blocks for `groups` groups of
// `perGroup` blocks, padded to a multiple of 8 groups.
inline uint32_t xcdGridBlocks(uint32_t groups, uint32_t perGroup)
{
    return (groups + kXcds - 1) / kXcds * kXcds * perGroup;
}

// ---------------------------------------------------------------------------------------
// Rank order and initialisation
// ---------------------------------------------------------------------------------------

// keys[i] = R^3 - popcount(mask i), vals[i] = i. G lanes cooperate on one item.
template <int G>
__global__ __launch_bounds__(256) void kPopKeys(const uint64_t* __restrict__ masks, uint32_t n, uint32_t W, uint32_t R3, uint32_t* __restrict__ keys, uint32_t* __restrict__ vals)
{
    const uint32_t gid = blockIdx.x * 256u + threadIdx.x;
    const uint32_t item = gid / G;
    const uint32_t sub = gid % G;
    uint32_t pop = 0;
    if (item < n) {
        const uint64_t* m = masks + (size_t)item * W;
        for (uint32_t w = sub; w < W; w += G)
            pop += (uint32_t)__popcll(m[w]);
    }
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1)
        pop += __shfl_xor(pop, off, G);
    if (item < n && sub == 0) {
        keys[item] = R3 - pop;
        vals[item] = item;
    }
}

// sortedMasks[r] = masks[order[r]]; best[r] = self key (or the sentinel key for empty items).
__global__ __launch_bounds__(256) void kGatherInit(const uint64_t* __restrict__ masks, const uint32_t* __restrict__ order, const uint32_t* __restrict__ sortedKeys,
    uint32_t n, uint32_t W, uint32_t R3, uint32_t level, uint64_t* __restrict__ sortedMasks, uint64_t* __restrict__ best)
{
    const size_t total = (size_t)n * W;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < total; i += (size_t)gridDim.x * 256) {
        const uint32_t r = (uint32_t)(i / W);
        const uint32_t w = (uint32_t)(i % W);
        sortedMasks[i] = masks[(size_t)order[r] * W + w];
        if (w == 0)
            best[r] = (sortedKeys[r] == R3) ? sentinelKey(level) : selfKey(r, level);
    }
}

// ---------------------------------------------------------------------------------------
// Level 2 (4^3 leaves, one uint64 per mask): exact anchored-content keys.
// ---------------------------------------------------------------------------------------

// Corner k (bit a set = Positive along axis a) sees the grid mirrored on its Negative axes.
// childFlags (optional): only ranks with a non-zero flag enter the table as children.
__global__ __launch_bounds__(256) void kBuildTable2(const uint64_t* __restrict__ sortedMasks, uint32_t n, uint64_t* __restrict__ tabKeys, uint32_t* __restrict__ tabVals, uint32_t logCap,
    uint32_t* __restrict__ bloom, uint32_t logBloom, const uint32_t* __restrict__ childFlags)
{
    const uint32_t gid = blockIdx.x * 256u + threadIdx.x;
    if (gid >= n * 8u)
        return;
    const uint32_t r = gid >> 3;
    const uint32_t corner = gid & 7u;
    const uint64_t m = sortedMasks[r];
    if (m == 0 || (childFlags && childFlags[r] == 0u))
        return;
    const uint64_t key = mirror4(m, (~corner) & 7u);
    const uint32_t f = bloomMix64(key);
    atomicOr(&bloom[bloomWord(f, logBloom)], bloomBits(f));
    const uint32_t capMask = (1u << logCap) - 1u;
    uint32_t s = homeSlot64(key, logCap);
    for (;;) {
        const unsigned long long prev = atomicCAS((unsigned long long*)&tabKeys[s], 0ull, (unsigned long long)key);
        if (prev == 0ull) {
            tabVals[s] = (r << 3) | corner;
            return;
        }
        s = (s + 1) & capMask;
    }
}

// One wave per parent rank (four per block), each with its own work list in LDS.
__global__ __launch_bounds__(256, 8) void kSearch2(const uint64_t* __restrict__ sortedMasks, uint32_t n, const uint64_t* __restrict__ tabKeys, const uint32_t* __restrict__ tabVals,
    uint32_t logCap, const uint32_t* __restrict__ bloom, uint32_t logBloom, uint64_t* __restrict__ best, uint32_t cfgBits, uint32_t rankBase,
    unsigned long long* __restrict__ counters)
{
    __shared__ uint64_t s_qKey[4][kWalkQueue2];
    __shared__ uint16_t s_qMeta[4][kWalkQueue2];
    const uint32_t wave = threadIdx.x >> 6;
    const uint32_t r = rankBase + blockIdx.x * 4u + wave;
    if (r >= n)
        return;
    ProbeCounts cnt;
    search2ParentQueued(r, sortedMasks, tabKeys, tabVals, logCap, bloom, logBloom, best, cfgBits, s_qKey[wave], s_qMeta[wave], cnt);
    flushCounts(cnt, counters, cfgBits);
}

// ---------------------------------------------------------------------------------------
// Levels 3..5: 32-bit nested sub-box hashes computed in LDS, one plane per wave.
// ---------------------------------------------------------------------------------------
// In-register transpose of an R x R bit matrix (bit c of a[i] = M[i][c]; afterwards bit c of
// a[i] = M[c][i]): log2(R) rounds, each exchanging the off-diagonal j x j blocks of every
// 2j x 2j block with a shift, an XOR and a mask.
template <int R>
__device__ __forceinline__ void transposeBits(uint32_t (&a)[R])
{
    constexpr uint32_t kMasks[5] = { 0x55555555u, 0x33333333u, 0x0F0F0F0Fu, 0x00FF00FFu, 0x0000FFFFu };
#pragma unroll
    for (int lj = 4; lj >= 0; --lj) {
        const int j = 1 << lj;
        if (j >= R)
            continue;
        const uint32_t m = kMasks[lj];
#pragma unroll
        for (int i = 0; i < R; ++i) {
            if (i & j)
                continue;
            const uint32_t t = ((a[i] >> j) ^ a[i + j]) & m;
            a[i] ^= t << j;
            a[i + j] ^= t;
        }
    }
}

// Transposed bit layouts 1 and 2 of every item (layout 0 is the sorted masks themselves). One
// thread per (layout, item, slab), layout slowest so a wave makes one layout: layout 1's rows at
// z = slab are the transpose of layout 0's rows (y, z = slab) for y < R (R consecutive rows,
// read as 16-byte or 8-byte vectors); layout 2's rows at y = slab are the transpose of layout
// 0's rows (y = slab, z) for z < R (one row per z; the lanes of consecutive slabs read
// consecutive rows). Both write R consecutive rows as vectors.
template <int L>
__global__ __launch_bounds__(256) void kLayouts(uint64_t* __restrict__ layouts, uint32_t n)
{
    using G = Geo<L>;
    using Row = typename G::Row;
    constexpr int R = G::R;
    using Vec = std::conditional_t<L == 3, uint2, uint4>;
    constexpr int PerVec = (int)(sizeof(Vec) / sizeof(Row));
    static_assert(R % PerVec == 0, "a slab's rows are whole vectors");
    const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= (size_t)n * 2 * R)
        return;
    const int lay = t < (size_t)n * R ? 1 : 2;
    const size_t tl = lay == 1 ? t : t - (size_t)n * R;
    const uint32_t r = (uint32_t)(tl / R);
    const int slab = (int)(tl % R);
    const Row* src = reinterpret_cast<const Row*>(layouts) + (size_t)r * G::R2;
    uint32_t a[R];
    if (lay == 1) {
        const Vec* v = reinterpret_cast<const Vec*>(src + R * slab);
#pragma unroll
        for (int i = 0; i < R / PerVec; ++i) {
            const Vec x = v[i];
            Row rows[PerVec];
            __builtin_memcpy(rows, &x, sizeof(Vec));
#pragma unroll
            for (int e = 0; e < PerVec; ++e)
                a[i * PerVec + e] = (uint32_t)rows[e];
        }
    } else {
#pragma unroll
        for (int i = 0; i < R; ++i)
            a[i] = (uint32_t)src[slab + R * i];
    }
    transposeBits<R>(a);
    Vec* dst = reinterpret_cast<Vec*>(reinterpret_cast<Row*>(layouts) + ((size_t)lay * n + r) * G::R2 + (size_t)R * slab);
#pragma unroll
    for (int i = 0; i < R / PerVec; ++i) {
        Row rows[PerVec];
#pragma unroll
        for (int e = 0; e < PerVec; ++e)
            rows[e] = (Row)a[i * PerVec + e];
        Vec x;
        __builtin_memcpy(&x, rows, sizeof(Vec));
        dst[i] = x;
    }
}

// Shape bitmap of the children (the search's pre-filter, ShapeLds): each wave takes items in
// turn, reads an item's layout-0 rows, derives its occupancy along x, y and z, and sets the bits
// of its extent triple in all six axis orders in the block's LDS copy, which is then ORed into
// the global bitmap (one atomic per non-zero word and block, instead of six per item on a few
// hundred words). Empty and unflagged items set nothing. The flush's atomics all land on those
// few hundred words, so each block costs a fixed share of time: level 5 (4 KiB per item) wants
// the waves, levels 4 and 3 fewer blocks (city, us per call at 512 / 1024 / 2048 / 4096 blocks:
// level 5 155 / 89 / 67 / 75, level 4 54 / 42 / 46 / 63, level 3 30 / 31 / 45 / 77).
template <int L>
constexpr uint32_t kShapeBlocks = L == 5 ? 2048 : 1024;
template <int L>
__global__ __launch_bounds__(256) void kShapeBits(const uint64_t* __restrict__ layouts, uint32_t n, uint32_t* __restrict__ bits, const uint32_t* __restrict__ childFlags)
{
    using G = Geo<L>;
    using Row = typename G::Row;
    __shared__ uint32_t s_bits[shapeWords<L>()];
    for (int i = (int)threadIdx.x; i < shapeWords<L>(); i += 256)
        s_bits[i] = 0u;
    __syncthreads();
    constexpr uint8_t kPerm[6][3] = { { 0, 1, 2 }, { 0, 2, 1 }, { 1, 0, 2 }, { 1, 2, 0 }, { 2, 0, 1 }, { 2, 1, 0 } };
    const uint32_t lane = threadIdx.x & 63u;
    for (uint32_t r = blockIdx.x * 4u + (threadIdx.x >> 6); r < n; r += gridDim.x * 4u) {
        if (childFlags && childFlags[r] == 0u)
            continue;
        const Row* g = reinterpret_cast<const Row*>(layouts) + (size_t)r * G::R2;
        uint32_t ax = 0, ay = 0, az = 0;
        for (int j = (int)lane; j < G::R2; j += 64)
            occupancyAdd((uint32_t)g[j], j, L, ax, ay, az);
        ax = waveOr(ax);
        ay = waveOr(ay);
        az = waveOr(az);
        if (lane < 6u && ax != 0u) {
            const uint32_t e[3] = { spanOf(ax), spanOf(ay), spanOf(az) };
            atomicOr(&s_bits[shapeWord(L, e[kPerm[lane][1]], e[kPerm[lane][2]])], 1u << shapeBit(e[kPerm[lane][0]]));
        }
    }
    __syncthreads();
    for (int i = (int)threadIdx.x; i < shapeWords<L>(); i += 256)
        if (s_bits[i] != 0u)
            atomicOr(&bits[i], s_bits[i]);
}

// Corner table of levels 3..5: for every item and permutation code q < F (1 frame: q = 0 only;
// 6 frames: every q), the hash of the whole grid in the pass frame of q seen from each of its 8
// corners goes into the table and the Bloom filter. Pass a reads child axis sigma_a with
// sigma = q^-1, so the value equals what the search computes, at the voxel where every scan
// ends, for a parent whose content is the child read in that frame. Passes 1 and 2 depend on the
// corner's first two direction bits only, so they run once per (direction pos0 | pos1 << 1, line
// c2 along sigma_2) and leave their end states in LDS; pass 3 then runs per corner over them.

// Stages layout `a` of item r into the LDS rows (row j at j + j / R) with the 64 lanes of one
// wave, every lane's loads issued before any is used; a chunk of rows (4 at level 5, 8 below)
// stays within one line of R rows. Returns whether this lane saw a non-zero row.
template <int L>
__device__ __forceinline__ int stageRows(const uint64_t* __restrict__ layouts, uint32_t n, uint32_t r, uint32_t a, typename Geo<L>::Row* s_rows, uint32_t lane)
{
    using G = Geo<L>;
    using Row = typename G::Row;
    using Chunk = std::conditional_t<L == 3, uint2, uint4>;
    constexpr int ChunkRows = (int)(sizeof(Chunk) / sizeof(Row));
    constexpr int Chunks = G::R2 / ChunkRows;
    constexpr int PerLane = (Chunks + 63) / 64;
    static_assert(G::R % ChunkRows == 0, "a chunk stays within one line of rows");
    const Chunk* g = reinterpret_cast<const Chunk*>(layouts) + ((size_t)a * n + r) * Chunks;
    Chunk c[PerLane];
#pragma unroll
    for (int i = 0; i < PerLane; ++i) {
        const uint32_t k = lane + 64u * i;
        c[i] = (Chunks % 64 == 0 || k < (uint32_t)Chunks) ? g[k] : Chunk {};
    }
    int any = 0;
#pragma unroll
    for (int i = 0; i < PerLane; ++i) {
        const uint32_t k = lane + 64u * i;
        if (Chunks % 64 != 0 && k >= (uint32_t)Chunks)
            break;
        const uint32_t j = k * ChunkRows;
        Row rows[ChunkRows];
        __builtin_memcpy(rows, &c[i], sizeof(Chunk));
#pragma unroll
        for (int e = 0; e < ChunkRows; ++e) {
            s_rows[j + (j >> L) + e] = rows[e];
            any |= (rows[e] != 0);
        }
    }
    return any;
}

// Passes 1 and 2 of frame q for direction bits `dir` (pos0 | pos1 << 1) along line c2, over the
// LDS rows of layout sigma_0 (`rows`): the rows of layout sigma_0 are indexed by the coordinates
// along the other two axes in ascending axis order.
template <int L>
__device__ __forceinline__ uint32_t cornerPass12(const typename Geo<L>::Row* rows, uint32_t q, uint32_t dir, int c2)
{
    using G = Geo<L>;
    constexpr int R = G::R;
    // sigma = q^-1: pass a reads the child axis b with permAxis(q, b) == a.
    uint32_t sigma[3];
#pragma unroll
    for (uint32_t b = 0; b < 3; ++b)
        sigma[permAxis(q, b)] = b;
    const bool lowFirst = sigma[1] < sigma[2];
    const bool pos0 = dir & 1u, pos1 = dir & 2u;
    const int c0 = pos0 ? 0 : R - 1;
    const int sh = pos0 ? c0 : 31 - c0;
    uint32_t s = 0;
    for (int k = 0; k < R; ++k) {
        const int c1 = pos1 ? (R - 1 - k) : k;
        const uint32_t row = (uint32_t)rows[lowFirst ? c1 + G::Pitch * c2 : c2 + G::Pitch * c1];
        s = hashStep(s, pos0 ? rowPrefix<true>(row, sh) : rowPrefix<false>(row, sh));
    }
    return s;
}

// sigma_0 of frame q (the child axis its first pass reads).
__device__ __forceinline__ uint32_t firstPassAxis(uint32_t q)
{
    return permAxis(q, 0) == 0u ? 0u : (permAxis(q, 1) == 0u ? 1u : 2u);
}

// Pass 3 of (frame q, corner) over the pass-2 end states `p2` (R per direction), then the table
// entry (the Bloom filter is set afterwards from a sweep of the table, see launchCornerTable).
// The t-th entry of a home slot starts probing at home + t: slots home .. home + t - 1 are the
// starts of the entries before it, so each is taken by the time the walks end and the probe run
// from home stays unbroken. The entries of one hash (up to 48:
// the frames and corners of children that are transforms of one another) thus take about one
// attempt each instead of walking past all the earlier ones, and they sit in the run in the
// order of their tickets.
template <int L>
__device__ __forceinline__ void cornerInsert(const uint32_t (*p2)[Geo<L>::R], uint32_t r, uint32_t q, uint32_t corner, uint64_t* __restrict__ table,
    uint32_t logCap, uint32_t* __restrict__ tickets)
{
    constexpr int R = Geo<L>::R;
    const uint32_t dir = corner & 3u;
    const bool pos2 = corner & 4u;
    uint32_t s3 = 0;
    for (int k = 0; k < R; ++k)
        s3 = hashStep(s3, p2[dir][pos2 ? (R - 1 - k) : k]);
    const uint32_t h = finalHash(s3);
    if (h == 0)
        return;
    const uint64_t packed = packEntry(h, r, q, corner);
    const uint32_t capMask = (1u << logCap) - 1u;
    const uint32_t home = homeSlot32(h, logCap);
    uint32_t slot = (home + atomicAdd(&tickets[home], 1u)) & capMask;
    for (;;) {
        const unsigned long long prev = atomicCAS((unsigned long long*)&table[slot], 0ull, (unsigned long long)packed);
        if (prev == 0ull)
            break;
        slot = (slot + 1) & capMask;
    }
}

// One wave per (item, frame q = blockIdx.y): passes 1-2 run as 4 directions x R lines over the
// 64 lanes, then lane `corner` < 8 runs pass 3 and inserts the corner's entry.
// Frame-major dispatch hands the entries of one hash their tickets mostly in ascending frame,
// the order of their keys for parents read in frame 0, so a search verifies the best one first
// and skips the rest.
template <int L>
__global__ __launch_bounds__(64) void kCornerTable(const uint64_t* __restrict__ layouts, uint32_t n, uint64_t* __restrict__ table, uint32_t logCap,
    const uint32_t* __restrict__ childFlags, uint32_t* __restrict__ tickets)
{
    using G = Geo<L>;
    using Row = typename G::Row;
    constexpr int R = G::R;
    __shared__ Row s_rows[G::Rows];
    __shared__ uint32_t s_p2[4][R]; // end states of passes 1-2 per (pos0 | pos1 << 1) and line c2
    const uint32_t r = blockIdx.x;
    const uint32_t q = blockIdx.y;
    if (childFlags && childFlags[r] == 0u) // only flagged ranks enter the table as children
        return;
    const uint32_t lane = threadIdx.x;
    if (__syncthreads_or(stageRows<L>(layouts, n, r, firstPassAxis(q), s_rows, lane)) == 0)
        return;
    for (uint32_t t = lane; t < 4u * R; t += 64)
        s_p2[t >> L][t & (R - 1)] = cornerPass12<L>(s_rows, q, t >> L, (int)(t & (R - 1)));
    __syncthreads();
    if (lane < 8)
        cornerInsert<L>(s_p2, r, q, lane, table, logCap, tickets);
}

// Search kernel for levels 3..5. One block per (parent rank, unit, plane chunk) with `Waves`
// waves; every wave owns P = 64/R planes (fixed coordinate along the first pass axis) and runs
// all passes, the filter and the table walk for them on its own: planes are independent, so
// after the rows are loaded there is no block barrier. Units per parent (searchUnits):
//   kUnitCombo:  48 (permutation, direction) combos, or the 8 directions of permutation 0 when
//                the table holds the children in all six frames (kCfgChildFrames).
//   kUnitFamily: 6 families (first-pass axis a0, direction dx) of 8 combos each, or 2 families
//                of 4 directions; the rows are loaded once per family.
//   kUnitShared: as kUnitFamily, and passes 1 and 2 run once for both directions of pass 3
//                (their states stay in a second plane buffer).
// The grid is one-dimensional. XcdGrouped: decoded by xcdBlock, so all (unit, chunk) blocks of
// one parent run back to back on one XCD and share its L2 copy of the parent's rows; otherwise
// the parent varies fastest. There are MaxWaves / Waves plane chunks per unit. The autotuner
// picks the variant per level.
template <int L, int Waves, int Kind, bool XcdGrouped>
__global__ __launch_bounds__(64 * Waves) void kSearch(const uint64_t* __restrict__ layouts, uint32_t n, const uint64_t* __restrict__ table, uint32_t logCap,
    const uint32_t* __restrict__ bloom, uint32_t logBloom, uint64_t* __restrict__ best, uint32_t cfgBits, unsigned long long* __restrict__ counters,
    uint32_t rankBase, uint32_t numParents, const uint32_t* __restrict__ shapeBits)
{
    using G = Geo<L>;
    constexpr int Buffers = Kind == kUnitShared ? 2 : 1;
    __shared__ typename G::Row s_rows[G::Rows];
    __shared__ uint32_t s_planes[Buffers][Waves * G::P * G::Slab];
    __shared__ uint32_t s_occ[3 * Waves];
    __shared__ uint32_t s_qMeta[Waves][kProbeQueue];
    __shared__ uint32_t s_qHash[Waves][kProbeQueue];
    constexpr bool ShapeOn = shapeFits<L, Waves, Kind>();
    __shared__ typename G::Row s_pre[ShapeOn ? 2 * G::R * G::Pitch : 1];
    __shared__ uint32_t s_n1[ShapeOn ? Waves * G::P * G::R : 1];
    __shared__ uint32_t s_shape[ShapeOn ? shapeLdsWords<L>() : 1];
    ProbeCounts cnt;
    const uint32_t Units = searchUnits(Kind, cfgBits);
    constexpr uint32_t Chunks = (uint32_t)(G::MaxWaves / Waves);
    uint32_t parent, member;
    if constexpr (XcdGrouped) {
        const XcdBlock xb = xcdBlock(blockIdx.x, Units * Chunks);
        parent = xb.group;
        member = xb.member;
    } else { // parent fastest: consecutive blocks run the same (unit, chunk) for consecutive parents
        parent = blockIdx.x % numParents;
        member = blockIdx.x / numParents;
    }
    if (parent >= numParents)
        return;
    const uint32_t r = rankBase + parent;
    const uint32_t unit = member / Chunks;
    const uint32_t chunk = member % Chunks;
    if (!(cfgBits & kCfgAxis) && (Kind == kUnitCombo ? (unit >> 3) != 0 : (unit >> 1) != 0))
        return; // only permutation code 0 (p0 == 0) is allowed
    // p0 of the unit's permutation codes: the axis of the first pass and the one layout it reads.
    const uint32_t a0 = Kind == kUnitCombo ? (unit >> 3) >> 1 : unit >> 1;
    Occupancy occ;
    if (!loadRowsOcc<L, 64 * Waves>(layouts, n, r, a0, s_rows, s_occ, occ))
        return;
    const int wave = (int)(threadIdx.x >> 6);
    // Shape pre-filter: only where the parent runs permutation code 0 alone (a0 = 0; rows
    // (a1, a2) = (y, z) at a1 + Pitch * a2).
    if constexpr (ShapeOn) {
        const bool parentPerms = (cfgBits & kCfgAxis) != 0 && (cfgBits & kCfgChildFrames) == 0;
        if (shapeBits != nullptr && !parentPerms) {
            buildShapeTables<L>(s_rows, shapeBits, s_shape, s_pre, (int)threadIdx.x, 64 * Waves);
            __syncthreads();
            ShapeLds sh;
            sh.pre = s_pre;
            sh.n1 = s_n1;
            sh.bits = s_shape;
            searchUnit<L, Waves, Kind, true>(r, unit, chunk, wave, s_rows, occ, s_planes[0], s_planes[Buffers - 1], s_qMeta[wave], s_qHash[wave], layouts, n,
                table, logCap, bloom, logBloom, best, cfgBits, cnt, sh);
            flushCounts(cnt, counters, cfgBits);
            return;
        }
    }
    searchUnit<L, Waves, Kind>(r, unit, chunk, wave, s_rows, occ, s_planes[0], s_planes[Buffers - 1], s_qMeta[wave], s_qHash[wave], layouts, n, table, logCap,
        bloom, logBloom, best, cfgBits, cnt);
    flushCounts(cnt, counters, cfgBits);
}

__global__ __launch_bounds__(256) void kPrunePacked(uint64_t* __restrict__ table, uint64_t cap, const uint64_t* __restrict__ best, uint64_t minKey,
    uint32_t* __restrict__ bloom, uint32_t logBloom)
{
    prunePacked(table, cap, best, minKey, bloom, logBloom, (uint64_t)blockIdx.x * 256 + threadIdx.x, (uint64_t)gridDim.x * 256);
}

__global__ __launch_bounds__(256) void kPruneSplit(const uint64_t* __restrict__ tabKeys, uint32_t* __restrict__ tabVals, uint64_t cap, const uint64_t* __restrict__ best,
    uint64_t minKey, uint32_t* __restrict__ bloom, uint32_t logBloom)
{
    pruneSplit(tabKeys, tabVals, cap, best, minKey, bloom, logBloom, (uint64_t)blockIdx.x * 256 + threadIdx.x, (uint64_t)gridDim.x * 256);
}

// Blocks of 256 threads for a sweep of cap slots, at most 8192 (the sweeps stride over the rest):
// one slot per thread for the split table, four for the packed one.
unsigned pruneBlocks(uint64_t cap)
{
    return (unsigned)std::min<uint64_t>(ceilDiv(cap, 256), 8192);
}
unsigned pruneBlocksPacked(uint64_t cap)
{
    return (unsigned)std::min<uint64_t>(ceilDiv(cap, 1024), 8192);
}

// The table's entries, then its Bloom filter: a prune pass with key 0 keeps every entry and sets
// the filter bits of all of them, a few atomics per wave instead of one per entry inserted (the
// table's slots run in filter-word order; the filter was cleared by the caller).
template <int L>
void launchCornerTable(const MatchPlan& p, const LevelView& v, const uint32_t* childFlags, hipStream_t stream)
{
    const uint32_t n = (uint32_t)p.n;
    hipLaunchKernelGGL(kCornerTable<L>, dim3(n, p.childFrames), dim3(64), 0, stream, v.sortedMasks, n, v.tabKeys, p.logCap, childFlags, v.tickets);
    const uint64_t cap = 1ull << p.logCap;
    hipLaunchKernelGGL(kPrunePacked, dim3(pruneBlocksPacked(cap)), dim3(256), 0, stream, v.tabKeys, cap, v.best, 0ull, v.bloom, p.logBloom);
}


// ---------------------------------------------------------------------------------------
// Post-processing (find_translation.cpp:923-997).
// ---------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void kFinalParents(const uint64_t* __restrict__ best, const uint32_t* __restrict__ order, uint32_t n, uint32_t level,
    uint64_t* __restrict__ outEncoded, uint64_t* __restrict__ parentTp, uint32_t* __restrict__ isUnique, unsigned long long* __restrict__ counters)
{
    const uint32_t c = blockIdx.x * 256u + threadIdx.x;
    if (c >= n)
        return;
    const uint64_t key = best[c];
    const uint32_t orig = order[c];
    if (outEncoded)
        outEncoded[orig] = key;
    if (key == sentinelKey(level)) {
        parentTp[orig] = tpSentinel();
        isUnique[orig] = 1u;
        return;
    }
    const uint64_t pr = keyPtr(key, level);
    const uint32_t parentOrig = order[pr];
    parentTp[orig] = keyToTp(key, level, parentOrig);
    isUnique[parentOrig] = 1u;
    if (pr != c)
        atomicAdd(&counters[0], 1ull);
}

__global__ __launch_bounds__(256) void kFinalMapping(const uint64_t* __restrict__ parentTp, const uint32_t* __restrict__ isUnique, const uint32_t* __restrict__ newIdx, uint32_t n,
    uint64_t* __restrict__ mapping, uint32_t* __restrict__ uniqueIdx)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n)
        return;
    if (isUnique[i]) {
        if (mapping)
            mapping[i] = tpIdentity(newIdx[i]);
        if (uniqueIdx)
            uniqueIdx[newIdx[i]] = i;
    } else if (mapping) {
        const uint64_t p = parentTp[i];
        const uint32_t parentOrig = (uint32_t)(p & ((1ull << 37) - 1));
        mapping[i] = (p & ~((1ull << 37) - 1)) | (uint64_t)newIdx[parentOrig];
    }
}

#define TDAG_TRY(expr)                   \
    do {                                 \
        const hipError_t e_ = (expr);    \
        if (e_ != hipSuccess)            \
            return e_;                   \
    } while (0)

template <int L, int Waves, int Kind, bool XcdGrouped>
hipError_t launchSearchShape(uint32_t parents, uint32_t r0, uint32_t n, const uint64_t* layouts, const uint64_t* table, uint32_t logCap, const uint32_t* bloom,
    uint32_t logBloom, uint64_t* best, uint32_t cfgBits, unsigned long long* counters, const uint32_t* shape, hipStream_t stream)
{
    if constexpr (Waves >= 1 && Waves <= Geo<L>::MaxWaves && searchLds<L, Waves, Kind>() <= kLdsBytes) {
        const uint32_t perParent = searchUnits(Kind, cfgBits) * (uint32_t)(Geo<L>::MaxWaves / Waves);
        const dim3 grid(XcdGrouped ? xcdGridBlocks(parents, perParent) : parents * perParent);
        hipLaunchKernelGGL((kSearch<L, Waves, Kind, XcdGrouped>), grid, dim3(64 * Waves), 0, stream, layouts, n, table, logCap, bloom, logBloom, best, cfgBits,
            counters, r0, parents, shape);
        return hipGetLastError();
    } else {
        return hipErrorInvalidValue;
    }
}

// Variant v: unit kind v % 3, plane-chunk split 2^((v / 3) % 3) (a block covers
// MaxWaves >> ((v / 3) % 3) waves' worth of planes), block order v / 9 (0: parent fastest,
// 1: XCD-grouped).
template <int L>
hipError_t launchSearch(uint32_t variant, uint32_t n, uint32_t r0, uint32_t r1, const uint64_t* layouts, const uint64_t* table, uint32_t logCap, const uint32_t* bloom,
    uint32_t logBloom, uint64_t* best, uint32_t cfgBits, unsigned long long* counters, const uint32_t* shape, hipStream_t stream)
{
    constexpr int M = Geo<L>::MaxWaves;
    const uint32_t p = r1 - r0;
#define TDAG_SHAPE(W, K, X) launchSearchShape<L, W, K, X>(p, r0, n, layouts, table, logCap, bloom, logBloom, best, cfgBits, counters, shape, stream)
    switch (variant) {
    case 0:
        return TDAG_SHAPE(M, kUnitCombo, false);
    case 1:
        return TDAG_SHAPE(M, kUnitFamily, false);
    case 2:
        return TDAG_SHAPE(M, kUnitShared, false);
    case 3:
        return TDAG_SHAPE(M / 2, kUnitCombo, false);
    case 4:
        return TDAG_SHAPE(M / 2, kUnitFamily, false);
    case 5:
        return TDAG_SHAPE(M / 2, kUnitShared, false);
    case 6:
        return TDAG_SHAPE(M / 4, kUnitCombo, false);
    case 7:
        return TDAG_SHAPE(M / 4, kUnitFamily, false);
    case 8:
        return TDAG_SHAPE(M / 4, kUnitShared, false);
    case 9:
        return TDAG_SHAPE(M, kUnitCombo, true);
    case 10:
        return TDAG_SHAPE(M, kUnitFamily, true);
    case 11:
        return TDAG_SHAPE(M, kUnitShared, true);
    case 12:
        return TDAG_SHAPE(M / 2, kUnitCombo, true);
    case 13:
        return TDAG_SHAPE(M / 2, kUnitFamily, true);
    case 14:
        return TDAG_SHAPE(M / 2, kUnitShared, true);
    case 15:
        return TDAG_SHAPE(M / 4, kUnitCombo, true);
    case 16:
        return TDAG_SHAPE(M / 4, kUnitFamily, true);
    case 17:
        return TDAG_SHAPE(M / 4, kUnitShared, true);
    default:
        return hipErrorInvalidValue;
    }
#undef TDAG_SHAPE
}

uint32_t log2Ceil(uint64_t v)
{
    uint32_t l = 0;
    while ((1ull << l) < v)
        ++l;
    return l;
}

} // namespace

bool variantValid(uint32_t level, uint32_t variant)
{
    if (level == 2)
        return variant == 0;
    if (level < 2 || level > 5 || variant >= kNumVariants)
        return false;
    const uint32_t R = 1u << level;
    const uint32_t maxWaves = R * R / 64u;
    const uint32_t waves = maxWaves >> ((variant / 3) % 3);
    if (waves < 1u)
        return false;
    const uint32_t rowBytes = R / 8u;
    const size_t lds = 3ull * R * (R + 1) * rowBytes + (variant % 3 == 2 ? 2ull : 1ull) * waves * (64u / R) * R * (R + 1) * 4ull;
    return lds <= kLdsBytes;
}

uint32_t defaultVariant(uint32_t level)
{
    switch (level) {
    case 5:
        return 16; // family units, a quarter of the planes per block, XCD-grouped
    case 4:
        return 9; // combo units, XCD-grouped
    case 3:
        return 2; // family units sharing passes 1 and 2, parent fastest
    default:
        return 0;
    }
}

uint32_t defaultFilterBits(uint32_t level)
{
    return level == 3 ? 16u : 4u;
}

uint32_t defaultChildFrames(uint32_t level)
{
    return level >= 4 ? 6u : 1u;
}

uint32_t searchCfgBits(const MatchPlan& p, uint32_t cfgBits)
{
    return (p.level >= 3 && p.childFrames == 6) ? (cfgBits | kCfgChildFrames) : (cfgBits & ~kCfgChildFrames);
}

MatchPlan planMatch(uint32_t level, uint64_t n, uint32_t variant, uint32_t firstBatch, uint32_t batchGrowth, uint32_t bloomBitsPerEntry, uint32_t childFrames)
{
    MatchPlan p {};
    p.level = level;
    p.firstBatch = firstBatch ? firstBatch : (uint32_t)std::max<uint64_t>(64, n / 64);
    p.batchGrowth = std::max<uint32_t>(2, batchGrowth ? batchGrowth : 2);
    p.variant = variantValid(level, variant) ? variant : defaultVariant(level);
    p.n = n;
    p.R = 1u << level;
    p.R3 = p.R * p.R * p.R;
    p.W = p.R3 >= 64 ? p.R3 / 64 : 1;
    p.childFrames = level == 2 ? 1u : (childFrames == 1 || childFrames == 6 ? childFrames : defaultChildFrames(level));
    const uint64_t entries = 8ull * p.childFrames * (n > 0 ? n : 1);
    p.logCap = std::max<uint32_t>(6, log2Ceil(4 * entries)); // load factor <= 1/4
    const uint64_t cap = 1ull << p.logCap;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t o = off;
        off += alignUp(bytes > 0 ? bytes : 1, 256);
        return o;
    };
    p.offKeys = take(n * 4);
    p.offOrder = take(n * 4);
    p.offSortScratch = take(radixScratchBytes(n, 4));
    p.offSortedMasks = take((level == 2 ? 1 : 3) * n * p.W * 8); // layouts 0..2 for levels >= 3
    p.offBest = take(n * 8);
    p.offTabKeys = take(cap * 8);
    p.offTabVals = take(level == 2 ? cap * 4 : 0);
    p.offTickets = take(level >= 3 ? cap * 4 : 0); // entries per home slot while the table is built
    // Filter words = entries * bitsPerEntry / 32 (rounded up to a power of two).
    p.bloomBitsPerEntry = bloomBitsPerEntry ? bloomBitsPerEntry : defaultFilterBits(level);
    p.logBloom = std::min<uint32_t>(22, std::max<uint32_t>(6, log2Ceil((entries * p.bloomBitsPerEntry + 31) / 32)));
    p.offBloom = take((1ull << p.logBloom) * 4);
    p.offShape = take(level >= 3 ? ((size_t)4 << (2 * level)) : 0);
    p.offParentTp = take(n * 8);
    p.offIsUnique = take(n * 4);
    p.offNewIdx = take(n * 4);
    p.offScanScratch = take(scanScratchBytes(n));
    p.offCounters = take(64);
    p.scratchBytes = off;
    return p;
}

LevelView levelView(const MatchPlan& p, void* d_scratch)
{
    char* base = (char*)d_scratch;
    LevelView v {};
    v.level = p.level;
    v.n = (uint32_t)p.n;
    v.sortedMasks = (uint64_t*)(base + p.offSortedMasks);
    v.best = (uint64_t*)(base + p.offBest);
    v.tabKeys = (uint64_t*)(base + p.offTabKeys);
    v.tabVals = (uint32_t*)(base + p.offTabVals);
    v.tickets = p.level >= 3 ? (uint32_t*)(base + p.offTickets) : nullptr;
    v.bloom = (uint32_t*)(base + p.offBloom);
    v.shape = p.level >= 3 ? (uint32_t*)(base + p.offShape) : nullptr;
    v.counters = (unsigned long long*)(base + p.offCounters);
    v.logCap = p.logCap;
    v.logBloom = p.logBloom;
    v.childFrames = p.childFrames;
    return v;
}

std::vector<uint32_t> batchEnds(const MatchPlan& p)
{
    std::vector<uint32_t> ends;
    const uint32_t n = (uint32_t)p.n;
    uint32_t r0 = 0;
    uint32_t batch = std::max<uint32_t>(1u, std::min<uint32_t>(n, p.firstBatch));
    while (r0 < n) {
        const uint32_t r1 = (uint32_t)std::min<uint64_t>(n, (uint64_t)r0 + batch);
        ends.push_back(r1);
        r0 = r1;
        batch = (uint32_t)std::min<uint64_t>((uint64_t)batch * p.batchGrowth, n);
    }
    return ends;
}

uint64_t pruneKey(uint32_t level, uint32_t r1)
{
    const int R = 1 << level;
    return encodeKey(r1, 0, 0, 1 - R, 1 - R, 1 - R, level);
}

hipError_t enqueuePrepareOrder(const MatchPlan& p, const uint64_t* d_masks, void* d_scratch, hipStream_t stream)
{
    char* base = (char*)d_scratch;
    uint32_t* keys = (uint32_t*)(base + p.offKeys);
    uint32_t* order = (uint32_t*)(base + p.offOrder);
    const LevelView v = levelView(p, d_scratch);
    const uint32_t n = v.n;
    const uint32_t W = p.W;

    TDAG_TRY(fillU64((uint64_t*)v.counters, kNumCounters, 0ull, stream));
    if (n == 0)
        return hipSuccess;
    // 1. Popcount keys and the stable popcount-descending rank order.
    {
        const unsigned blocks256 = (unsigned)ceilDiv((size_t)n * std::min<uint32_t>(W, 64), 256);
        if (W >= 64)
            hipLaunchKernelGGL(kPopKeys<64>, dim3(blocks256), dim3(256), 0, stream, d_masks, n, W, p.R3, keys, order);
        else if (W >= 8)
            hipLaunchKernelGGL(kPopKeys<8>, dim3((unsigned)ceilDiv((size_t)n * 8, 256)), dim3(256), 0, stream, d_masks, n, W, p.R3, keys, order);
        else
            hipLaunchKernelGGL(kPopKeys<1>, dim3((unsigned)ceilDiv(n, 256)), dim3(256), 0, stream, d_masks, n, W, p.R3, keys, order);
        TDAG_TRY(hipGetLastError());
        TDAG_TRY(radixSortPairsU32(keys, order, n, 0, 16, base + p.offSortScratch, stream));
    }
    // 2. Masks in rank order; best[] seeded with the self-match.
    {
        const size_t total = (size_t)n * W;
        const unsigned blocks = (unsigned)std::min<size_t>(ceilDiv(total, 256), 65536);
        hipLaunchKernelGGL(kGatherInit, dim3(blocks), dim3(256), 0, stream, d_masks, order, keys, n, W, p.R3, p.level, v.sortedMasks, v.best);
        TDAG_TRY(hipGetLastError());
    }
    return hipSuccess;
}

hipError_t enqueueBuildTable(const MatchPlan& p, void* d_scratch, hipStream_t stream, const uint32_t* childFlags)
{
    const LevelView v = levelView(p, d_scratch);
    const uint32_t n = v.n;
    const uint64_t cap = 1ull << p.logCap;
    if (n == 0)
        return hipSuccess;
    // 3. Child table and its Bloom filter (levels >= 3 first build the transposed bit layouts).
    TDAG_TRY(fillU64(v.tabKeys, cap, 0ull, stream));
    TDAG_TRY(fillU32(v.bloom, 1ull << p.logBloom, 0u, stream));
    switch (p.level) {
    case 2:
        hipLaunchKernelGGL(kBuildTable2, dim3((unsigned)ceilDiv((size_t)n * 8, 256)), dim3(256), 0, stream, v.sortedMasks, n, v.tabKeys, v.tabVals, p.logCap, v.bloom,
            p.logBloom, childFlags);
        break;
    case 3:
        hipLaunchKernelGGL(kLayouts<3>, dim3((unsigned)ceilDiv((size_t)n * 2 * 8, 256)), dim3(256), 0, stream, v.sortedMasks, n);
        TDAG_TRY(fillU32(v.tickets, cap, 0u, stream));
        launchCornerTable<3>(p, v, childFlags, stream);
        TDAG_TRY(fillU32(v.shape, shapeWords<3>(), 0u, stream));
        hipLaunchKernelGGL(kShapeBits<3>, dim3(std::min<uint32_t>(kShapeBlocks<3>, (n + 3) / 4)), dim3(256), 0, stream, v.sortedMasks, n, v.shape, childFlags);
        break;
    case 4:
        hipLaunchKernelGGL(kLayouts<4>, dim3((unsigned)ceilDiv((size_t)n * 2 * 16, 256)), dim3(256), 0, stream, v.sortedMasks, n);
        TDAG_TRY(fillU32(v.tickets, cap, 0u, stream));
        launchCornerTable<4>(p, v, childFlags, stream);
        TDAG_TRY(fillU32(v.shape, shapeWords<4>(), 0u, stream));
        hipLaunchKernelGGL(kShapeBits<4>, dim3(std::min<uint32_t>(kShapeBlocks<4>, (n + 3) / 4)), dim3(256), 0, stream, v.sortedMasks, n, v.shape, childFlags);
        break;
    case 5:
        hipLaunchKernelGGL(kLayouts<5>, dim3((unsigned)ceilDiv((size_t)n * 2 * 32, 256)), dim3(256), 0, stream, v.sortedMasks, n);
        TDAG_TRY(fillU32(v.tickets, cap, 0u, stream));
        launchCornerTable<5>(p, v, childFlags, stream);
        TDAG_TRY(fillU32(v.shape, shapeWords<5>(), 0u, stream));
        hipLaunchKernelGGL(kShapeBits<5>, dim3(std::min<uint32_t>(kShapeBlocks<5>, (n + 3) / 4)), dim3(256), 0, stream, v.sortedMasks, n, v.shape, childFlags);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t enqueuePrepare(const MatchPlan& p, const uint64_t* d_masks, void* d_scratch, hipStream_t stream)
{
    TDAG_TRY(enqueuePrepareOrder(p, d_masks, d_scratch, stream));
    return enqueueBuildTable(p, d_scratch, stream, nullptr);
}

hipError_t enqueueSearch(const MatchPlan& p, uint32_t cfgBitsIn, void* d_scratch, hipStream_t stream)
{
    const uint32_t cfgBits = searchCfgBits(p, cfgBitsIn);
    const LevelView v = levelView(p, d_scratch);
    const uint32_t n = v.n;
    const uint64_t cap = 1ull << p.logCap;
    // The batches of batchEnds(p), walked in place (no host allocation on this path).
    uint32_t r0 = 0;
    uint32_t batch = std::max<uint32_t>(1u, std::min<uint32_t>(n, p.firstBatch));
    while (r0 < n) {
        const uint32_t r1 = (uint32_t)std::min<uint64_t>(n, (uint64_t)r0 + batch);
        batch = (uint32_t)std::min<uint64_t>((uint64_t)batch * p.batchGrowth, n);
        switch (p.level) {
        case 2:
            hipLaunchKernelGGL(kSearch2, dim3((unsigned)ceilDiv(r1 - r0, 4)), dim3(256), 0, stream, v.sortedMasks, r1, v.tabKeys, v.tabVals, p.logCap, v.bloom,
                p.logBloom, v.best, cfgBits, r0, v.counters);
            TDAG_TRY(hipGetLastError());
            break;
        case 3:
            TDAG_TRY(launchSearch<3>(p.variant, n, r0, r1, v.sortedMasks, v.tabKeys, p.logCap, v.bloom, p.logBloom, v.best, cfgBits, v.counters, v.shape, stream));
            break;
        case 4:
            TDAG_TRY(launchSearch<4>(p.variant, n, r0, r1, v.sortedMasks, v.tabKeys, p.logCap, v.bloom, p.logBloom, v.best, cfgBits, v.counters, v.shape, stream));
            break;
        case 5:
            TDAG_TRY(launchSearch<5>(p.variant, n, r0, r1, v.sortedMasks, v.tabKeys, p.logCap, v.bloom, p.logBloom, v.best, cfgBits, v.counters, v.shape, stream));
            break;
        default:
            return hipErrorInvalidValue;
        }
        r0 = r1;
        if (r0 < n) {
            const uint64_t minKey = pruneKey(p.level, r0);
            TDAG_TRY(fillU32(v.bloom, 1ull << p.logBloom, 0u, stream));
            if (p.level == 2)
                hipLaunchKernelGGL(kPruneSplit, dim3(pruneBlocks(cap)), dim3(256), 0, stream, v.tabKeys, v.tabVals, cap, v.best, minKey, v.bloom, p.logBloom);
            else
                hipLaunchKernelGGL(kPrunePacked, dim3(pruneBlocksPacked(cap)), dim3(256), 0, stream, v.tabKeys, cap, v.best, minKey, v.bloom, p.logBloom);
            TDAG_TRY(hipGetLastError());
        }
    }
    return hipSuccess;
}

hipError_t enqueueFinish(const MatchPlan& p, void* d_scratch, hipStream_t stream, uint64_t* d_encoded, uint64_t* d_mapping, uint32_t* d_unique)
{
    char* base = (char*)d_scratch;
    const LevelView v = levelView(p, d_scratch);
    const uint32_t n = v.n;
    if (n == 0)
        return hipSuccess;
    uint32_t* order = (uint32_t*)(base + p.offOrder);
    uint64_t* parentTp = (uint64_t*)(base + p.offParentTp);
    uint32_t* isUnique = (uint32_t*)(base + p.offIsUnique);
    uint32_t* newIdx = (uint32_t*)(base + p.offNewIdx);
    uint32_t* total = (uint32_t*)(v.counters + 1);
    TDAG_TRY(fillU32(isUnique, n, 0u, stream));
    hipLaunchKernelGGL(kFinalParents, dim3((unsigned)ceilDiv(n, 256)), dim3(256), 0, stream, v.best, order, n, p.level, d_encoded, parentTp, isUnique, v.counters);
    TDAG_TRY(hipGetLastError());
    TDAG_TRY(exclusiveScanU32(isUnique, newIdx, n, total, base + p.offScanScratch, stream));
    hipLaunchKernelGGL(kFinalMapping, dim3((unsigned)ceilDiv(n, 256)), dim3(256), 0, stream, parentTp, isUnique, newIdx, n, d_mapping, d_unique);
    return hipGetLastError();
}

hipError_t enqueueMatch(const MatchPlan& p, const uint64_t* d_masks, uint32_t cfgBits, void* d_scratch, hipStream_t stream, uint64_t* d_encoded, uint64_t* d_mapping,
    uint32_t* d_unique, hipEvent_t evSearch0, hipEvent_t evSearch1)
{
    TDAG_TRY(enqueuePrepare(p, d_masks, d_scratch, stream));
    if (evSearch0)
        TDAG_TRY(hipEventRecord(evSearch0, stream));
    if (p.n > 0)
        TDAG_TRY(enqueueSearch(p, cfgBits, d_scratch, stream));
    if (evSearch1)
        TDAG_TRY(hipEventRecord(evSearch1, stream));
    return enqueueFinish(p, d_scratch, stream, d_encoded, d_mapping, d_unique);
}

} // namespace tdag
