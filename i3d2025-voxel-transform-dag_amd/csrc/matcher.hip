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

#include "common.hpp"
#include "matcher.hpp"
#include "primitives.hpp"

namespace tdag {

namespace {

inline size_t ceilDiv(size_t a, size_t b) { return (a + b - 1) / b; }
inline size_t alignUp(size_t v, size_t a) { return (v + a - 1) / a * a; }

__device__ __forceinline__ uint64_t loadRelaxed(const uint64_t* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ bool configAllows(uint32_t cfgBits, uint32_t permCode, uint32_t sym, int tx, int ty, int tz)
{
    if (!(cfgBits & kCfgAxis) && permCode != 0)
        return false;
    if (!(cfgBits & kCfgSymmetry) && sym != 0)
        return false;
    if (!(cfgBits & kCfgTranslation) && (tx | ty | tz) != 0)
        return false;
    return true;
}

// LDS bytes per workgroup on gfx950.
constexpr size_t kLdsBytes = 160 * 1024;

// Workgroups of a launch are dealt to the 8 XCDs round-robin (block b runs on XCD b mod 8), and
// each XCD has its own L2. A launch of `groups` groups of `perGroup` blocks that share data uses
// a 1-D grid of xcdGridBlocks(groups, perGroup) blocks, decoded so that the k-th block of XCD x
// is member k mod perGroup of group (k div perGroup) * 8 + x: a group's blocks run back to back
// on one XCD. Blocks whose group is >= groups (padding to a multiple of 8) exit at once.
constexpr uint32_t kXcds = 8;

struct XcdBlock {
    uint32_t group, member;
};

__device__ __forceinline__ XcdBlock xcdBlock(uint32_t b, uint32_t perGroup)
{
    const uint32_t k = b / kXcds;
    return { (k / perGroup) * kXcds + b % kXcds, k % perGroup };
}

inline uint32_t xcdGridBlocks(uint32_t groups, uint32_t perGroup)
{
    return (groups + kXcds - 1) / kXcds * kXcds * perGroup;
}

// Child rank stored in a table entry whose child can no longer improve (tombstone; the key stays
// so linear-probing clusters remain intact). Ranks are < 2^28, so this never names an item.
constexpr uint32_t kDeadRank = 0x1FFFFFFFu;

// Per-thread probe statistics, summed per wave into the call's counters when kCfgCount is set.
struct ProbeCounts {
    uint32_t probes = 0, filterPass = 0, keyHits = 0, candidates = 0, verifications = 0, verified = 0;
};

__device__ __forceinline__ uint32_t waveSum(uint32_t v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ void flushCounts(const ProbeCounts& c, unsigned long long* counters, uint32_t cfgBits)
{
    if (!(cfgBits & kCfgCount))
        return;
    const uint32_t v[6] = { waveSum(c.probes), waveSum(c.filterPass), waveSum(c.keyHits), waveSum(c.candidates), waveSum(c.verifications), waveSum(c.verified) };
    if ((threadIdx.x & 63u) == 0)
        for (int i = 0; i < 6; ++i)
            if (v[i])
                atomicAdd(&counters[2 + i], (unsigned long long)v[i]);
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
__global__ __launch_bounds__(256) void kBuildTable2(const uint64_t* __restrict__ sortedMasks, uint32_t n, uint64_t* __restrict__ tabKeys, uint32_t* __restrict__ tabVals, uint32_t logCap,
    uint32_t* __restrict__ bloom, uint32_t logBloom)
{
    const uint32_t gid = blockIdx.x * 256u + threadIdx.x;
    if (gid >= n * 8u)
        return;
    const uint32_t r = gid >> 3;
    const uint32_t corner = gid & 7u;
    const uint64_t m = sortedMasks[r];
    if (m == 0)
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

// One wave per parent rank; lane = anchored voxel v' = (x, y, z) of the sub-box. For each
// permutation the 8 direction keys are formed first and their table loads issued together.
__global__ __launch_bounds__(256) void kSearch2(const uint64_t* __restrict__ sortedMasks, uint32_t n, const uint64_t* __restrict__ tabKeys, const uint32_t* __restrict__ tabVals,
    uint32_t logCap, const uint32_t* __restrict__ bloom, uint32_t logBloom, uint64_t* __restrict__ best, uint32_t cfgBits, uint32_t rankBase,
    unsigned long long* __restrict__ counters)
{
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t r = rankBase + blockIdx.x * 4u + (threadIdx.x >> 6);
    if (r >= n)
        return;
    const uint64_t P = sortedMasks[r];
    if (P == 0)
        return;
    ProbeCounts cnt;
    const int vx = (int)(lane & 3u), vy = (int)((lane >> 2) & 3u), vz = (int)(lane >> 4);
    const uint64_t clip = rangeMaskX4(0, 4 - vx) & rangeMaskY4(0, 4 - vy) & rangeMaskZ4(0, 4 - vz);
    const uint32_t capMask = (1u << logCap) - 1u;
    const uint32_t numPerms = (cfgBits & kCfgAxis) ? 6u : 1u;
    for (uint32_t pc = 0; pc < numPerms; ++pc) {
        // permute (find_translation.cpp:109-125): out(o) = P(v) with v[p[a]] = o[a].
        int v[3];
        v[permAxis(pc, 0)] = vx;
        v[permAxis(pc, 1)] = vy;
        v[permAxis(pc, 2)] = vz;
        const uint32_t src = (uint32_t)(v[0] + 4 * v[1] + 16 * v[2]);
        const uint64_t Q = (uint64_t)__ballot((P >> src) & 1ull);
        uint64_t key[8];
        uint32_t bw[8];
#pragma unroll
        for (uint32_t d = 0; d < 8; ++d)
            key[d] = (mirror4(Q, (~d) & 7u) >> lane) & clip;
#pragma unroll
        for (uint32_t d = 0; d < 8; ++d)
            bw[d] = key[d] ? bloom[bloomWord(bloomMix64(key[d]), logBloom)] : 0u;
#pragma unroll
        for (uint32_t d = 0; d < 8; ++d) {
            const uint32_t need = bloomBits(bloomMix64(key[d]));
            cnt.probes += key[d] != 0;
            if (key[d] == 0 || (bw[d] & need) != need)
                continue;
            ++cnt.filterPass;
            uint32_t s = homeSlot64(key[d], logCap);
            uint64_t k = tabKeys[s];
            while (k != 0) {
                if (k == key[d]) {
                    ++cnt.keyHits;
                    const uint32_t e = tabVals[s];
                    const uint32_t c = e >> 3;
                    const uint32_t kc = e & 7u;
                    if (c > r && c != kDeadRank) {
                        const uint32_t sym = d ^ kc;
                        const int tx = (kc & 1u) ? -vx : vx;
                        const int ty = (kc & 2u) ? -vy : vy;
                        const int tz = (kc & 4u) ? -vz : vz;
                        if (configAllows(cfgBits, pc, sym, tx, ty, tz)) {
                            ++cnt.candidates;
                            const uint64_t k64 = encodeKey(r, pc, sym, tx, ty, tz, 2);
                            if (k64 < loadRelaxed(&best[c])) {
                                ++cnt.verifications;
                                // Bitwise verification of the derived transform.
                                const uint64_t T = translate4(mirror4(Q, sym), tx, ty, tz);
                                if (T == sortedMasks[c]) {
                                    ++cnt.verified;
                                    atomicMin((unsigned long long*)&best[c], (unsigned long long)k64);
                                }
                            }
                        }
                    }
                }
                s = (s + 1) & capMask;
                k = tabKeys[s];
            }
        }
    }
    flushCounts(cnt, counters, cfgBits);
}

// ---------------------------------------------------------------------------------------
// Levels 3..5: 32-bit nested sub-box hashes computed in LDS, one line per thread per pass.
// ---------------------------------------------------------------------------------------
template <int L>
struct RowType;
template <>
struct RowType<3> {
    using T = uint8_t;
};
template <>
struct RowType<4> {
    using T = uint16_t;
};
template <>
struct RowType<5> {
    using T = uint32_t;
};

// Geometry of the level-L (R = 2^L) search. A "row" is the R bits of one line of the grid along
// one axis, stored as an R-bit unsigned integer (bit i = coordinate i).
template <int L>
struct Geo {
    using Row = typename RowType<L>::T;
    static constexpr int R = 1 << L;
    static constexpr int R2 = R * R;
    static constexpr int R3 = R2 * R;
    static constexpr int W = R3 / 64; // u64 words per bit layout
    static constexpr int Pitch = R + 1; // padded pitch of rows in LDS and of lines in a plane
    static constexpr int Rows = R * Pitch; // padded rows per layout in LDS
    static constexpr int P = 64 / R; // planes per wave
    static constexpr int Slab = R * Pitch; // words per plane in LDS
    static constexpr int MaxWaves = R / P; // waves that cover all R planes
};

// Three bit layouts of one item in global memory: layout a has axis a fastest, so the line along
// axis a at (u, w) — its coordinates along the other two axes in ascending axis order — is row
// j = u + R*w, bits [R*j, R*j + R).
//   layout 0: x + R*y + R*R*z (the LargeSubGrid layout)
//   layout 1: y + R*x + R*R*z
//   layout 2: z + R*x + R*R*y
// In LDS row (u, w) of layout a sits at a*Rows + u + Pitch*w; the odd pitch keeps lanes that
// read rows at stride Pitch on distinct banks.
template <int L>
__device__ __forceinline__ uint32_t lbit(const uint64_t* s_bits, int x, int y, int z)
{
    constexpr int R = Geo<L>::R;
    const uint32_t i = (uint32_t)(x + R * y + R * R * z);
    return (uint32_t)((s_bits[i >> 6] >> (i & 63u)) & 1ull);
}

// Loads the three layouts of item r into LDS rows. Returns whether the item is non-empty (the
// result is uniform over the block; contains a block barrier).
template <int L, int Threads>
__device__ __forceinline__ bool loadRows(const uint64_t* __restrict__ layouts, uint32_t n, uint32_t r, typename Geo<L>::Row* s_rows, int numLayouts)
{
    using G = Geo<L>;
    using Row = typename G::Row;
    const Row* g = reinterpret_cast<const Row*>(layouts);
    int any = 0;
    for (int i = (int)threadIdx.x; i < numLayouts * G::R2; i += Threads) {
        const int lay = i >> (2 * L), j = i & (G::R2 - 1);
        const Row v = g[((size_t)lay * n + r) * G::R2 + j];
        s_rows[lay * G::Rows + j + (j >> L)] = v;
        any |= (v != 0);
    }
    return __syncthreads_or(any) != 0;
}

// Orders LDS accesses of the lanes of one wave (LDS operations of a wave execute in order; this
// keeps the compiler from moving them across).
__device__ __forceinline__ void waveSync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Value of the first pass (along A0) at coordinate c0 of a row: the trimmed bit sequence of the
// row segment read from the scan start toward c0 is the binary number formed by those bits, so
// it needs no hashing. Positive (scan from R-1 down to c0): bits [c0, R-1] with bit R-1 most
// significant, i.e. row >> c0. Negative (scan from 0 up to c0): bits [0, c0] with bit 0 most
// significant, i.e. bitreverse32(row) >> (31 - c0). 0 iff the segment is empty. `sh` is c0 or
// 31 - c0 respectively.
template <bool Pos>
__device__ __forceinline__ uint32_t rowPrefix(uint32_t row, int sh)
{
    return Pos ? (row >> sh) : (__brev(row) >> sh);
}

// Passes 1 and 2 of the nested hash on plane c0 (the coordinate along the first pass axis A0):
// the lane owns the line along A1 at c[A2] = u and reads, at each step, the row along A0
// through that voxel. The pass-2 state of voxel (c1, c2) goes to plane[c1 + Pitch*c2].
template <int L, int A0, int A1, bool Pos0, bool Pos1>
__device__ __forceinline__ void pass12(const typename Geo<L>::Row* s_rows, int c0, int u, uint32_t* plane)
{
    using G = Geo<L>;
    constexpr int R = G::R;
    constexpr int A2 = 3 - A0 - A1;
    constexpr int Step = (A1 < A2) ? 1 : G::Pitch; // row (lo, hi) = (c[A1], c[A2]) or (c[A2], c[A1])
    const typename G::Row* rows = s_rows + A0 * G::Rows + ((A1 < A2) ? G::Pitch * u : u);
    uint32_t* line = plane + G::Pitch * u;
    const int sh = Pos0 ? c0 : 31 - c0;
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const int c1 = Pos1 ? (R - 1 - k) : k;
        s = hashStep(s, rowPrefix<Pos0>((uint32_t)rows[Step * c1], sh));
        line[c1] = s;
    }
}

template <int L, int A0, int A1>
__device__ __forceinline__ void pass12Dirs(uint32_t d, const typename Geo<L>::Row* s_rows, int c0, int u, uint32_t* plane)
{
    switch (d & 3u) {
    case 0:
        pass12<L, A0, A1, false, false>(s_rows, c0, u, plane);
        break;
    case 1:
        pass12<L, A0, A1, true, false>(s_rows, c0, u, plane);
        break;
    case 2:
        pass12<L, A0, A1, false, true>(s_rows, c0, u, plane);
        break;
    default:
        pass12<L, A0, A1, true, true>(s_rows, c0, u, plane);
        break;
    }
}

// Passes 1 and 2 for permutation code pc (axes p0, p1) and direction bits d (bit 0: pass 1,
// bit 1: pass 2).
template <int L>
__device__ __forceinline__ void passes12(uint32_t pc, uint32_t d, const typename Geo<L>::Row* s_rows, int c0, int u, uint32_t* plane)
{
    switch (pc) { // transform_dag.h:138-145
    case 0:
        pass12Dirs<L, 0, 1>(d, s_rows, c0, u, plane);
        break;
    case 1:
        pass12Dirs<L, 0, 2>(d, s_rows, c0, u, plane);
        break;
    case 2:
        pass12Dirs<L, 1, 0>(d, s_rows, c0, u, plane);
        break;
    case 3:
        pass12Dirs<L, 1, 2>(d, s_rows, c0, u, plane);
        break;
    case 4:
        pass12Dirs<L, 2, 0>(d, s_rows, c0, u, plane);
        break;
    default:
        pass12Dirs<L, 2, 1>(d, s_rows, c0, u, plane);
        break;
    }
}

// Pass 3 along the lane's line (c1 = u) of its plane over the pass-2 states in[Pitch*c2], fused
// with the Bloom filter: stores the final hash of every voxel at out[Pitch*c2] (out may equal in)
// and returns the bit set of voxels c2 that are non-empty and pass the filter. The filter word is
// loaded unconditionally (every hash indexes a valid word) so the sweep has no branches and its
// loads overlap the hashing.
template <int L, bool Pos2>
__device__ __forceinline__ uint32_t pass3Filter(const uint32_t* in, uint32_t* out, const uint32_t* __restrict__ bloom, uint32_t logBloom, ProbeCounts& cnt)
{
    constexpr int R = Geo<L>::R;
    const uint32_t shift = 32u - logBloom;
    uint32_t st[R];
#pragma unroll
    for (int c2 = 0; c2 < R; ++c2)
        st[c2] = in[Geo<L>::Pitch * c2];
    uint32_t s = 0, pending = 0, nonEmpty = 0;
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const int c2 = Pos2 ? (R - 1 - k) : k;
        s = hashStep(s, st[c2]);
        const uint32_t f = finalHash(s);
        out[Geo<L>::Pitch * c2] = f;
        const uint32_t bw = bloom[f >> shift];
        const uint32_t need = bloomBits(f);
        const uint32_t miss = need & ~bw;
        const uint32_t nz = f < 1u ? f : 1u;
        const uint32_t ok = nz & ((miss < 1u ? miss : 1u) ^ 1u);
        pending |= ok << c2;
        nonEmpty += nz;
    }
    cnt.probes += nonEmpty;
    cnt.filterPass += (uint32_t)__builtin_popcount(pending);
    return pending;
}

// Expected child row (yo, zo) (bits along x) for the transform encoded in `key` applied to the
// parent rows: out[o] = M[o - t] with M = mirror(permute(P)) (find_translation.cpp:66-84 for the
// translation). The source of an output row is one line of the parent along axis p0, read from
// layout p0.
template <int L>
__device__ __forceinline__ uint32_t transformedRow(const typename Geo<L>::Row* s_rows, int tx, int ty, int tz, uint32_t sym, uint32_t pc, int yo, int zo)
{
    using G = Geo<L>;
    constexpr int R = G::R;
    const int q1 = yo - ty, q2 = zo - tz;
    if (q1 < 0 || q1 >= R || q2 < 0 || q2 >= R)
        return 0u;
    const int m1 = (sym & 2u) ? R - 1 - q1 : q1;
    const int m2 = (sym & 4u) ? R - 1 - q2 : q2;
    const int a1 = (int)permAxis(pc, 1), a2 = (int)permAxis(pc, 2);
    // Parent coordinates along a1 and a2; the row along a0 is indexed by them in ascending axis order.
    const int lo = a1 < a2 ? m1 : m2, hi = a1 < a2 ? m2 : m1;
    uint32_t line = (uint32_t)s_rows[(int)permAxis(pc, 0) * G::Rows + lo + G::Pitch * hi];
    if (sym & 1u)
        line = __brev(line) >> (32 - R);
    const uint32_t rowMask = R == 32 ? 0xFFFFFFFFu : ((1u << R) - 1u);
    return (tx >= 0 ? (line << tx) : (line >> -tx)) & rowMask;
}

// Wave-cooperative bitwise check child bc == T(parent) for the transform in key bk.
template <int L>
__device__ __forceinline__ bool verifyWave(const typename Geo<L>::Row* s_rows, const uint64_t* __restrict__ sortedMasks, uint32_t bc, uint64_t bk, uint32_t lane)
{
    using G = Geo<L>;
    constexpr int R = G::R;
    constexpr uint32_t b = L + 1;
    constexpr uint64_t m = (1ull << b) - 1ull;
    const int tx = (int)(bk & m) - R;
    const int ty = (int)((bk >> b) & m) - R;
    const int tz = (int)((bk >> (2 * b)) & m) - R;
    const uint32_t sym = (uint32_t)((bk >> (3 * b)) & 7u);
    const uint32_t pc = (uint32_t)((bk >> (3 * b + 3)) & 7u);
    const typename G::Row* child = reinterpret_cast<const typename G::Row*>(sortedMasks) + (size_t)bc * G::R2;
    bool ok = true;
    for (int j = (int)lane; j < G::R2; j += 64)
        ok = ok && (transformedRow<L>(s_rows, tx, ty, tz, sym, pc, j & (R - 1), j >> L) == (uint32_t)child[j]);
    return __ballot(!ok) == 0ull;
}

// Walks the child table for the voxels in `pending` (bit c2 = voxel (c0, u, c2) of the lane's
// line passed the filter; its hash is in the plane) and merges verified hits into best. Lanes stop
// at each candidate that beats best[c]; the wave then verifies the stopped lanes' candidates
// together and resumes.
template <int L>
__device__ __forceinline__ void probeLine(uint32_t pending, const uint32_t* line, const typename Geo<L>::Row* s_rows,
    const uint64_t* __restrict__ sortedMasks, const uint64_t* __restrict__ table, uint32_t logCap,
    uint64_t* __restrict__ best, uint32_t cfgBits, uint32_t r, uint32_t pc, uint32_t d, int c0, int u, ProbeCounts& cnt)
{
    using G = Geo<L>;
    constexpr int R = G::R;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t capMask = (1u << logCap) - 1u;
    bool active = false;
    uint32_t h = 0, slot = 0;
    int c2 = 0;
    while (true) {
        bool found = false;
        uint32_t hitC = 0;
        uint64_t hitKey = 0;
        while (true) {
            if (!active) {
                if (pending == 0)
                    break;
                c2 = __ffs(pending) - 1;
                pending &= pending - 1u;
                h = line[G::Pitch * c2];
                slot = homeSlot32(h, logCap);
                active = true;
            }
            const uint64_t e = table[slot];
            if (e == 0) {
                active = false;
                continue;
            }
            slot = (slot + 1) & capMask;
            if ((uint32_t)(e >> 32) != h)
                continue;
            ++cnt.keyHits;
            const uint32_t cc = (uint32_t)e >> 3;
            const uint32_t kc = (uint32_t)e & 7u;
            if (cc <= r || cc == kDeadRank)
                continue;
            // Translation of the anchored match (find_translation.cpp:835-837); the voxel's
            // coordinates along the three pass axes are (c0, u, c2).
            const int voxel[3] = { c0, u, c2 };
            const uint32_t sym = d ^ kc;
            int tt[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const int anchor = ((d >> a) & 1u) ? 0 : R - 1;
                tt[a] = ((sym >> a) & 1u) ? voxel[a] - anchor : anchor - voxel[a];
            }
            if (!configAllows(cfgBits, pc, sym, tt[0], tt[1], tt[2]))
                continue;
            ++cnt.candidates;
            const uint64_t k64 = encodeKey(r, pc, sym, tt[0], tt[1], tt[2], L);
            if (k64 >= loadRelaxed(&best[cc]))
                continue;
            ++cnt.verifications;
            found = true;
            hitC = cc;
            hitKey = k64;
            break;
        }
        uint64_t stopped = __ballot(found);
        if (stopped == 0)
            break; // no lane found a candidate, so every lane ran out of work
        while (stopped) {
            const int leader = __ffsll((unsigned long long)stopped) - 1;
            const uint32_t bc = (uint32_t)__shfl((int)hitC, leader, 64);
            const uint64_t bk = (uint64_t)__shfl((long long)hitKey, leader, 64);
            const bool allOk = verifyWave<L>(s_rows, sortedMasks, bc, bk, lane);
            if ((int)lane == leader && allOk) {
                ++cnt.verified;
                atomicMin((unsigned long long*)&best[bc], (unsigned long long)bk);
            }
            stopped &= stopped - 1ull;
        }
    }
}

// Pass 3 in direction dz with the filter (pass-2 states read from `in`, final hashes written to
// `out`), then the table walk of the lane's line.
template <int L>
__device__ __forceinline__ void lastPassProbe(uint32_t dz, const uint32_t* in, uint32_t* out, const typename Geo<L>::Row* s_rows,
    const uint64_t* __restrict__ sortedMasks, const uint64_t* __restrict__ table, uint32_t logCap, const uint32_t* __restrict__ bloom, uint32_t logBloom,
    uint64_t* __restrict__ best, uint32_t cfgBits, uint32_t r, uint32_t pc, uint32_t d, int c0, int u, ProbeCounts& cnt)
{
    const uint32_t pending = dz ? pass3Filter<L, true>(in, out, bloom, logBloom, cnt) : pass3Filter<L, false>(in, out, bloom, logBloom, cnt);
    probeLine<L>(pending, out, s_rows, sortedMasks, table, logCap, best, cfgBits, r, pc, d, c0, u, cnt);
}

// Transposed bit layouts 1 and 2 of every item (layout 0 is the sorted masks themselves).
template <int L>
__global__ __launch_bounds__(256) void kLayouts(uint64_t* __restrict__ layouts, uint32_t n)
{
    constexpr int R = Geo<L>::R;
    constexpr int W = Geo<L>::W;
    __shared__ uint64_t s_bits[W];
    const uint32_t r = blockIdx.x;
    for (int i = (int)threadIdx.x; i < W; i += 256)
        s_bits[i] = layouts[(size_t)r * W + i];
    __syncthreads();
    for (int i = (int)threadIdx.x; i < 2 * W; i += 256) {
        const int lay = 1 + i / W, wi = i % W;
        uint64_t out = 0;
        for (int j = 0; j < 64; ++j) {
            const uint32_t k = (uint32_t)(wi * 64 + j);
            const int f = (int)(k & (R - 1)), a = (int)((k >> L) & (R - 1)), b = (int)(k >> (2 * L));
            // layout 1: (y, x, z) = (f, a, b); layout 2: (z, x, y) = (f, a, b)
            const int x = a;
            const int y = lay == 1 ? f : b;
            const int z = lay == 1 ? b : f;
            out |= (uint64_t)lbit<L>(s_bits, x, y, z) << j;
        }
        layouts[((size_t)lay * n + r) * W + wi] = out;
    }
}

// One wave per item: the hash of the whole grid seen from each of its 8 corners (identity
// permutation, direction = corner) goes into the table and the Bloom filter. It is the value the
// search computes at the voxel where every scan ends, computed with the same functions: R lanes
// per corner run passes 1 and 2 along y for one z each; the corner's first lane folds pass 3.
template <int L>
__global__ __launch_bounds__(64) void kCornerTable(const uint64_t* __restrict__ layouts, uint32_t n, uint64_t* __restrict__ table, uint32_t logCap,
    uint32_t* __restrict__ bloom, uint32_t logBloom)
{
    using G = Geo<L>;
    constexpr int R = G::R;
    __shared__ typename G::Row s_rows[G::Rows];
    const uint32_t r = blockIdx.x;
    if (!loadRows<L, 64>(layouts, n, r, s_rows, 1))
        return;
    const uint32_t capMask = (1u << logCap) - 1u;
    const int lane = (int)threadIdx.x;
    const int c2 = lane & (R - 1), sub = lane >> L;
    for (int cb = 0; cb < 8; cb += G::P) {
        const uint32_t corner = (uint32_t)(cb + sub);
        const bool pos0 = corner & 1u, pos1 = corner & 2u, pos2 = corner & 4u;
        const int c0 = pos0 ? 0 : R - 1;
        const int sh = pos0 ? c0 : 31 - c0;
        uint32_t s = 0;
        for (int k = 0; k < R; ++k) {
            const int c1 = pos1 ? (R - 1 - k) : k;
            const uint32_t row = (uint32_t)s_rows[c1 + G::Pitch * c2]; // layout 0: row along x at (y, z) = (c1, c2)
            s = hashStep(s, pos0 ? rowPrefix<true>(row, sh) : rowPrefix<false>(row, sh));
        }
        uint32_t s3 = 0;
        for (int k = 0; k < R; ++k) {
            const int z = pos2 ? (R - 1 - k) : k;
            s3 = hashStep(s3, (uint32_t)__shfl((int)s, sub * R + z, 64));
        }
        const uint32_t h = finalHash(s3);
        if (c2 == 0 && h != 0) {
            atomicOr(&bloom[bloomWord(h, logBloom)], bloomBits(h));
            const uint64_t packed = ((uint64_t)h << 32) | (uint64_t)((r << 3) | corner);
            uint32_t slot = homeSlot32(h, logCap);
            for (;;) {
                const unsigned long long prev = atomicCAS((unsigned long long*)&table[slot], 0ull, (unsigned long long)packed);
                if (prev == 0ull)
                    break;
                slot = (slot + 1) & capMask;
            }
        }
    }
}

// Search-unit kinds of the level-3..5 kernel.
constexpr int kUnitCombo = 0; // one (permutation, direction) combo per block
constexpr int kUnitFamily = 1; // a family of 8 combos per block, each run from scratch
constexpr int kUnitShared = 2; // a family of 8 combos; passes 1 and 2 shared by both pass-3 directions

// LDS bytes of one kSearch block.
template <int L, int Waves, int Kind>
constexpr size_t searchLds()
{
    return 3 * Geo<L>::Rows * sizeof(typename Geo<L>::Row) + (Kind == kUnitShared ? 2 : 1) * (size_t)Waves * Geo<L>::P * Geo<L>::Slab * 4;
}

// Search kernel for levels 3..5. One block per (parent rank, unit, plane chunk) with `Waves`
// waves; every wave owns P = 64/R planes (fixed coordinate along the first pass axis) and runs
// all passes, the filter and the table walk for them on its own: planes are independent, so
// after the rows are loaded there is no block barrier.
//   kUnitCombo:  grid.y = 48 combos.
//   kUnitFamily: grid.y = 6 families (first-pass axis a0, direction dx) of 8 combos each; the
//                rows are loaded once per family.
//   kUnitShared: as kUnitFamily, and passes 1 and 2 run once for both directions of pass 3
//                (their states stay in a second plane buffer).
// The grid is one-dimensional. XcdGrouped: decoded by xcdBlock, so all (unit, chunk) blocks of
// one parent run back to back on one XCD and share its L2 copy of the parent's rows; otherwise
// the parent varies fastest. There are MaxWaves / Waves plane chunks per unit. The autotuner
// picks the variant per level.
template <int L, int Waves, int Kind, bool XcdGrouped>
__global__ __launch_bounds__(64 * Waves) void kSearch(const uint64_t* __restrict__ layouts, uint32_t n, const uint64_t* __restrict__ table, uint32_t logCap,
    const uint32_t* __restrict__ bloom, uint32_t logBloom, uint64_t* __restrict__ best, uint32_t cfgBits, unsigned long long* __restrict__ counters,
    uint32_t rankBase, uint32_t numParents)
{
    using G = Geo<L>;
    constexpr int R = G::R;
    constexpr int Buffers = Kind == kUnitShared ? 2 : 1;
    __shared__ typename G::Row s_rows[3 * G::Rows];
    __shared__ uint32_t s_planes[Buffers][Waves * G::P * G::Slab];
    ProbeCounts cnt;
    constexpr uint32_t Units = Kind == kUnitCombo ? 48u : 6u;
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
    const bool axisOn = (cfgBits & kCfgAxis) != 0;
    if (!axisOn && (Kind == kUnitCombo ? (unit >> 3) != 0 : (unit >> 1) != 0))
        return; // only permutation code 0 (p0 == 0) is allowed
    if (!loadRows<L, 64 * Waves>(layouts, n, r, s_rows, 3))
        return;
    const int lane = (int)(threadIdx.x & 63u);
    const int wave = (int)(threadIdx.x >> 6);
    const int u = lane & (R - 1), w = lane >> L;
    const int c0 = ((int)chunk * Waves + wave) * G::P + w;
    const int planeOff = (wave * G::P + w) * G::Slab;
    uint32_t* plane = s_planes[0] + planeOff;
    uint32_t* line = plane + u; // the lane's pass-3 line: c1 = u, stride Pitch along c2
    uint32_t* outLine = s_planes[Buffers - 1] + planeOff + u;
    if constexpr (Kind == kUnitCombo) {
        const uint32_t pc = unit >> 3, d = unit & 7u;
        passes12<L>(pc, d, s_rows, c0, u, plane);
        waveSync();
        lastPassProbe<L>((d >> 2) & 1u, line, line, s_rows, layouts, table, logCap, bloom, logBloom, best, cfgBits, r, pc, d, c0, u, cnt);
    } else {
        const uint32_t a0 = unit >> 1, dx = unit & 1u;
#pragma unroll 1
        for (uint32_t sel = 0; sel < 2; ++sel) {
            const uint32_t pc = 2u * a0 + sel; // the permutation codes with p0 == a0
            if (!axisOn && pc != 0)
                continue;
#pragma unroll 1
            for (uint32_t dy = 0; dy < 2; ++dy) {
                if constexpr (Kind == kUnitShared) {
                    waveSync();
                    passes12<L>(pc, dx | (dy << 1), s_rows, c0, u, plane);
                    waveSync();
                }
#pragma unroll 1
                for (uint32_t dz = 0; dz < 2; ++dz) {
                    const uint32_t d = dx | (dy << 1) | (dz << 2);
                    if constexpr (Kind == kUnitFamily) {
                        waveSync();
                        passes12<L>(pc, d, s_rows, c0, u, plane);
                        waveSync();
                    }
                    lastPassProbe<L>(dz, line, outLine, s_rows, layouts, table, logCap, bloom, logBloom, best, cfgBits, r, pc, d, c0, u, cnt);
                }
            }
        }
    }
    flushCounts(cnt, counters, cfgBits);
}

// ---------------------------------------------------------------------------------------
// Rank-batched search: after parents [0, r1) have been searched, every child whose best key is
// below the smallest key a parent >= r1 can produce is final. Its table entries become
// tombstones and the Bloom filter is rebuilt from the live entries, so later batches neither
// probe nor load best[] for it. (The reference prunes its candidate ranges between search
// batches for the same reason, find_translation.cpp:887-915.)
// ---------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void kPrunePacked(uint64_t* __restrict__ table, uint64_t cap, const uint64_t* __restrict__ best, uint64_t minKey,
    uint32_t* __restrict__ bloom, uint32_t logBloom)
{
    for (uint64_t s = (uint64_t)blockIdx.x * 256 + threadIdx.x; s < cap; s += (uint64_t)gridDim.x * 256) {
        const uint64_t e = table[s];
        if (e == 0)
            continue;
        const uint32_t c = (uint32_t)e >> 3;
        if (c == kDeadRank)
            continue;
        if (best[c] < minKey) {
            table[s] = (e & 0xFFFFFFFF00000007ull) | ((uint64_t)kDeadRank << 3);
            continue;
        }
        const uint32_t h = (uint32_t)(e >> 32);
        atomicOr(&bloom[bloomWord(h, logBloom)], bloomBits(h));
    }
}

__global__ __launch_bounds__(256) void kPruneSplit(const uint64_t* __restrict__ tabKeys, uint32_t* __restrict__ tabVals, uint64_t cap, const uint64_t* __restrict__ best,
    uint64_t minKey, uint32_t* __restrict__ bloom, uint32_t logBloom)
{
    for (uint64_t s = (uint64_t)blockIdx.x * 256 + threadIdx.x; s < cap; s += (uint64_t)gridDim.x * 256) {
        const uint64_t k = tabKeys[s];
        if (k == 0)
            continue;
        const uint32_t e = tabVals[s];
        const uint32_t c = e >> 3;
        if (c == kDeadRank)
            continue;
        if (best[c] < minKey) {
            tabVals[s] = (kDeadRank << 3) | (e & 7u);
            continue;
        }
        const uint32_t f = bloomMix64(k);
        atomicOr(&bloom[bloomWord(f, logBloom)], bloomBits(f));
    }
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
    uint32_t logBloom, uint64_t* best, uint32_t cfgBits, unsigned long long* counters, hipStream_t stream)
{
    if constexpr (Waves >= 1 && Waves <= Geo<L>::MaxWaves && searchLds<L, Waves, Kind>() <= kLdsBytes) {
        const uint32_t perParent = (Kind == kUnitCombo ? 48u : 6u) * (uint32_t)(Geo<L>::MaxWaves / Waves);
        const dim3 grid(XcdGrouped ? xcdGridBlocks(parents, perParent) : parents * perParent);
        hipLaunchKernelGGL((kSearch<L, Waves, Kind, XcdGrouped>), grid, dim3(64 * Waves), 0, stream, layouts, n, table, logCap, bloom, logBloom, best, cfgBits,
            counters, r0, parents);
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
    uint32_t logBloom, uint64_t* best, uint32_t cfgBits, unsigned long long* counters, hipStream_t stream)
{
    constexpr int M = Geo<L>::MaxWaves;
    const uint32_t p = r1 - r0;
#define TDAG_SHAPE(W, K, X) launchSearchShape<L, W, K, X>(p, r0, n, layouts, table, logCap, bloom, logBloom, best, cfgBits, counters, stream)
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

MatchPlan planMatch(uint32_t level, uint64_t n, uint32_t variant, uint32_t firstBatch, uint32_t batchGrowth, uint32_t bloomBitsPerEntry)
{
    MatchPlan p {};
    p.level = level;
    p.firstBatch = firstBatch ? firstBatch : (uint32_t)std::max<uint64_t>(256, n / 64);
    p.batchGrowth = std::max<uint32_t>(2, batchGrowth ? batchGrowth : 2);
    p.variant = variantValid(level, variant) ? variant : defaultVariant(level);
    p.n = n;
    p.R = 1u << level;
    p.R3 = p.R * p.R * p.R;
    p.W = p.R3 >= 64 ? p.R3 / 64 : 1;
    const uint64_t entries = 8 * (n > 0 ? n : 1);
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
    // Filter words = entries * bitsPerEntry / 32 (rounded up to a power of two).
    p.bloomBitsPerEntry = bloomBitsPerEntry ? bloomBitsPerEntry : defaultFilterBits(level);
    p.logBloom = std::min<uint32_t>(22, std::max<uint32_t>(6, log2Ceil((entries * p.bloomBitsPerEntry + 31) / 32)));
    p.offBloom = take((1ull << p.logBloom) * 4);
    p.offParentTp = take(n * 8);
    p.offIsUnique = take(n * 4);
    p.offNewIdx = take(n * 4);
    p.offScanScratch = take(scanScratchBytes(n));
    p.offCounters = take(64);
    p.scratchBytes = off;
    return p;
}

hipError_t enqueueMatch(const MatchPlan& p, const uint64_t* d_masks, uint32_t cfgBits, void* d_scratch, hipStream_t stream,
    uint64_t* d_encoded, uint64_t* d_mapping, uint32_t* d_unique, hipEvent_t evSearch0, hipEvent_t evSearch1)
{
    char* base = (char*)d_scratch;
    uint32_t* keys = (uint32_t*)(base + p.offKeys);
    uint32_t* order = (uint32_t*)(base + p.offOrder);
    uint64_t* sortedMasks = (uint64_t*)(base + p.offSortedMasks);
    uint64_t* best = (uint64_t*)(base + p.offBest);
    uint64_t* tabKeys = (uint64_t*)(base + p.offTabKeys);
    uint32_t* bloom = (uint32_t*)(base + p.offBloom);
    uint32_t* tabVals = (uint32_t*)(base + p.offTabVals);
    uint64_t* parentTp = (uint64_t*)(base + p.offParentTp);
    uint32_t* isUnique = (uint32_t*)(base + p.offIsUnique);
    uint32_t* newIdx = (uint32_t*)(base + p.offNewIdx);
    unsigned long long* counters = (unsigned long long*)(base + p.offCounters);
    uint32_t* total = (uint32_t*)(counters + 1);
    const uint32_t n = (uint32_t)p.n;
    const uint32_t W = p.W;
    const uint64_t cap = 1ull << p.logCap;

    TDAG_TRY(fillU64((uint64_t*)counters, kNumCounters, 0ull, stream));
    if (n == 0) {
        if (evSearch0)
            TDAG_TRY(hipEventRecord(evSearch0, stream));
        if (evSearch1)
            TDAG_TRY(hipEventRecord(evSearch1, stream));
        return hipSuccess;
    }

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
        hipLaunchKernelGGL(kGatherInit, dim3(blocks), dim3(256), 0, stream, d_masks, order, keys, n, W, p.R3, p.level, sortedMasks, best);
        TDAG_TRY(hipGetLastError());
    }
    // 3. Child table and its Bloom filter (levels >= 3 first build the transposed bit layouts).
    TDAG_TRY(fillU64(tabKeys, cap, 0ull, stream));
    TDAG_TRY(fillU32(bloom, 1ull << p.logBloom, 0u, stream));
    switch (p.level) {
    case 2:
        hipLaunchKernelGGL(kBuildTable2, dim3((unsigned)ceilDiv((size_t)n * 8, 256)), dim3(256), 0, stream, sortedMasks, n, tabKeys, tabVals, p.logCap, bloom, p.logBloom);
        break;
    case 3:
        hipLaunchKernelGGL(kLayouts<3>, dim3(n), dim3(256), 0, stream, sortedMasks, n);
        hipLaunchKernelGGL(kCornerTable<3>, dim3(n), dim3(64), 0, stream, sortedMasks, n, tabKeys, p.logCap, bloom, p.logBloom);
        break;
    case 4:
        hipLaunchKernelGGL(kLayouts<4>, dim3(n), dim3(256), 0, stream, sortedMasks, n);
        hipLaunchKernelGGL(kCornerTable<4>, dim3(n), dim3(64), 0, stream, sortedMasks, n, tabKeys, p.logCap, bloom, p.logBloom);
        break;
    case 5:
        hipLaunchKernelGGL(kLayouts<5>, dim3(n), dim3(256), 0, stream, sortedMasks, n);
        hipLaunchKernelGGL(kCornerTable<5>, dim3(n), dim3(64), 0, stream, sortedMasks, n, tabKeys, p.logCap, bloom, p.logBloom);
        break;
    default:
        return hipErrorInvalidValue;
    }
    TDAG_TRY(hipGetLastError());
    // 4. Search, in rank-ordered batches of growing size with table pruning in between.
    if (evSearch0)
        TDAG_TRY(hipEventRecord(evSearch0, stream));
    {
        uint32_t r0 = 0;
        uint32_t batch = std::max<uint32_t>(1u, std::min<uint32_t>(n, p.firstBatch));
        while (r0 < n) {
            const uint32_t r1 = (uint32_t)std::min<uint64_t>(n, (uint64_t)r0 + batch);
            switch (p.level) {
            case 2:
                hipLaunchKernelGGL(kSearch2, dim3((unsigned)ceilDiv(r1 - r0, 4)), dim3(256), 0, stream, sortedMasks, r1, tabKeys, tabVals, p.logCap, bloom, p.logBloom, best,
                    cfgBits, r0, counters);
                TDAG_TRY(hipGetLastError());
                break;
            case 3:
                TDAG_TRY(launchSearch<3>(p.variant, n, r0, r1, sortedMasks, tabKeys, p.logCap, bloom, p.logBloom, best, cfgBits, counters, stream));
                break;
            case 4:
                TDAG_TRY(launchSearch<4>(p.variant, n, r0, r1, sortedMasks, tabKeys, p.logCap, bloom, p.logBloom, best, cfgBits, counters, stream));
                break;
            case 5:
                TDAG_TRY(launchSearch<5>(p.variant, n, r0, r1, sortedMasks, tabKeys, p.logCap, bloom, p.logBloom, best, cfgBits, counters, stream));
                break;
            }
            r0 = r1;
            if (r0 < n) {
                const uint64_t minKey = encodeKey(r0, 0, 0, 1 - (int)p.R, 1 - (int)p.R, 1 - (int)p.R, p.level);
                const unsigned pruneBlocks = (unsigned)std::min<uint64_t>(ceilDiv(cap, 256), 8192);
                TDAG_TRY(fillU32(bloom, 1ull << p.logBloom, 0u, stream));
                if (p.level == 2)
                    hipLaunchKernelGGL(kPruneSplit, dim3(pruneBlocks), dim3(256), 0, stream, tabKeys, tabVals, cap, best, minKey, bloom, p.logBloom);
                else
                    hipLaunchKernelGGL(kPrunePacked, dim3(pruneBlocks), dim3(256), 0, stream, tabKeys, cap, best, minKey, bloom, p.logBloom);
                TDAG_TRY(hipGetLastError());
                batch = (uint32_t)std::min<uint64_t>((uint64_t)batch * p.batchGrowth, n);
            }
        }
    }
    if (evSearch1)
        TDAG_TRY(hipEventRecord(evSearch1, stream));
    // 5. Post-processing.
    TDAG_TRY(fillU32(isUnique, n, 0u, stream));
    hipLaunchKernelGGL(kFinalParents, dim3((unsigned)ceilDiv(n, 256)), dim3(256), 0, stream, best, order, n, p.level, d_encoded, parentTp, isUnique, counters);
    TDAG_TRY(hipGetLastError());
    TDAG_TRY(exclusiveScanU32(isUnique, newIdx, n, total, base + p.offScanScratch, stream));
    hipLaunchKernelGGL(kFinalMapping, dim3((unsigned)ceilDiv(n, 256)), dim3(256), 0, stream, parentTp, isUnique, newIdx, n, d_mapping, d_unique);
    TDAG_TRY(hipGetLastError());
    return hipSuccess;
}

} // namespace tdag
