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
struct Geo {
    static constexpr int R = 1 << L;
    static constexpr int R2 = R * R;
    static constexpr int R3 = R2 * R;
    static constexpr int W = R3 / 64;
    static constexpr int Threads = R2;
    static constexpr int Slab = R2 + R; // padded z-slab pitch of a hash buffer (words)
    static constexpr int Buf = R * Slab; // words per hash buffer
};

// LDS index of voxel (x, y, z) in a hash buffer: x is XOR-swizzled by y and every z-slab is
// padded by R words, which keeps the 64 lanes of a wave on 64 distinct banks in all three pass
// orientations (lanes own lines indexed (t mod R, t div R)).
template <int L>
__device__ __forceinline__ uint32_t sidx(int x, int y, int z)
{
    constexpr int R = Geo<L>::R;
    return (uint32_t)((x ^ (y & (R - 1))) + R * y + Geo<L>::Slab * z);
}

template <int L>
__device__ __forceinline__ uint32_t lbit(const uint64_t* s_bits, int x, int y, int z)
{
    constexpr int R = Geo<L>::R;
    const uint32_t i = (uint32_t)(x + R * y + R * R * z);
    return (uint32_t)((s_bits[i >> 6] >> (i & 63u)) & 1ull);
}

// Three bit layouts of one item: layout a has axis a fastest, so a line along axis a is R
// consecutive bits starting at bit R * (u + R * w), where (u, w) are the line's coordinates
// along the other two axes in ascending axis order.
//   layout 0: x + R*y + R*R*z (the LargeSubGrid layout)
//   layout 1: y + R*x + R*R*z
//   layout 2: z + R*x + R*R*y
template <int L>
__device__ __forceinline__ uint32_t lineBits(const uint64_t* s_layouts, int A, int u, int w)
{
    constexpr int R = Geo<L>::R;
    constexpr int W = Geo<L>::W;
    const uint32_t start = (uint32_t)(R * (u + R * w));
    const uint64_t word = s_layouts[A * W + (start >> 6)];
    if constexpr (R == 32)
        return (uint32_t)(word >> (start & 63u));
    else
        return (uint32_t)(word >> (start & 63u)) & ((1u << R) - 1u);
}

// Exact value of the first pass: the trimmed bit sequence of a row segment, read from the scan
// start toward v, is the binary number formed by those bits, so it needs no hashing. Positive
// (scan from R-1 down to v): bits [v, R-1] with bit R-1 most significant, i.e. row >> v.
// Negative (scan from 0 up to v): bits [0, v] with bit 0 most significant. 0 iff empty.
template <int L, bool Pos>
__device__ __forceinline__ uint32_t rowPrefix(uint32_t row, int v)
{
    constexpr int R = Geo<L>::R;
    if constexpr (Pos)
        return row >> v;
    else
        return (__brev(row) >> (32 - R)) >> (R - 1 - v);
}

// Passes 1 and 2 of the nested hash fused into one sweep: every thread owns a line along A1 and
// reads, at each step, the bit row along A0 through that voxel from the transposed layouts.
// Axes and directions are compile time; the result (pass-2 state) goes to `out`.
template <int L, int A0, int A1, bool Pos0, bool Pos1>
__device__ __forceinline__ void pass12(const uint64_t* s_layouts, uint32_t* out)
{
    constexpr int R = Geo<L>::R;
    constexpr int A2 = 3 - A0 - A1;
    constexpr int B = A0 < A2 ? A0 : A2; // the thread's line coordinates, ascending axis order
    constexpr int C = A0 < A2 ? A2 : A0;
    constexpr int RB = A1 < A2 ? A1 : A2; // coordinates of a row along A0, ascending axis order
    constexpr int RC = A1 < A2 ? A2 : A1;
    const int t = (int)threadIdx.x;
    const int u = t & (R - 1), w = t >> L;
    int c[3];
    c[B] = u;
    c[C] = w;
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < R; ++k) {
        c[A1] = Pos1 ? (R - 1 - k) : k;
        const uint32_t row = lineBits<L>(s_layouts, A0, c[RB], c[RC]);
        s = hashStep(s, rowPrefix<L, Pos0>(row, c[A0]));
        out[sidx<L>(c[0], c[1], c[2])] = s;
    }
}

// Pass 3 along A2 (compile time) over the pass-2 states; stores finalHash(state) per voxel.
template <int L, int A, bool Pos>
__device__ __forceinline__ void pass3(const uint32_t* in, uint32_t* out)
{
    constexpr int R = Geo<L>::R;
    constexpr int B = (A == 0) ? 1 : 0;
    constexpr int C = (A == 2) ? 1 : 2;
    const int t = (int)threadIdx.x;
    const int u = t & (R - 1), w = t >> L;
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < R; ++k) {
        int c[3];
        c[A] = Pos ? (R - 1 - k) : k;
        c[B] = u;
        c[C] = w;
        const uint32_t i = sidx<L>(c[0], c[1], c[2]);
        s = hashStep(s, in[i]);
        out[i] = finalHash(s);
    }
}

template <int L, int A0, int A1>
__device__ __forceinline__ void pass12Dirs(uint32_t d, const uint64_t* s_layouts, uint32_t* out)
{
    switch (d & 3u) {
    case 0:
        pass12<L, A0, A1, false, false>(s_layouts, out);
        break;
    case 1:
        pass12<L, A0, A1, true, false>(s_layouts, out);
        break;
    case 2:
        pass12<L, A0, A1, false, true>(s_layouts, out);
        break;
    default:
        pass12<L, A0, A1, true, true>(s_layouts, out);
        break;
    }
}

// First two passes for permutation code pc (axes p0, p1) and direction bits d (bit 0: pass 1,
// bit 1: pass 2).
template <int L>
__device__ __forceinline__ void passes12(uint32_t pc, uint32_t d, const uint64_t* s_layouts, uint32_t* out)
{
    switch (pc) { // transform_dag.h:138-145
    case 0:
        pass12Dirs<L, 0, 1>(d, s_layouts, out);
        break;
    case 1:
        pass12Dirs<L, 0, 2>(d, s_layouts, out);
        break;
    case 2:
        pass12Dirs<L, 1, 0>(d, s_layouts, out);
        break;
    case 3:
        pass12Dirs<L, 1, 2>(d, s_layouts, out);
        break;
    case 4:
        pass12Dirs<L, 2, 0>(d, s_layouts, out);
        break;
    default:
        pass12Dirs<L, 2, 1>(d, s_layouts, out);
        break;
    }
}

template <int L>
__device__ __forceinline__ void passLast(int A, uint32_t pos, const uint32_t* in, uint32_t* out)
{
    switch (A * 2 + (int)(pos & 1u)) {
    case 0:
        pass3<L, 0, false>(in, out);
        break;
    case 1:
        pass3<L, 0, true>(in, out);
        break;
    case 2:
        pass3<L, 1, false>(in, out);
        break;
    case 3:
        pass3<L, 1, true>(in, out);
        break;
    case 4:
        pass3<L, 2, false>(in, out);
        break;
    default:
        pass3<L, 2, true>(in, out);
        break;
    }
}

// Full nested hash of every voxel for (pc, d) into h (final values).
template <int L>
__device__ __forceinline__ void hashGrid(uint32_t pc, uint32_t d, const uint64_t* s_layouts, uint32_t* h)
{
    passes12<L>(pc, d, s_layouts, h);
    __syncthreads();
    passLast<L>((int)permAxis(pc, 2), d >> 2, h, h);
    __syncthreads();
}

// Expected child word `wi` for the transform encoded in `key` applied to the parent bits.
template <int L>
__device__ __forceinline__ uint64_t transformedWord(const uint64_t* s_bits, uint64_t key, uint32_t wi)
{
    constexpr int R = Geo<L>::R;
    constexpr uint32_t b = L + 1;
    constexpr uint64_t m = (1ull << b) - 1ull;
    const int tx = (int)(key & m) - R;
    const int ty = (int)((key >> b) & m) - R;
    const int tz = (int)((key >> (2 * b)) & m) - R;
    const uint32_t sym = (uint32_t)((key >> (3 * b)) & 7u);
    const uint32_t pc = (uint32_t)((key >> (3 * b + 3)) & 7u);
    const int a0 = (int)permAxis(pc, 0), a1 = (int)permAxis(pc, 1), a2 = (int)permAxis(pc, 2);
    uint64_t out = 0;
    for (int j = 0; j < 64; ++j) {
        const uint32_t i = wi * 64u + (uint32_t)j;
        int q[3] = { (int)(i & (R - 1)) - tx, (int)((i >> L) & (R - 1)) - ty, (int)(i >> (2 * L)) - tz }; // out[o] = M[o - t]
        if (q[0] < 0 || q[1] < 0 || q[2] < 0 || q[0] >= R || q[1] >= R || q[2] >= R)
            continue;
        if (sym & 1u)
            q[0] = R - 1 - q[0];
        if (sym & 2u)
            q[1] = R - 1 - q[1];
        if (sym & 4u)
            q[2] = R - 1 - q[2];
        int v[3];
        v[a0] = q[0];
        v[a1] = q[1];
        v[a2] = q[2];
        out |= (uint64_t)lbit<L>(s_bits, v[0], v[1], v[2]) << j;
    }
    return out;
}

template <int L>
__device__ __forceinline__ bool loadLayouts(const uint64_t* __restrict__ layouts, uint32_t n, uint32_t r, uint64_t* s_layouts, uint32_t* s_any)
{
    constexpr int W = Geo<L>::W;
    if (threadIdx.x == 0)
        *s_any = 0;
    __syncthreads();
    for (int i = (int)threadIdx.x; i < 3 * W; i += Geo<L>::Threads) {
        const int lay = i / W, wi = i % W;
        const uint64_t v = layouts[((size_t)lay * n + r) * W + wi];
        s_layouts[i] = v;
        if (v)
            atomicOr(s_any, 1u);
    }
    __syncthreads();
    return *s_any != 0;
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

// One block per item; the hash of the whole grid seen from each of its 8 corners goes into the
// table and the Bloom filter (identity permutation, direction = corner, value at the corner).
template <int L>
__global__ __launch_bounds__(Geo<L>::Threads) void kCornerTable(const uint64_t* __restrict__ layouts, uint32_t n, uint64_t* __restrict__ table, uint32_t logCap,
    uint32_t* __restrict__ bloom, uint32_t logBloom)
{
    constexpr int R = Geo<L>::R;
    constexpr int W = Geo<L>::W;
    __shared__ uint64_t s_layouts[3 * W];
    __shared__ uint32_t s_h[Geo<L>::Buf];
    __shared__ uint32_t s_any;
    const uint32_t r = blockIdx.x;
    if (!loadLayouts<L>(layouts, n, r, s_layouts, &s_any))
        return;
    const uint32_t capMask = (1u << logCap) - 1u;
    for (uint32_t corner = 0; corner < 8; ++corner) {
        hashGrid<L>(0, corner, s_layouts, s_h);
        if (threadIdx.x == 0) {
            const int ax = (corner & 1u) ? 0 : R - 1;
            const int ay = (corner & 2u) ? 0 : R - 1;
            const int az = (corner & 4u) ? 0 : R - 1;
            const uint32_t h = s_h[sidx<L>(ax, ay, az)];
            if (h != 0) {
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
        __syncthreads();
    }
}

// Probes every voxel's final hash in s_h for combo (pc, d) and merges verified hits into best.
// Up to 8 Bloom-filter words per thread are in flight at a time; only hashes that pass the filter
// touch the table. Hits that beat best[c] are verified cooperatively by the wave that found them.
template <int L>
__device__ __forceinline__ void probeGrid(const uint32_t* s_h, const uint64_t* s_layouts, const uint64_t* __restrict__ sortedMasks,
    const uint64_t* __restrict__ table, uint32_t logCap, const uint32_t* __restrict__ bloom, uint32_t logBloom,
    uint64_t* __restrict__ best, uint32_t cfgBits, uint32_t r, uint32_t pc, uint32_t d, ProbeCounts& cnt)
{
    constexpr int R = Geo<L>::R;
    constexpr int W = Geo<L>::W;
    constexpr int T = Geo<L>::Threads;
    constexpr int PerThread = Geo<L>::R3 / T; // == R
    constexpr int Batch = PerThread < 8 ? PerThread : 8;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t capMask = (1u << logCap) - 1u;
    const int a0 = (int)permAxis(pc, 0), a1 = (int)permAxis(pc, 1), a2 = (int)permAxis(pc, 2);
    // Filter phase: bit j of `pending` = the thread's voxel in slab j is non-empty and passes
    // the Bloom filter. Loads are issued in batches for memory-level parallelism.
    uint32_t pendingZ = 0;
    for (int j0 = 0; j0 < PerThread; j0 += Batch) {
        uint32_t hv[Batch];
        uint32_t bw[Batch];
#pragma unroll
        for (int b = 0; b < Batch; ++b)
            hv[b] = s_h[(j0 + b) * Geo<L>::Slab + threadIdx.x]; // slab z = j0 + b, in-slab index = thread
#pragma unroll
        for (int b = 0; b < Batch; ++b)
            bw[b] = hv[b] ? bloom[bloomWord(hv[b], logBloom)] : 0u;
#pragma unroll
        for (int b = 0; b < Batch; ++b) {
            const uint32_t need = bloomBits(hv[b]);
            const bool pass = hv[b] != 0 && (bw[b] & need) == need;
            cnt.probes += hv[b] != 0;
            cnt.filterPass += pass;
            pendingZ |= (uint32_t)pass << (j0 + b);
        }
    }
    // Walk phase: every lane walks the table for its pending voxels one after another and
    // stops at each candidate that beats best[]; the wave then verifies the stopped lanes'
    // candidates together and resumes.
    const int y = (int)(threadIdx.x >> L);
    const int x = (int)(threadIdx.x & (R - 1)) ^ y; // undo the LDS swizzle
    bool active = false;
    uint32_t h = 0, slot = 0;
    int z = 0;
    while (true) {
        bool found = false;
        uint32_t hitC = 0;
        uint64_t hitKey = 0;
        while (true) {
            if (!active) {
                if (pendingZ == 0)
                    break;
                z = __ffs(pendingZ) - 1;
                pendingZ &= pendingZ - 1u;
                h = s_h[z * Geo<L>::Slab + threadIdx.x];
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
            // Voxel of this hash in the permuted frame.
            const int c3[3] = { x, y, z };
            const int voxel[3] = { c3[a0], c3[a1], c3[a2] };
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
            bool ok = true;
            for (uint32_t wi = lane; wi < (uint32_t)W; wi += 64u)
                ok = ok && (transformedWord<L>(s_layouts, bk, wi) == sortedMasks[(size_t)bc * W + wi]);
            const bool allOk = (__ballot(!ok) == 0ull);
            if ((int)lane == leader && allOk) {
                ++cnt.verified;
                atomicMin((unsigned long long*)&best[bc], (unsigned long long)bk);
            }
            stopped &= stopped - 1ull;
        }
    }
}

// Search kernel. One block per (parent rank, unit); R^2 threads own one line each per sweep.
//   Units == 48: a unit is one combo (perm, direction): two sweeps, one hash buffer.
//   Units == 6:  a unit is a family (first-pass axis a0 and direction dx) covering 8 combos.
//                Buffers = 1: the fused first/second sweep is redone per combo;
//                Buffers = 2: it is shared by the two directions of the third pass.
// The autotuner picks per level.
template <int L, int Units, int Buffers>
__global__ __launch_bounds__(Geo<L>::Threads) void kSearch(const uint64_t* __restrict__ layouts, uint32_t n, const uint64_t* __restrict__ table, uint32_t logCap,
    const uint32_t* __restrict__ bloom, uint32_t logBloom, uint64_t* __restrict__ best, uint32_t cfgBits, unsigned long long* __restrict__ counters,
    uint32_t rankBase)
{
    ProbeCounts cnt;
    constexpr int W = Geo<L>::W;
    __shared__ uint64_t s_layouts[3 * W];
    __shared__ uint32_t s_buf[Buffers][Geo<L>::Buf];
    __shared__ uint32_t s_any;
    const uint32_t r = rankBase + blockIdx.x;
    const bool axisOn = (cfgBits & kCfgAxis) != 0;
    if constexpr (Units == 48) {
        const uint32_t combo = blockIdx.y;
        const uint32_t pc = combo >> 3;
        const uint32_t d = combo & 7u;
        if (!axisOn && pc != 0)
            return;
        if (!loadLayouts<L>(layouts, n, r, s_layouts, &s_any))
            return;
        hashGrid<L>(pc, d, s_layouts, s_buf[0]);
        probeGrid<L>(s_buf[0], s_layouts, layouts, table, logCap, bloom, logBloom, best, cfgBits, r, pc, d, cnt);
    } else {
        const uint32_t family = blockIdx.y;
        const int a0 = (int)(family >> 1);
        const uint32_t dx = family & 1u;
        if (!axisOn && a0 != 0)
            return; // permutation code 0 is the only one left and has p0 == 0
        if (!loadLayouts<L>(layouts, n, r, s_layouts, &s_any))
            return;
        for (uint32_t sel = 0; sel < 2; ++sel) {
            const uint32_t pc = 2u * (uint32_t)a0 + sel; // permutation codes with p0 == a0
            if (!axisOn && pc != 0)
                continue;
            const int a2 = (int)permAxis(pc, 2);
            for (uint32_t dy = 0; dy < 2; ++dy) {
                if constexpr (Buffers == 2) {
                    passes12<L>(pc, dx | (dy << 1), s_layouts, s_buf[1]);
                    __syncthreads();
                }
                for (uint32_t dz = 0; dz < 2; ++dz) {
                    const uint32_t d = dx | (dy << 1) | (dz << 2);
                    if constexpr (Buffers == 1) {
                        hashGrid<L>(pc, d, s_layouts, s_buf[0]);
                    } else {
                        passLast<L>(a2, dz, s_buf[1], s_buf[0]);
                        __syncthreads();
                    }
                    probeGrid<L>(s_buf[0], s_layouts, layouts, table, logCap, bloom, logBloom, best, cfgBits, r, pc, d, cnt);
                    __syncthreads();
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

template <int L>
hipError_t launchSearch(uint32_t variant, uint32_t n, uint32_t r0, uint32_t r1, const uint64_t* layouts, const uint64_t* table, uint32_t logCap, const uint32_t* bloom,
    uint32_t logBloom, uint64_t* best, uint32_t cfgBits, unsigned long long* counters, hipStream_t stream)
{
    const dim3 block(Geo<L>::Threads);
    const uint32_t parents = r1 - r0;
    switch (variant) {
    case kVariantCombo:
        hipLaunchKernelGGL((kSearch<L, 48, 1>), dim3(parents, 48), block, 0, stream, layouts, n, table, logCap, bloom, logBloom, best, cfgBits, counters, r0);
        break;
    case kVariantFamily1:
        hipLaunchKernelGGL((kSearch<L, 6, 1>), dim3(parents, 6), block, 0, stream, layouts, n, table, logCap, bloom, logBloom, best, cfgBits, counters, r0);
        break;
    case kVariantFamily2:
        if constexpr (L < 5)
            hipLaunchKernelGGL((kSearch<L, 6, 2>), dim3(parents, 6), block, 0, stream, layouts, n, table, logCap, bloom, logBloom, best, cfgBits, counters, r0);
        else
            return hipErrorInvalidValue;
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
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
        return variant == kVariantCombo;
    if (level < 2 || level > 5 || variant >= kNumVariants)
        return false;
    return level < 5 || variant <= kVariantFamily1;
}

uint32_t defaultVariant(uint32_t level)
{
    switch (level) {
    case 5:
        return kVariantFamily1;
    default:
        return kVariantCombo;
    }
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
    p.bloomBitsPerEntry = bloomBitsPerEntry ? bloomBitsPerEntry : 16;
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
        hipLaunchKernelGGL(kCornerTable<3>, dim3(n), dim3(Geo<3>::Threads), 0, stream, sortedMasks, n, tabKeys, p.logCap, bloom, p.logBloom);
        break;
    case 4:
        hipLaunchKernelGGL(kLayouts<4>, dim3(n), dim3(256), 0, stream, sortedMasks, n);
        hipLaunchKernelGGL(kCornerTable<4>, dim3(n), dim3(Geo<4>::Threads), 0, stream, sortedMasks, n, tabKeys, p.logCap, bloom, p.logBloom);
        break;
    case 5:
        hipLaunchKernelGGL(kLayouts<5>, dim3(n), dim3(256), 0, stream, sortedMasks, n);
        hipLaunchKernelGGL(kCornerTable<5>, dim3(n), dim3(Geo<5>::Threads), 0, stream, sortedMasks, n, tabKeys, p.logCap, bloom, p.logBloom);
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
