// Device-side building blocks of the exhaustive matcher's search, shared by the per-level
// kernels and the batched persistent kernel: filter/table probing, the level-2 wave search, the
// level-3..5 plane-per-wave nested-hash search, bitwise verification and table pruning.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "common.hpp"
#include "matcher.hpp"

namespace tdag {
namespace dev {

// best[c] as a filter before verification: a plain load, which may hit in the XCD's L2 and
// may return an older (larger) value than another XCD's latest atomicMin. That only lets a
// candidate through to verification; the atomicMin after it decides. (An agent-scope atomic
// load would bypass the L2 on every candidate.)
__device__ __forceinline__ uint64_t bestHint(const uint64_t* p)
{
    return *p;
}

// threadIdx.x through an empty asm: to the compiler a new value at every call, so values derived
// from the lane are computed where they are used. In the persistent batched kernel they would
// otherwise be hoisted out of its tile loop and stay live in registers across every tile.
__device__ __forceinline__ uint32_t threadX()
{
    uint32_t t = threadIdx.x;
    asm volatile("" : "+v"(t));
    return t;
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


// Child rank stored in a table entry whose child can no longer improve (tombstone; the key stays
// so linear-probing clusters remain intact). Ranks are < 2^28, so this never names an item.
constexpr uint32_t kDeadRank = 0x1FFFFFFFu;

// Level-3..5 table entries pack (hash << 32) | (rank << 6) | (perm << 3) | corner: the child's
// grid hashed in the pass frame of axis-permutation code `perm` (pass a reads child axis
// sigma_a with perm = sigma^-1, so child axis b is read by pass permAxis(perm, b)) from `corner`
// (bit a: direction of pass a). Ranks are < 2^26; kPackedDead marks a tombstone.
constexpr uint32_t kPackedDead = 0x3FFFFFFu;
__device__ __forceinline__ uint64_t packEntry(uint32_t h, uint32_t rank, uint32_t perm, uint32_t corner)
{
    return ((uint64_t)h << 32) | (uint64_t)((rank << 6) | (perm << 3) | corner);
}

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

__device__ __forceinline__ uint32_t waveOr(uint32_t v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        v |= (uint32_t)__shfl_xor((int)v, off, 64);
    return v;
}

__device__ __forceinline__ void flushCounts(const ProbeCounts& c, unsigned long long* counters, uint32_t cfgBits)
{
    if (!(cfgBits & kCfgCount))
        return;
    const uint32_t v[6] = { waveSum(c.probes), waveSum(c.filterPass), waveSum(c.keyHits), waveSum(c.candidates), waveSum(c.verifications), waveSum(c.verified) };
    if ((threadX() & 63u) == 0)
        for (int i = 0; i < 6; ++i)
            if (v[i])
                atomicAdd(&counters[2 + i], (unsigned long long)v[i]);
}

// Level-2 search: lane = anchored voxel v' = (x, y, z) of the sub-box. For each permutation the
// direction keys are formed kDirGroup at a time and their filter loads issued together (a group
// of 2 keeps the kernel at 8 waves per SIMD).
constexpr uint32_t kDirGroup = 2;

// Orders the calling wave's LDS accesses without waiting for its global loads.
__device__ __forceinline__ void waveSyncLds()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
}

// Capacity of the level-2 work list of one wave (entries of 8 + 2 bytes in LDS).
constexpr uint32_t kWalkQueue2 = 128;

// Table slots a walk reads per round trip (consecutive, issued together): the probe run of one
// key is contiguous and often holds several entries.
constexpr int kWalkSlots = 4;
// The same at level 2, whose kernel runs 8 waves per SIMD within 64 VGPRs.
constexpr int kWalkSlots2 = 2;

// Table walks of `count` queued level-2 keys of parent r under permutation pc (its permuted mask
// Q): lane j walks entry base + j, so the walks of all directions run side by side. Entry meta:
// source lane (anchored voxel) | direction << 6.
__device__ __forceinline__ void walkQueue2(const uint64_t* qKey, const uint16_t* qMeta, uint32_t count, uint32_t r, uint32_t pc, uint64_t Q,
    const uint64_t* __restrict__ sortedMasks, const uint64_t* __restrict__ tabKeys, const uint32_t* __restrict__ tabVals, uint32_t logCap,
    uint64_t* __restrict__ best, uint32_t cfgBits, ProbeCounts& cnt)
{
    const uint32_t lane = threadX() & 63u;
    const uint32_t capMask = (1u << logCap) - 1u;
    for (uint32_t base = 0; base < count; base += 64) {
        if (base + lane >= count)
            continue;
        const uint64_t key = qKey[base + lane];
        const uint32_t meta = qMeta[base + lane];
        const uint32_t src = meta & 63u, d = meta >> 6;
        const int vx = (int)(src & 3u), vy = (int)((src >> 2) & 3u), vz = (int)(src >> 4);
        // kWalkSlots2 consecutive slots (keys and values) per round trip.
        uint32_t s = homeSlot64(key, logCap);
        bool walking = true;
        while (walking) {
            uint64_t kv[kWalkSlots2];
            uint32_t vv[kWalkSlots2];
#pragma unroll
            for (int j = 0; j < kWalkSlots2; ++j) {
                kv[j] = tabKeys[(s + (uint32_t)j) & capMask];
                vv[j] = tabVals[(s + (uint32_t)j) & capMask];
            }
#pragma unroll
            for (int j = 0; j < kWalkSlots2; ++j) {
                if (kv[j] == 0) {
                    walking = false;
                    break;
                }
                if (kv[j] != key)
                    continue;
                ++cnt.keyHits;
                const uint32_t c = vv[j] >> 3;
                const uint32_t kc = vv[j] & 7u;
                if (c <= r || c == kDeadRank)
                    continue;
                const uint32_t sym = d ^ kc;
                const int tx = (kc & 1u) ? -vx : vx;
                const int ty = (kc & 2u) ? -vy : vy;
                const int tz = (kc & 4u) ? -vz : vz;
                if (!configAllows(cfgBits, pc, sym, tx, ty, tz))
                    continue;
                ++cnt.candidates;
                const uint64_t k64 = encodeKey(r, pc, sym, tx, ty, tz, 2);
                if (k64 >= bestHint(&best[c]))
                    continue;
                ++cnt.verifications;
                // Bitwise verification of the derived transform.
                const uint64_t T = translate4(mirror4(Q, sym), tx, ty, tz);
                if (T == sortedMasks[c]) {
                    ++cnt.verified;
                    atomicMin((unsigned long long*)&best[c], (unsigned long long)k64);
                }
            }
            s = (s + (uint32_t)kWalkSlots2) & capMask;
        }
    }
}

// Level-2 search of parent r by one wave with a work list: for each permutation the 8 direction
// keys of every lane (anchored voxel) are filtered, the passes are appended to the wave's LDS
// queue (qKey/qMeta, kWalkQueue2 entries), and full queues, and the rest at the end of the
// permutation, are walked by walkQueue2 with one entry per lane. Compared with walking each
// direction's passes in place, the wave pays one chain of table round trips per 64 passes
// instead of one per direction. Same candidates, same verification, same minimum.
__device__ __forceinline__ void search2ParentQueued(uint32_t r, const uint64_t* __restrict__ sortedMasks, const uint64_t* __restrict__ tabKeys,
    const uint32_t* __restrict__ tabVals, uint32_t logCap, const uint32_t* __restrict__ bloom, uint32_t logBloom, uint64_t* __restrict__ best, uint32_t cfgBits,
    uint64_t* qKey, uint16_t* qMeta, ProbeCounts& cnt)
{
    const uint32_t lane = threadX() & 63u;
    const uint64_t P = sortedMasks[r];
    if (P == 0)
        return;
    const int vx = (int)(lane & 3u), vy = (int)((lane >> 2) & 3u), vz = (int)(lane >> 4);
    const uint64_t clip = rangeMaskX4(0, 4 - vx) & rangeMaskY4(0, 4 - vy) & rangeMaskZ4(0, 4 - vz);
    const uint32_t numPerms = (cfgBits & kCfgAxis) ? 6u : 1u;
    for (uint32_t pc = 0; pc < numPerms; ++pc) {
        // permute (find_translation.cpp:109-125): out(o) = P(v) with v[p[a]] = o[a].
        int v[3];
        v[permAxis(pc, 0)] = vx;
        v[permAxis(pc, 1)] = vy;
        v[permAxis(pc, 2)] = vz;
        const uint32_t src = (uint32_t)(v[0] + 4 * v[1] + 16 * v[2]);
        const uint64_t Q = (uint64_t)__ballot((P >> src) & 1ull);
        uint32_t count = 0;
#pragma unroll 1
        for (uint32_t d0 = 0; d0 < 8; d0 += kDirGroup) {
            uint64_t key[kDirGroup];
            uint32_t bw[kDirGroup];
#pragma unroll
            for (uint32_t j = 0; j < kDirGroup; ++j)
                key[j] = (mirror4(Q, (~(d0 + j)) & 7u) >> lane) & clip;
#pragma unroll
            for (uint32_t j = 0; j < kDirGroup; ++j)
                bw[j] = key[j] ? bloom[bloomWord(bloomMix64(key[j]), logBloom)] : 0u;
#pragma unroll
            for (uint32_t j = 0; j < kDirGroup; ++j) {
                const uint32_t need = bloomBits(bloomMix64(key[j]));
                cnt.probes += key[j] != 0;
                const bool pass = key[j] != 0 && (bw[j] & need) == need;
                const uint64_t m = __ballot(pass);
                if (m == 0ull)
                    continue;
                const uint32_t np = (uint32_t)__popcll(m);
                if (count + np > kWalkQueue2) {
                    waveSyncLds();
                    walkQueue2(qKey, qMeta, count, r, pc, Q, sortedMasks, tabKeys, tabVals, logCap, best, cfgBits, cnt);
                    waveSyncLds();
                    count = 0;
                }
                if (pass) {
                    ++cnt.filterPass;
                    const uint32_t pos = count + __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
                    qKey[pos] = key[j];
                    qMeta[pos] = (uint16_t)(lane | ((d0 + j) << 6));
                }
                count += np;
            }
        }
        if (count > 0) {
            waveSyncLds();
            walkQueue2(qKey, qMeta, count, r, pc, Q, sortedMasks, tabKeys, tabVals, logCap, best, cfgBits, cnt);
            waveSyncLds();
        }
    }
}

// ---------------------------------------------------------------------------------------
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
    for (int i = (int)threadX(); i < numLayouts * G::R2; i += Threads) {
        const int lay = i >> (2 * L), j = i & (G::R2 - 1);
        const Row v = g[((size_t)lay * n + r) * G::R2 + j];
        s_rows[lay * G::Rows + j + (j >> L)] = v;
        any |= (v != 0);
    }
    return __syncthreads_or(any) != 0;
}

// Occupancy of an item along the three axes: bit i of a[k] is set iff some voxel has coordinate
// i along axis k, i.e. the OR of the rows of layout k. Values are uniform (scalar registers).
struct Occupancy {
    uint32_t a[3];
    // Selects without indexing, so the values stay in registers.
    __device__ __forceinline__ uint32_t axis(uint32_t k) const { return k == 0u ? a[0] : (k == 1u ? a[1] : a[2]); }
};

// Occupancy contribution of row (u, w) of layout a0 (rows along a0; u and w index the lower and
// the higher of the other two axes).
__device__ __forceinline__ void occupancyAdd(uint32_t row, int j, int L, uint32_t& accA0, uint32_t& accLo, uint32_t& accHi)
{
    const uint32_t any = row != 0u ? 1u : 0u;
    accA0 |= row;
    accLo |= any << (j & ((1 << L) - 1));
    accHi |= any << (j >> L);
}

__device__ __forceinline__ Occupancy occupancyOf(uint32_t a0, uint32_t occA0, uint32_t occLo, uint32_t occHi)
{
    // The other two axes in ascending order: a0 = 0 -> (1, 2), 1 -> (0, 2), 2 -> (0, 1).
    Occupancy o;
    o.a[0] = a0 == 0u ? occA0 : occLo;
    o.a[1] = a0 == 1u ? occA0 : (a0 == 0u ? occLo : occHi);
    o.a[2] = a0 == 2u ? occA0 : occHi;
    return o;
}

// Loads the rows of item r's layout a0 (rows along axis a0, the only layout a search unit whose
// first pass runs along a0 reads) into s_rows and returns the item's occupancy along all three
// axes, derived from them. s_occ holds 3 words per wave of the block. Contains one block barrier.
template <int L, int Threads>
__device__ __forceinline__ bool loadRowsOcc(const uint64_t* __restrict__ layouts, uint32_t n, uint32_t r, uint32_t a0, typename Geo<L>::Row* s_rows,
    uint32_t* s_occ, Occupancy& occ)
{
    using G = Geo<L>;
    using Row = typename G::Row;
    constexpr int Waves = Threads / 64;
    const Row* g = reinterpret_cast<const Row*>(layouts) + ((size_t)a0 * n + r) * G::R2;
    uint32_t acc0 = 0, accLo = 0, accHi = 0;
    for (int j = (int)threadX(); j < G::R2; j += Threads) {
        const Row v = g[j];
        s_rows[j + (j >> L)] = v;
        occupancyAdd((uint32_t)v, j, L, acc0, accLo, accHi);
    }
    acc0 = waveOr(acc0);
    accLo = waveOr(accLo);
    accHi = waveOr(accHi);
    const int wave = (int)(threadX() >> 6);
    if ((threadX() & 63u) == 0) {
        s_occ[3 * wave + 0] = acc0;
        s_occ[3 * wave + 1] = accLo;
        s_occ[3 * wave + 2] = accHi;
    }
    __syncthreads();
    uint32_t o[3] = { 0, 0, 0 };
#pragma unroll
    for (int w = 0; w < Waves; ++w)
#pragma unroll
        for (int a = 0; a < 3; ++a)
            o[a] |= s_occ[3 * w + a];
    occ = occupancyOf(a0, __builtin_amdgcn_readfirstlane(o[0]), __builtin_amdgcn_readfirstlane(o[1]), __builtin_amdgcn_readfirstlane(o[2]));
    return o[0] != 0u;
}

// Orders LDS accesses of the lanes of one wave (LDS operations of a wave execute in order; this
// keeps the compiler from moving them across).
__device__ __forceinline__ void waveSync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Wave-level loadRowsOcc: the calling wave alone loads the rows of item r's layout a0 into its
// own s_rows and returns its occupancy. Returns whether the item is non-empty (uniform over the
// wave).
template <int L>
__device__ __forceinline__ bool loadRowsWave(const uint64_t* __restrict__ layouts, uint32_t n, uint32_t r, uint32_t a0, typename Geo<L>::Row* s_rows,
    Occupancy& occ)
{
    using G = Geo<L>;
    using Row = typename G::Row;
    const Row* g = reinterpret_cast<const Row*>(layouts) + ((size_t)a0 * n + r) * G::R2;
    uint32_t acc0 = 0, accLo = 0, accHi = 0;
    for (int j = (int)(threadX() & 63u); j < G::R2; j += 64) {
        const Row v = g[j];
        s_rows[j + (j >> L)] = v;
        occupancyAdd((uint32_t)v, j, L, acc0, accLo, accHi);
    }
    const uint32_t o0 = __builtin_amdgcn_readfirstlane(waveOr(acc0));
    occ = occupancyOf(a0, o0, __builtin_amdgcn_readfirstlane(waveOr(accLo)), __builtin_amdgcn_readfirstlane(waveOr(accHi)));
    waveSync();
    return o0 != 0u;
}

// Steps at the start of a scan along an axis with occupancy `occ` that read only empty rows: a
// positive scan runs from coordinate R-1 down, a negative one from 0 up. Their states stay 0.
template <int L>
__device__ __forceinline__ int emptyLead(uint32_t occ, uint32_t pos)
{
    if (occ == 0u)
        return Geo<L>::R;
    return pos ? __builtin_clz(occ) - (32 - Geo<L>::R) : __builtin_ctz(occ);
}

// Shape pre-filter of the Bloom lookups (levels 3..5). A verified match makes the child equal
// to the transformed content of the parent's anchored sub-box, so the sub-box content has the
// child's extents (the span of occupied coordinates along each axis) up to an axis permutation.
// The extents of every child in the table, inserted under all six axis orders, form a bitmap of
// R^2 words of R bits (the bit of spans e = extent - 1 at shapeWord and shapeBit); a sub-box whose extent triple is not in it
// cannot lead to a verified match, so its Bloom word is not loaded. The extents of a sub-box are
// derived along the three passes: along A2 from the pass-2 states read by pass 3, along A1 from
// per-slab masks of the non-empty row segments that pass 2 read, along A0 from the OR of the
// rows over the box's A1 range (a per-block table) masked to its A0 range.
template <int L>
constexpr int shapeWords()
{
    return 1 << (2 * L);
}
// LDS words of a block's copy of the bitmap: the R^2 words, then one word that stays 0 (an
// empty sub-box reads it, so it never passes).
template <int L>
constexpr int shapeLdsWords()
{
    return shapeWords<L>() + 1;
}

// Positions of a triple (e0, e1, e2) of spans (extent - 1, each < R) in the bitmap: bit 31 - e0
// of word R*R - 1 - (e1 + R*e2). The complemented positions let the search form them from
// leading plus trailing zero counts: for the occupancy x != 0 of an axis, 31 - span(x) =
// clz(x) + ctz(x), so no subtraction is needed.
__host__ __device__ __forceinline__ uint32_t shapeWord(uint32_t L, uint32_t e1, uint32_t e2)
{
    return (1u << (2 * L)) - 1u - (e1 + (e2 << L));
}
__host__ __device__ __forceinline__ uint32_t shapeBit(uint32_t e0)
{
    return 31u - e0;
}

// Extent - 1 of the set bits of x != 0 (bare v_ffbh / v_ffbl: for x == 0 the value is
// meaningless, and callers mask the key it enters).
__device__ __forceinline__ uint32_t spanOf(uint32_t x)
{
    return 31u - (uint32_t)__builtin_clz(x) - (uint32_t)__builtin_ctz(x);
}

// 31 - spanOf(x), the complemented span, for x != 0; 0xFFFFFFFE for x == 0 (v_ffbh_u32 and
// v_ffbl_b32 both return 0xFFFFFFFF for 0, a defined result here, unlike __builtin_clz(0)).
__device__ __forceinline__ uint32_t spanCompl(uint32_t x)
{
    uint32_t hi, lo;
    asm("v_ffbh_u32 %0, %1" : "=v"(hi) : "v"(x));
    asm("v_ffbl_b32 %0, %1" : "=v"(lo) : "v"(x));
    return hi + lo;
}

// LDS views of the shape pre-filter of one block (all null: the filter is off).
struct ShapeLds {
    const void* pre = nullptr; // Geo<L>::Row [2][R * Pitch]: OR over a1 in [0, c1] (index 0) / [c1, R) (index 1) of row (a1, a2), at c1 * Pitch + a2
    uint32_t* n1 = nullptr; // [Waves][P][R]: for plane w of a wave and line a2, bit k set iff pass-2 step k read a non-empty segment
    const uint32_t* bits = nullptr; // shapeLdsWords<L>() words
};

// LDS bytes of the pre tables (in the row type: an OR of rows fits it, so level 4 needs half and
// level 3 a quarter of 32-bit entries) and words of the n1 masks of `waves` waves.
template <int L>
constexpr int shapePreBytes()
{
    return 2 * Geo<L>::R * Geo<L>::Pitch * (int)sizeof(typename Geo<L>::Row);
}
template <int L>
constexpr int shapeN1Words(int waves)
{
    return waves * Geo<L>::P * Geo<L>::R;
}

// Fills the shape pre-filter's per-parent LDS: the extent bitmap of the level (copied from HBM)
// and the pre tables of the parent whose layout-0 rows (rows along x, (a1, a2) = (y, z)) are in
// s_rows. Threads tid = 0..threads-1 of the caller share the work; the caller synchronises
// before the tables are read.
template <int L>
__device__ __forceinline__ void buildShapeTables(const typename Geo<L>::Row* s_rows, const uint32_t* __restrict__ shapeBits, uint32_t* s_shape, typename Geo<L>::Row* s_pre,
    int tid, int threads)
{
    using G = Geo<L>;
    for (int i = tid; i < shapeWords<L>(); i += threads)
        s_shape[i] = shapeBits[i];
    if (tid == 0)
        s_shape[shapeWords<L>()] = 0u;
    for (int i = tid; i < 2 * G::R; i += threads) {
        const int dy = i / G::R, a2 = i % G::R;
        uint32_t acc = 0;
        for (int k = 0; k < G::R; ++k) {
            const int a1 = dy ? G::R - 1 - k : k;
            acc |= (uint32_t)s_rows[a1 + G::Pitch * a2];
            s_pre[dy * G::R * G::Pitch + a1 * G::Pitch + a2] = (typename G::Row)acc;
        }
    }
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

// 1 if x != 0, else 0, as one v_min_u32 (written as a compare, the compiler selects between
// constants, and the shifted ones of an unrolled sweep end up in registers of their own).
__device__ __forceinline__ uint32_t nonZeroBit(uint32_t x)
{
    uint32_t b;
    asm("v_min_u32 %0, 1, %1" : "=v"(b) : "v"(x));
    return b;
}

// Passes 1 and 2 of the nested hash on plane c0 (the coordinate along the first pass axis A0):
// the lane owns the line along A1 at c[A2] = u and reads, at each step, the row along A0
// through that voxel. The pass-2 state of voxel (c1, c2) goes to plane[c1 + Pitch*c2].
// The first `lead` steps (a wave-uniform count, see emptyLead) read only empty rows: chunks of
// R/4 steps inside them store the state 0 without hashing. All rows are read up front so the
// chunk branches do not wait on LDS.
template <int L, int A0, int A1, bool Pos0, bool Pos1, bool Shape = false>
__device__ __forceinline__ void pass12(const typename Geo<L>::Row* s_rows, int c0, int u, uint32_t* plane, int lead, uint32_t* n1Line = nullptr)
{
    using G = Geo<L>;
    constexpr int R = G::R;
    constexpr int C = R / 4;
    constexpr int A2 = 3 - A0 - A1;
    constexpr int Step = (A1 < A2) ? 1 : G::Pitch; // row (lo, hi) = (c[A1], c[A2]) or (c[A2], c[A1])
    const typename G::Row* rows = s_rows + ((A1 < A2) ? G::Pitch * u : u); // s_rows holds layout A0
    uint32_t* line = plane + G::Pitch * u;
    const int sh = Pos0 ? c0 : 31 - c0;
    uint32_t v[R];
#pragma unroll
    for (int k = 0; k < R; ++k)
        v[k] = (uint32_t)rows[Step * (Pos1 ? (R - 1 - k) : k)];
    uint32_t s = 0, nz = 0;
#pragma unroll
    for (int k0 = 0; k0 < R; k0 += C) {
        if (k0 + C <= lead) {
#pragma unroll
            for (int k = k0; k < k0 + C; ++k)
                line[Pos1 ? (R - 1 - k) : k] = 0u;
            continue;
        }
#pragma unroll
        for (int k = k0; k < k0 + C; ++k) {
            const uint32_t seg = rowPrefix<Pos0>(v[k], sh);
            s = hashStep(s, seg);
            line[Pos1 ? (R - 1 - k) : k] = s;
            if constexpr (Shape)
                nz |= nonZeroBit(seg) << k;
        }
    }
    if constexpr (Shape)
        *n1Line = nz;
}

template <int L, int A0, int A1, bool Shape = false>
__device__ __forceinline__ void pass12Dirs(uint32_t d, const typename Geo<L>::Row* s_rows, int c0, int u, uint32_t* plane, int lead, uint32_t* n1Line = nullptr)
{
    switch (d & 3u) {
    case 0:
        pass12<L, A0, A1, false, false, Shape>(s_rows, c0, u, plane, lead, n1Line);
        break;
    case 1:
        pass12<L, A0, A1, true, false, Shape>(s_rows, c0, u, plane, lead, n1Line);
        break;
    case 2:
        pass12<L, A0, A1, false, true, Shape>(s_rows, c0, u, plane, lead, n1Line);
        break;
    default:
        pass12<L, A0, A1, true, true, Shape>(s_rows, c0, u, plane, lead, n1Line);
        break;
    }
}

// Bits [0, k] set (k < 32).
__device__ __forceinline__ uint32_t lowMask(uint32_t k)
{
    return k >= 31u ? 0xFFFFFFFFu : (2u << k) - 1u;
}

// Passes 1 and 2 for permutation code pc (axes p0, p1) and direction bits d (bit 0: pass 1,
// bit 1: pass 2); `lead` as in pass12.
template <int L>
__device__ __forceinline__ void passes12(uint32_t pc, uint32_t d, const typename Geo<L>::Row* s_rows, int c0, int u, uint32_t* plane, int lead)
{
    switch (pc) { // transform_dag.h:138-145
    case 0:
        pass12Dirs<L, 0, 1>(d, s_rows, c0, u, plane, lead);
        break;
    case 1:
        pass12Dirs<L, 0, 2>(d, s_rows, c0, u, plane, lead);
        break;
    case 2:
        pass12Dirs<L, 1, 0>(d, s_rows, c0, u, plane, lead);
        break;
    case 3:
        pass12Dirs<L, 1, 2>(d, s_rows, c0, u, plane, lead);
        break;
    case 4:
        pass12Dirs<L, 2, 0>(d, s_rows, c0, u, plane, lead);
        break;
    default:
        pass12Dirs<L, 2, 1>(d, s_rows, c0, u, plane, lead);
        break;
    }
}

// Buffer resource over the Bloom filter's 2^logBloom words: a raw buffer load at an offset past
// the end returns 0 without touching memory.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t bloomRsrc(const uint32_t* bloom, uint32_t logBloom)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<uint32_t*>(bloom), (short)0, (int)(4u << logBloom), 0x00020000);
}

// Filter word of nested hash f (bloomWordNested), or 0 for f == 0: an empty region's hash is
// never in the table, and its lane then issues no memory request (the offset is out of range),
// which keeps the filter's cache-line traffic to the non-empty voxels.
__device__ __forceinline__ uint32_t bloomWordOrZero(__amdgpu_buffer_rsrc_t rs, uint32_t f, uint32_t shift)
{
    const uint32_t off = f != 0u ? (f >> shift) * 4u : 0xFFFFFFF0u;
    return __builtin_amdgcn_raw_buffer_load_b32(rs, off, 0, 0);
}

// Filter word of nested hash f where m (0 or all ones) says whether the lane loads it: the byte
// offset (f >> shift) * 4 with bit 31 set when m == 0, out of range (no memory request, 0
// returned). Offsets are below 2^24, so one bitwise select of three operands forms it: the
// offset bits from f >> (shift - 2), bit 31 and the two low bits from ~m.
__device__ __forceinline__ uint32_t bloomWordMasked(__amdgpu_buffer_rsrc_t rs, uint32_t f, uint32_t shift, uint32_t m)
{
    // v_bitop3_b32 with truth table 0xB1 over (a, b, c) = (f >> (shift - 2), m, mask): c ? a : ~b.
    const uint32_t off = __builtin_amdgcn_bitop3_b32(f >> (shift - 2u), m, 0x7FFFFFFCu, 0xB1);
    return __builtin_amdgcn_raw_buffer_load_b32(rs, off, 0, 0);
}

// acc | (x << k) kept as a chain of v_lshl_or_b32: the empty asm stops the compiler from
// regrouping an unrolled sweep's ORs into shifts merged three at a time (one instruction more per
// two bits).
__device__ __forceinline__ uint32_t lshlOr(uint32_t x, int k, uint32_t acc)
{
    uint32_t r = (x << k) | acc;
    asm volatile("" : "+v"(r));
    return r;
}

// Pass 3 along the lane's line (c1 = u) of its plane over the pass-2 states in[Pitch*c2], fused
// with the Bloom filter: stores the final hash of every voxel at out[Pitch*c2] (out may equal in)
// and returns the bit set of voxels c2 that are non-empty and pass the filter. Filter words are
// loaded without branches (empty voxels read 0 through an out-of-range buffer offset), one chunk
// of R/4 steps at a time, and tested while the next chunk hashes. Chunks inside the first `lead`
// steps (wave-uniform; their states are 0) are skipped: their voxels are empty, never pending,
// and out keeps its old words.
// Shape (see ShapeLds): preLine = the pre table row of the lane's c1 for the combo's A1
// direction, n1Plane = the n1 masks of the lane's plane, mask0 / mask1 = the box's range along A0
// (coordinate bits) and A1 (pass-2 step bits); sub-boxes whose extents are not in the bitmap
// load no filter word.
template <int L, bool Pos2, bool Shape = false>
__device__ __forceinline__ uint32_t pass3Filter(const uint32_t* in, uint32_t* out, const uint32_t* __restrict__ bloom, uint32_t logBloom, uint32_t cfgBits,
    int lead, ProbeCounts& cnt, const typename Geo<L>::Row* preLine = nullptr, const uint32_t* n1Plane = nullptr, const uint32_t* shapeBits = nullptr, uint32_t mask0 = 0u,
    uint32_t mask1 = 0u)
{
    constexpr int R = Geo<L>::R;
    constexpr int C = R / 4;
    const uint32_t shift = 32u - logBloom;
    const __amdgpu_buffer_rsrc_t rs = bloomRsrc(bloom, logBloom);
    uint32_t st[R];
#pragma unroll
    for (int k = 0; k < R; ++k)
        st[k] = in[Geo<L>::Pitch * (Pos2 ? (R - 1 - k) : k)];
    uint32_t s = 0, pending = 0;
    // Shape state of the box so far: its occupancy along A0 and along A1 (pass-2 steps), both
    // masked to the box's range; 4R times its first non-empty slab along A2; and 4 times the
    // bitmap word index of the box without the A1 term, set at every non-empty slab (the span
    // along A2 ends there): word = clz(occ1) + ctz(occ1) + R*first2 + R*R - 32 - R*last2 (see
    // shapeWord). Before the first non-empty slab occ1 == 0 (its spanCompl is 0xFFFFFFFE), and
    // word4 = 8 + 4R^2 makes the byte offset 4R^2: an empty box reads the bitmap's zero word and
    // never passes. A slab is non-empty exactly when some row segment of it in the box's A1
    // range is, so occ1 != 0 from then on and the offset stays below 4R^2. A box that passes is
    // thus non-empty (f != 0), and the filter load takes the pass bit alone (bloomWordMasked).
    uint32_t occ0 = 0, occ1 = 0, first4 = 4u * R * R, word4 = 8u + 4u * R * R;
    uint32_t pf[C], pbw[C]; // the previous chunk's hashes and filter words (0: nothing passes)
#pragma unroll
    for (int k = 0; k < C; ++k)
        pf[k] = pbw[k] = 0u;
#pragma unroll
    for (int k0 = 0; k0 < R; k0 += C) {
        if (k0 + C <= lead)
            continue;
        uint32_t f[C], bw[C];
#pragma unroll
        for (int k = 0; k < C; ++k) {
            s = hashStep(s, st[k0 + k]);
            f[k] = finalHash(s);
            out[Geo<L>::Pitch * (Pos2 ? (R - 1 - k0 - k) : k0 + k)] = f[k];
            if constexpr (Shape) {
                const int kk = k0 + k;
                const int a2 = Pos2 ? (R - 1 - kk) : kk;
                occ0 |= (uint32_t)preLine[a2] & mask0;
                occ1 |= n1Plane[a2] & mask1;
                // Selects between registers only (a select between two constants would keep one
                // of them in a register of its own for every step of the unrolled sweep).
                const bool slab = st[kk] != 0u;
                first4 = slab ? min(first4, 4u * (uint32_t)(R * kk)) : first4;
                word4 = slab ? first4 + 4u * (uint32_t)(R * R - 32 - R * kk) : word4;
                // Bit 31 - e0 of the box's word, sign-extended to 0 or all ones (v_bfe_i32).
                const uint32_t off = (spanCompl(occ1) << 2) + word4;
                const uint32_t wordBits = *reinterpret_cast<const uint32_t*>(reinterpret_cast<const char*>(shapeBits) + off);
                const uint32_t pass = (uint32_t)__builtin_amdgcn_sbfe((int)wordBits, spanCompl(occ0), 1u);
                bw[k] = bloomWordMasked(rs, f[k], shift, pass);
            } else {
                bw[k] = bloomWordOrZero(rs, f[k], shift);
            }
        }
        if (k0 > 0) {
#pragma unroll
            for (int k = 0; k < C; ++k)
                pending = lshlOr(bloomHit(pbw[k], pf[k]), Pos2 ? (R - 1 - (k0 - C) - k) : (k0 - C) + k, pending);
        }
#pragma unroll
        for (int k = 0; k < C; ++k) {
            pf[k] = f[k];
            pbw[k] = bw[k];
        }
    }
#pragma unroll
    for (int k = 0; k < C; ++k)
        pending = lshlOr(bloomHit(pbw[k], pf[k]), Pos2 ? (C - 1 - k) : (R - C) + k, pending);
    if (cfgBits & kCfgCount) {
        uint32_t acc = 0, nonEmpty = 0;
#pragma unroll
        for (int k = 0; k < R; ++k) {
            acc |= st[k];
            nonEmpty += acc != 0u;
        }
        cnt.probes += nonEmpty;
        cnt.filterPass += (uint32_t)__builtin_popcount(pending);
    }
    return pending;
}

// Expected child row (yo, zo) (bits along x) for the transform encoded in `key` applied to the
// parent rows: out[o] = M[o - t] with M = mirror(permute(P)) (find_translation.cpp:66-84 for the
// translation). The source of an output row is one line of the parent along axis p0, read from
// the parent's layout p0 (`prow`, rows (lo, hi) at lo + R*hi).
template <int L>
__device__ __forceinline__ uint32_t transformedRow(const typename Geo<L>::Row* __restrict__ prow, int tx, int ty, int tz, uint32_t sym, uint32_t pc, int yo,
    int zo)
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
    uint32_t line = (uint32_t)prow[lo + R * hi];
    if (sym & 1u)
        line = __brev(line) >> (32 - R);
    const uint32_t rowMask = R == 32 ? 0xFFFFFFFFu : ((1u << R) - 1u);
    return (tx >= 0 ? (line << tx) : (line >> -tx)) & rowMask;
}

// Wave-cooperative bitwise check child bc == T(parent r) for the transform in key bk; the n
// items' three layouts follow each other in `layouts` (layout 0 = the rank-ordered masks).
template <int L>
__device__ __forceinline__ bool verifyWave(const uint64_t* __restrict__ layouts, uint32_t n, uint32_t r, uint32_t bc, uint64_t bk, uint32_t lane)
{
    using G = Geo<L>;
    using Row = typename G::Row;
    constexpr int R = G::R;
    constexpr uint32_t b = L + 1;
    constexpr uint64_t m = (1ull << b) - 1ull;
    const int tx = (int)(bk & m) - R;
    const int ty = (int)((bk >> b) & m) - R;
    const int tz = (int)((bk >> (2 * b)) & m) - R;
    const uint32_t sym = (uint32_t)((bk >> (3 * b)) & 7u);
    const uint32_t pc = (uint32_t)((bk >> (3 * b + 3)) & 7u);
    const Row* child = reinterpret_cast<const Row*>(layouts) + (size_t)bc * G::R2;
    const Row* prow = reinterpret_cast<const Row*>(layouts) + ((size_t)permAxis(pc, 0) * n + r) * G::R2;
    bool ok = true;
    for (int j = (int)lane; j < G::R2; j += 64)
        ok = ok && (transformedRow<L>(prow, tx, ty, tz, sym, pc, j & (R - 1), j >> L) == (uint32_t)child[j]);
    return __ballot(!ok) == 0ull;
}

// A wave's list of filter passes in LDS, kept across the (permutation, direction) combos of a
// search unit so the table is walked once per full list instead of once per combo. Entry i:
// meta[i] = lane | c2 << 6 | d << 11 | pc << 14 for voxel (c0Base + lane / R, lane % R, c2) of
// combo (pc, d), and hash[i] = its final nested hash (the planes are reused by the next combo).
// `count` is uniform over the wave.
constexpr uint32_t kProbeQueue = 64;
struct ProbeList {
    uint32_t* meta;
    uint32_t* hash;
    uint32_t count;
};

// Walks the child table for the listed voxels and merges verified hits into best. Lane j takes
// entry base + j, so the walks of a round run side by side whatever combo or line the voxels came
// from. Lanes stop at each candidate that beats best[c]; the wave then verifies the stopped
// lanes' candidates together and resumes.
template <int L>
__device__ __forceinline__ void probeQueue(const ProbeList& list, int c0Base, const uint64_t* __restrict__ layouts, uint32_t n,
    const uint64_t* __restrict__ table, uint32_t logCap, uint64_t* __restrict__ best, uint32_t cfgBits, uint32_t r, ProbeCounts& cnt)
{
    using G = Geo<L>;
    constexpr int R = G::R;
    const uint32_t lane = threadX() & 63u;
    const uint32_t count = list.count;
    const uint32_t capMask = (1u << logCap) - 1u;
    for (uint32_t base = 0; base < count; base += 64) {
        bool active = base + lane < count;
        uint32_t h = 0, slot = 0, pc = 0, d = 0;
        int c0 = 0, u = 0, c2 = 0;
        if (active) {
            const uint32_t e = list.meta[base + lane];
            const int src = (int)(e & 63u), w = src >> L;
            c2 = (int)((e >> 6) & 31u);
            d = (e >> 11) & 7u;
            pc = (e >> 14) & 7u;
            u = src & (R - 1);
            c0 = c0Base + w;
            h = list.hash[base + lane];
            slot = homeSlot32(h, logCap);
        }
        while (true) {
            bool found = false;
            uint32_t hitC = 0;
            uint64_t hitKey = 0;
            while (active) {
                // kWalkSlots consecutive slots per round trip: the probe run of one hash is
                // contiguous, and at level 3 it often holds several entries.
                uint64_t ev[kWalkSlots];
#pragma unroll
                for (int k = 0; k < kWalkSlots; ++k)
                    ev[k] = table[(slot + (uint32_t)k) & capMask];
                // Levels 4 and 5: best[] of every key hit among them, the loads issued together (a
                // hit past the end of the run loads a value that is never used). At level 3 the
                // registers would cost a wave per SIMD.
                uint64_t bv[kWalkSlots] = {};
                if constexpr (L >= 4) {
#pragma unroll
                    for (int k = 0; k < kWalkSlots; ++k) {
                        const uint32_t ck = (uint32_t)ev[k] >> 6;
                        bv[k] = ((uint32_t)(ev[k] >> 32) == h && ck > r && ck < n) ? bestHint(&best[ck]) : 0ull;
                    }
                }
#pragma unroll
                for (int k = 0; k < kWalkSlots; ++k) {
                    const uint64_t e = ev[k];
                    if (e == 0) {
                        active = false;
                        break;
                    }
                    slot = (slot + 1) & capMask;
                    if ((uint32_t)(e >> 32) != h)
                        continue;
                    ++cnt.keyHits;
                    const uint32_t cc = (uint32_t)e >> 6;
                    const uint32_t q = ((uint32_t)e >> 3) & 7u, kc = (uint32_t)e & 7u;
                    if (cc <= r || cc == kPackedDead)
                        continue;
                    // Per pass index a: mirror bit and translation of the anchored match
                    // (find_translation.cpp:835-837); the voxel's coordinates along the three pass
                    // axes are (c0, u, c2).
                    const int voxel[3] = { c0, u, c2 };
                    const uint32_t sa = d ^ kc;
                    int ta[3];
#pragma unroll
                    for (int a = 0; a < 3; ++a) {
                        const int anchor = ((d >> a) & 1u) ? 0 : R - 1;
                        ta[a] = ((sa >> a) & 1u) ? voxel[a] - anchor : anchor - voxel[a];
                    }
                    // Pass a reads parent axis permAxis(pc, a) and child axis sigma_a; child axis b
                    // is read by pass a = permAxis(q, b). Output axis b of the transform thus takes
                    // parent axis permAxis(pc, a), mirror bit sa_a and translation ta_a.
                    uint32_t pb[3], sym = 0;
                    int tt[3];
#pragma unroll
                    for (int bo = 0; bo < 3; ++bo) {
                        const uint32_t a = permAxis(q, bo);
                        pb[bo] = permAxis(pc, a);
                        sym |= ((sa >> a) & 1u) << bo;
                        tt[bo] = a == 0 ? ta[0] : (a == 1 ? ta[1] : ta[2]);
                    }
                    const uint32_t tc = permCode(pb[0], pb[1], pb[2]);
                    if (!configAllows(cfgBits, tc, sym, tt[0], tt[1], tt[2]))
                        continue;
                    ++cnt.candidates;
                    const uint64_t k64 = encodeKey(r, tc, sym, tt[0], tt[1], tt[2], L);
                    uint64_t bestNow;
                    if constexpr (L >= 4)
                        bestNow = bv[k];
                    else
                        bestNow = bestHint(&best[cc]);
                    if (k64 >= bestNow)
                        continue;
                    ++cnt.verifications;
                    found = true;
                    hitC = cc;
                    hitKey = k64;
                    break;
                }
                if (found)
                    break;
            }
            uint64_t stopped = __ballot(found);
            if (stopped == 0)
                break; // no lane found a candidate, so every walk has ended
            while (stopped) {
                const int leader = __ffsll((unsigned long long)stopped) - 1;
                const uint32_t bc = (uint32_t)__shfl((int)hitC, leader, 64);
                const uint64_t bk = (uint64_t)__shfl((long long)hitKey, leader, 64);
                const bool allOk = verifyWave<L>(layouts, n, r, bc, bk, lane);
                if ((int)lane == leader && allOk) {
                    ++cnt.verified;
                    atomicMin((unsigned long long*)&best[bc], (unsigned long long)bk);
                }
                stopped &= stopped - 1ull;
            }
        }
    }
}

// Walks and empties the list (no-op when empty).
template <int L>
__device__ __forceinline__ void flushProbes(ProbeList& list, int c0Base, const uint64_t* __restrict__ layouts, uint32_t n, const uint64_t* __restrict__ table,
    uint32_t logCap, uint64_t* __restrict__ best, uint32_t cfgBits, uint32_t r, ProbeCounts& cnt)
{
    if (list.count == 0)
        return;
    waveSync();
    probeQueue<L>(list, c0Base, layouts, n, table, logCap, best, cfgBits, r, cnt);
    waveSync();
    list.count = 0;
}

// Appends the lanes' filter passes of combo (pc, d) (bit c2 of `pending` = voxel (c0, u, c2);
// final hashes in the wave's planes fWave, plane w at fWave + w*Slab, voxel (u, c2) at
// u + Pitch*c2) to the list, one voxel per lane and step, walking the list whenever the next step
// might not fit.
template <int L>
__device__ __forceinline__ void probePending(uint32_t pending, ProbeList& list, const uint32_t* fWave, int c0Base, const uint64_t* __restrict__ layouts,
    uint32_t n, const uint64_t* __restrict__ table, uint32_t logCap, uint64_t* __restrict__ best, uint32_t cfgBits, uint32_t r, uint32_t pc, uint32_t d,
    ProbeCounts& cnt)
{
    using G = Geo<L>;
    const uint32_t lane = threadX() & 63u;
    const uint32_t u = lane & (G::R - 1), w = lane >> L;
    for (;;) {
        const uint64_t has = __ballot(pending != 0u);
        const uint32_t pushed = (uint32_t)__popcll(has);
        if (pushed == 0)
            break;
        if (list.count + pushed > kProbeQueue)
            flushProbes<L>(list, c0Base, layouts, n, table, logCap, best, cfgBits, r, cnt);
        if (pending != 0u) {
            const uint32_t pos = list.count + __builtin_amdgcn_mbcnt_hi((uint32_t)(has >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)has, 0u));
            const uint32_t c2 = (uint32_t)__ffs(pending) - 1u;
            pending &= pending - 1u;
            list.meta[pos] = lane | (c2 << 6) | (d << 11) | (pc << 14);
            list.hash[pos] = fWave[w * G::Slab + u + G::Pitch * c2];
        }
        list.count += pushed;
    }
}

// Pass 3 in direction dz with the filter (pass-2 states read from `in`, final hashes written to
// `out`, which is the lane's line of the wave's planes outWave), then the wave's filter passes go
// to its probe list.
template <int L, bool Shape = false>
__device__ __forceinline__ void lastPassProbe(uint32_t dz, int lead, const uint32_t* in, uint32_t* out, const uint32_t* outWave, ProbeList& list, int c0Base,
    const uint64_t* __restrict__ layouts, uint32_t n, const uint64_t* __restrict__ table, uint32_t logCap, const uint32_t* __restrict__ bloom,
    uint32_t logBloom, uint64_t* __restrict__ best, uint32_t cfgBits, uint32_t r, uint32_t pc, uint32_t d, ProbeCounts& cnt, const typename Geo<L>::Row* preLine = nullptr,
    const uint32_t* n1Plane = nullptr, const uint32_t* shapeBits = nullptr, uint32_t mask0 = 0u, uint32_t mask1 = 0u)
{
    const uint32_t pending = dz ? pass3Filter<L, true, Shape>(in, out, bloom, logBloom, cfgBits, lead, cnt, preLine, n1Plane, shapeBits, mask0, mask1)
                                : pass3Filter<L, false, Shape>(in, out, bloom, logBloom, cfgBits, lead, cnt, preLine, n1Plane, shapeBits, mask0, mask1);
    waveSync();
    probePending<L>(pending, list, outWave, c0Base, layouts, n, table, logCap, best, cfgBits, r, pc, d, cnt);
}

// Search-unit kinds of the level-3..5 kernel.
constexpr int kUnitCombo = 0; // one (permutation, direction) combo per block
constexpr int kUnitFamily = 1; // a family of 8 combos per block, each run from scratch
constexpr int kUnitShared = 2; // a family of 8 combos; passes 1 and 2 shared by both pass-3 directions

// Search units per parent: combos (pc, d), or families (p0, dx) of 8 combos; with
// kCfgChildFrames the parent runs permutation code 0 only (combos 0..7, families 0 and 1).
__host__ __device__ __forceinline__ uint32_t searchUnits(int kind, uint32_t cfgBits)
{
    const bool childFrames = (cfgBits & kCfgChildFrames) != 0;
    return kind == kUnitCombo ? (childFrames ? 8u : 48u) : (childFrames ? 2u : 6u);
}

// LDS bytes of one kSearch block.
// LDS bytes of one kSearch block: rows, planes, occupancy and probe lists, plus the shape
// pre-filter's tables where they still fit (shapeFits; the largest blocks run without it).
template <int L, int Waves, int Kind>
constexpr size_t searchLdsBase()
{
    return Geo<L>::Rows * sizeof(typename Geo<L>::Row) + (Kind == kUnitShared ? 2 : 1) * (size_t)Waves * Geo<L>::P * Geo<L>::Slab * 4 + 3 * Waves * 4 +
        (size_t)Waves * kProbeQueue * 8;
}
template <int L, int Waves>
constexpr size_t shapeLds()
{
    return (size_t)shapePreBytes<L>() + (size_t)Waves * Geo<L>::P * Geo<L>::R * 4 + (size_t)shapeLdsWords<L>() * 4;
}
template <int L, int Waves, int Kind>
constexpr bool shapeFits()
{
    return searchLdsBase<L, Waves, Kind>() + shapeLds<L, Waves>() <= kLdsBytes;
}
template <int L, int Waves, int Kind>
constexpr size_t searchLds()
{
    return searchLdsBase<L, Waves, Kind>() + (shapeFits<L, Waves, Kind>() ? shapeLds<L, Waves>() : 0);
}

// Search of one unit of parent rank r (rows already in s_rows, occupancy occ) by `Waves` waves:
// `wave` is the caller's wave index within them and covers planes ((chunk * Waves + wave) * P + w)
// for w < P. planes0 (and planes1 for kUnitShared) hold Waves * P * Slab words each; qMeta and
// qHash are the wave's probe list (kProbeQueue entries each), walked when full and once at the
// end of the unit. Scan steps over empty rows are skipped (emptyLead), and so is the whole unit
// for a wave whose planes hold no voxel. With Shape (only permutation code 0; `shape` filled by
// the block), pass 3 loads no filter word for sub-boxes whose extents no child has.
template <int L, int Waves, int Kind, bool Shape = false>
__device__ __forceinline__ void searchUnit(uint32_t r, uint32_t unit, uint32_t chunk, int wave, const typename Geo<L>::Row* s_rows, const Occupancy& occ,
    uint32_t* planes0, uint32_t* planes1, uint32_t* qMeta, uint32_t* qHash, const uint64_t* __restrict__ layouts, uint32_t n, const uint64_t* __restrict__ table,
    uint32_t logCap, const uint32_t* __restrict__ bloom, uint32_t logBloom, uint64_t* __restrict__ best, uint32_t cfgBits, ProbeCounts& cnt,
    const ShapeLds& shape = ShapeLds {})
{
    using G = Geo<L>;
    constexpr int R = G::R;
    // Parent-side permutations: all six, unless the table holds the children under all six
    // (kCfgChildFrames) or permutations are off. The shape pre-filter runs only without them
    // (permutation code 0: A0, A1, A2 = x, y, z).
    const bool parentPerms = (cfgBits & kCfgAxis) != 0 && (cfgBits & kCfgChildFrames) == 0;
    const int lane = (int)(threadX() & 63u);
    const int u = lane & (R - 1), w = lane >> L;
    const int c0Base = ((int)chunk * Waves + wave) * G::P;
    const int c0 = c0Base + w;
    const int planeOff = (wave * G::P + w) * G::Slab;
    uint32_t* plane = planes0 + planeOff;
    uint32_t* line = plane + u; // the lane's pass-3 line: c1 = u, stride Pitch along c2
    uint32_t* outLine = planes1 + planeOff + u;
    const uint32_t* outWave = planes1 + wave * G::P * G::Slab;
    // Pass 1 along p0 in direction bit 0 of d reads the rows' segments toward c0; plane c0 is
    // empty iff that segment of the occupancy along p0 is. p0 and d's bit 0 are fixed per unit.
    const uint32_t p0 = Kind == kUnitCombo ? (unit >> 3) >> 1 : unit >> 1;
    const uint32_t dx = unit & 1u;
    const uint32_t segment = dx ? rowPrefix<true>(occ.axis(p0), c0) : rowPrefix<false>(occ.axis(p0), 31 - c0);
    if (__ballot(segment != 0u) == 0ull)
        return;
    ProbeList list { qMeta, qHash, 0u };
    // Shape state of the lane: its n1 slot (written by pass 2) and its plane's n1 masks (read by
    // pass 3), and the box's A0 range (fixed per unit: dx and c0).
    uint32_t* n1Line = nullptr;
    const uint32_t* n1Plane = nullptr;
    uint32_t mask0 = 0u;
    if constexpr (Shape) {
        n1Plane = shape.n1 + (wave * G::P + w) * R;
        n1Line = const_cast<uint32_t*>(n1Plane) + u;
        mask0 = dx ? (0xFFFFFFFFu << c0) : lowMask((uint32_t)c0);
    }
    // The pre table row of the lane's c1 and its A1 range in pass-2 steps, for A1 direction dy.
    const auto* pre = static_cast<const typename G::Row*>(shape.pre);
    auto preLineOf = [&](uint32_t dy) { return Shape ? pre + dy * (uint32_t)(R * G::Pitch) + (uint32_t)(u * G::Pitch) : nullptr; };
    auto mask1Of = [&](uint32_t dy) { return Shape ? lowMask(dy ? (uint32_t)(R - 1 - u) : (uint32_t)u) : 0u; };
    if constexpr (Kind == kUnitCombo) {
        const uint32_t pc = unit >> 3, d = unit & 7u;
        const int lead2 = emptyLead<L>(occ.axis(permAxis(pc, 1)), (d >> 1) & 1u);
        if constexpr (Shape)
            pass12Dirs<L, 0, 1, true>(d, s_rows, c0, u, plane, lead2, n1Line);
        else
            passes12<L>(pc, d, s_rows, c0, u, plane, lead2);
        waveSync();
        lastPassProbe<L, Shape>((d >> 2) & 1u, emptyLead<L>(occ.axis(permAxis(pc, 2)), (d >> 2) & 1u), line, line, planes0 + wave * G::P * G::Slab, list, c0Base,
            layouts, n, table, logCap, bloom, logBloom, best, cfgBits, r, pc, d, cnt, preLineOf((d >> 1) & 1u), n1Plane, shape.bits, mask0, mask1Of((d >> 1) & 1u));
    } else {
#pragma unroll 1
        for (uint32_t sel = 0; sel < 2; ++sel) {
            const uint32_t pc = 2u * p0 + sel; // the permutation codes with p0
            if (!parentPerms && pc != 0)
                continue;
            const uint32_t occ1 = occ.axis(permAxis(pc, 1)), occ2 = occ.axis(permAxis(pc, 2));
#pragma unroll 1
            for (uint32_t dy = 0; dy < 2; ++dy) {
                const int lead2 = emptyLead<L>(occ1, dy);
                const typename G::Row* preLine = preLineOf(dy);
                const uint32_t mask1 = mask1Of(dy);
                if constexpr (Kind == kUnitShared) {
                    waveSync();
                    if constexpr (Shape)
                        pass12Dirs<L, 0, 1, true>(dx | (dy << 1), s_rows, c0, u, plane, lead2, n1Line);
                    else
                        passes12<L>(pc, dx | (dy << 1), s_rows, c0, u, plane, lead2);
                    waveSync();
                }
#pragma unroll 1
                for (uint32_t dz = 0; dz < 2; ++dz) {
                    const uint32_t d = dx | (dy << 1) | (dz << 2);
                    if constexpr (Kind == kUnitFamily) {
                        waveSync();
                        if constexpr (Shape)
                            pass12Dirs<L, 0, 1, true>(d, s_rows, c0, u, plane, lead2, n1Line);
                        else
                            passes12<L>(pc, d, s_rows, c0, u, plane, lead2);
                        waveSync();
                    }
                    lastPassProbe<L, Shape>(dz, emptyLead<L>(occ2, dz), line, outLine, outWave, list, c0Base, layouts, n, table, logCap, bloom, logBloom, best,
                        cfgBits, r, pc, d, cnt, preLine, n1Plane, shape.bits, mask0, mask1);
                }
            }
        }
    }
    flushProbes<L>(list, c0Base, layouts, n, table, logCap, best, cfgBits, r, cnt);
}

// Rank-batched search: after parents [0, r1) have been searched, every child whose best key is
// below the smallest key a parent >= r1 can produce is final. Its table entries become
// tombstones and the Bloom filter is rebuilt from the live entries, so later batches neither
// probe nor load best[] for it. (The reference prunes its candidate ranges between search
// batches for the same reason, find_translation.cpp:887-915.) Slots [start, cap) step `stride`;
// start and stride are multiples of 64 per wave (lane l takes slot base + l), so a wave's lanes
// hold 64 consecutive slots. Home slots and filter words are both top bits of the hash, so those
// slots hold few distinct filter words: the wave ORs the bits of equal words together (a scan
// over the lanes whose combining steps join equal words only) and the last lane of each run of
// equal words sets them with one atomic.
// One sweep of the packed table: entries whose child settled below minKey become tombstones (the
// hash stays, so walks still step over them), and the Bloom filter (cleared by the caller) gets
// the bits of every live entry. Each lane takes 4 consecutive slots (two 16-byte loads), so a
// wave covers 256 slots per step with four times the bytes in flight of one slot per lane. The
// slots run in filter-word order: a lane ORs its slots' bits per run of equal words, runs that
// end inside the lane go out at once, and the lane's last run joins a segmented OR scan across
// the wave, one atomic per run. A run split over two lanes costs one extra atomic, no lost bits:
// the scan only ever adds bits of the same word. minKey 0 settles nothing, so that sweep reads
// no best[]. start and stride count 4-slot groups; stride is a multiple of 64 and start lies in
// lane (start & 63) of its wave, so whole waves enter each step and the shuffles see every lane.
__device__ __forceinline__ void prunePacked(uint64_t* __restrict__ table, uint64_t cap, const uint64_t* __restrict__ best, uint64_t minKey, uint32_t* __restrict__ bloom,
    uint32_t logBloom, uint64_t start, uint64_t stride)
{
    constexpr uint32_t kNone = 0xFFFFFFFFu; // no word: filter words stay below 2^22
    const int lane = (int)(threadX() & 63u);
    const uint64_t groups = cap >> 2; // cap is a power of two >= 64
    for (uint64_t g = start; (g & ~63ull) < groups; g += stride) {
        uint64_t e[4] = {0ull, 0ull, 0ull, 0ull};
        if (g < groups) {
            const ulonglong2* quad = reinterpret_cast<const ulonglong2*>(table + 4 * g);
            const ulonglong2 lo = quad[0];
            const ulonglong2 hi = quad[1];
            e[0] = lo.x;
            e[1] = lo.y;
            e[2] = hi.x;
            e[3] = hi.y;
        }
        bool live[4];
        uint64_t bk[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const uint32_t c = (uint32_t)e[j] >> 6;
            live[j] = e[j] != 0ull && c != kPackedDead;
            bk[j] = (live[j] && minKey != 0ull) ? best[c] : ~0ull;
        }
        uint32_t word = kNone;
        uint32_t bits = 0u;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (live[j] && bk[j] < minKey) {
                table[4 * g + j] = (e[j] & 0xFFFFFFFF0000003Full) | ((uint64_t)kPackedDead << 6);
                live[j] = false;
            }
            if (live[j]) {
                const uint32_t h = (uint32_t)(e[j] >> 32);
                const uint32_t w = bloomWordNested(h, logBloom);
                if (w != word) {
                    if (word != kNone)
                        atomicOr(&bloom[word], bits);
                    word = w;
                    bits = 0u;
                }
                bits |= bloomBits(h);
            }
        }
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t w = (uint32_t)__shfl_up((int)word, off, 64);
            const uint32_t b = (uint32_t)__shfl_up((int)bits, off, 64);
            if (lane >= off && w == word)
                bits |= b;
        }
        const uint32_t next = (uint32_t)__shfl_down((int)word, 1, 64);
        if (word != kNone && (lane == 63 || next != word))
            atomicOr(&bloom[word], bits);
    }
}

__device__ __forceinline__ void pruneSplit(const uint64_t* __restrict__ tabKeys, uint32_t* __restrict__ tabVals, uint64_t cap, const uint64_t* __restrict__ best,
    uint64_t minKey, uint32_t* __restrict__ bloom, uint32_t logBloom, uint64_t start, uint64_t stride)
{
    for (uint64_t s = start; s < cap; s += stride) {
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

} // namespace dev
} // namespace tdag
