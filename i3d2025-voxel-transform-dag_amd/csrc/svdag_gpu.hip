// SVDAG producer on the device: EditStructure<void, uint32_t>(VoxelGrid) followed by toDAG()
// (structure.h:423-429, 658-701, 703-763) for a dense grid already resident in HBM.
//
// The reference's post-order DFS visits children in Morton order and drops empty subtrees, so the
// non-empty nodes of every level appear in ascending Morton code of their position, and toDAG
// keeps the first occurrence of every distinct node. The device build computes the same indices:
//   1. one thread per 4^3 block (block = Morton code) gathers its 64-bit leaf mask (Morton bits);
//      a scan compacts the non-empty blocks, which keeps them in Morton order;
//   2. per level, first-occurrence dedup: an open-addressing table keyed by the node's content
//      keeps, per distinct key, the minimum item index (CAS into an empty slot, then atomicMin);
//      a scan over "item is its key's minimum" numbers the uniques in first-occurrence order;
//   3. parents are runs of equal code >> 3; a scan numbers them and each child writes its
//      deduplicated id into slot code & 7 of its parent.
// The per-level unique lists are copied back into the host SVDAG (a few MB at 1024^3).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "pipeline.hpp"
#include "primitives.hpp"

namespace tdag {

namespace {

constexpr uint32_t kThreads = 256;
constexpr uint32_t kNone = 0xFFFFFFFFu;

__device__ inline uint32_t compact3(uint32_t x)
{
    x &= 0x09249249u;
    x = (x ^ (x >> 2)) & 0x030C30C3u;
    x = (x ^ (x >> 4)) & 0x0300F00Fu;
    x = (x ^ (x >> 8)) & 0x030000FFu;
    x = (x ^ (x >> 16)) & 0x000003FFu;
    return x;
}

// Morton position of a 2-bit coordinate: bit 0 stays, bit 1 moves to bit 3.
__device__ inline uint32_t spread2(uint32_t v) { return (v & 1u) | ((v & 2u) << 2); }

// A row of 4 voxels (x = 0..3 as nibble bits) -> its Morton bits {0, 1, 8, 9}.
__device__ inline uint64_t nibbleToMorton(uint32_t v)
{
    return (uint64_t)(v & 1u) | (uint64_t)((v >> 1) & 1u) << 1 | (uint64_t)((v >> 2) & 1u) << 8 | (uint64_t)((v >> 3) & 1u) << 9;
}

__global__ __launch_bounds__(kThreads) void kLeafMasks(const uint8_t* __restrict__ bits, uint32_t R, uint64_t numBlocks, uint64_t* __restrict__ masks,
    uint32_t* __restrict__ flags)
{
    const uint64_t code = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
    if (code >= numBlocks)
        return;
    const uint32_t c = (uint32_t)code;
    const uint64_t x0 = 4ull * compact3(c), y0 = 4ull * compact3(c >> 1), z0 = 4ull * compact3(c >> 2);
    uint64_t mask = 0;
#pragma unroll
    for (uint32_t z = 0; z < 4; ++z)
#pragma unroll
        for (uint32_t y = 0; y < 4; ++y) {
            const uint64_t i = x0 + (uint64_t)R * (y0 + y) + (uint64_t)R * R * (z0 + z); // x0 is a multiple of 4: one nibble
            const uint32_t nib = (bits[i >> 3] >> (i & 7)) & 0xFu;
            mask |= nibbleToMorton(nib) << (spread2(y) << 1 | spread2(z) << 2);
        }
    masks[code] = mask;
    flags[code] = mask != 0 ? 1u : 0u;
}

// One workgroup per 32^3-voxel cube = 8^3 blocks = the aligned Morton range [cube * 512, +512):
// the cube's 32 x 32 rows of 32 voxels are 1024 aligned 32-bit words (bit k = voxel x0 + k,
// little-endian bytes), staged in LDS, then each thread builds two blocks' masks from nibbles.
// Needs R >= 32 and a 4-byte aligned grid.
__global__ __launch_bounds__(kThreads) void kLeafMasksCube(const uint32_t* __restrict__ words, uint32_t R, uint64_t* __restrict__ masks,
    uint32_t* __restrict__ flags)
{
    __shared__ uint32_t rows[32][33]; // [z][y], padded against bank conflicts
    const uint32_t cube = blockIdx.x;
    const uint32_t cx = compact3(cube), cy = compact3(cube >> 1), cz = compact3(cube >> 2);
    const uint64_t wordsPerRow = R / 32;
    for (uint32_t k = threadIdx.x; k < 1024; k += kThreads) {
        const uint32_t y = k & 31u, z = k >> 5;
        rows[z][y] = words[((uint64_t)(32 * cz + z) * R + (32 * cy + y)) * wordsPerRow + cx];
    }
    __syncthreads();
    for (uint32_t j = threadIdx.x; j < 512; j += kThreads) {
        const uint32_t bx = compact3(j), by = compact3(j >> 1), bz = compact3(j >> 2);
        uint64_t mask = 0;
#pragma unroll
        for (uint32_t z = 0; z < 4; ++z)
#pragma unroll
            for (uint32_t y = 0; y < 4; ++y) {
                const uint32_t nib = (rows[4 * bz + z][4 * by + y] >> (4 * bx)) & 0xFu;
                mask |= nibbleToMorton(nib) << (spread2(y) << 1 | spread2(z) << 2);
            }
        const uint64_t code = (uint64_t)cube << 9 | j;
        masks[code] = mask;
        flags[code] = mask != 0 ? 1u : 0u;
    }
}

__global__ __launch_bounds__(kThreads) void kCompactLeaves(const uint64_t* __restrict__ masks, const uint32_t* __restrict__ flags,
    const uint32_t* __restrict__ pos, uint64_t numBlocks, uint32_t* __restrict__ codes, uint32_t* __restrict__ keys)
{
    const uint64_t code = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
    if (code >= numBlocks || !flags[code])
        return;
    const uint32_t p = pos[code];
    codes[p] = (uint32_t)code;
    keys[2 * (size_t)p] = (uint32_t)masks[code];
    keys[2 * (size_t)p + 1] = (uint32_t)(masks[code] >> 32);
}

template <int K>
__device__ inline uint32_t keyHash(const uint32_t* k)
{
    uint64_t h = 0x9E3779B97F4A7C15ull;
#pragma unroll
    for (int j = 0; j < K; ++j) {
        h ^= k[j];
        h *= 0xBF58476D1CE4E5B9ull;
        h ^= h >> 29;
    }
    return (uint32_t)(h ^ (h >> 32));
}

template <int K>
__device__ inline bool keyEq(const uint32_t* a, const uint32_t* b)
{
#pragma unroll
    for (int j = 0; j < K; ++j)
        if (a[j] != b[j])
            return false;
    return true;
}

// Slot of each distinct key ends up holding the minimum index of the items with that key. Slots
// never empty again and only ever move to a smaller index with the same key, so comparing
// against the key of whatever index a slot holds is always meaningful. At most cap probes.
template <int K>
__global__ __launch_bounds__(kThreads) void kInsertFirst(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ d_n, uint32_t* table, uint32_t capMask)
{
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= *d_n)
        return;
    const uint32_t* k = keys + (size_t)i * K;
    uint32_t h = keyHash<K>(k) & capMask;
    for (uint32_t probe = 0; probe <= capMask; ++probe) {
        uint32_t cur = __hip_atomic_load(&table[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (cur == kNone) {
            cur = atomicCAS(&table[h], kNone, i);
            if (cur == kNone)
                return;
        }
        if (keyEq<K>(keys + (size_t)cur * K, k)) {
            if (i < cur)
                atomicMin(&table[h], i);
            return;
        }
        h = (h + 1) & capMask;
    }
}

// Over the whole bound: items past the count get first = 0, so the scan of `first` numbers the
// uniques only.
template <int K>
__global__ __launch_bounds__(kThreads) void kRepresentative(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ d_n, uint32_t bound,
    const uint32_t* __restrict__ table, uint32_t capMask, uint32_t* __restrict__ rep, uint32_t* __restrict__ first)
{
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= bound)
        return;
    if (i >= *d_n) {
        first[i] = 0;
        return;
    }
    const uint32_t* k = keys + (size_t)i * K;
    uint32_t h = keyHash<K>(k) & capMask;
    uint32_t r = i; // every key was inserted, so the loop finds it; i is only a guard value
    for (uint32_t probe = 0; probe <= capMask; ++probe) {
        const uint32_t cur = table[h];
        if (cur == kNone)
            break;
        if (keyEq<K>(keys + (size_t)cur * K, k)) {
            r = cur;
            break;
        }
        h = (h + 1) & capMask;
    }
    rep[i] = r;
    first[i] = r == i ? 1u : 0u;
}

template <int K>
__global__ __launch_bounds__(kThreads) void kAssignIds(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ d_n, const uint32_t* __restrict__ rep,
    const uint32_t* __restrict__ uid, uint32_t* __restrict__ id, uint32_t* __restrict__ uniqueKeys)
{
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= *d_n)
        return;
    const uint32_t r = rep[i];
    id[i] = uid[r];
    if (r == i)
        for (int j = 0; j < K; ++j)
            uniqueKeys[(size_t)uid[i] * K + j] = keys[(size_t)i * K + j];
}

// Over the whole bound (items past the count start nothing).
__global__ __launch_bounds__(kThreads) void kParentStarts(const uint32_t* __restrict__ codes, const uint32_t* __restrict__ d_n, uint32_t bound,
    uint32_t* __restrict__ start)
{
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= bound)
        return;
    start[i] = (i < *d_n && (i == 0 || (codes[i] >> 3) != (codes[i - 1] >> 3))) ? 1u : 0u;
}

// pstart = exclusive scan of start: the parent of item i is pstart[i] + start[i] - 1.
__global__ __launch_bounds__(kThreads) void kParents(const uint32_t* __restrict__ codes, const uint32_t* __restrict__ id, const uint32_t* __restrict__ start,
    const uint32_t* __restrict__ pstart, const uint32_t* __restrict__ d_n, uint32_t* __restrict__ pkeys, uint32_t* __restrict__ pcodes)
{
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= *d_n)
        return;
    const uint32_t p = pstart[i] + start[i] - 1;
    pkeys[(size_t)p * 8 + (codes[i] & 7u)] = id[i];
    if (start[i])
        pcodes[p] = codes[i] >> 3;
}

// The upper levels, at most kTail nodes each, run in one workgroup: the dedup table lives in LDS,
// block-wide scans number the uniques and the parents, and level after level follows without a
// launch. The per-level counts land in the same places as for the multi-kernel levels.
constexpr uint32_t kTail = 4096;
constexpr uint32_t kTailThreads = 1024;
constexpr uint32_t kTailCap = 8192; // >= 2 * kTail
constexpr uint32_t kMaxLevels = 13;

struct TailPlan {
    uint32_t first, last; // levels first..last
    const uint32_t* keys0; // keys and codes of level `first`
    const uint32_t* codes0;
    uint32_t* uniq[kMaxLevels]; // unique keys of each level
    uint32_t* pkeys[kMaxLevels]; // keys / codes of level l + 1, written at level l
    uint32_t* pcodes[kMaxLevels];
    uint32_t* counts; // [2l] items, [2l + 1] uniques
    uint32_t* rep; // kTail words each
    uint32_t* uid;
    uint32_t* pstart;
};

// Exclusive scan of one value per thread over the 1024-thread block; *total gets the sum.
__device__ inline uint32_t blockExclusiveScan(uint32_t v, uint32_t* part, uint32_t* total)
{
    const uint32_t t = threadIdx.x;
    part[t] = v;
    __syncthreads();
    for (uint32_t off = 1; off < kTailThreads; off <<= 1) {
        const uint32_t add = t >= off ? part[t - off] : 0u;
        __syncthreads();
        part[t] += add;
        __syncthreads();
    }
    *total = part[kTailThreads - 1];
    const uint32_t excl = part[t] - v;
    __syncthreads();
    return excl;
}

template <int K>
__device__ inline void tailInsert(const uint32_t* keys, uint32_t i, uint32_t* table)
{
    const uint32_t* k = keys + (size_t)i * K;
    uint32_t h = keyHash<K>(k) & (kTailCap - 1);
    for (uint32_t probe = 0; probe < kTailCap; ++probe) {
        uint32_t cur = __hip_atomic_load(&table[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (cur == kNone) {
            cur = atomicCAS(&table[h], kNone, i);
            if (cur == kNone)
                return;
        }
        if (keyEq<K>(keys + (size_t)cur * K, k)) {
            if (i < cur)
                atomicMin(&table[h], i);
            return;
        }
        h = (h + 1) & (kTailCap - 1);
    }
}

template <int K>
__device__ inline uint32_t tailLookup(const uint32_t* keys, uint32_t i, const uint32_t* table)
{
    const uint32_t* k = keys + (size_t)i * K;
    uint32_t h = keyHash<K>(k) & (kTailCap - 1);
    for (uint32_t probe = 0; probe < kTailCap; ++probe) {
        const uint32_t cur = table[h];
        if (cur == kNone)
            break;
        if (keyEq<K>(keys + (size_t)cur * K, k))
            return cur;
        h = (h + 1) & (kTailCap - 1);
    }
    return i; // not reached: every key was inserted
}

__global__ __launch_bounds__(kTailThreads) void kTailLevels(TailPlan P)
{
    __shared__ uint32_t table[kTailCap];
    __shared__ uint32_t part[kTailThreads];
    __shared__ uint32_t s_n;
    const uint32_t t = threadIdx.x;
    const uint32_t* keys = P.keys0;
    const uint32_t* codes = P.codes0;
    for (uint32_t level = P.first; level <= P.last; ++level) {
        const bool leafLevel = level == 2;
        const uint32_t K = leafLevel ? 2u : 8u;
        if (t == 0)
            s_n = P.counts[2 * level];
        for (uint32_t i = t; i < kTailCap; i += kTailThreads)
            table[i] = kNone;
        __syncthreads();
        const uint32_t n = s_n; // <= kTail: at most one node per position of this level
        for (uint32_t i = t; i < n; i += kTailThreads)
            leafLevel ? tailInsert<2>(keys, i, table) : tailInsert<8>(keys, i, table);
        __syncthreads();
        // Items 4t .. 4t + 3 belong to thread t, so the scans number them in index order.
        uint32_t firstBits = 0, startBits = 0;
        for (uint32_t j = 0; j < 4; ++j) {
            const uint32_t i = 4 * t + j;
            if (i >= n)
                break;
            const uint32_t r = leafLevel ? tailLookup<2>(keys, i, table) : tailLookup<8>(keys, i, table);
            P.rep[i] = r;
            firstBits |= (r == i ? 1u : 0u) << j;
            startBits |= ((i == 0 || (codes[i] >> 3) != (codes[i - 1] >> 3)) ? 1u : 0u) << j;
        }
        uint32_t nu = 0, np = 0;
        uint32_t ubase = blockExclusiveScan((uint32_t)__builtin_popcount(firstBits), part, &nu);
        uint32_t pbase = blockExclusiveScan((uint32_t)__builtin_popcount(startBits), part, &np);
        for (uint32_t j = 0; j < 4; ++j) {
            const uint32_t i = 4 * t + j;
            if (i >= n)
                break;
            P.uid[i] = ubase; // exclusive count of uniques before item i
            P.pstart[i] = pbase; // exclusive count of parent starts before item i
            ubase += (firstBits >> j) & 1u;
            pbase += (startBits >> j) & 1u;
        }
        if (level < P.last)
            for (uint32_t i = t; i < np * 8; i += kTailThreads)
                P.pkeys[level][i] = kNone;
        __syncthreads();
        for (uint32_t i = t; i < n; i += kTailThreads) {
            const uint32_t r = P.rep[i];
            const uint32_t id = P.uid[r];
            if (r == i)
                for (uint32_t w = 0; w < K; ++w)
                    P.uniq[level][(size_t)id * K + w] = keys[(size_t)i * K + w];
            if (level < P.last) {
                const uint32_t st = (i == 0 || (codes[i] >> 3) != (codes[i - 1] >> 3)) ? 1u : 0u;
                const uint32_t p = P.pstart[i] + st - 1;
                P.pkeys[level][(size_t)p * 8 + (codes[i] & 7u)] = id;
                if (st)
                    P.pcodes[level][p] = codes[i] >> 3;
            }
        }
        if (t == 0) {
            P.counts[2 * level + 1] = nu;
            if (level < P.last)
                P.counts[2 * (level + 1)] = np;
        }
        __syncthreads();
        if (level < P.last) {
            keys = P.pkeys[level];
            codes = P.pcodes[level];
        }
    }
}

// Carves the build's buffers out of one workspace, 256-byte aligned. With a null base it only
// adds up the size, so the same code path first sizes the workspace, then runs.
struct Carver {
    char* base = nullptr;
    size_t off = 0;
    template <typename T>
    T* get(size_t count)
    {
        off = (off + 255) & ~size_t(255);
        T* p = base ? (T*)(base + off) : nullptr;
        off += std::max<size_t>(count * sizeof(T), 4);
        return p;
    }
};

inline uint32_t gridFor(uint64_t n) { return (uint32_t)((n + kThreads - 1) / kThreads); }

// First-occurrence dedup of the d_n[0] (<= bound) keys of K words: id[i] is the unique number of
// item i, uniqueKeys receives the uniques in first-occurrence order and d_nu their count. No host
// synchronisation: every launch covers the bound and reads the count on the device.
inline uint32_t tableCap(uint32_t items)
{
    uint32_t cap = 64;
    while (cap < 2 * (uint64_t)items)
        cap <<= 1;
    return cap;
}

// Buffers are carved for `bound` items; the launches, the scan and the table cover `work` items
// (work <= bound: the level's count is at most work, which the host learned after the leaves).
template <int K>
hipError_t dedupLevel(Carver& A, const uint32_t* keys, const uint32_t* d_n, uint32_t bound, uint32_t work, uint32_t* id, uint32_t* uniqueKeys, uint32_t* d_nu,
    hipStream_t st)
{
    uint32_t* table = A.get<uint32_t>(tableCap(bound));
    uint32_t* rep = A.get<uint32_t>(bound);
    uint32_t* first = A.get<uint32_t>(bound);
    uint32_t* uid = A.get<uint32_t>(bound);
    void* scan = A.get<uint8_t>(scanScratchBytes(bound));
    if (!A.base || work == 0)
        return hipSuccess;
    const uint32_t cap = tableCap(work);
    hipError_t e;
    if ((e = fillU32(table, cap, kNone, st)) != hipSuccess)
        return e;
    hipLaunchKernelGGL(kInsertFirst<K>, dim3(gridFor(work)), dim3(kThreads), 0, st, keys, d_n, table, cap - 1);
    hipLaunchKernelGGL(kRepresentative<K>, dim3(gridFor(work)), dim3(kThreads), 0, st, keys, d_n, work, table, cap - 1, rep, first);
    if ((e = hipGetLastError()) != hipSuccess)
        return e;
    if ((e = exclusiveScanU32(first, uid, work, d_nu, scan, st)) != hipSuccess)
        return e;
    hipLaunchKernelGGL(kAssignIds<K>, dim3(gridFor(work)), dim3(kThreads), 0, st, keys, d_n, rep, uid, id, uniqueKeys);
    return hipGetLastError();
}

// The whole build. Buffers are carved against the per-level position bounds; with a null carver
// base nothing is launched and the carver only sizes the workspace. The launches cover at most the
// number of non-empty leaves, read back once after the leaf compaction.
hipError_t enqueueBuild(Carver& A, const uint8_t* d_bits, uint32_t R, uint32_t levels, hipStream_t st, uint32_t*& d_counts, std::vector<uint32_t*>& uniqueKeys)
{
    const uint64_t B = R / 4;
    const uint64_t numBlocks = B * B * B;
    // Level l holds at most one node per position: bound[l] = numBlocks / 8^(l - 2). The counts
    // stay on the device until the end.
    std::vector<uint32_t> bound(levels + 1, 0);
    for (uint32_t l = 2; l <= levels; ++l)
        bound[l] = (uint32_t)(numBlocks >> (3 * (l - 2)));
    const bool run = A.base != nullptr;
    d_counts = A.get<uint32_t>(2 * (levels + 1)); // [2l] items, [2l + 1] uniques of level l
    uint64_t* masks = A.get<uint64_t>(numBlocks);
    uint32_t* flags = A.get<uint32_t>(numBlocks);
    uint32_t* pos = A.get<uint32_t>(numBlocks);
    void* scan = A.get<uint8_t>(scanScratchBytes(numBlocks));
    uint32_t* codes = A.get<uint32_t>(numBlocks);
    uint32_t* keys = A.get<uint32_t>(numBlocks * 2);
    hipError_t e;
    uint32_t leaves = 0; // non-empty blocks: every level has at most this many nodes
    if (run) {
        if ((e = fillU32(d_counts, 2 * (levels + 1), 0u, st)) != hipSuccess)
            return e;
        if (R >= 32 && ((uintptr_t)d_bits & 3u) == 0)
            hipLaunchKernelGGL(kLeafMasksCube, dim3((uint32_t)(numBlocks / 512)), dim3(kThreads), 0, st, (const uint32_t*)d_bits, R, masks, flags);
        else
            hipLaunchKernelGGL(kLeafMasks, dim3(gridFor(numBlocks)), dim3(kThreads), 0, st, d_bits, R, numBlocks, masks, flags);
        if ((e = hipGetLastError()) != hipSuccess)
            return e;
        if ((e = exclusiveScanU32(flags, pos, numBlocks, d_counts + 4, scan, st)) != hipSuccess)
            return e;
        hipLaunchKernelGGL(kCompactLeaves, dim3(gridFor(numBlocks)), dim3(kThreads), 0, st, masks, flags, pos, numBlocks, codes, keys);
        if ((e = hipGetLastError()) != hipSuccess)
            return e;
        // The one mid-build readback: it sizes the work of every later launch.
        if ((e = hipMemcpyAsync(&leaves, d_counts + 4, 4, hipMemcpyDeviceToHost, st)) != hipSuccess || (e = hipStreamSynchronize(st)) != hipSuccess)
            return e;
    }
    uniqueKeys.assign(levels + 1, nullptr);
    for (uint32_t level = 2;; ++level) {
        if (bound[level] <= kTail) {
            // This level and all above it: one workgroup (the choice depends on the bounds only,
            // so the sizing pass carves the same buffers).
            TailPlan P {};
            P.first = level;
            P.last = levels;
            P.keys0 = keys;
            P.codes0 = codes;
            P.counts = d_counts;
            P.rep = A.get<uint32_t>(kTail);
            P.uid = A.get<uint32_t>(kTail);
            P.pstart = A.get<uint32_t>(kTail);
            for (uint32_t l = level; l <= levels; ++l) {
                uniqueKeys[l] = A.get<uint32_t>((size_t)bound[l] * (l == 2 ? 2 : 8));
                P.uniq[l] = uniqueKeys[l];
                if (l < levels) {
                    P.pkeys[l] = A.get<uint32_t>((size_t)bound[l + 1] * 8);
                    P.pcodes[l] = A.get<uint32_t>(bound[l + 1]);
                }
            }
            if (run) {
                hipLaunchKernelGGL(kTailLevels, dim3(1), dim3(kTailThreads), 0, st, P);
                if ((e = hipGetLastError()) != hipSuccess)
                    return e;
            }
            break;
        }
        const uint32_t K = level == 2 ? 2 : 8;
        const uint32_t* d_n = run ? d_counts + 2 * level : nullptr;
        uint32_t* d_nu = run ? d_counts + 2 * level + 1 : nullptr;
        uint32_t* id = A.get<uint32_t>(bound[level]);
        uniqueKeys[level] = A.get<uint32_t>((size_t)bound[level] * K);
        const uint32_t work = std::min(bound[level], leaves);
        e = level == 2 ? dedupLevel<2>(A, keys, d_n, bound[level], work, id, uniqueKeys[level], d_nu, st)
                       : dedupLevel<8>(A, keys, d_n, bound[level], work, id, uniqueKeys[level], d_nu, st);
        if (e != hipSuccess)
            return e;
        if (level == levels)
            break;
        // Parents at level + 1: runs of equal code >> 3.
        const uint32_t nb = bound[level], pb = bound[level + 1];
        const uint32_t pwork = std::min(pb, leaves);
        uint32_t* start = A.get<uint32_t>(nb);
        uint32_t* pstart = A.get<uint32_t>(nb);
        void* pscan = A.get<uint8_t>(scanScratchBytes(nb));
        uint32_t* pkeys = A.get<uint32_t>((size_t)pb * 8);
        uint32_t* pcodes = A.get<uint32_t>(pb);
        if (run && work > 0) {
            hipLaunchKernelGGL(kParentStarts, dim3(gridFor(work)), dim3(kThreads), 0, st, codes, d_n, work, start);
            if ((e = hipGetLastError()) != hipSuccess)
                return e;
            if ((e = exclusiveScanU32(start, pstart, work, d_counts + 2 * (level + 1), pscan, st)) != hipSuccess)
                return e;
            if ((e = fillU32(pkeys, (size_t)pwork * 8, kNone, st)) != hipSuccess)
                return e;
            hipLaunchKernelGGL(kParents, dim3(gridFor(work)), dim3(kThreads), 0, st, codes, id, start, pstart, d_n, pkeys, pcodes);
            if ((e = hipGetLastError()) != hipSuccess)
                return e;
        }
        codes = pcodes;
        keys = pkeys;
    }
    return hipSuccess;
}

} // namespace

hipError_t svdagFromDenseDevice(const uint8_t* d_bits, uint32_t R, hipStream_t st, void*& scratch, size_t& scratchBytes, Svdag& out)
{
    uint32_t levels = 0;
    while ((1u << levels) < R)
        ++levels;
    out = Svdag {};
    out.resolution = R;
    out.rootLevel = levels;
    out.nodes.resize(levels + 1);
    out.structureType = 1;
    uint32_t* d_counts = nullptr;
    std::vector<uint32_t*> uniqueKeys;
    Carver sizing;
    hipError_t e = enqueueBuild(sizing, d_bits, R, levels, st, d_counts, uniqueKeys);
    if (e != hipSuccess)
        return e;
    if (scratchBytes < sizing.off || !scratch) {
        if (scratch) {
            if ((e = hipStreamSynchronize(st)) != hipSuccess)
                return e;
            (void)hipFree(scratch);
            scratch = nullptr;
            scratchBytes = 0;
        }
        if ((e = hipMalloc(&scratch, sizing.off)) != hipSuccess)
            return e;
        scratchBytes = sizing.off;
    }
    Carver A { (char*)scratch, 0 };
    if ((e = enqueueBuild(A, d_bits, R, levels, st, d_counts, uniqueKeys)) != hipSuccess)
        return e;
    // One readback of the counts, then the unique lists.
    std::vector<uint32_t> counts(2 * (levels + 1));
    if ((e = hipMemcpyAsync(counts.data(), d_counts, counts.size() * 4, hipMemcpyDeviceToHost, st)) != hipSuccess || (e = hipStreamSynchronize(st)) != hipSuccess)
        return e;
    out.leaves.resize(counts[5]);
    if (counts[5] && (e = hipMemcpyAsync(out.leaves.data(), uniqueKeys[2], (size_t)counts[5] * 8, hipMemcpyDeviceToHost, st)) != hipSuccess)
        return e;
    for (uint32_t level = 3; level <= levels; ++level) {
        const uint32_t nu = counts[2 * level + 1];
        out.nodes[level].resize(nu);
        if (nu && (e = hipMemcpyAsync(out.nodes[level].data(), uniqueKeys[level], (size_t)nu * 32, hipMemcpyDeviceToHost, st)) != hipSuccess)
            return e;
    }
    return hipStreamSynchronize(st);
}

} // namespace tdag
