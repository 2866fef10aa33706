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
__global__ __launch_bounds__(kThreads) void kInsertFirst(const uint32_t* __restrict__ keys, uint32_t n, uint32_t* table, uint32_t capMask)
{
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n)
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

template <int K>
__global__ __launch_bounds__(kThreads) void kRepresentative(const uint32_t* __restrict__ keys, uint32_t n, const uint32_t* __restrict__ table, uint32_t capMask,
    uint32_t* __restrict__ rep, uint32_t* __restrict__ first)
{
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n)
        return;
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
__global__ __launch_bounds__(kThreads) void kAssignIds(const uint32_t* __restrict__ keys, uint32_t n, const uint32_t* __restrict__ rep, const uint32_t* __restrict__ uid,
    uint32_t* __restrict__ id, uint32_t* __restrict__ uniqueKeys)
{
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n)
        return;
    const uint32_t r = rep[i];
    id[i] = uid[r];
    if (r == i)
        for (int j = 0; j < K; ++j)
            uniqueKeys[(size_t)uid[i] * K + j] = keys[(size_t)i * K + j];
}

__global__ __launch_bounds__(kThreads) void kParentStarts(const uint32_t* __restrict__ codes, uint32_t n, uint32_t* __restrict__ start)
{
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n)
        return;
    start[i] = (i == 0 || (codes[i] >> 3) != (codes[i - 1] >> 3)) ? 1u : 0u;
}

// pstart = exclusive scan of start: the parent of item i is pstart[i] + start[i] - 1.
__global__ __launch_bounds__(kThreads) void kParents(const uint32_t* __restrict__ codes, const uint32_t* __restrict__ id, const uint32_t* __restrict__ start,
    const uint32_t* __restrict__ pstart, uint32_t n, uint32_t* __restrict__ pkeys, uint32_t* __restrict__ pcodes)
{
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n)
        return;
    const uint32_t p = pstart[i] + start[i] - 1;
    pkeys[(size_t)p * 8 + (codes[i] & 7u)] = id[i];
    if (start[i])
        pcodes[p] = codes[i] >> 3;
}

// Stream-ordered device allocations of one build, released together.
struct Arena {
    hipStream_t stream;
    std::vector<void*> blocks;
    hipError_t err = hipSuccess;
    template <typename T>
    T* get(size_t count)
    {
        void* p = nullptr;
        if (err == hipSuccess)
            err = hipMallocAsync(&p, std::max<size_t>(count * sizeof(T), 256), stream);
        if (err != hipSuccess)
            return nullptr;
        blocks.push_back(p);
        return (T*)p;
    }
    void release(void* p)
    {
        auto it = std::find(blocks.begin(), blocks.end(), p);
        if (it != blocks.end()) {
            (void)hipFreeAsync(p, stream);
            blocks.erase(it);
        }
    }
    ~Arena()
    {
        for (void* p : blocks)
            (void)hipFreeAsync(p, stream);
        (void)hipStreamSynchronize(stream);
    }
};

inline uint32_t gridFor(uint64_t n) { return (uint32_t)((n + kThreads - 1) / kThreads); }

// First-occurrence dedup of n keys of K words: id[i] (unique number of item i) and the unique keys
// in first-occurrence order, copied to `hostOut` (nu * K words).
template <int K>
hipError_t dedupLevel(Arena& A, const uint32_t* keys, uint32_t n, uint32_t* id, std::vector<uint32_t>& hostOut, uint32_t* d_total, hipStream_t st)
{
    uint32_t cap = 64;
    while (cap < 2 * (uint64_t)n)
        cap <<= 1;
    uint32_t* table = A.get<uint32_t>(cap);
    uint32_t* rep = A.get<uint32_t>(n);
    uint32_t* first = A.get<uint32_t>(n);
    uint32_t* uid = A.get<uint32_t>(n);
    void* scan = A.get<uint8_t>(scanScratchBytes(n));
    if (A.err != hipSuccess)
        return A.err;
    hipError_t e;
    if ((e = fillU32(table, cap, kNone, st)) != hipSuccess)
        return e;
    hipLaunchKernelGGL(kInsertFirst<K>, dim3(gridFor(n)), dim3(kThreads), 0, st, keys, n, table, cap - 1);
    hipLaunchKernelGGL(kRepresentative<K>, dim3(gridFor(n)), dim3(kThreads), 0, st, keys, n, table, cap - 1, rep, first);
    if ((e = hipGetLastError()) != hipSuccess)
        return e;
    if ((e = exclusiveScanU32(first, uid, n, d_total, scan, st)) != hipSuccess)
        return e;
    uint32_t nu = 0;
    if ((e = hipMemcpyAsync(&nu, d_total, 4, hipMemcpyDeviceToHost, st)) != hipSuccess || (e = hipStreamSynchronize(st)) != hipSuccess)
        return e;
    uint32_t* uniqueKeys = A.get<uint32_t>((size_t)nu * K);
    if (A.err != hipSuccess)
        return A.err;
    hipLaunchKernelGGL(kAssignIds<K>, dim3(gridFor(n)), dim3(kThreads), 0, st, keys, n, rep, uid, id, uniqueKeys);
    if ((e = hipGetLastError()) != hipSuccess)
        return e;
    hostOut.resize((size_t)nu * K);
    if ((e = hipMemcpyAsync(hostOut.data(), uniqueKeys, (size_t)nu * K * 4, hipMemcpyDeviceToHost, st)) != hipSuccess)
        return e;
    if ((e = hipStreamSynchronize(st)) != hipSuccess)
        return e;
    A.release(table);
    A.release(rep);
    A.release(first);
    A.release(uid);
    A.release(scan);
    A.release(uniqueKeys);
    return hipSuccess;
}

} // namespace

hipError_t svdagFromDenseDevice(const uint8_t* d_bits, uint32_t R, hipStream_t st, Svdag& out)
{
    uint32_t levels = 0;
    while ((1u << levels) < R)
        ++levels;
    out = Svdag {};
    out.resolution = R;
    out.rootLevel = levels;
    out.nodes.resize(levels + 1);
    out.structureType = 1;
    Arena A { st };
    const uint64_t B = R / 4;
    const uint64_t numBlocks = B * B * B;
    uint32_t* d_total = A.get<uint32_t>(1);
    uint64_t* masks = A.get<uint64_t>(numBlocks);
    uint32_t* flags = A.get<uint32_t>(numBlocks);
    uint32_t* pos = A.get<uint32_t>(numBlocks);
    void* scan = A.get<uint8_t>(scanScratchBytes(numBlocks));
    if (A.err != hipSuccess)
        return A.err;
    hipError_t e;
    hipLaunchKernelGGL(kLeafMasks, dim3(gridFor(numBlocks)), dim3(kThreads), 0, st, d_bits, R, numBlocks, masks, flags);
    if ((e = hipGetLastError()) != hipSuccess)
        return e;
    if ((e = exclusiveScanU32(flags, pos, numBlocks, d_total, scan, st)) != hipSuccess)
        return e;
    uint32_t n = 0;
    if ((e = hipMemcpyAsync(&n, d_total, 4, hipMemcpyDeviceToHost, st)) != hipSuccess || (e = hipStreamSynchronize(st)) != hipSuccess)
        return e;
    if (n == 0)
        return hipSuccess;
    uint32_t* codes = A.get<uint32_t>(n);
    uint32_t* keys = A.get<uint32_t>((size_t)n * 2);
    if (A.err != hipSuccess)
        return A.err;
    hipLaunchKernelGGL(kCompactLeaves, dim3(gridFor(numBlocks)), dim3(kThreads), 0, st, masks, flags, pos, numBlocks, codes, keys);
    if ((e = hipGetLastError()) != hipSuccess)
        return e;
    A.release(masks);
    A.release(flags);
    A.release(pos);
    A.release(scan);
    std::vector<uint32_t> host;
    for (uint32_t level = 2;; ++level) {
        uint32_t* id = A.get<uint32_t>(n);
        if (A.err != hipSuccess)
            return A.err;
        if (level == 2) {
            if ((e = dedupLevel<2>(A, keys, n, id, host, d_total, st)) != hipSuccess)
                return e;
            out.leaves.resize(host.size() / 2);
            std::copy((const uint8_t*)host.data(), (const uint8_t*)(host.data() + host.size()), (uint8_t*)out.leaves.data());
        } else {
            if ((e = dedupLevel<8>(A, keys, n, id, host, d_total, st)) != hipSuccess)
                return e;
            auto& nodes = out.nodes[level];
            nodes.resize(host.size() / 8);
            std::copy(host.begin(), host.end(), &nodes.data()[0][0]);
        }
        if (level == levels)
            break;
        // Parents at level + 1: runs of equal code >> 3.
        uint32_t* start = A.get<uint32_t>(n);
        uint32_t* pstart = A.get<uint32_t>(n);
        void* pscan = A.get<uint8_t>(scanScratchBytes(n));
        if (A.err != hipSuccess)
            return A.err;
        hipLaunchKernelGGL(kParentStarts, dim3(gridFor(n)), dim3(kThreads), 0, st, codes, n, start);
        if ((e = hipGetLastError()) != hipSuccess)
            return e;
        if ((e = exclusiveScanU32(start, pstart, n, d_total, pscan, st)) != hipSuccess)
            return e;
        uint32_t np = 0;
        if ((e = hipMemcpyAsync(&np, d_total, 4, hipMemcpyDeviceToHost, st)) != hipSuccess || (e = hipStreamSynchronize(st)) != hipSuccess)
            return e;
        uint32_t* pkeys = A.get<uint32_t>((size_t)np * 8);
        uint32_t* pcodes = A.get<uint32_t>(np);
        if (A.err != hipSuccess)
            return A.err;
        if ((e = fillU32(pkeys, (size_t)np * 8, kNone, st)) != hipSuccess)
            return e;
        hipLaunchKernelGGL(kParents, dim3(gridFor(n)), dim3(kThreads), 0, st, codes, id, start, pstart, n, pkeys, pcodes);
        if ((e = hipGetLastError()) != hipSuccess)
            return e;
        A.release(start);
        A.release(pstart);
        A.release(pscan);
        A.release(id);
        A.release(codes);
        A.release(keys);
        codes = pcodes;
        keys = pkeys;
        n = np;
    }
    return hipSuccess;
}

} // namespace tdag
