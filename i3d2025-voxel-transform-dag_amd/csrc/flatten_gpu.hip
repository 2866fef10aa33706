// createLargeSubGrid on the device (large_sub_grid.h:127-141 over EditStructure::get,
// structure.h:334-355): the linear R^3 masks (bit i = x + R*y + R*R*z, W = max(1, R^3/64) words) of
// level-L SVDAG nodes, written straight into device memory for the matcher.
//   1. kBlockLeaves: one thread per (item, 4^3 block), block = Morton code within the item, walks
//      from the item's node down the Morton child order (child index = x | y << 1 | z << 2 of the
//      coordinate bit at that level) and stores the block's leaf mask (0 for an empty subtree);
//   2. kAssembleWords: one thread per output word; a word covers 64 / R rows of R voxels, i.e.
//      exactly 16 row nibbles, each read from one block's leaf mask (Morton bits {0, 1, 8, 9}
//      shifted by the row's y, z within the block).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "pipeline.hpp"

namespace tdag {

namespace {

constexpr uint32_t kFlatThreads = 256;
constexpr uint32_t kNoChild = 0xFFFFFFFFu;

__device__ inline uint32_t spread2f(uint32_t v) { return (v & 1u) | ((v & 2u) << 2); }

__device__ inline uint32_t spread3f(uint32_t v)
{
    uint32_t x = v & 0x3FFu;
    x = (x | (x << 16)) & 0x030000FFu;
    x = (x | (x << 8)) & 0x0300F00Fu;
    x = (x | (x << 4)) & 0x030C30C3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}

// An item index outside the level sets *badItem and reads as empty.
__global__ __launch_bounds__(kFlatThreads) void kBlockLeaves(DeviceSvdag d, uint32_t level, const uint32_t* __restrict__ items, uint64_t n,
    uint64_t* __restrict__ blockMasks, uint32_t* __restrict__ badItem)
{
    const uint32_t blocksLog = 3 * (level - 2); // (R / 4)^3 blocks per item
    const uint64_t t = (uint64_t)blockIdx.x * kFlatThreads + threadIdx.x;
    if (t >= (n << blocksLog))
        return;
    const uint64_t item = t >> blocksLog;
    const uint32_t blk = (uint32_t)(t & ((1ull << blocksLog) - 1));
    uint32_t node = items[item];
    if (node >= d.count[level]) {
        if (badItem)
            *badItem = 1u;
        blockMasks[t] = 0;
        return;
    }
    for (uint32_t l = level; l > 2; --l) {
        const uint32_t child = (blk >> (3 * (l - 3))) & 7u;
        node = d.nodes[l][(size_t)node * 8 + child];
        if (node == kNoChild) {
            blockMasks[t] = 0;
            return;
        }
    }
    blockMasks[t] = d.leaves[node];
}

__global__ __launch_bounds__(kFlatThreads) void kAssembleWords(uint32_t level, uint64_t n, const uint64_t* __restrict__ blockMasks, uint64_t* __restrict__ out)
{
    const uint32_t R = 1u << level;
    const uint32_t wordsLog = level == 2 ? 0 : 3 * level - 6; // W = R^3 / 64
    const uint32_t blocksLog = 3 * (level - 2);
    const uint64_t t = (uint64_t)blockIdx.x * kFlatThreads + threadIdx.x;
    if (t >= (n << wordsLog))
        return;
    const uint64_t item = t >> wordsLog;
    const uint32_t w = (uint32_t)(t & ((1ull << wordsLog) - 1));
    const uint64_t* bm = blockMasks + (item << blocksLog);
    const uint32_t rowsPerWord = 64u / R;
    uint64_t word = 0;
    for (uint32_t r = 0; r < rowsPerWord; ++r) {
        const uint32_t g = w * rowsPerWord + r; // global row = y + R * z
        const uint32_t y = g & (R - 1), z = g >> level;
        const uint32_t o = spread2f(y & 3u) << 1 | spread2f(z & 3u) << 2;
        const uint32_t blockYZ = spread3f(y >> 2) << 1 | spread3f(z >> 2) << 2;
        for (uint32_t bx = 0; bx < R / 4; ++bx) {
            const uint64_t leaf = bm[spread3f(bx) | blockYZ];
            const uint64_t nib = ((leaf >> o) & 3ull) | (((leaf >> (o + 8)) & 3ull) << 2);
            word |= nib << (r * R + 4 * bx);
        }
    }
    out[t] = word;
}

} // namespace

size_t flattenTempBytes(uint32_t level, uint64_t n) { return (size_t)(n << (3 * (level - 2))) * 8; }

hipError_t flattenDevice(const DeviceSvdag& d, uint32_t level, const uint32_t* d_items, uint64_t n, uint64_t* d_out, uint64_t* d_temp, hipStream_t stream,
    uint32_t* d_badItem)
{
    if (n == 0)
        return hipSuccess;
    if (level < 2 || level > d.maxLevel)
        return hipErrorInvalidValue;
    const uint64_t blocks = n << (3 * (level - 2));
    const uint64_t words = level == 2 ? n : n << (3 * level - 6);
    hipLaunchKernelGGL(kBlockLeaves, dim3((uint32_t)((blocks + kFlatThreads - 1) / kFlatThreads)), dim3(kFlatThreads), 0, stream, d, level, d_items, n, d_temp,
        d_badItem);
    hipLaunchKernelGGL(kAssembleWords, dim3((uint32_t)((words + kFlatThreads - 1) / kFlatThreads)), dim3(kFlatThreads), 0, stream, level, n, d_temp, d_out);
    return hipGetLastError();
}

hipError_t uploadSvdag(const Svdag& s, uint32_t maxLevel, DeviceSvdag& d)
{
    d = DeviceSvdag {};
    hipError_t e;
    d.count[2] = (uint32_t)s.leaves.size();
    if ((e = hipMalloc((void**)&d.leaves, std::max<size_t>(8, s.leaves.size() * 8))) != hipSuccess)
        return e;
    if (!s.leaves.empty() && (e = hipMemcpy((void*)d.leaves, s.leaves.data(), s.leaves.size() * 8, hipMemcpyHostToDevice)) != hipSuccess)
        return e;
    for (uint32_t l = 3; l <= maxLevel && l <= s.rootLevel && l < 6; ++l) {
        const auto& nodes = s.nodes[l];
        d.count[l] = (uint32_t)nodes.size();
        if ((e = hipMalloc((void**)&d.nodes[l], std::max<size_t>(32, nodes.size() * 32))) != hipSuccess)
            return e;
        if (!nodes.empty() && (e = hipMemcpy((void*)d.nodes[l], nodes.data(), nodes.size() * 32, hipMemcpyHostToDevice)) != hipSuccess)
            return e;
    }
    d.maxLevel = std::min<uint32_t>(std::min(maxLevel, s.rootLevel), 5);
    return hipSuccess;
}

void freeDeviceSvdag(DeviceSvdag& d)
{
    if (d.leaves)
        (void)hipFree((void*)d.leaves);
    for (auto& p : d.nodes)
        if (p)
            (void)hipFree((void*)p);
    d = DeviceSvdag {};
}

} // namespace tdag
