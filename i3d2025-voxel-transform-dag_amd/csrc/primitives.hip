// Device scan and stable LSD radix sort for gfx950 (wave64, 256-thread blocks).
#include "primitives.hpp"

#include <algorithm>

namespace tdag {

namespace {

constexpr int kBlock = 256;
constexpr int kItems = 16;
constexpr int kTile = kBlock * kItems; // 4096 elements per block

inline size_t alignUp(size_t v, size_t a) { return (v + a - 1) / a * a; }
inline size_t ceilDiv(size_t a, size_t b) { return (a + b - 1) / b; }

__device__ __forceinline__ uint32_t waveInclusiveScan(uint32_t x)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t v = __shfl_up(x, off, 64);
        if (lane >= off)
            x += v;
    }
    return x;
}

// Exclusive scan of one value per thread across a 256-thread block; returns the block total.
__device__ __forceinline__ uint32_t blockExclusiveScan(uint32_t x, uint32_t* s_waves, uint32_t& total)
{
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const uint32_t inc = waveInclusiveScan(x);
    if (lane == 63)
        s_waves[wave] = inc;
    __syncthreads();
    uint32_t waveOff = 0;
    total = 0;
#pragma unroll
    for (int w = 0; w < kBlock / 64; ++w) {
        const uint32_t v = s_waves[w];
        if (w < wave)
            waveOff += v;
        total += v;
    }
    __syncthreads();
    return waveOff + inc - x;
}

__global__ __launch_bounds__(kBlock) void kScanReduce(const uint32_t* __restrict__ in, size_t n, uint32_t* __restrict__ sums)
{
    __shared__ uint32_t s_waves[kBlock / 64];
    const size_t base = (size_t)blockIdx.x * kTile;
    uint32_t acc = 0;
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
        const size_t idx = base + (size_t)i * kBlock + threadIdx.x;
        if (idx < n)
            acc += in[idx];
    }
    uint32_t total;
    (void)blockExclusiveScan(acc, s_waves, total);
    if (threadIdx.x == 0)
        sums[blockIdx.x] = total;
}

// Exclusive scan of a tile, offset by offsets[blockIdx.x] (or 0 when offsets == nullptr).
__global__ __launch_bounds__(kBlock) void kScanDown(const uint32_t* in, uint32_t* out, size_t n, const uint32_t* __restrict__ offsets, uint32_t* d_total)
{
    __shared__ uint32_t s_tile[kTile];
    __shared__ uint32_t s_waves[kBlock / 64];
    const size_t base = (size_t)blockIdx.x * kTile;
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
        const size_t idx = base + (size_t)i * kBlock + threadIdx.x;
        s_tile[i * kBlock + threadIdx.x] = idx < n ? in[idx] : 0u;
    }
    __syncthreads();
    uint32_t local[kItems];
    uint32_t acc = 0;
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
        local[i] = acc;
        acc += s_tile[threadIdx.x * kItems + i];
    }
    uint32_t total;
    const uint32_t threadOff = blockExclusiveScan(acc, s_waves, total);
    const uint32_t blockOff = offsets ? offsets[blockIdx.x] : 0u;
#pragma unroll
    for (int i = 0; i < kItems; ++i)
        s_tile[threadIdx.x * kItems + i] = blockOff + threadOff + local[i];
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
        const size_t idx = base + (size_t)i * kBlock + threadIdx.x;
        if (idx < n)
            out[idx] = s_tile[i * kBlock + threadIdx.x];
    }
    if (d_total && blockIdx.x == gridDim.x - 1 && threadIdx.x == 0)
        *d_total = blockOff + total;
}

__global__ void kFillU32(uint32_t* p, size_t n, uint32_t v)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void kFillU64(uint64_t* p, size_t n, uint64_t v)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

template <typename K>
__global__ __launch_bounds__(kBlock) void kRadixHist(const K* __restrict__ keys, size_t n, int shift, uint32_t* __restrict__ hist, uint32_t numTiles)
{
    __shared__ uint32_t s_hist[256];
    s_hist[threadIdx.x] = 0;
    __syncthreads();
    const size_t base = (size_t)blockIdx.x * kTile;
#pragma unroll 4
    for (int i = 0; i < kItems; ++i) {
        const size_t idx = base + (size_t)i * kBlock + threadIdx.x;
        if (idx < n) {
            const uint32_t d = (uint32_t)((keys[idx] >> shift) & 255u);
            atomicAdd(&s_hist[d], 1u);
        }
    }
    __syncthreads();
    hist[(size_t)threadIdx.x * numTiles + blockIdx.x] = s_hist[threadIdx.x];
}

template <typename K>
__global__ __launch_bounds__(kBlock) void kRadixScatter(const K* __restrict__ keysIn, const uint32_t* __restrict__ valsIn, size_t n, int shift,
    const uint32_t* __restrict__ histScan, uint32_t numTiles, K* __restrict__ keysOut, uint32_t* __restrict__ valsOut)
{
    __shared__ uint32_t s_base[256];
    __shared__ uint32_t s_waveCnt[kBlock / 64][256];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    s_base[threadIdx.x] = histScan[(size_t)threadIdx.x * numTiles + blockIdx.x];
#pragma unroll
    for (int w = 0; w < kBlock / 64; ++w)
        s_waveCnt[w][threadIdx.x] = 0;
    __syncthreads();
    const uint64_t lanesBelow = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const size_t base = (size_t)blockIdx.x * kTile;
    for (int i = 0; i < kItems; ++i) {
        const size_t idx = base + (size_t)i * kBlock + threadIdx.x;
        const bool valid = idx < n;
        K key = 0;
        uint32_t val = 0;
        if (valid) {
            key = keysIn[idx];
            val = valsIn[idx];
        }
        const uint32_t d = (uint32_t)((key >> shift) & 255u);
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        const uint32_t rankInWave = (uint32_t)__popcll(peers & lanesBelow);
        if (valid && rankInWave == 0)
            s_waveCnt[wave][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        uint32_t pos = 0;
        if (valid) {
            pos = s_base[d] + rankInWave;
            for (int w = 0; w < wave; ++w)
                pos += s_waveCnt[w][d];
        }
        __syncthreads();
        {
            uint32_t sum = 0;
#pragma unroll
            for (int w = 0; w < kBlock / 64; ++w) {
                sum += s_waveCnt[w][threadIdx.x];
                s_waveCnt[w][threadIdx.x] = 0;
            }
            s_base[threadIdx.x] += sum;
        }
        __syncthreads();
        if (valid) {
            keysOut[pos] = key;
            valsOut[pos] = val;
        }
    }
}

struct ScanPlan {
    // Sizes of the recursion levels of the block-sum arrays.
    size_t levels[8];
    int numLevels;
};

ScanPlan makeScanPlan(size_t n)
{
    ScanPlan p {};
    size_t m = n;
    while (m > (size_t)kTile && p.numLevels < 8) {
        m = ceilDiv(m, kTile);
        p.levels[p.numLevels++] = m;
    }
    return p;
}

hipError_t scanRecursive(const uint32_t* in, uint32_t* out, size_t n, uint32_t* d_total, uint32_t* scratch, hipStream_t stream)
{
    if (n == 0) {
        if (d_total) {
            hipLaunchKernelGGL(kFillU32, dim3(1), dim3(64), 0, stream, d_total, (size_t)1, 0u);
            return hipGetLastError();
        }
        return hipSuccess;
    }
    const size_t tiles = ceilDiv(n, kTile);
    if (tiles == 1) {
        hipLaunchKernelGGL(kScanDown, dim3(1), dim3(kBlock), 0, stream, in, out, n, (const uint32_t*)nullptr, d_total);
        return hipGetLastError();
    }
    uint32_t* sums = scratch;
    uint32_t* rest = scratch + alignUp(tiles, 64);
    hipLaunchKernelGGL(kScanReduce, dim3((unsigned)tiles), dim3(kBlock), 0, stream, in, n, sums);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return e;
    e = scanRecursive(sums, sums, tiles, nullptr, rest, stream);
    if (e != hipSuccess)
        return e;
    hipLaunchKernelGGL(kScanDown, dim3((unsigned)tiles), dim3(kBlock), 0, stream, in, out, n, (const uint32_t*)sums, d_total);
    return hipGetLastError();
}

template <typename K>
hipError_t radixSortImpl(K* d_keys, uint32_t* d_vals, size_t n, int beginBit, int endBit, void* d_scratch, hipStream_t stream)
{
    if (n <= 1 || beginBit >= endBit)
        return hipSuccess;
    const size_t tiles = ceilDiv(n, kTile);
    char* p = (char*)d_scratch;
    K* altKeys = (K*)p;
    p += alignUp(n * sizeof(K), 256);
    uint32_t* altVals = (uint32_t*)p;
    p += alignUp(n * sizeof(uint32_t), 256);
    uint32_t* hist = (uint32_t*)p;
    p += alignUp(256 * tiles * sizeof(uint32_t), 256);
    uint32_t* scanScratch = (uint32_t*)p;

    K* srcK = d_keys;
    uint32_t* srcV = d_vals;
    K* dstK = altKeys;
    uint32_t* dstV = altVals;
    int passes = 0;
    for (int shift = beginBit; shift < endBit; shift += 8) {
        hipLaunchKernelGGL(kRadixHist<K>, dim3((unsigned)tiles), dim3(kBlock), 0, stream, (const K*)srcK, n, shift, hist, (uint32_t)tiles);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess)
            return e;
        e = scanRecursive(hist, hist, 256 * tiles, nullptr, scanScratch, stream);
        if (e != hipSuccess)
            return e;
        hipLaunchKernelGGL(kRadixScatter<K>, dim3((unsigned)tiles), dim3(kBlock), 0, stream, (const K*)srcK, (const uint32_t*)srcV, n, shift,
            (const uint32_t*)hist, (uint32_t)tiles, dstK, dstV);
        e = hipGetLastError();
        if (e != hipSuccess)
            return e;
        K* tk = srcK;
        srcK = dstK;
        dstK = tk;
        uint32_t* tv = srcV;
        srcV = dstV;
        dstV = tv;
        ++passes;
    }
    if (passes & 1) {
        hipError_t e = hipMemcpyAsync(d_keys, srcK, n * sizeof(K), hipMemcpyDeviceToDevice, stream);
        if (e != hipSuccess)
            return e;
        e = hipMemcpyAsync(d_vals, srcV, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, stream);
        if (e != hipSuccess)
            return e;
    }
    return hipSuccess;
}

} // namespace

hipError_t fillU32(uint32_t* d_p, size_t n, uint32_t v, hipStream_t stream)
{
    if (n == 0)
        return hipSuccess;
    const unsigned blocks = (unsigned)std::min<size_t>(ceilDiv(n, 256), 4096);
    hipLaunchKernelGGL(kFillU32, dim3(blocks), dim3(256), 0, stream, d_p, n, v);
    return hipGetLastError();
}

hipError_t fillU64(uint64_t* d_p, size_t n, uint64_t v, hipStream_t stream)
{
    if (n == 0)
        return hipSuccess;
    const unsigned blocks = (unsigned)std::min<size_t>(ceilDiv(n, 256), 4096);
    hipLaunchKernelGGL(kFillU64, dim3(blocks), dim3(256), 0, stream, d_p, n, v);
    return hipGetLastError();
}

size_t scanScratchBytes(size_t n)
{
    const ScanPlan p = makeScanPlan(n);
    size_t words = 0;
    for (int i = 0; i < p.numLevels; ++i)
        words += alignUp(p.levels[i], 64);
    return (words + 64) * sizeof(uint32_t);
}

hipError_t exclusiveScanU32(const uint32_t* d_in, uint32_t* d_out, size_t n, uint32_t* d_total, void* d_scratch, hipStream_t stream)
{
    return scanRecursive(d_in, d_out, n, d_total, (uint32_t*)d_scratch, stream);
}

size_t radixScratchBytes(size_t n, int keyBytes)
{
    const size_t tiles = ceilDiv(n > 0 ? n : 1, kTile);
    return alignUp(n * keyBytes, 256) + alignUp(n * sizeof(uint32_t), 256) + alignUp(256 * tiles * sizeof(uint32_t), 256) + scanScratchBytes(256 * tiles);
}

hipError_t radixSortPairsU32(uint32_t* d_keys, uint32_t* d_vals, size_t n, int beginBit, int endBit, void* d_scratch, hipStream_t stream)
{
    return radixSortImpl<uint32_t>(d_keys, d_vals, n, beginBit, endBit, d_scratch, stream);
}

hipError_t radixSortPairsU64(uint64_t* d_keys, uint32_t* d_vals, size_t n, int beginBit, int endBit, void* d_scratch, hipStream_t stream)
{
    return radixSortImpl<uint64_t>(d_keys, d_vals, n, beginBit, endBit, d_scratch, stream);
}

} // namespace tdag
