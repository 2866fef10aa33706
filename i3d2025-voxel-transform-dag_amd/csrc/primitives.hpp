// Device primitives written for gfx950 (wave64): exclusive scan of uint32 and a stable LSD
// radix sort of (key, value) pairs. Used for the popcount rank order of the matcher, the
// unique-item compaction, and the hierarchical dedup's key sort.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

namespace tdag {

// Fill kernels (used instead of hipMemsetAsync so that callers may capture the work in graphs).
hipError_t fillU32(uint32_t* d_p, size_t n, uint32_t v, hipStream_t stream);
hipError_t fillU64(uint64_t* d_p, size_t n, uint64_t v, hipStream_t stream);

// Scratch needed by exclusiveScanU32 for n elements (bytes).
size_t scanScratchBytes(size_t n);
// out[i] = sum(in[0..i)); *d_total (optional, device) = sum(in). in may equal out.
hipError_t exclusiveScanU32(const uint32_t* d_in, uint32_t* d_out, size_t n, uint32_t* d_total, void* d_scratch, hipStream_t stream);

// Scratch needed by radixSortPairs for n elements with the given key width (bytes).
size_t radixScratchBytes(size_t n, int keyBytes);
// Stable sort of (keys, vals) by the bit range [beginBit, endBit) of the keys.
// keys/vals are sorted in place (ping-pong through scratch).
hipError_t radixSortPairsU32(uint32_t* d_keys, uint32_t* d_vals, size_t n, int beginBit, int endBit, void* d_scratch, hipStream_t stream);
hipError_t radixSortPairsU64(uint64_t* d_keys, uint32_t* d_vals, size_t n, int beginBit, int endBit, void* d_scratch, hipStream_t stream);

} // namespace tdag
