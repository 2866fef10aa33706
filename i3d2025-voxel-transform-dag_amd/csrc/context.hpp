// Internal view of tdag_context and the helpers every C-ABI source file shares (error
// reporting, argument checks, scratch growth).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <string>

#include "tdag.h"

struct tdag_context {
    int device = 0;
    hipStream_t stream = nullptr;
    void* scratch = nullptr;
    size_t scratchBytes = 0;
    // Host-buffer entry points stage their inputs/outputs here.
    void* io = nullptr;
    size_t ioBytes = 0;
    uint64_t* pinned = nullptr; // 4 words of pinned host memory for small readbacks
    hipEvent_t ev[4] = { nullptr, nullptr, nullptr, nullptr };
    tdag_match_stats last {};
    int32_t variant[6] = { -1, -1, -1, -1, -1, -1 }; // per level; -1 = library default
    bool counters = false; // collect per-stage probe counters
    uint32_t firstBatch[6] = { 0, 0, 0, 0, 0, 0 }; // rank-batched search schedule per level (0 = default)
    uint32_t batchGrowth[6] = { 0, 0, 0, 0, 0, 0 };
    uint32_t bloomBits[6] = { 0, 0, 0, 0, 0, 0 }; // Bloom filter bits per table entry (0 = default)
    uint32_t childFrames[6] = { 0, 0, 0, 0, 0, 0 }; // levels 3..5: child-table permutation frames, 1 or 6 (0 = default)
    void* svdagScratch = nullptr; // workspace of the device SVDAG build (grown on demand)
    size_t svdagScratchBytes = 0;
};

namespace tdag {

// Looks up, once per device, the device address of the allocation-free matcher's scratch array
// (device must be current). Later calls return the first call's result.
hipError_t noallocScratchInit(int device);

namespace capi {

// Records `msg` as the calling thread's last error and returns `s`.
tdag_status fail(tdag_status s, const std::string& msg);
// Records a HIP failure (ENOMEM for out-of-memory, EDEVICE otherwise).
tdag_status failHip(hipError_t e, const char* what);
// Grows a device buffer to at least `need` bytes (contents are not kept).
tdag_status ensureBuffer(void*& buf, size_t& have, size_t need, const char* what);
// TransformDAGConfig flags as the matcher's configuration bits.
uint32_t cfgBitsOf(const tdag_config* cfg);
// Common argument checks of the matcher entry points.
tdag_status checkArgs(tdag_context* ctx, uint32_t level, uint64_t n, const tdag_config* cfg);

} // namespace capi
} // namespace tdag

#define CAPI_HIP(expr, what)                          \
    do {                                              \
        const hipError_t e_ = (expr);                 \
        if (e_ != hipSuccess)                         \
            return ::tdag::capi::failHip(e_, what);   \
    } while (0)
