// Hierarchical sym+perm dedup (constructTransformDAGHierarchical, transform_dag_hierarchical.cpp
// :304-523; with staticOffsets the StaticStructure pass, :531-743) on the device. Same TDag and
// mid-level mapping as hierarchicalDedup, bit for bit.
//
// The reference scans each group of nodes (same sorted child pointers) in index order and keeps
// a node unless the first transform ID, in ID order, that maps it onto an earlier kept node
// exists. That is a sequential scan; the device restates it per node:
//   * child equality (TaggedPointer::operator==) is equality of the pointer and of the state's
//     coset under the child's invariance subgroup H, so replacing each child state by the
//     smallest state of its coset turns node equality into plain equality;
//   * the canonical key of a node is its smallest normalised transformed copy over the allowed
//     transforms (a subgroup of the 48), so two nodes are equivalent exactly when their keys are
//     equal;
//   * equivalent nodes share a group (their child pointers are the same multiset), so the node the
//     reference keeps for a class is the class's smallest index, found with a min-index hash
//     table on the keys; every other member takes the first allowed ID that maps it onto it;
//   * kept nodes are numbered in the reference's emission order, ascending (group hash, index),
//     by a radix sort and a scan.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "pipeline.hpp"
#include "primitives.hpp"

namespace tdag {

namespace {

constexpr uint32_t kHT = 256;
constexpr uint32_t kEmptyPtr = 0xFFFFFFFFu;
constexpr uint32_t kNoSlot = 0xFFFFFFFFu;

// Transform tables in device memory; kernels stage them in LDS.
struct DevTables {
    uint8_t compose[48 * 48];
    uint8_t childPos[48 * 8];
    uint8_t required[48 * 48];
    uint8_t cosetOf[48 * 48]; // cosetOf[s * 48 + h]: the state t with required[s][t] == h
    uint8_t inverseId[48];
    uint8_t ids[48];
    uint32_t numIds;
};

struct LdsTables {
    uint8_t compose[48 * 48];
    uint8_t childPos[48 * 8];
    uint8_t required[48 * 48];
    uint8_t cosetOf[48 * 48];
    uint8_t ids[48];
};

__device__ inline void stageTables(const DevTables* g, LdsTables* s)
{
    for (uint32_t i = threadIdx.x; i < 48 * 48; i += blockDim.x) {
        s->compose[i] = g->compose[i];
        s->required[i] = g->required[i];
        s->cosetOf[i] = g->cosetOf[i];
    }
    for (uint32_t i = threadIdx.x; i < 48 * 8; i += blockDim.x)
        s->childPos[i] = g->childPos[i];
    for (uint32_t i = threadIdx.x; i < 48; i += blockDim.x)
        s->ids[i] = g->ids[i];
    __syncthreads();
}

inline uint32_t gridOf(uint64_t n) { return (uint32_t)((n + kHT - 1) / kHT); }

__global__ __launch_bounds__(kHT) void kLeafNodes(const uint64_t* __restrict__ leaves, uint32_t n, const uint32_t* __restrict__ lptr,
    const uint64_t* __restrict__ linv, const uint8_t* __restrict__ lst, uint32_t* __restrict__ cptr, uint64_t* __restrict__ cinv, uint8_t* __restrict__ cst)
{
    const uint64_t i = (uint64_t)blockIdx.x * kHT + threadIdx.x;
    if (i >= (uint64_t)n * 8)
        return;
    const uint32_t b = (uint32_t)(leaves[i >> 3] >> (8 * (i & 7))) & 0xFFu;
    cptr[i] = lptr[b];
    cinv[i] = linv[b];
    cst[i] = lst[b];
}

__global__ __launch_bounds__(kHT) void kInnerNodes(const uint32_t* __restrict__ child, uint32_t n, uint32_t below, const uint32_t* __restrict__ mptr,
    const uint64_t* __restrict__ minv, const uint8_t* __restrict__ mst, uint32_t* __restrict__ cptr, uint64_t* __restrict__ cinv, uint8_t* __restrict__ cst,
    uint32_t* __restrict__ bad)
{
    const uint64_t i = (uint64_t)blockIdx.x * kHT + threadIdx.x;
    if (i >= (uint64_t)n * 8)
        return;
    const uint32_t ch = child[i];
    if (ch != kEmptyChild && ch >= below) { // a child index outside the level below
        atomicOr(bad, 1u);
        cptr[i] = kEmptyPtr;
        cinv[i] = ~0ull;
        cst[i] = 0;
        return;
    }
    if (ch == kEmptyChild) {
        cptr[i] = kEmptyPtr;
        cinv[i] = ~0ull;
        cst[i] = 0;
    } else {
        cptr[i] = mptr[ch];
        cinv[i] = minv[ch];
        cst[i] = mst[ch];
    }
}

// Group key: hash_combine over the 8 child pointers sorted ascending (with hashOf, a pointer p
// enters as hashOf[p]; the empty child as itself).
__global__ __launch_bounds__(kHT) void kSeeds(const uint32_t* __restrict__ cptr, uint32_t n, const uint32_t* __restrict__ hashOf, uint64_t* __restrict__ seed,
    uint32_t* __restrict__ idx)
{
    const uint32_t i = blockIdx.x * kHT + threadIdx.x;
    if (i >= n)
        return;
    uint64_t v[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const uint32_t p = cptr[(size_t)i * 8 + c];
        v[c] = (hashOf && p != kEmptyPtr) ? (uint64_t)hashOf[p] : (uint64_t)p;
    }
    for (int a = 1; a < 8; ++a) { // insertion sort of 8
        const uint64_t x = v[a];
        int b = a - 1;
        while (b >= 0 && v[b] > x) {
            v[b + 1] = v[b];
            --b;
        }
        v[b + 1] = x;
    }
    uint64_t s = 0;
    for (int c = 0; c < 8; ++c)
        s ^= v[c] + 0x9e3779b9ull + (s << 6) + (s >> 2);
    seed[i] = s;
    idx[i] = i;
}

// Canonical key: 8 words ptr | normalised state << 40, smallest over the allowed IDs. The
// transformed copy is assembled in a per-thread LDS row (child c goes to slot childPos[id][c]);
// private arrays are only indexed by unrolled loop counters.
__global__ __launch_bounds__(kHT) void kCanonKeys(const DevTables* __restrict__ gt, const uint32_t* __restrict__ cptr, const uint8_t* __restrict__ normTab,
    uint32_t normRows, const uint8_t* __restrict__ cst, uint32_t n, uint32_t numIds, uint64_t* __restrict__ keys, uint32_t* __restrict__ bad)
{
    __shared__ LdsTables T;
    __shared__ uint64_t rows[kHT * 8];
    stageTables(gt, &T);
    const uint32_t i = blockIdx.x * kHT + threadIdx.x;
    if (i >= n)
        return;
    uint64_t* row = rows + threadIdx.x * 8;
    uint32_t p[8];
    uint32_t st[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        p[c] = cptr[(size_t)i * 8 + c];
        st[c] = cst[(size_t)i * 8 + c];
        // Pointers index the normalisation table's rows and states its 48 columns; neither can be
        // out of range when the previous level was consistent, and a violation is reported
        // instead of read.
        if ((p[c] != kEmptyPtr && p[c] >= normRows) || st[c] >= 48u) {
            atomicOr(bad, 2u);
            p[c] = kEmptyPtr;
            st[c] = 0;
        }
    }
    uint64_t best[8];
#pragma unroll
    for (int c = 0; c < 8; ++c)
        best[c] = ~0ull;
    for (uint32_t k = 0; k < numIds; ++k) {
        const uint32_t id = T.ids[k];
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const uint32_t s = T.compose[id * 48 + st[c]];
            // smallest state of s's coset under the child's invariance subgroup (per pointer)
            const uint32_t m = p[c] == kEmptyPtr ? 0u : normTab[(size_t)p[c] * 48 + s];
            row[T.childPos[id * 8 + c]] = (uint64_t)p[c] | (uint64_t)m << 40;
        }
        int order = 0; // -1: row < best, 1: row > best
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const uint64_t v = row[c];
            if (order == 0 && v != best[c])
                order = v < best[c] ? -1 : 1;
        }
        if (order < 0) {
#pragma unroll
            for (int c = 0; c < 8; ++c)
                best[c] = row[c];
        }
    }
#pragma unroll
    for (int c = 0; c < 8; ++c)
        keys[(size_t)i * 8 + c] = best[c];
}

// norm[h * 48 + s]: the smallest state of s's coset under the invariance subgroup inv[h] (the
// states t with required[s][t] in inv[h]).
__global__ __launch_bounds__(kHT) void kNormTable(const DevTables* __restrict__ gt, const uint64_t* __restrict__ inv, uint32_t count, uint8_t* __restrict__ norm)
{
    const uint64_t t = (uint64_t)blockIdx.x * kHT + threadIdx.x;
    if (t >= (uint64_t)count * 48)
        return;
    const uint32_t h = (uint32_t)(t / 48), s = (uint32_t)(t % 48);
    uint64_t bits = inv[h] & ((1ull << 48) - 1);
    uint32_t m = s;
    while (bits) {
        const uint32_t e = (uint32_t)__builtin_ctzll(bits);
        bits &= bits - 1;
        m = min(m, (uint32_t)gt->cosetOf[s * 48 + e]);
    }
    norm[t] = (uint8_t)m;
}

__device__ inline uint32_t hashKey(const uint64_t* k)
{
    uint64_t h = 0x9E3779B97F4A7C15ull;
    for (int j = 0; j < 8; ++j) {
        h ^= k[j];
        h *= 0xBF58476D1CE4E5B9ull;
        h ^= h >> 29;
    }
    return (uint32_t)(h ^ (h >> 32));
}

__device__ inline bool keyEq(const uint64_t* a, const uint64_t* b)
{
    for (int j = 0; j < 8; ++j)
        if (a[j] != b[j])
            return false;
    return true;
}

// Min-index table over the keys (slots only move to a smaller index with the same key; at most
// cap probes).
__global__ __launch_bounds__(kHT) void kInsertMin(const uint64_t* __restrict__ keys, uint32_t n, uint32_t* table, uint32_t capMask)
{
    const uint32_t i = blockIdx.x * kHT + threadIdx.x;
    if (i >= n)
        return;
    const uint64_t* k = keys + (size_t)i * 8;
    uint32_t h = hashKey(k) & capMask;
    for (uint32_t probe = 0; probe <= capMask; ++probe) {
        uint32_t cur = __hip_atomic_load(&table[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (cur == kNoSlot) {
            cur = atomicCAS(&table[h], kNoSlot, i);
            if (cur == kNoSlot)
                return;
        }
        if (keyEq(keys + (size_t)cur * 8, k)) {
            if (i < cur)
                atomicMin(&table[h], i);
            return;
        }
        h = (h + 1) & capMask;
    }
}

__global__ __launch_bounds__(kHT) void kRepOf(const uint64_t* __restrict__ keys, uint32_t n, const uint32_t* __restrict__ table, uint32_t capMask,
    uint32_t* __restrict__ rep)
{
    const uint32_t i = blockIdx.x * kHT + threadIdx.x;
    if (i >= n)
        return;
    const uint64_t* k = keys + (size_t)i * 8;
    uint32_t h = hashKey(k) & capMask;
    uint32_t r = i;
    for (uint32_t probe = 0; probe <= capMask; ++probe) {
        const uint32_t cur = table[h];
        if (cur == kNoSlot)
            break;
        if (keyEq(keys + (size_t)cur * 8, k)) {
            r = cur;
            break;
        }
        h = (h + 1) & capMask;
    }
    rep[i] = r;
}

// nodeEq(transformNode(x, id), y), child by child; y's children are read from memory at the
// landing slot.
__device__ inline bool transformedEqDev(const LdsTables& T, const uint32_t* xp, const uint64_t* xi, const uint32_t* xs, const uint32_t* __restrict__ cptr,
    const uint8_t* __restrict__ cst, uint32_t y, uint32_t id)
{
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const uint32_t pc = T.childPos[id * 8 + c];
        if (xp[c] != cptr[(size_t)y * 8 + pc])
            return false;
        const uint32_t s = T.compose[id * 48 + xs[c]];
        if (!((xi[c] >> T.required[s * 48 + cst[(size_t)y * 8 + pc]]) & 1ull))
            return false;
    }
    return true;
}

// Kept node: its invariance mask. Other node: the first allowed ID mapping it onto its class's
// kept node (0xFF and *bad set if none, which the key construction rules out).
__global__ __launch_bounds__(kHT) void kResolve(const DevTables* __restrict__ gt, const uint32_t* __restrict__ cptr, const uint64_t* __restrict__ cinv,
    const uint8_t* __restrict__ cst, uint32_t n, uint32_t numIds, const uint32_t* __restrict__ rep, uint8_t* __restrict__ hitId, uint64_t* __restrict__ nodeInv,
    uint32_t* __restrict__ bad)
{
    __shared__ LdsTables T;
    stageTables(gt, &T);
    const uint32_t i = blockIdx.x * kHT + threadIdx.x;
    if (i >= n)
        return;
    uint32_t xp[8], xs[8];
    uint64_t xi[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        xp[c] = cptr[(size_t)i * 8 + c];
        xi[c] = cinv[(size_t)i * 8 + c];
        xs[c] = cst[(size_t)i * 8 + c];
    }
    const uint32_t r = rep[i];
    if (r == i) {
        uint64_t inv = 0;
        for (uint32_t k = 0; k < numIds; ++k)
            if (transformedEqDev(T, xp, xi, xs, cptr, cst, i, T.ids[k]))
                inv |= 1ull << T.ids[k];
        nodeInv[i] = inv;
        hitId[i] = 0xFF;
        return;
    }
    for (uint32_t k = 0; k < numIds; ++k)
        if (transformedEqDev(T, xp, xi, xs, cptr, cst, r, T.ids[k])) {
            hitId[i] = T.ids[k];
            return;
        }
    hitId[i] = 0xFF;
    *bad = 1u;
}

__global__ __launch_bounds__(kHT) void kKeptSorted(const uint32_t* __restrict__ sortedIdx, const uint32_t* __restrict__ rep, uint32_t n, uint32_t* __restrict__ kf)
{
    const uint32_t j = blockIdx.x * kHT + threadIdx.x;
    if (j >= n)
        return;
    const uint32_t i = sortedIdx[j];
    kf[j] = rep[i] == i ? 1u : 0u;
}

__global__ __launch_bounds__(kHT) void kHandles(const uint32_t* __restrict__ sortedIdx, const uint32_t* __restrict__ kf, const uint32_t* __restrict__ pos, uint32_t n,
    uint32_t* __restrict__ handle)
{
    const uint32_t j = blockIdx.x * kHT + threadIdx.x;
    if (j >= n || !kf[j])
        return;
    handle[sortedIdx[j]] = pos[j];
}

__global__ __launch_bounds__(kHT) void kMapping(const DevTables* __restrict__ gt, const uint32_t* __restrict__ rep, const uint8_t* __restrict__ hitId,
    const uint64_t* __restrict__ nodeInv, const uint32_t* __restrict__ handle, uint32_t n, uint32_t* __restrict__ mptr, uint64_t* __restrict__ minv,
    uint8_t* __restrict__ mst)
{
    const uint32_t i = blockIdx.x * kHT + threadIdx.x;
    if (i >= n)
        return;
    const uint32_t r = rep[i];
    mptr[i] = handle[r];
    minv[i] = nodeInv[r];
    mst[i] = r == i ? 0 : gt->inverseId[hitId[i]];
}

__device__ inline uint64_t stateToTpDev(uint32_t ptr, uint32_t state)
{
    // kPermTab rows (0,1,2) (0,2,1) (1,0,2) (1,2,0) (2,0,1) (2,1,0), packed 2 bits per axis.
    constexpr uint32_t perm[6] = { 0u | 1u << 2 | 2u << 4, 0u | 2u << 2 | 1u << 4, 1u | 0u << 2 | 2u << 4, 1u | 2u << 2 | 0u << 4, 2u | 0u << 2 | 1u << 4,
        2u | 1u << 2 | 0u << 4 };
    uint64_t v = (uint64_t)ptr & ((1ull << 37) - 1);
    v |= (uint64_t)(state & 7u) << 37;
    v |= (uint64_t)perm[state >> 3] << 40;
    v |= 32ull << 46 | 32ull << 52 | 32ull << 58;
    return v;
}

// Kept node i is written at its handle: its 8 children as TransformPointers (inner levels) or its
// leaf mask; `size` gets 1 + its child count (the packed node size of the static form).
__global__ __launch_bounds__(kHT) void kOutput(const uint32_t* __restrict__ rep, const uint32_t* __restrict__ handle, uint32_t n, const uint32_t* __restrict__ cptr,
    const uint8_t* __restrict__ cst, const uint64_t* __restrict__ leaves, const uint64_t* __restrict__ nodeInv, uint64_t* __restrict__ outNodes,
    uint64_t* __restrict__ outLeaves, uint32_t* __restrict__ size, uint64_t* __restrict__ invByHandle)
{
    const uint32_t i = blockIdx.x * kHT + threadIdx.x;
    if (i >= n || rep[i] != i)
        return;
    const uint32_t h = handle[i];
    invByHandle[h] = nodeInv[i];
    if (leaves) {
        outLeaves[h] = leaves[i];
        return;
    }
    uint32_t cnt = 0;
    for (int c = 0; c < 8; ++c) {
        const uint32_t p = cptr[(size_t)i * 8 + c];
        outNodes[(size_t)h * 8 + c] = stateToTpDev(p, p == kEmptyPtr ? 0u : cst[(size_t)i * 8 + c]);
        cnt += p != kEmptyPtr ? 1u : 0u;
    }
    size[h] = 1 + cnt;
}

// transform_dag_hierarchical.cpp:311-318 in index order: the last node with an untransformed
// pointer to a handle is the one recorded, i.e. the largest such index.
__global__ __launch_bounds__(kHT) void kMid(const uint32_t* __restrict__ mptr, const uint8_t* __restrict__ mst, uint32_t n, uint32_t* __restrict__ mid)
{
    const uint32_t i = blockIdx.x * kHT + threadIdx.x;
    if (i >= n || mst[i] != 0)
        return;
    atomicMax(&mid[mptr[i]], i);
}

// Carves the pass's buffers out of one workspace, 256-byte aligned; with a null base it only adds
// up the size.
struct Carve {
    char* base = nullptr;
    size_t off = 0;
    template <typename T>
    T* get(size_t count)
    {
        off = (off + 255) & ~size_t(255);
        T* p = base ? (T*)(base + off) : nullptr;
        off += std::max<size_t>(count * sizeof(T), 8);
        return p;
    }
};

} // namespace

hipError_t hierarchicalDedupDevice(const Svdag& in, uint32_t midLevel, std::vector<uint32_t>& outMidInv, const tdag_config& cfg, bool staticOffsets,
    hipStream_t st, void*& work, size_t& workBytes, TDag& out, std::string* failure)
{
    out = TDag {};
    out.resolution = in.resolution;
    out.rootLevel = in.rootLevel;
    out.structureType = 0; // default-constructed output (transform_dag_hierarchical.cpp:307)
    out.nodes.resize(in.nodes.size());
    outMidInv.clear();
    HierTables H;
    hierarchicalTables(cfg, H);
    DevTables t {};
    std::memcpy(t.compose, H.compose, sizeof(t.compose));
    std::memcpy(t.childPos, H.childPos, sizeof(t.childPos));
    std::memcpy(t.required, H.required, sizeof(t.required));
    std::memcpy(t.inverseId, H.inverseId, sizeof(t.inverseId));
    std::memcpy(t.ids, H.ids, sizeof(t.ids));
    t.numIds = H.numIds;
    for (uint32_t s = 0; s < 48; ++s)
        for (uint32_t u = 0; u < 48; ++u)
            t.cosetOf[s * 48 + H.required[s][u]] = (uint8_t)u;
    size_t maxN = in.leaves.size();
    for (uint32_t l = 3; l <= in.rootLevel; ++l)
        maxN = std::max(maxN, in.nodes[l].size());
    maxN = std::max<size_t>(maxN, 1);
    uint32_t cap = 64;
    while (cap < 2 * maxN)
        cap <<= 1;
    uint32_t* lptr;
    uint64_t* linv;
    uint8_t* lst;
    uint64_t* leaves;
    uint32_t *child, *cptr, *idx, *table, *rep, *kf, *pos, *handle, *sizes, *offsets, *mid, *flags;
    uint64_t *cinv, *seed, *keys, *nodeInv, *outBuf, *invByHandle;
    uint8_t *cst, *hitId, *normTab;
    void *sortScratch, *scanScratch;
    DevTables* dT;
    uint32_t* mptr[2];
    uint64_t* minv[2];
    uint8_t* mst[2];
    auto carveAll = [&](Carve& B) {
        dT = B.get<DevTables>(1);
        lptr = B.get<uint32_t>(256);
        linv = B.get<uint64_t>(256);
        lst = B.get<uint8_t>(256);
        leaves = B.get<uint64_t>(in.leaves.size());
        child = B.get<uint32_t>(maxN * 8);
        cptr = B.get<uint32_t>(maxN * 8);
        cinv = B.get<uint64_t>(maxN * 8);
        cst = B.get<uint8_t>(maxN * 8);
        seed = B.get<uint64_t>(maxN);
        idx = B.get<uint32_t>(maxN);
        sortScratch = B.get<uint8_t>(radixScratchBytes(maxN, 8));
        keys = B.get<uint64_t>(maxN * 8);
        table = B.get<uint32_t>(cap);
        rep = B.get<uint32_t>(maxN);
        hitId = B.get<uint8_t>(maxN);
        nodeInv = B.get<uint64_t>(maxN);
        kf = B.get<uint32_t>(maxN);
        pos = B.get<uint32_t>(maxN);
        scanScratch = B.get<uint8_t>(scanScratchBytes(maxN));
        handle = B.get<uint32_t>(maxN);
        for (int q = 0; q < 2; ++q) {
            mptr[q] = B.get<uint32_t>(maxN);
            minv[q] = B.get<uint64_t>(maxN);
            mst[q] = B.get<uint8_t>(maxN);
        }
        outBuf = B.get<uint64_t>(maxN * 8);
        sizes = B.get<uint32_t>(maxN);
        offsets = B.get<uint32_t>(maxN);
        mid = B.get<uint32_t>(maxN);
        normTab = B.get<uint8_t>(std::max<size_t>(maxN, 256) * 48);
        invByHandle = B.get<uint64_t>(maxN);
        flags = B.get<uint32_t>(4); // [0] scan total, [1] inconsistent-class flag, [2] offsets total, [3] index checks
    };
    Carve sizing;
    carveAll(sizing);
    if (workBytes < sizing.off || !work) {
        if (work) {
            if (hipError_t e0 = hipStreamSynchronize(st); e0 != hipSuccess)
                return e0;
            (void)hipFree(work);
            work = nullptr;
            workBytes = 0;
        }
        if (hipError_t e0 = hipMalloc(&work, sizing.off); e0 != hipSuccess)
            return e0;
        workBytes = sizing.off;
    }
    Carve B { (char*)work, 0 };
    carveAll(B);
    hipError_t e;
#define HD_TRY(x)                                                                                                                                               \
    do {                                                                                                                                                        \
        if ((e = (x)) != hipSuccess)                                                                                                                            \
            return e;                                                                                                                                           \
    } while (0)
    HD_TRY(hipMemcpyAsync(dT, &t, sizeof(t), hipMemcpyHostToDevice, st));
    HD_TRY(hipMemcpyAsync(lptr, H.leafPtr, sizeof(H.leafPtr), hipMemcpyHostToDevice, st));
    HD_TRY(hipMemcpyAsync(linv, H.leafInv, sizeof(H.leafInv), hipMemcpyHostToDevice, st));
    HD_TRY(hipMemcpyAsync(lst, H.leafState, sizeof(H.leafState), hipMemcpyHostToDevice, st));
    if (!in.leaves.empty())
        HD_TRY(hipMemcpyAsync(leaves, in.leaves.data(), in.leaves.size() * 8, hipMemcpyHostToDevice, st));
    HD_TRY(fillU32(flags, 4, 0u, st));
    // Leaf children point at canonical 2x2x2 bytes: their normalisation table comes from the
    // 256-entry table's invariance masks.
    std::vector<uint8_t> leafNorm(256 * 48);
    for (uint32_t b = 0; b < 256; ++b)
        for (uint32_t s0 = 0; s0 < 48; ++s0) {
            uint32_t m = s0;
            for (uint32_t e0 = 0; e0 < 48; ++e0)
                if ((H.leafInv[b] >> e0) & 1ull)
                    m = std::min<uint32_t>(m, t.cosetOf[s0 * 48 + e0]);
            leafNorm[b * 48 + s0] = (uint8_t)m;
        }
    HD_TRY(hipMemcpyAsync(normTab, leafNorm.data(), leafNorm.size(), hipMemcpyHostToDevice, st));
    bool haveOffsets = false; // offsets of the previous level's output nodes are valid
    uint32_t normRows = 256; // rows of normTab: the 256 leaf bytes, then the previous level's kept nodes
    uint32_t prevN = 0; // items of the previous level (the range of this level's child indices)
    int cur = 0; // mapping buffers of the level just processed
    for (uint32_t level = 2; level <= std::max(2u, in.rootLevel); ++level) {
        const uint32_t n = (uint32_t)(level == 2 ? in.leaves.size() : in.nodes[level].size());
        const int nxt = cur ^ 1;
        uint32_t nk = 0;
        if (n > 0) {
            if (level == 2) {
                hipLaunchKernelGGL(kLeafNodes, dim3(gridOf((uint64_t)n * 8)), dim3(kHT), 0, st, leaves, n, lptr, linv, lst, cptr, cinv, cst);
            } else {
                HD_TRY(hipMemcpyAsync(child, in.nodes[level].data(), (size_t)n * 32, hipMemcpyHostToDevice, st));
                hipLaunchKernelGGL(kInnerNodes, dim3(gridOf((uint64_t)n * 8)), dim3(kHT), 0, st, child, n, prevN, mptr[cur], minv[cur], mst[cur], cptr, cinv, cst,
                    flags + 3);
            }
            const uint32_t* hashOf = (staticOffsets && level >= 4 && haveOffsets) ? offsets : nullptr;
            hipLaunchKernelGGL(kSeeds, dim3(gridOf(n)), dim3(kHT), 0, st, cptr, n, hashOf, seed, idx);
            HD_TRY(hipGetLastError());
            HD_TRY(radixSortPairsU64(seed, idx, n, 0, 64, sortScratch, st));
            hipLaunchKernelGGL(kCanonKeys, dim3(gridOf(n)), dim3(kHT), 0, st, dT, cptr, normTab, normRows, cst, n, t.numIds, keys, flags + 3);
            uint32_t lcap = 64;
            while (lcap < 2 * (uint64_t)n)
                lcap <<= 1;
            HD_TRY(fillU32(table, lcap, kNoSlot, st));
            hipLaunchKernelGGL(kInsertMin, dim3(gridOf(n)), dim3(kHT), 0, st, keys, n, table, lcap - 1);
            hipLaunchKernelGGL(kRepOf, dim3(gridOf(n)), dim3(kHT), 0, st, keys, n, table, lcap - 1, rep);
            hipLaunchKernelGGL(kResolve, dim3(gridOf(n)), dim3(kHT), 0, st, dT, cptr, cinv, cst, n, t.numIds, rep, hitId, nodeInv, flags + 1);
            hipLaunchKernelGGL(kKeptSorted, dim3(gridOf(n)), dim3(kHT), 0, st, idx, rep, n, kf);
            HD_TRY(hipGetLastError());
            HD_TRY(exclusiveScanU32(kf, pos, n, flags, scanScratch, st));
            hipLaunchKernelGGL(kHandles, dim3(gridOf(n)), dim3(kHT), 0, st, idx, kf, pos, n, handle);
            hipLaunchKernelGGL(kMapping, dim3(gridOf(n)), dim3(kHT), 0, st, dT, rep, hitId, nodeInv, handle, n, mptr[nxt], minv[nxt], mst[nxt]);
            hipLaunchKernelGGL(kOutput, dim3(gridOf(n)), dim3(kHT), 0, st, rep, handle, n, cptr, cst, level == 2 ? leaves : nullptr, nodeInv, outBuf, outBuf, sizes,
                invByHandle);
            HD_TRY(hipGetLastError());
            uint32_t hk[4] = { 0, 0, 0, 0 };
            HD_TRY(hipMemcpyAsync(hk, flags, 16, hipMemcpyDeviceToHost, st));
            HD_TRY(hipStreamSynchronize(st));
            if (hk[3]) {
                if (failure)
                    *failure = "a child index or state outside its table at level " + std::to_string(level);
                return hipErrorInvalidValue;
            }
            if (hk[1]) {
                // A node the keys put in a class whose kept node no allowed transform maps it onto:
                // the keys contradict the reference's equality (an internal invariant).
                if (failure)
                    *failure = "inconsistent canonical class at level " + std::to_string(level);
                return hipErrorInvalidValue;
            }
            nk = hk[0];
            // The next level's children point at these nk kept nodes.
            hipLaunchKernelGGL(kNormTable, dim3(gridOf((uint64_t)nk * 48)), dim3(kHT), 0, st, dT, invByHandle, nk, normTab);
            normRows = nk;
            HD_TRY(hipGetLastError());
            if (level == 2) {
                out.leaves.resize(nk);
                HD_TRY(hipMemcpyAsync(out.leaves.data(), outBuf, (size_t)nk * 8, hipMemcpyDeviceToHost, st));
            } else {
                out.nodes[level].resize(nk);
                HD_TRY(hipMemcpyAsync(out.nodes[level].data(), outBuf, (size_t)nk * 64, hipMemcpyDeviceToHost, st));
                if (staticOffsets)
                    HD_TRY(exclusiveScanU32(sizes, offsets, nk, flags + 2, scanScratch, st));
            }
            if (level == midLevel) {
                HD_TRY(fillU32(mid, std::max<uint32_t>(nk, 1), 0u, st));
                hipLaunchKernelGGL(kMid, dim3(gridOf(n)), dim3(kHT), 0, st, mptr[nxt], mst[nxt], n, mid);
                HD_TRY(hipGetLastError());
                outMidInv.resize(nk);
                if (nk)
                    HD_TRY(hipMemcpyAsync(outMidInv.data(), mid, (size_t)nk * 4, hipMemcpyDeviceToHost, st));
            }
            HD_TRY(hipStreamSynchronize(st));
        } else if (level == midLevel) {
            outMidInv.clear();
        }
        haveOffsets = staticOffsets && level >= 3 && n > 0;
        prevN = n;
        cur = nxt;
        if (level >= in.rootLevel)
            break;
    }
#undef HD_TRY
    return hipSuccess;
}

} // namespace tdag
