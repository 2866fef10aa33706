// constructTransformDAG driver and the pipeline entry points of the C-ABI (host code).
//
// Driver semantics (transform_dag.cpp:138-240): the hierarchical pass runs with
// midLevel = maxTranslationLevel + 1; levels <= maxTranslationLevel are rebuilt top-down, each
// from the original nodes referenced by the surviving parents, with the GPU matcher deciding
// which items survive and how every parent pointer is rewritten.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <memory>
#include <string>
#include <thread>

#include "context.hpp"
#include "device_build.hpp"
#include "errors.hpp"
#include "pipeline.hpp"
#include "primitives.hpp"
#include "tdag.h"

struct tdag_svdag {
    tdag::Svdag s;
    // Device copy of levels 2..5 for tdag_svdag_flatten_device and the driver, made on first use.
    tdag::DeviceSvdag dev;
    int devDevice = -1;
    ~tdag_svdag()
    {
        if (devDevice < 0)
            return;
        int current = -1;
        const bool haveCurrent = hipGetDevice(&current) == hipSuccess;
        if (hipSetDevice(devDevice) == hipSuccess)
            tdag::freeDeviceSvdag(dev);
        if (haveCurrent)
            (void)hipSetDevice(current); // leave the caller's current device as it was
    }
};

struct tdag_tdag {
    tdag::TDag t;
    std::vector<tdag::LevelRecord> records;
    bool isStatic = false; // built by constructStaticTransformDAG: read back in the packed form
    tdag_config cfg {}; // the TransformDAGConfig it was built with
};

namespace {
// The serializers and the encoder take the config again (as the reference's writers do); it must
// be the one the structure was built with, or the pointer types they write would not describe the
// pointers (e.g. translation bits for a level the build did not translate).
bool sameConfig(const tdag_config& a, const tdag_config& b)
{
    return a.symmetry == b.symmetry && a.axis_permutation == b.axis_permutation && a.translation == b.translation &&
        a.max_translation_level == b.max_translation_level;
}
} // namespace

namespace {

int hostThreads()
{
    const unsigned hc = std::thread::hardware_concurrency();
    return (int)std::max(1u, std::min(16u, hc ? hc : 1u));
}

std::array<uint64_t, 8> toTransformNode(const std::array<uint32_t, 8>& n) // transform_dag.cpp:123-135
{
    std::array<uint64_t, 8> o;
    for (int c = 0; c < 8; ++c)
        o[c] = n[c] == tdag::kEmptyChild ? 0x82082400FFFFFFFFull : (0x8208240000000000ull | n[c]);
    return o;
}

tdag_status setErr(tdag_status s, const std::string& m) { return tdag::setError(s, m); }

tdag_status setHipErr(hipError_t e, const std::string& what)
{
    return tdag::setError(e == hipErrorOutOfMemory ? TDAG_ENOMEM : TDAG_EDEVICE, what + ": " + hipGetErrorName(e) + " (" + hipGetErrorString(e) + ")");
}

// The SVDAG's device copy on ctx's device (levels 2..min(5, root)).
tdag_status ensureDeviceCopy(tdag_context* ctx, const tdag_svdag* h)
{
    tdag_svdag* m = const_cast<tdag_svdag*>(h);
    if (m->devDevice == ctx->device)
        return TDAG_OK;
    if (m->devDevice >= 0) {
        if (hipError_t e = hipSetDevice(m->devDevice); e == hipSuccess)
            tdag::freeDeviceSvdag(m->dev);
        m->devDevice = -1;
    }
    if (hipError_t e = hipSetDevice(ctx->device); e != hipSuccess)
        return setHipErr(e, "hipSetDevice");
    if (hipError_t e = tdag::uploadSvdag(m->s, 5, m->dev); e != hipSuccess) {
        tdag::freeDeviceSvdag(m->dev);
        return setHipErr(e, "SVDAG upload");
    }
    m->devDevice = ctx->device;
    return TDAG_OK;
}

inline size_t align256(size_t v) { return (v + 255) & ~size_t(255); }

} // namespace

extern "C" {

tdag_status tdag_scene_dense(uint32_t kind, uint32_t resolution, uint64_t seed, uint8_t* out_bits, uint64_t out_bytes)
{
    if (!out_bits)
        return setErr(TDAG_EINVAL, "out_bits is null");
    if (resolution < 4 || (resolution & (resolution - 1)) || resolution > 4096)
        return setErr(TDAG_EINVAL, "resolution must be a power of two in [4, 4096]");
    if (out_bytes < (uint64_t)resolution * resolution * resolution / 8)
        return setErr(TDAG_EINVAL, "out_bytes is smaller than resolution^3 / 8");
    if (!tdag::generateSceneDense(kind, resolution, seed, out_bits, hostThreads()))
        return setErr(TDAG_EINVAL, "unknown scene kind " + std::to_string(kind));
    return TDAG_OK;
}

tdag_status tdag_svdag_from_dense(const uint8_t* bits, uint32_t resolution, tdag_svdag** out)
{
    if (!bits || !out)
        return setErr(TDAG_EINVAL, "null argument");
    if (resolution < 4 || (resolution & (resolution - 1)) || resolution > 4096)
        return setErr(TDAG_EINVAL, "resolution must be a power of two in [4, 4096]");
    tdag_svdag* h = new tdag_svdag();
    h->s = tdag::svdagFromDense(bits, resolution, hostThreads());
    *out = h;
    return TDAG_OK;
}

tdag_status tdag_svdag_from_dense_device(tdag_context* ctx, const uint8_t* d_bits, uint32_t resolution, tdag_svdag** out, float* out_ms)
{
    if (!ctx || !d_bits || !out)
        return setErr(TDAG_EINVAL, "null argument");
    if (resolution < 4 || (resolution & (resolution - 1)) || resolution > 4096)
        return setErr(TDAG_EINVAL, "resolution must be a power of two in [4, 4096]");
    if (hipError_t e = hipSetDevice(ctx->device); e != hipSuccess)
        return setErr(TDAG_EDEVICE, std::string("hipSetDevice: ") + hipGetErrorName(e));
    const auto t0 = std::chrono::steady_clock::now();
    tdag_svdag* h = new tdag_svdag();
    if (hipError_t e = tdag::svdagFromDenseDevice(d_bits, resolution, ctx->stream, ctx->svdagScratch, ctx->svdagScratchBytes, h->s); e != hipSuccess) {
        delete h;
        return setErr(e == hipErrorOutOfMemory ? TDAG_ENOMEM : TDAG_EDEVICE, std::string("device SVDAG build: ") + hipGetErrorName(e) + " (" + hipGetErrorString(e) + ")");
    }
    if (out_ms)
        *out_ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
    *out = h;
    return TDAG_OK;
}

tdag_status tdag_svdag_from_view(const tdag_svdag_view* v, tdag_svdag** out)
{
    if (!v || !out)
        return setErr(TDAG_EINVAL, "null argument");
    *out = nullptr;
    if (v->root_level < 3 || v->root_level > 12 || v->resolution != (1u << v->root_level))
        return setErr(TDAG_EINVAL, "resolution must be 2^root_level with root_level in [3, 12]");
    if (v->structure_type != 0 && v->structure_type != 1)
        return setErr(TDAG_EINVAL, "structure_type " + std::to_string(v->structure_type) + " is neither Tree (0) nor DAG (1)");
    if (!v->nodes_per_level || !v->node_counts)
        return setErr(TDAG_EINVAL, "nodes_per_level / node_counts are null");
    if (v->num_subgrids && !v->subgrids)
        return setErr(TDAG_EINVAL, "subgrids is null");
    if (v->num_subgrids >= tdag::kEmptyChild)
        return setErr(TDAG_EINVAL, "too many subgrids for 32-bit child indices");
    std::unique_ptr<tdag_svdag> h(new tdag_svdag());
    tdag::Svdag& s = h->s;
    s.resolution = v->resolution;
    s.rootLevel = v->root_level;
    s.structureType = v->structure_type;
    s.nodes.resize(v->root_level + 1);
    s.leaves.assign(v->subgrids, v->subgrids + v->num_subgrids);
    uint64_t below = v->num_subgrids;
    for (uint32_t level = 3; level <= v->root_level; ++level) {
        const uint64_t count = v->node_counts[level];
        const uint32_t* nodes = v->nodes_per_level[level];
        if (count && !nodes)
            return setErr(TDAG_EINVAL, "level " + std::to_string(level) + " has " + std::to_string(count) + " nodes but a null array");
        if (count >= tdag::kEmptyChild)
            return setErr(TDAG_EINVAL, "level " + std::to_string(level) + " has too many nodes for 32-bit child indices");
        auto& dst = s.nodes[level];
        dst.resize(count);
        for (uint64_t i = 0; i < count; ++i)
            for (int c = 0; c < 8; ++c) {
                const uint32_t child = nodes[i * 8 + c];
                if (child != tdag::kEmptyChild && child >= below)
                    return setErr(TDAG_EINVAL, "level " + std::to_string(level) + " node " + std::to_string(i) + " child " + std::to_string(c) + ": index " +
                            std::to_string(child) + " outside level " + std::to_string(level - 1) + " (" + std::to_string(below) + " entries)");
                dst[i][c] = child;
            }
        below = count;
    }
    *out = h.release();
    return TDAG_OK;
}

tdag_status tdag_svdag_info(const tdag_svdag* h, uint32_t* resolution, uint32_t* root_level, uint64_t* level_sizes, uint32_t cap)
{
    if (!h)
        return setErr(TDAG_EINVAL, "null handle");
    if (resolution)
        *resolution = h->s.resolution;
    if (root_level)
        *root_level = h->s.rootLevel;
    if (level_sizes)
        for (uint32_t l = 0; l < cap; ++l)
            level_sizes[l] = l == 2 ? h->s.leaves.size() : (l < h->s.nodes.size() ? h->s.nodes[l].size() : 0);
    return TDAG_OK;
}

tdag_status tdag_svdag_copy_level(const tdag_svdag* h, uint32_t level, void* out)
{
    if (!h || !out)
        return setErr(TDAG_EINVAL, "null argument");
    if (level == 2) {
        std::memcpy(out, h->s.leaves.data(), h->s.leaves.size() * 8);
        return TDAG_OK;
    }
    if (level < 3 || level > h->s.rootLevel)
        return setErr(TDAG_EINVAL, "level out of range");
    const auto& n = h->s.nodes[level];
    std::memcpy(out, n.data(), n.size() * 32);
    return TDAG_OK;
}

tdag_status tdag_svdag_flatten(const tdag_svdag* h, uint32_t level, const uint32_t* nodes, uint64_t n, uint64_t* out_masks)
{
    if (!h || (n && (!nodes || !out_masks)))
        return setErr(TDAG_EINVAL, "null argument");
    if (level < 2 || level > h->s.rootLevel || level > 5)
        return setErr(TDAG_EINVAL, "level must be in [2, min(5, rootLevel)]");
    const uint64_t size = level == 2 ? h->s.leaves.size() : h->s.nodes[level].size();
    for (uint64_t i = 0; i < n; ++i)
        if (nodes[i] >= size)
            return setErr(TDAG_EINVAL, "node index out of range");
    tdag::flattenNodes(h->s, level, nodes, n, out_masks, hostThreads());
    return TDAG_OK;
}

tdag_status tdag_svdag_flatten_device(tdag_context* ctx, const tdag_svdag* h, uint32_t level, const uint32_t* d_nodes, uint64_t n, uint64_t* d_out_masks)
{
    if (!ctx || !h || (n && (!d_nodes || !d_out_masks)))
        return setErr(TDAG_EINVAL, "null argument");
    if (level < 2 || level > h->s.rootLevel || level > 5)
        return setErr(TDAG_EINVAL, "level must be in [2, min(5, rootLevel)]");
    if (n == 0)
        return TDAG_OK;
    if (tdag_status s = ensureDeviceCopy(ctx, h); s != TDAG_OK)
        return s;
    const size_t temp = align256(tdag::flattenTempBytes(level, n));
    if (tdag_status s = tdag::capi::ensureBuffer(ctx->svdagScratch, ctx->svdagScratchBytes, temp + 256, "flatten scratch"); s != TDAG_OK)
        return s;
    uint64_t* d_temp = (uint64_t*)ctx->svdagScratch;
    uint32_t* d_bad = (uint32_t*)((char*)ctx->svdagScratch + temp);
    hipError_t e = tdag::fillU32(d_bad, 1, 0u, ctx->stream);
    if (e == hipSuccess)
        e = tdag::flattenDevice(h->dev, level, d_nodes, n, d_out_masks, d_temp, ctx->stream, d_bad);
    uint32_t bad = 0;
    if (e == hipSuccess)
        e = hipMemcpyAsync(&bad, d_bad, 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess)
        return setHipErr(e, "device flatten");
    if (bad)
        return setErr(TDAG_EINVAL, "node index out of range");
    return TDAG_OK;
}

tdag_status tdag_svdag_free(tdag_svdag* h)
{
    delete h;
    return TDAG_OK;
}

// One exhaustive matcher level of the driver with its input flattened on the device: `items` are
// the referenced level-`level` nodes in ascending order. Its device buffers live in the context's
// staging buffer. The matcher's mapping and unique list come back to the host; the masks only when
// they are kept.
static tdag_status matchLevelOnDevice(tdag_context* ctx, const tdag_svdag* in, uint32_t level, const tdag_config* cfg, const std::vector<uint32_t>& items,
    uint64_t W, bool keepMasks, std::vector<uint64_t>& outMasks, std::vector<uint64_t>& mapping, std::vector<uint32_t>& uniq, uint64_t& nu,
    tdag_match_stats& st, tdag_mlp* mlp = nullptr, uint32_t k = 0, tdag_learned_stats* ls = nullptr)
{
    const uint64_t n = items.size();
    // With mlp: the learned level-2 matcher (proposals, verification, exhaustive fallback).
    auto match = [&](const uint64_t* d_masks, uint64_t* d_map, uint32_t* d_uniq) -> tdag_status {
        if (!mlp)
            return tdag_match_level_device(ctx, level, d_masks, n, cfg, nullptr, d_map, d_uniq, &nu, nullptr, &st);
        const tdag_status s = tdag_match_level_learned_device(ctx, mlp, level, d_masks, n, cfg, k, nullptr, d_map, d_uniq, ls);
        nu = ls->num_unique;
        st.num_matched = ls->num_matched;
        st.ms_total = ls->ms_total;
        return s;
    };
    if (n == 0)
        return match(nullptr, nullptr, nullptr);
    if (tdag_status s = ensureDeviceCopy(ctx, in); s != TDAG_OK)
        return s;
    const size_t oMasks = align256(n * 4), oMap = align256(oMasks + n * W * 8), oUniq = align256(oMap + n * 8), oTemp = align256(oUniq + n * 4);
    const size_t total = oTemp + align256(tdag::flattenTempBytes(level, n));
    if (tdag_status s = tdag::capi::ensureBuffer(ctx->io, ctx->ioBytes, total, "driver level buffers"); s != TDAG_OK)
        return s;
    char* b = (char*)ctx->io;
    uint32_t* d_items = (uint32_t*)b;
    uint64_t* d_masks = (uint64_t*)(b + oMasks);
    uint64_t* d_map = (uint64_t*)(b + oMap);
    uint32_t* d_uniq = (uint32_t*)(b + oUniq);
    hipError_t e = hipMemcpyAsync(d_items, items.data(), n * 4, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess)
        e = tdag::flattenDevice(in->dev, level, d_items, n, d_masks, (uint64_t*)(b + oTemp), ctx->stream);
    if (e != hipSuccess)
        return setHipErr(e, "device flatten");
    if (tdag_status s = match(d_masks, d_map, d_uniq); s != TDAG_OK)
        return s;
    mapping.resize(n);
    uniq.resize(n);
    e = hipMemcpyAsync(mapping.data(), d_map, n * 8, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(uniq.data(), d_uniq, nu * 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess && keepMasks) {
        outMasks.resize(n * W);
        e = hipMemcpyAsync(outMasks.data(), d_masks, n * W * 8, hipMemcpyDeviceToHost, ctx->stream);
    }
    if (e == hipSuccess)
        e = hipStreamSynchronize(ctx->stream);
    return e == hipSuccess ? TDAG_OK : setHipErr(e, "driver readback");
}

// The context's matcher tuning per level, for the device translation stage.
static void contextTuning(const tdag_context* ctx, tdag::LevelTuning* tune)
{
    for (uint32_t l = 0; l < 6; ++l) {
        tune[l].variant = ctx->variant[l];
        tune[l].firstBatch = ctx->firstBatch[l];
        tune[l].batchGrowth = ctx->batchGrowth[l];
        tune[l].bloomBits = ctx->bloomBits[l];
        tune[l].childFrames = ctx->childFrames[l];
    }
}

// The level maxT + 1 parents of the translation stage: the hierarchical pass's survivors as
// TransformNodes over the input level below (transform_dag.cpp:165-170).
static std::vector<std::array<uint64_t, 8>> midParents(const tdag::Svdag& dag, uint32_t maxT, const std::vector<uint32_t>& translationParents)
{
    std::vector<std::array<uint64_t, 8>> mid(translationParents.size());
    for (size_t i = 0; i < translationParents.size(); ++i)
        mid[i] = toTransformNode(dag.nodes[maxT + 1][translationParents[i]]);
    return mid;
}

static tdag_status translationRecords(const tdag::DeviceTranslation& dt, uint32_t maxT, std::vector<tdag::LevelRecord>& records)
{
    for (uint32_t level = maxT; level >= 2; --level) {
        tdag::LevelRecord rec;
        rec.level = level;
        rec.items = dt.n[level];
        rec.unique = dt.nu[level];
        rec.matched = dt.matched[level];
        records.push_back(std::move(rec));
        if (level == 2)
            break;
    }
    return TDAG_OK;
}

// The translation stage with every step on the device (csrc/device_build.hip), run once.
static tdag_status deviceTranslation(tdag_context* ctx, const tdag_svdag* in, const tdag_config& cfg, const std::vector<uint32_t>& translationParents,
    tdag::TDag& t, std::vector<tdag::LevelRecord>& records)
{
    const uint32_t maxT = cfg.max_translation_level;
    if (t.nodes[maxT + 1].size() != translationParents.size())
        return setErr(TDAG_EINTERNAL, "mid level size mismatch");
    if (tdag_status s = ensureDeviceCopy(ctx, in); s != TDAG_OK)
        return s;
    tdag::LevelTuning tune[6];
    contextTuning(ctx, tune);
    tdag::DeviceTranslation dt;
    hipError_t e = dt.init(in->dev, in->s, midParents(in->s, maxT, translationParents), maxT, tune);
    if (e == hipSuccess)
        e = dt.runEager(tdag::capi::cfgBitsOf(&cfg), ctx->stream);
    if (e == hipErrorInvalidValue)
        return setErr(TDAG_EINVAL, "device translation stage: a child pointer or item outside its level");
    if (e == hipSuccess)
        e = dt.readResult(t, ctx->stream);
    if (e != hipSuccess)
        return setHipErr(e, "device translation stage");
    return translationRecords(dt, maxT, records);
}

// Both drivers: constructTransformDAG (transform_dag.cpp:138-240) and, with staticForm,
// constructStaticTransformDAG (transform_dag.cpp:242-390). The static driver makes the same
// matcher calls on the same items: its hierarchical pass keeps the same nodes (only the order of
// levels >= 4 differs) and it re-reads the same referenced nodes level by level in ascending
// order; its packed output is produced from this node-index result on demand.
static tdag_status constructTransformDag(tdag_context* ctx, const tdag_svdag* in, const tdag_config* cfg, uint32_t flags, bool learned, tdag_mlp* mlp, uint32_t k,
    bool staticForm, tdag_tdag** out)
{
    if (!in || !cfg || !out)
        return setErr(TDAG_EINVAL, "null argument");
    if (cfg->translation && !ctx)
        return setErr(TDAG_EINVAL, "a context is required when translation is enabled");
    if (flags & ~(TDAG_KEEP_LEVEL_INPUTS | TDAG_DEVICE_TRANSLATION))
        return setErr(TDAG_EINVAL, "unknown flag");
    if ((flags & TDAG_DEVICE_TRANSLATION) && ((flags & TDAG_KEEP_LEVEL_INPUTS) || learned || !ctx))
        return setErr(TDAG_EINVAL, "the device translation stage needs a context and keeps no level inputs, and the learned driver has its own level 2");
    const tdag::Svdag& dag = in->s;
    if (dag.structureType != 1)
        return setErr(TDAG_EINVAL, "input must be a DAG (transform_dag.cpp:141)");
    const uint32_t maxT = cfg->max_translation_level;
    if (maxT < 2 || maxT > 5)
        return setErr(TDAG_EINVAL, "max_translation_level " + std::to_string(maxT) + " outside [2, 5]");
    if (maxT + 1 > dag.rootLevel)
        return setErr(TDAG_EINVAL, "max_translation_level + 1 exceeds the root level");
    const int threads = hostThreads();
    tdag_tdag* res = new tdag_tdag();
    res->isStatic = staticForm;
    res->cfg = *cfg;
    std::vector<uint32_t> translationParents;
    if (ctx) {
        // The hierarchical pass on the device (same output as the host pass).
        if (hipError_t e = hipSetDevice(ctx->device); e != hipSuccess) {
            delete res;
            return setHipErr(e, "hipSetDevice");
        }
        std::string failure;
        if (hipError_t e = tdag::hierarchicalDedupDevice(dag, maxT + 1, translationParents, *cfg, staticForm, ctx->stream, ctx->svdagScratch, ctx->svdagScratchBytes, res->t,
                &failure);
            e != hipSuccess) {
            delete res;
            if (!failure.empty())
                return setErr(TDAG_EINTERNAL, "device hierarchical pass: " + failure);
            return setHipErr(e, "device hierarchical pass");
        }
    } else {
        res->t = tdag::hierarchicalDedup(dag, maxT + 1, translationParents, *cfg, threads, staticForm);
    }
    if (cfg->translation && (flags & TDAG_DEVICE_TRANSLATION)) {
        // Every step of the translation stage on the device; only counts come back per level.
        tdag::TDag& t = res->t;
        if (tdag_status s = deviceTranslation(ctx, in, *cfg, translationParents, t, res->records); s != TDAG_OK) {
            delete res;
            return s;
        }
        *out = res;
        return TDAG_OK;
    }
    if (cfg->translation) {
        tdag::TDag& t = res->t;
        t.leaves.clear();
        for (uint32_t level = 3; level <= maxT; ++level)
            t.nodes[level].clear();
        if (t.nodes[maxT + 1].size() != translationParents.size()) {
            delete res;
            return setErr(TDAG_EINTERNAL, "mid level size mismatch");
        }
        for (size_t i = 0; i < translationParents.size(); ++i)
            t.nodes[maxT + 1][i] = toTransformNode(dag.nodes[maxT + 1][translationParents[i]]);
        for (uint32_t level = maxT; level >= 2; --level) {
            const size_t numItems = level == 2 ? dag.leaves.size() : dag.nodes[level].size();
            std::vector<uint8_t> referenced(numItems, 0);
            auto& parents = t.nodes[level + 1];
            for (const auto& node : parents)
                for (uint64_t c : node)
                    if (c != 0x82082400FFFFFFFFull)
                        referenced[c & ((1ull << 37) - 1)] = 1;
            std::vector<uint32_t> indices;
            for (uint32_t i = 0; i < numItems; ++i)
                if (referenced[i])
                    indices.push_back(i);
            const uint32_t R = 1u << level;
            const uint64_t W = std::max<uint64_t>(1, (uint64_t)R * R * R / 64);
            tdag::LevelRecord rec;
            rec.level = level;
            rec.items = indices.size();
            std::vector<uint64_t> mapping(indices.size());
            std::vector<uint32_t> uniq(indices.size());
            uint64_t nu = 0;
            tdag_match_stats st {};
            tdag_status s;
            // Flatten on the device straight into the matcher's input (FlatGridGenerator,
            // transform_dag.cpp:196-204), match there (the learned matcher at level 2 of the
            // learned driver), read back mapping and uniques.
            const bool useLearned = learned && level == 2;
            tdag_learned_stats ls {};
            s = matchLevelOnDevice(ctx, in, level, cfg, indices, W, (flags & TDAG_KEEP_LEVEL_INPUTS) != 0, rec.masks, mapping, uniq, nu, st,
                useLearned ? mlp : nullptr, k, &ls);
            if (useLearned) {
                rec.learned = true;
                rec.proposalHits = ls.proposal_hits;
                rec.fallback = ls.fallback_queries;
            }
            if (s != TDAG_OK) {
                delete res;
                char msg[512];
                tdag_last_error(msg, sizeof(msg));
                return setErr(s, "matcher failed at level " + std::to_string(level) + ": " + msg);
            }
            rec.unique = nu;
            rec.matched = st.num_matched;
            rec.msMatch = st.ms_total;
            if (!(flags & TDAG_KEEP_LEVEL_INPUTS))
                std::vector<uint64_t>().swap(rec.masks);
            if (level == 2) {
                t.leaves.resize(nu);
                for (uint64_t ui = 0; ui < nu; ++ui)
                    t.leaves[ui] = dag.leaves[indices[uniq[ui]]];
            } else {
                t.nodes[level].resize(nu);
                for (uint64_t ui = 0; ui < nu; ++ui)
                    t.nodes[level][ui] = toTransformNode(dag.nodes[level][indices[uniq[ui]]]);
            }
            for (auto& node : parents)
                for (uint64_t& c : node) {
                    if (c == 0x82082400FFFFFFFFull)
                        continue;
                    const uint32_t ptr = (uint32_t)(c & ((1ull << 37) - 1));
                    const auto it = std::lower_bound(indices.begin(), indices.end(), ptr);
                    c = mapping[it - indices.begin()];
                }
            res->records.push_back(std::move(rec));
            if (level == 2)
                break;
        }
    }
    *out = res;
    return TDAG_OK;
}

tdag_status tdag_construct_transform_dag(tdag_context* ctx, const tdag_svdag* in, const tdag_config* cfg, uint32_t flags, tdag_tdag** out)
{
    return constructTransformDag(ctx, in, cfg, flags, false, nullptr, 0, false, out);
}

} // extern "C"

// A build's translation stage captured as one hipGraph (tdag_build_graph_*): the hierarchical
// pass's result (the levels above the translated ones) and the device stage that replays the rest.
struct tdag_build_graph {
    tdag_context* ctx = nullptr;
    tdag_config cfg {};
    tdag::TDag base; // hierarchical result; its translated levels are replaced on readout
    tdag::DeviceSvdag dev {}; // the graph's own device copy of the input levels it reads
    bool devOwned = false;
    tdag::DeviceTranslation dt;
    uint32_t cfgBits = 0;
    ~tdag_build_graph()
    {
        dt.release();
        if (devOwned)
            tdag::freeDeviceSvdag(dev);
    }
};

extern "C" {

tdag_status tdag_build_graph_create(tdag_context* ctx, const tdag_svdag* in, const tdag_config* cfg, tdag_build_graph** out)
{
    if (!ctx || !in || !cfg || !out)
        return setErr(TDAG_EINVAL, "null argument");
    *out = nullptr;
    if (!cfg->translation)
        return setErr(TDAG_EINVAL, "the build graph replays the translation stage: translation must be on");
    const tdag::Svdag& dag = in->s;
    if (dag.structureType != 1)
        return setErr(TDAG_EINVAL, "input must be a DAG (transform_dag.cpp:141)");
    const uint32_t maxT = cfg->max_translation_level;
    if (maxT < 2 || maxT > 5 || maxT + 1 > dag.rootLevel)
        return setErr(TDAG_EINVAL, "max_translation_level outside [2, min(5, root level - 1)]");
    if (hipError_t e = hipSetDevice(ctx->device); e != hipSuccess)
        return setHipErr(e, "hipSetDevice");
    std::unique_ptr<tdag_build_graph> g(new tdag_build_graph());
    g->ctx = ctx;
    g->cfg = *cfg;
    g->cfgBits = tdag::capi::cfgBitsOf(cfg);
    std::vector<uint32_t> translationParents;
    std::string failure;
    if (hipError_t e = tdag::hierarchicalDedupDevice(dag, maxT + 1, translationParents, *cfg, false, ctx->stream, ctx->svdagScratch, ctx->svdagScratchBytes, g->base,
            &failure);
        e != hipSuccess)
        return failure.empty() ? setHipErr(e, "device hierarchical pass") : setErr(TDAG_EINTERNAL, "device hierarchical pass: " + failure);
    if (g->base.nodes[maxT + 1].size() != translationParents.size())
        return setErr(TDAG_EINTERNAL, "mid level size mismatch");
    // The graph reads its own copy of the SVDAG's levels, so it does not depend on the handle.
    if (hipError_t e = tdag::uploadSvdag(dag, 5, g->dev); e != hipSuccess) {
        tdag::freeDeviceSvdag(g->dev);
        return setHipErr(e, "SVDAG upload");
    }
    g->devOwned = true;
    tdag::LevelTuning tune[6];
    contextTuning(ctx, tune);
    hipError_t e = g->dt.init(g->dev, dag, midParents(dag, maxT, translationParents), maxT, tune);
    if (e == hipSuccess)
        e = g->dt.runEager(g->cfgBits, ctx->stream);
    if (e == hipErrorInvalidValue)
        return setErr(TDAG_EINVAL, "device translation stage: a child pointer or item outside its level");
    if (e == hipSuccess)
        e = g->dt.capture(g->cfgBits, ctx->stream);
    if (e != hipSuccess)
        return setHipErr(e, "device translation stage capture");
    *out = g.release();
    return TDAG_OK;
}

tdag_status tdag_build_graph_run(tdag_build_graph* g, float* out_ms)
{
    if (!g)
        return setErr(TDAG_EINVAL, "null argument");
    tdag_context* ctx = g->ctx;
    if (hipError_t e = hipSetDevice(ctx->device); e != hipSuccess)
        return setHipErr(e, "hipSetDevice");
    hipError_t e = hipEventRecord(ctx->ev[0], ctx->stream);
    if (e == hipSuccess)
        e = g->dt.replay(ctx->stream);
    if (e == hipSuccess)
        e = hipEventRecord(ctx->ev[3], ctx->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess)
        return setHipErr(e, "build graph replay");
    if (out_ms)
        (void)hipEventElapsedTime(out_ms, ctx->ev[0], ctx->ev[3]);
    return TDAG_OK;
}

tdag_status tdag_build_graph_result(tdag_build_graph* g, tdag_tdag** out)
{
    if (!g || !out)
        return setErr(TDAG_EINVAL, "null argument");
    std::unique_ptr<tdag_tdag> res(new tdag_tdag());
    res->cfg = g->cfg;
    res->t = g->base;
    if (hipError_t e = g->dt.readResult(res->t, g->ctx->stream); e != hipSuccess)
        return setHipErr(e, "build graph readout");
    translationRecords(g->dt, g->cfg.max_translation_level, res->records);
    *out = res.release();
    return TDAG_OK;
}

tdag_status tdag_build_graph_destroy(tdag_build_graph* g)
{
    if (g) {
        (void)hipSetDevice(g->ctx->device);
        delete g;
    }
    return TDAG_OK;
}

tdag_status tdag_svdag_static_words(const tdag_svdag* h, uint32_t level, void* out, uint64_t cap, uint64_t* out_count)
{
    if (!h || !out_count)
        return setErr(TDAG_EINVAL, "null argument");
    if (level < 2 || level > std::max(2u, h->s.rootLevel))
        return setErr(TDAG_EINVAL, "level " + std::to_string(level) + " outside [2, root level]");
    if (level == 2) {
        *out_count = h->s.leaves.size();
        if (out) {
            if (cap < h->s.leaves.size())
                return setErr(TDAG_EINVAL, "buffer too small");
            std::memcpy(out, h->s.leaves.data(), h->s.leaves.size() * 8);
        }
        return TDAG_OK;
    }
    const std::vector<uint32_t> w = tdag::staticSvdagWords(h->s, level);
    *out_count = w.size();
    if (out) {
        if (cap < w.size())
            return setErr(TDAG_EINVAL, "buffer too small");
        std::memcpy(out, w.data(), w.size() * 4);
    }
    return TDAG_OK;
}

tdag_status tdag_construct_static_transform_dag(tdag_context* ctx, const tdag_static_svdag_view* in, const tdag_config* cfg, uint32_t flags, tdag_tdag** out)
{
    if (!in || !cfg || !out)
        return setErr(TDAG_EINVAL, "null argument");
    if (in->root_level < 2 || in->root_level > 12 || in->resolution != (1u << in->root_level))
        return setErr(TDAG_EINVAL, "resolution must be 2^root_level with root_level in [2, 12]");
    if (in->root_level >= 3 && (!in->level_words || !in->level_word_counts))
        return setErr(TDAG_EINVAL, "level_words / level_word_counts are null");
    if (in->num_leaves && !in->leaves)
        return setErr(TDAG_EINVAL, "leaves is null");
    tdag_svdag dag;
    std::string err;
    if (!tdag::svdagFromStatic(in->resolution, in->root_level, in->level_words, in->level_word_counts, in->leaves, in->num_leaves, dag.s, err))
        return setErr(TDAG_EINVAL, "static SVDAG: " + err);
    return constructTransformDag(ctx, &dag, cfg, flags, false, nullptr, 0, true, out);
}

tdag_status tdag_construct_transform_dag_learned(tdag_context* ctx, const tdag_svdag* in, const tdag_config* cfg, uint32_t flags, tdag_mlp* mlp, uint32_t k,
    tdag_tdag** out)
{
    if (!ctx || !mlp)
        return setErr(TDAG_EINVAL, "the learned driver needs a context and a network");
    if (k == 0 || k > 64)
        return setErr(TDAG_EINVAL, "k must be in [1, 64]");
    return constructTransformDag(ctx, in, cfg, flags, true, mlp, k, false, out);
}

tdag_status tdag_tdag_serialize(const tdag_tdag* t, const tdag_config* cfg, uint8_t* buf, uint64_t cap, uint64_t* out_size)
{
    if (!t || !cfg || !out_size)
        return setErr(TDAG_EINVAL, "null argument");
    if (t->isStatic)
        return setErr(TDAG_EINVAL, "a static-structure result is written by tdag_tdag_static_serialize");
    if (!sameConfig(*cfg, t->cfg))
        return setErr(TDAG_EINVAL, "config differs from the one the structure was built with");
    const std::vector<uint8_t> bytes = tdag::serializeTdag(t->t, *cfg);
    *out_size = bytes.size();
    if (buf) {
        if (cap < bytes.size())
            return setErr(TDAG_EINVAL, "buffer too small");
        std::memcpy(buf, bytes.data(), bytes.size());
    }
    return TDAG_OK;
}

tdag_status tdag_tdag_static_words(const tdag_tdag* t, uint32_t level, uint64_t* out, uint64_t cap, uint64_t* out_count)
{
    if (!t || !out_count)
        return setErr(TDAG_EINVAL, "null argument");
    if (!t->isStatic)
        return setErr(TDAG_EINVAL, "not a static-structure result (tdag_construct_static_transform_dag)");
    if (level < 2 || level > std::max(2u, t->t.rootLevel))
        return setErr(TDAG_EINVAL, "level " + std::to_string(level) + " outside [2, root level]");
    const std::vector<uint64_t> w = tdag::staticTdagWords(t->t, level);
    *out_count = w.size();
    if (out) {
        if (cap < w.size())
            return setErr(TDAG_EINVAL, "buffer too small");
        std::memcpy(out, w.data(), w.size() * 8);
    }
    return TDAG_OK;
}

tdag_status tdag_tdag_static_serialize(const tdag_tdag* t, uint8_t* buf, uint64_t cap, uint64_t* out_size)
{
    if (!t || !out_size)
        return setErr(TDAG_EINVAL, "null argument");
    if (!t->isStatic)
        return setErr(TDAG_EINVAL, "not a static-structure result (tdag_construct_static_transform_dag)");
    const std::vector<uint8_t> bytes = tdag::serializeStaticTdag(t->t);
    *out_size = bytes.size();
    if (buf) {
        if (cap < bytes.size())
            return setErr(TDAG_EINVAL, "buffer too small");
        std::memcpy(buf, bytes.data(), bytes.size());
    }
    return TDAG_OK;
}

tdag_status tdag_tdag_level_record(const tdag_tdag* t, uint32_t level, tdag_level_record* out)
{
    if (!t || !out)
        return setErr(TDAG_EINVAL, "null argument");
    for (const auto& r : t->records)
        if (r.level == level) {
            out->level = r.level;
            out->items = r.items;
            out->unique = r.unique;
            out->matched = r.matched;
            out->ms_match = r.msMatch;
            out->has_masks = r.masks.empty() ? 0u : 1u;
            out->learned = r.learned ? 1u : 0u;
            out->proposal_hits = r.proposalHits;
            out->fallback_queries = r.fallback;
            return TDAG_OK;
        }
    return setErr(TDAG_EINVAL, "no matcher record for level " + std::to_string(level));
}

tdag_status tdag_tdag_copy_level(const tdag_tdag* t, uint32_t level, uint64_t* out, uint64_t cap, uint64_t* out_count)
{
    if (!t || !out_count)
        return setErr(TDAG_EINVAL, "null argument");
    if (t->isStatic)
        return setErr(TDAG_EINVAL, "a static-structure result is read with tdag_tdag_static_words");
    if (level < 2 || level > t->t.rootLevel)
        return setErr(TDAG_EINVAL, "level " + std::to_string(level) + " outside [2, root level]");
    const uint64_t words = level == 2 ? t->t.leaves.size() : t->t.nodes[level].size() * 8;
    *out_count = words;
    if (!out)
        return TDAG_OK;
    if (cap < words)
        return setErr(TDAG_EINVAL, "buffer too small");
    if (level == 2)
        std::memcpy(out, t->t.leaves.data(), words * 8);
    else
        std::memcpy(out, t->t.nodes[level].data(), words * 8);
    return TDAG_OK;
}

tdag_status tdag_tdag_level_masks(const tdag_tdag* t, uint32_t level, uint64_t* out, uint64_t cap_words)
{
    if (!t || !out)
        return setErr(TDAG_EINVAL, "null argument");
    for (const auto& r : t->records)
        if (r.level == level) {
            if (r.masks.size() > cap_words)
                return setErr(TDAG_EINVAL, "buffer too small");
            std::memcpy(out, r.masks.data(), r.masks.size() * 8);
            return TDAG_OK;
        }
    return setErr(TDAG_EINVAL, "no matcher record for level " + std::to_string(level));
}

tdag_status tdag_tdag_info(const tdag_tdag* t, uint32_t* root_level, uint64_t* level_sizes, uint32_t cap)
{
    if (!t)
        return setErr(TDAG_EINVAL, "null handle");
    if (root_level)
        *root_level = t->t.rootLevel;
    if (level_sizes)
        for (uint32_t l = 0; l < cap; ++l)
            level_sizes[l] = l == 2 ? t->t.leaves.size() : (l < t->t.nodes.size() ? t->t.nodes[l].size() : 0);
    return TDAG_OK;
}

tdag_status tdag_tdag16_encode(const tdag_tdag* t, const tdag_config* cfg, uint32_t flags, tdag_tdag16** out)
{
    if (!t || !cfg || !out)
        return setErr(TDAG_EINVAL, "null argument");
    if (t->isStatic)
        return setErr(TDAG_EINVAL, "encode the result of tdag_construct_transform_dag (edit-structure form)");
    if (!sameConfig(*cfg, t->cfg))
        return setErr(TDAG_EINVAL, "config differs from the one the structure was built with");
    if (flags & ~(TDAG16_POINTER_TABLES | TDAG16_HUFFMAN))
        return setErr(TDAG_EINVAL, "unknown flag");
    try {
        *out = tdag::newTdag16Handle(tdag::encodeTdag16(t->t, *cfg, (flags & TDAG16_POINTER_TABLES) != 0, (flags & TDAG16_HUFFMAN) != 0));
    } catch (const std::bad_alloc&) {
        return setErr(TDAG_ENOMEM, "out of host memory");
    } catch (const std::exception& e) {
        return setErr(TDAG_EINVAL, e.what());
    }
    return TDAG_OK;
}

tdag_status tdag_tdag_free(tdag_tdag* t)
{
    delete t;
    return TDAG_OK;
}

} // extern "C"
