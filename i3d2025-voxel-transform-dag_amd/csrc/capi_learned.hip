// C-ABI of the learned proposal (include/tdag.h, "Learned proposal" section).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "context.hpp"
#include "learned.hpp"
#include "matcher.hpp"
#include "tdag.h"

using namespace tdag::capi;


extern "C" {

uint64_t tdag_mlp_param_count(void) { return tdag::learn::kParams; }

tdag_status tdag_mlp_create(tdag_context* ctx, uint64_t seed, tdag_mlp** out)
{
    if (!ctx || !out)
        return fail(TDAG_EINVAL, "null argument");
    CAPI_HIP(hipSetDevice(ctx->device), "hipSetDevice");
    auto* m = new tdag_mlp();
    m->ctx = ctx;
    hipError_t e = tdag::learn::mlpAlloc(m->mlp);
    if (e == hipSuccess)
        e = tdag::learn::mlpInit(m->mlp, seed, ctx->stream);
    if (e != hipSuccess) {
        tdag::learn::mlpFree(m->mlp);
        delete m;
        return failHip(e, "mlp allocation / initialisation");
    }
    *out = m;
    return TDAG_OK;
}

tdag_status tdag_mlp_destroy(tdag_mlp* m)
{
    if (!m)
        return TDAG_OK;
    tdag::learn::mlpFree(m->mlp);
    delete m;
    return TDAG_OK;
}

tdag_status tdag_mlp_get_params(tdag_mlp* m, float* out)
{
    if (!m || !out)
        return fail(TDAG_EINVAL, "null argument");
    CAPI_HIP(hipSetDevice(m->ctx->device), "hipSetDevice");
    CAPI_HIP(tdag::learn::mlpGetParams(m->mlp, out, m->ctx->stream), "mlp parameter readback");
    return TDAG_OK;
}

tdag_status tdag_mlp_set_params(tdag_mlp* m, const float* in)
{
    if (!m || !in)
        return fail(TDAG_EINVAL, "null argument");
    CAPI_HIP(hipSetDevice(m->ctx->device), "hipSetDevice");
    CAPI_HIP(tdag::learn::mlpSetParams(m->mlp, in, m->ctx->stream), "mlp parameter upload");
    return TDAG_OK;
}

tdag_status tdag_mlp_loss_grad(tdag_mlp* m, const uint64_t* d_q, const uint64_t* d_p, uint32_t batch, float tau, float* out_loss, float* out_grad)
{
    if (!m || !d_q || !d_p || batch < 2 || batch > 4096 || !(tau > 0.f))
        return fail(TDAG_EINVAL, "bad argument (batch must be in [2, 4096], tau > 0)");
    CAPI_HIP(hipSetDevice(m->ctx->device), "hipSetDevice");
    CAPI_HIP(tdag::learn::mlpLossGrad(m->mlp, d_q, d_p, batch, tau, out_loss, m->ctx->stream), "mlp loss / gradient");
    if (out_grad) {
        CAPI_HIP(hipMemcpyAsync(out_grad, m->mlp.g, tdag::learn::kParams * 4, hipMemcpyDeviceToHost, m->ctx->stream), "gradient readback");
        CAPI_HIP(hipStreamSynchronize(m->ctx->stream), "hipStreamSynchronize");
    }
    return TDAG_OK;
}

tdag_status tdag_mlp_train(tdag_mlp* m, const uint64_t* d_q, const uint64_t* d_p, uint64_t n_pairs, uint32_t steps, uint32_t batch, float lr, float tau, uint64_t seed,
    float* out_losses)
{
    if (!m || !d_q || !d_p || n_pairs < 2 || batch < 2 || batch > 4096 || !(tau > 0.f) || !(lr > 0.f))
        return fail(TDAG_EINVAL, "bad argument (need >= 2 pairs, batch in [2, 4096], tau > 0, lr > 0)");
    if (n_pairs >= (1ull << 32))
        return fail(TDAG_EINVAL, "too many pairs");
    CAPI_HIP(hipSetDevice(m->ctx->device), "hipSetDevice");
    CAPI_HIP(tdag::learn::mlpTrain(m->mlp, d_q, d_p, n_pairs, steps, batch, lr, tau, seed, out_losses, m->ctx->stream), "mlp training");
    return TDAG_OK;
}

tdag_status tdag_mlp_embed(tdag_mlp* m, const uint64_t* d_masks, uint64_t n, float* d_out)
{
    if (!m || (n > 0 && (!d_masks || !d_out)))
        return fail(TDAG_EINVAL, "null argument");
    CAPI_HIP(hipSetDevice(m->ctx->device), "hipSetDevice");
    CAPI_HIP(tdag::learn::mlpEmbed(m->mlp, d_masks, n, d_out, m->ctx->stream), "mlp embedding");
    CAPI_HIP(hipStreamSynchronize(m->ctx->stream), "hipStreamSynchronize");
    return TDAG_OK;
}

tdag_status tdag_gemm_bf16(tdag_context* ctx, int trans_a, int trans_b, const uint16_t* d_a, int lda, const uint16_t* d_b, int ldb, float* d_c, int ldc, int m, int n, int k)
{
    if (!ctx || !d_a || !d_b || !d_c || m < 0 || n < 0 || k < 0)
        return fail(TDAG_EINVAL, "bad argument");
    CAPI_HIP(hipSetDevice(ctx->device), "hipSetDevice");
    CAPI_HIP(tdag::learn::gemmBf16(trans_a != 0, trans_b != 0, d_a, lda, d_b, ldb, d_c, ldc, m, n, k, ctx->stream), "bf16 gemm");
    CAPI_HIP(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    return TDAG_OK;
}

tdag_status tdag_mlp_topk(tdag_context* ctx, const uint16_t* d_emb, uint64_t n, uint32_t k, uint32_t* d_cand)
{
    if (!ctx || (n > 0 && (!d_emb || !d_cand)))
        return fail(TDAG_EINVAL, "null argument");
    if (k == 0 || k > tdag::learn::kMaxProposals)
        return fail(TDAG_EINVAL, "k must be in [1, 64]");
    if (n >= (1ull << 31))
        return fail(TDAG_EINVAL, "n must be below 2^31");
    CAPI_HIP(hipSetDevice(ctx->device), "hipSetDevice");
    if (tdag_status s = ensureBuffer(ctx->scratch, ctx->scratchBytes, tdag::learn::topkWorkBytes(n), "retrieval workspace allocation"); s != TDAG_OK)
        return s;
    CAPI_HIP(tdag::learn::topkPropose(d_emb, (uint32_t)n, k, d_cand, ctx->scratch, ctx->stream), "top-k retrieval");
    CAPI_HIP(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    return TDAG_OK;
}

tdag_status tdag_mlp_set_retrieval(tdag_mlp* m, uint32_t mode)
{
    if (!m)
        return fail(TDAG_EINVAL, "null argument");
    if (mode != TDAG_RETRIEVAL_ALL && mode != TDAG_RETRIEVAL_BUCKETED)
        return fail(TDAG_EINVAL, "unknown retrieval mode " + std::to_string(mode));
    m->mlp.retrieval = mode;
    return TDAG_OK;
}

tdag_status tdag_mlp_get_retrieval(const tdag_mlp* m, uint32_t* out_mode)
{
    if (!m || !out_mode)
        return fail(TDAG_EINVAL, "null argument");
    *out_mode = m->mlp.retrieval;
    return TDAG_OK;
}

tdag_status tdag_mlp_topk_bucketed(tdag_context* ctx, const uint64_t* d_masks, const uint16_t* d_emb, uint64_t n, uint32_t k, uint32_t* d_cand)
{
    if (!ctx || (n > 0 && (!d_masks || !d_emb || !d_cand)))
        return fail(TDAG_EINVAL, "null argument");
    if (k == 0 || k > tdag::learn::kMaxProposals)
        return fail(TDAG_EINVAL, "k must be in [1, 64]");
    if (n >= (1ull << 31))
        return fail(TDAG_EINVAL, "n must be below 2^31");
    CAPI_HIP(hipSetDevice(ctx->device), "hipSetDevice");
    if (tdag_status s = ensureBuffer(ctx->scratch, ctx->scratchBytes, tdag::learn::bucketWorkBytes(n), "retrieval workspace allocation"); s != TDAG_OK)
        return s;
    CAPI_HIP(tdag::learn::bucketPropose(d_masks, d_emb, (uint32_t)n, k, d_cand, ctx->scratch, ctx->stream), "bucketed retrieval");
    CAPI_HIP(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    return TDAG_OK;
}

tdag_status tdag_match_level_learned_device(tdag_context* ctx, tdag_mlp* mlp, uint32_t level, const uint64_t* d_masks, uint64_t n, const tdag_config* cfg, uint32_t k,
    uint64_t* d_encoded, uint64_t* d_mapping, uint32_t* d_unique, tdag_learned_stats* stats)
{
    if (tdag_status s = checkArgs(ctx, level, n, cfg); s != TDAG_OK)
        return s;
    if (level != 2)
        return fail(TDAG_EINVAL, "the learned proposal is implemented for level 2 (4^3 leaves)");
    if (k == 0 || k > tdag::learn::kMaxProposals)
        return fail(TDAG_EINVAL, "k must be in [1, 64]");
    if (n > 0 && !d_masks)
        return fail(TDAG_EINVAL, "d_masks is null");
    if (mlp && mlp->ctx != ctx)
        return fail(TDAG_EINVAL, "the network belongs to another context");
    CAPI_HIP(hipSetDevice(ctx->device), "hipSetDevice");
    const tdag::MatchPlan plan = tdag::planMatch(level, n, 0, ctx->firstBatch[level], ctx->batchGrowth[level], ctx->bloomBits[level]);
    const size_t workOff = (plan.scratchBytes + 255) / 256 * 256;
    if (tdag_status s = ensureBuffer(ctx->scratch, ctx->scratchBytes, workOff + tdag::learn::learnedWorkBytes(n), "matcher scratch allocation"); s != TDAG_OK)
        return s;
    const uint32_t cfgBits = cfgBitsOf(cfg);
    tdag::learn::LearnedStats ls;
    CAPI_HIP(hipEventRecord(ctx->ev[0], ctx->stream), "hipEventRecord");
    CAPI_HIP(tdag::learn::matchLearned2(plan, mlp ? &mlp->mlp : nullptr, k, d_masks, cfgBits, ctx->scratch, (char*)ctx->scratch + workOff, ctx->stream, d_encoded, d_mapping,
                 d_unique, &ls),
        "learned matcher");
    CAPI_HIP(hipEventRecord(ctx->ev[3], ctx->stream), "hipEventRecord");
    CAPI_HIP(hipMemcpyAsync(ctx->pinned, (char*)ctx->scratch + plan.offCounters, 16, hipMemcpyDeviceToHost, ctx->stream), "counter readback");
    CAPI_HIP(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    if (stats) {
        tdag_learned_stats s {};
        s.num_items = n;
        s.num_matched = ctx->pinned[0];
        s.num_unique = n > 0 ? (ctx->pinned[1] & 0xFFFFFFFFull) : 0;
        s.proposal_hits = ls.proposalHits;
        s.fallback_queries = ls.fallback;
        s.ms_embed = ls.msEmbed;
        s.ms_propose = ls.msPropose;
        s.ms_verify = ls.msVerify;
        s.ms_fallback = ls.msFallback;
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[3]);
        s.ms_total = ms;
        *stats = s;
    }
    return TDAG_OK;
}

tdag_status tdag_match_level_learned(tdag_context* ctx, tdag_mlp* mlp, uint32_t level, const uint64_t* masks, uint64_t n, const tdag_config* cfg, uint32_t k,
    uint64_t* out_encoded, uint64_t* out_mapping, uint32_t* out_unique, uint64_t* out_num_unique, tdag_learned_stats* stats)
{
    if (tdag_status s = checkArgs(ctx, level, n, cfg); s != TDAG_OK)
        return s;
    if (n > 0 && (!masks || !out_mapping || !out_unique))
        return fail(TDAG_EINVAL, "masks, out_mapping and out_unique are required");
    CAPI_HIP(hipSetDevice(ctx->device), "hipSetDevice");
    const size_t W = level >= 2 ? std::max<size_t>(1, ((size_t)1 << (3 * level)) / 64) : 1;
    const size_t a = ((size_t)n * W * 8 + 255) / 256 * 256;
    const size_t b = ((size_t)n * 8 + 255) / 256 * 256;
    const size_t c = ((size_t)n * 4 + 255) / 256 * 256;
    if (tdag_status s = ensureBuffer(ctx->io, ctx->ioBytes, a + 2 * b + c, "io buffer allocation"); s != TDAG_OK)
        return s;
    uint64_t* dMasks = (uint64_t*)ctx->io;
    uint64_t* dEncoded = (uint64_t*)((char*)ctx->io + a);
    uint64_t* dMapping = (uint64_t*)((char*)ctx->io + a + b);
    uint32_t* dUnique = (uint32_t*)((char*)ctx->io + a + 2 * b);
    if (n > 0)
        CAPI_HIP(hipMemcpyAsync(dMasks, masks, (size_t)n * W * 8, hipMemcpyHostToDevice, ctx->stream), "mask upload");
    tdag_learned_stats st {};
    if (tdag_status s = tdag_match_level_learned_device(ctx, mlp, level, dMasks, n, cfg, k, out_encoded ? dEncoded : nullptr, dMapping, dUnique, &st); s != TDAG_OK)
        return s;
    if (n > 0) {
        if (out_encoded)
            CAPI_HIP(hipMemcpyAsync(out_encoded, dEncoded, (size_t)n * 8, hipMemcpyDeviceToHost, ctx->stream), "encoded readback");
        CAPI_HIP(hipMemcpyAsync(out_mapping, dMapping, (size_t)n * 8, hipMemcpyDeviceToHost, ctx->stream), "mapping readback");
        CAPI_HIP(hipMemcpyAsync(out_unique, dUnique, (size_t)st.num_unique * 4, hipMemcpyDeviceToHost, ctx->stream), "unique readback");
        CAPI_HIP(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    }
    if (out_num_unique)
        *out_num_unique = st.num_unique;
    if (stats)
        *stats = st;
    return TDAG_OK;
}

} // extern "C"
