// Learned proposal for the level-2 matcher (4^3 leaves): an MLP embeds each leaf together with its
// transform-invariant occupancy descriptors; for every query the 64 most similar earlier-ranked
// leaves are proposed, each proposal is verified bitwise (so every accepted match is exact), and
// queries without a verified proposal fall back to the exhaustive search. No reference
// counterpart.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "matcher.hpp"
#include "tdag.h"

namespace tdag {
namespace learn {

constexpr int kIn = 80; // 64 mask bits + 12 descriptor bytes / 16, zero-padded
constexpr int kEmb = 128;
constexpr int kLayers = 4;
constexpr int kLayerIn[kLayers] = { 80, 256, 256, 256 };
constexpr int kLayerOut[kLayers] = { 256, 256, 256, 128 };
constexpr uint32_t kMaxProposals = 64;

constexpr size_t paramOffsetW(int l)
{
    size_t off = 0;
    for (int i = 0; i < l; ++i)
        off += (size_t)kLayerIn[i] * kLayerOut[i] + kLayerOut[i];
    return off;
}
constexpr size_t paramOffsetB(int l) { return paramOffsetW(l) + (size_t)kLayerIn[l] * kLayerOut[l]; }
constexpr size_t kParams = paramOffsetW(kLayers);

// Device state of one network: fp32 master parameters (W_l row-major in x out, then b_l), their
// bf16 copy used by the MFMA kernels, gradients and Adam moments, plus a workspace that grows
// with the largest batch seen.
constexpr uint32_t kRetrieveAll = 0; // TDAG_RETRIEVAL_ALL
constexpr uint32_t kRetrieveBucketed = 1; // TDAG_RETRIEVAL_BUCKETED
struct Mlp {
    uint32_t retrieval = kRetrieveAll; // how the matcher retrieves its proposals
    float* w = nullptr;
    uint16_t* wbf = nullptr;
    float* g = nullptr;
    float* m = nullptr;
    float* v = nullptr;
    uint32_t step = 0;
    void* ws = nullptr;
    size_t wsBytes = 0;
};

hipError_t mlpAlloc(Mlp& mlp);
void mlpFree(Mlp& mlp);
// Uploads fp32 parameters (kParams floats, host) and refreshes the bf16 copy; resets Adam.
hipError_t mlpSetParams(Mlp& mlp, const float* host, hipStream_t stream);
hipError_t mlpGetParams(const Mlp& mlp, float* host, hipStream_t stream);
// He-uniform weights from a seeded generator, zero biases.
hipError_t mlpInit(Mlp& mlp, uint64_t seed, hipStream_t stream);

// Loss and gradient (into mlp.g) of InfoNCE over B pairs (d_q[i], d_p[i]) at temperature tau.
hipError_t mlpLossGrad(Mlp& mlp, const uint64_t* d_q, const uint64_t* d_p, uint32_t B, float tau, float* out_loss, hipStream_t stream);
// `steps` Adam steps on batches of `batch` pairs drawn (without replacement per epoch, seeded)
// from the n pairs; writes each step's loss when out_losses is not null.
hipError_t mlpTrain(Mlp& mlp, const uint64_t* d_q, const uint64_t* d_p, uint64_t n, uint32_t steps, uint32_t batch, float lr, float tau, uint64_t seed,
    float* out_losses, hipStream_t stream);
// L2-normalised embeddings (n x 128, fp32) of n leaves.
hipError_t mlpEmbed(Mlp& mlp, const uint64_t* d_masks, uint64_t n, float* d_out, hipStream_t stream);

struct LearnedStats {
    uint64_t proposalHits = 0; // queries resolved by a verified proposal
    uint64_t fallback = 0; // non-empty queries sent to the exhaustive search
    float msEmbed = 0, msPropose = 0, msVerify = 0, msFallback = 0;
};

// Level-2 matcher with proposals: `mlp` null selects popcount-neighbour proposals (the k
// leaves ranked just before the query). Outputs and counters as enqueueMatch; synchronises.
hipError_t matchLearned2(const MatchPlan& plan, Mlp* mlp, uint32_t k, const uint64_t* d_masks, uint32_t cfgBits, void* d_scratch, void* d_work, hipStream_t stream,
    uint64_t* d_encoded, uint64_t* d_mapping, uint32_t* d_unique, LearnedStats* stats);
// Device workspace bytes matchLearned2 needs for n leaves (besides the plan's scratch).
size_t learnedWorkBytes(uint64_t n);
// Grows the network's embedding workspace for n leaves (allocates; call before enqueueLearned2).
hipError_t mlpPrepareEmbed(Mlp& mlp, uint64_t n);
// matchLearned2 without allocation or synchronisation, so a stream capture can record it: the
// network's workspace must have been prepared for p.n (mlpPrepareEmbed). ev (optional, 5 events)
// marks the stage boundaries. The proposal-hit and fallback counters are at learnedCounters().
hipError_t enqueueLearned2(const MatchPlan& plan, Mlp* mlp, uint32_t k, const uint64_t* d_masks, uint32_t cfgBits, void* d_scratch, void* d_work, hipStream_t stream,
    uint64_t* d_encoded, uint64_t* d_mapping, uint32_t* d_unique, hipEvent_t* ev);
const unsigned long long* learnedCounters(const MatchPlan& plan, void* d_work);

// Retrieval of the learned proposal: for each of the n leaves (rank order, bf16 embeddings n x 128
// row-major) the k <= 64 best earlier leaves under "higher score first, then lower rank", best
// first, into cand[q * k ..] (0xFFFFFFFF where fewer exist). `work` holds topkWorkBytes(n) bytes.
size_t topkWorkBytes(uint64_t n);
hipError_t topkPropose(const uint16_t* Ebf, uint32_t n, uint32_t k, uint32_t* cand, void* work, hipStream_t stream);
// Bucketed retrieval: as topkPropose, with the candidates of each leaf restricted to the earlier
// leaves of its exact transform-invariant descriptor (sortedMasks: the n leaves in rank order).
// `work` holds bucketWorkBytes(n) bytes.
size_t bucketWorkBytes(uint64_t n);
hipError_t bucketPropose(const uint64_t* sortedMasks, const uint16_t* Ebf, uint32_t n, uint32_t k, uint32_t* cand, void* work, hipStream_t stream);

// Test hook: C = opA x opB with the bf16 MFMA kernel (row-major storage; TA/TB transpose).
hipError_t gemmBf16(bool ta, bool tb, const uint16_t* A, int lda, const uint16_t* B, int ldb, float* C, int ldc, int M, int N, int K, hipStream_t stream);

} // namespace learn
} // namespace tdag

// The C-ABI handle of a network (tdag_mlp in tdag.h): the context it belongs to and its state.
struct tdag_mlp {
    tdag_context* ctx = nullptr;
    tdag::learn::Mlp mlp {};
};
