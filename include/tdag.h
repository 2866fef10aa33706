/*
 * tdag.h — C-ABI of the MI355X-native Transform-Aware SVDAG matching path.
 *
 * Plain pointers and sizes only. Every entry point returns a tdag_status; on failure a
 * thread-local message is available through tdag_last_error(). No exceptions cross this
 * boundary (the C++ wrappers in tdag.hpp rethrow them as std::runtime_error, matching the
 * reference's assert_always behaviour).
 *
 * Masks ("grids") use the layout of voxcom::LargeSubGrid<R> (large_sub_grid.h:80-97):
 * R = 2^level, bit i = x + R*y + R*R*z lives in word i >> 6, W = max(1, R^3/64) words per mask.
 * TransformPointer values are the 64-bit bit patterns of voxcom::TransformPointer
 * (transform_dag.h:19-54, GCC x86-64 bit-field layout): ptr bits 0-36, sx/sy/sz bits 37-39,
 * p0/p1/p2 bits 40-45, tx/ty/tz (biased by 32) bits 46-63.
 */
#ifndef TDAG_H_
#define TDAG_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum tdag_status {
    TDAG_OK = 0,
    TDAG_EINVAL = 1, /* level outside [2,5], null pointers, sizes out of range */
    TDAG_ENOMEM = 2, /* device or host allocation failed */
    TDAG_EDEVICE = 3, /* a HIP runtime call failed or no device is present */
    TDAG_EINTERNAL = 4 /* an internal invariant did not hold */
} tdag_status;

/* Byte layout of voxcom::TransformDAGConfig (transform_dag.h:157-160): three bools, one
 * padding byte, uint32 maxTranslationLevel. sizeof == 8. */
typedef struct tdag_config {
    uint8_t symmetry;
    uint8_t axis_permutation;
    uint8_t translation;
    uint8_t pad;
    uint32_t max_translation_level;
} tdag_config;

/* Library context: owns a HIP stream, cached device scratch and timing events. One call per
 * context at a time (thread-compatible, not thread-safe). */
typedef struct tdag_context tdag_context;

/* Per-call statistics of the last matcher call on a context. */
typedef struct tdag_match_stats {
    uint64_t num_items;
    uint64_t num_unique;
    uint64_t num_matched; /* items whose minimum-key parent is another item ("verified matches") */
    uint64_t table_entries;
    float ms_total; /* device time of the whole call (hipEvents on the call's stream) */
    float ms_search; /* device time of the search kernel(s) alone */
    /* Per-stage probe counters, filled only when tdag_context_set_counters(ctx, 1) is on:
     * non-empty sub-boxes probed, Bloom-filter passes, table entries with an equal key,
     * candidates (child ranked after the parent, transform allowed by the config),
     * candidates that beat the child's current best and were verified, and verified matches. */
    uint64_t probes, filter_pass, key_hits, candidates, verifications, verified;
} tdag_match_stats;

/* Copies the calling thread's last error message into buf (NUL-terminated, truncated to
 * cap). Returns the full message length. */
size_t tdag_last_error(char* buf, size_t cap);

/* Library version string, e.g. "tdag-mi355x 0.1 gfx950". */
const char* tdag_version(void);

/* Creates a context on HIP device `device`. Replaces the per-call thrust::device_vector
 * ownership of find_translation.cu:1204-1205, 1286-1287 with cached scratch. */
tdag_status tdag_context_create(int device, tdag_context** out_ctx);
tdag_status tdag_context_destroy(tdag_context* ctx);

/*
 * Device-resident matcher — the hot path. Replaces the search of
 * findDuplicatesAmongFlattenedItems_gpu / _cpp2 (transform_dag.h:307-311;
 * find_translation.cpp:616-1000) for one level.
 *
 *   level      2..5 (R = 4, 8, 16, 32)
 *   d_masks    device pointer, n * W uint64 words, items in their original order
 *   cfg        host pointer, TransformDAGConfig layout (filters candidate transforms)
 *   d_encoded  optional device output (n uint64): the reference's encodedParents array —
 *              the minimum encodeFixed64 key per item, whose ptr field is the parent's rank in
 *              the stable popcount-descending order (find_translation.cpp:631-643, 864)
 *   d_mapping  optional device output (n uint64): outPrevLevelMapping, TransformPointer bits
 *   d_unique   optional device output (n uint32): original indices of the unique items, in
 *              original order; the first *out_num_unique entries are valid
 *   out_num_unique  optional host output
 *   stream     hipStream_t to enqueue on (NULL = the context's own stream). The call
 *              synchronises this stream before returning when out_num_unique or stats is set.
 */
tdag_status tdag_match_level_device(tdag_context* ctx, uint32_t level, const uint64_t* d_masks, uint64_t n,
    const tdag_config* cfg, uint64_t* d_encoded, uint64_t* d_mapping, uint32_t* d_unique,
    uint64_t* out_num_unique, void* stream, tdag_match_stats* stats);

/*
 * Host-buffer matcher: findDuplicatesAmongFlattenedItems_gpu drop-in
 * (transform_dag.h:310-311). Copies masks host->device, matches, copies results back.
 *   out_encoded may be NULL; out_mapping (n), out_unique (n) required.
 */
tdag_status tdag_match_level(tdag_context* ctx, uint32_t level, const uint64_t* masks, uint64_t n,
    const tdag_config* cfg, uint64_t* out_encoded, uint64_t* out_mapping, uint32_t* out_unique,
    uint64_t* out_num_unique, tdag_match_stats* stats);

/* Device scratch of tdag_match_level_device_noalloc on each GPU: 64 MiB. */
#define TDAG_NOALLOC_SCRATCH_BYTES (64u << 20)

/*
 * Allocation-free form of tdag_match_level_device. No counterpart in the reference: it is for
 * callers that cannot allocate device memory during the call (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc). Same kernels, arguments and results as
 * tdag_match_level_device with the context's tuning; the call allocates nothing and does not
 * synchronise. Its scratch is one fixed array of TDAG_NOALLOC_SCRATCH_BYTES per GPU, held from
 * the first context created on that GPU until the process ends; a level / item count whose
 * scratch needs more returns TDAG_EINVAL (at default tuning about 131 K items fit at level 2,
 * 65 K at level 3, 10.9 K at level 4 and 4.4 K at level 5). Every call on one GPU shares that
 * array, so calls that overlap on one GPU (any contexts, any streams) race; consecutive calls
 * are independent.
 *   d_counts  optional device output, 2 uint64: verified matches, number of unique items
 *   stream    hipStream_t (NULL = the context's stream)
 */
tdag_status tdag_match_level_device_noalloc(tdag_context* ctx, uint32_t level, const uint64_t* d_masks, uint64_t n,
    const tdag_config* cfg, uint64_t* d_encoded, uint64_t* d_mapping, uint32_t* d_unique, uint64_t* d_counts, void* stream);

/*
 * The exhaustive search restricted to a subset of children: only items whose flag is non-zero
 * (d_child_flags: n uint32 in original order, device) are looked up; the others keep their
 * identity key and are reported unmatched. This is the fallback stage of a proposal scheme
 * (queries without a verified proposal are searched exhaustively); with the flags set to the
 * items that have no match at all it gives the least time any proposal scheme with an
 * exhaustive fallback can reach at this level. Same parents, scratch and tuning as
 * tdag_match_level_device; d_encoded (optional) as there. Synchronises. No reference counterpart.
 */
tdag_status tdag_match_level_subset_device(tdag_context* ctx, uint32_t level, const uint64_t* d_masks, uint64_t n, const tdag_config* cfg,
    const uint32_t* d_child_flags, uint64_t* d_encoded, tdag_match_stats* stats);

/* Statistics of the last matcher call on ctx. */
tdag_status tdag_last_stats(const tdag_context* ctx, tdag_match_stats* out);

/* Enables (1) or disables (0) the per-stage probe counters of tdag_match_stats for this context
 * (diagnostics; costs a few percent). No reference counterpart. */
tdag_status tdag_context_set_counters(tdag_context* ctx, int enable);

/* Search-kernel variant used at `level` (3..5) by this context. variant % 3: a block runs one
 * (permutation, direction) combo (0), a family of 8 combos (1), or a family of 8 combos sharing
 * the first two hash passes between both directions of the third (2). (variant / 3) % 3: log2
 * of how many blocks share the R planes of one unit (more blocks = less LDS per block = more
 * blocks per CU). variant / 9: block order (0: parent fastest; 1: all blocks of a parent together
 * on one XCD, sharing its L2). Variants whose blocks exceed the LDS are rejected with EINVAL.
 * -1 restores the library default.
 * Level 2 has a single kernel (variant 0). No reference counterpart: a tuning knob. */
tdag_status tdag_context_set_variant(tdag_context* ctx, uint32_t level, int32_t variant);
tdag_status tdag_context_get_variant(const tdag_context* ctx, uint32_t level, int32_t* variant);

/* Rank-batched search schedule at `level`: parents are searched in rank order, first
 * `first_batch` of them, then batches growing by `growth`; between batches the child table drops
 * every child that can no longer improve. 0 selects the defaults (max(64, n/64), 2). A
 * tuning knob with no reference counterpart (the reference prunes between fixed-size batches,
 * find_translation.cpp:771, 887-915). */
tdag_status tdag_context_set_batching(tdag_context* ctx, uint32_t level, uint32_t first_batch, uint32_t growth);
tdag_status tdag_context_get_batching(const tdag_context* ctx, uint32_t level, uint32_t* first_batch, uint32_t* growth);

/* Size of the Bloom filter in front of the child table, in bits per table entry (0 = the library
 * default: 16 at level 3, 4 elsewhere).
 * Small filters stay in the L1 cache at the cost of more false positives. A tuning knob with no
 * reference counterpart. */
tdag_status tdag_context_set_filter_bits(tdag_context* ctx, uint32_t level, uint32_t bits_per_entry);
tdag_status tdag_context_get_filter_bits(const tdag_context* ctx, uint32_t level, uint32_t* bits_per_entry);

/* Where the 6 axis permutations are enumerated at `level` (3..5): frames = 1 hashes every
 * parent sub-box under all 48 (permutation, direction) combos against a table holding each
 * child from its 8 corners; frames = 6 hashes the parent under its 8 directions only against a
 * table holding each child from its 8 corners in all 6 permutation frames (48 entries per
 * child). Both find the same exhaustive minimum. 0 selects the default (6 at levels 4 and 5,
 * 1 at level 3). A tuning knob with no reference counterpart. */
tdag_status tdag_context_set_child_frames(tdag_context* ctx, uint32_t level, uint32_t frames);
tdag_status tdag_context_get_child_frames(const tdag_context* ctx, uint32_t level, uint32_t* frames);

/* Number of per-variant slots in the autotuner's timing output. */
#define TDAG_VARIANT_SLOTS 32

/* Autotuner: runs the level-`level` matcher on the device masks once per valid variant at the
 * default schedule (after one untimed warm-up run each; at levels 3..5 for both child-frame
 * counts, keeping the one with the fastest variant), then tries four first-batch sizes
 * (n/256, n/64, n/16, n) with each of the three fastest variants and keeps the fastest pair,
 * then filter sizes of 4, 8, 16 and 32 bits per table entry, then batch growths of 3, 4, 6, 8
 * and 12 (each batch that many times the previous one; the default is 2) with the first batch
 * halved, kept and doubled; the winners become this context's settings for `level`. Every trial
 * times the whole matcher call (preparation, search, post-processing), since the frame count and
 * the filter size also change the table build.
 * out_ms (TDAG_VARIANT_SLOTS + 1 floats, may be NULL) receives the average call time per variant
 * at the default schedule (-1 for invalid or unused slots) followed by the tuned time; also
 * writes the winning variant and first batch. No reference counterpart. */
tdag_status tdag_autotune_level(tdag_context* ctx, uint32_t level, const uint64_t* d_masks, uint64_t n, const tdag_config* cfg, uint32_t repeats,
    float* out_ms, int32_t* out_best, uint32_t* out_first_batch);

/* ------------------------------------------------------------------------------------------
 * Batched form: several levels whose inputs are known (one job per level) are matched in one
 * call, with every level's parent search in ONE persistent kernel fed by a device work queue,
 * optionally captured once as a hipGraph and replayed. Each level's rank-batched schedule (the
 * popcount buckets of parents) comes from the context's per-level settings (the autotuner's
 * picks). Results are bit-identical to tdag_match_level_device on each job. No reference
 * counterpart: the reference launches its GPU matcher per level and per batch
 * (find_translation.cu:1159-1563).
 * ---------------------------------------------------------------------------------------- */
typedef struct tdag_level_job {
    uint32_t level; /* 2..5; at most one job per level */
    uint32_t reserved;
    uint64_t n;
    const uint64_t* d_masks; /* device, n x R^3/64 words, linear layout */
    uint64_t* d_encoded; /* optional device outputs, as in tdag_match_level_device */
    uint64_t* d_mapping;
    uint32_t* d_unique;
} tdag_level_job;

typedef struct tdag_batch tdag_batch;
/* A network of the learned proposal (below). */
typedef struct tdag_mlp tdag_mlp;

/* Flags of tdag_batch_create (its `use_graph` argument; 0 and 1 keep their old meaning).
 * TDAG_BATCH_GRAPH: capture the whole call into a hipGraph and replay it on every run.
 * TDAG_BATCH_CONCURRENT: instead of the persistent kernel, each level runs the matcher that
 * tdag_match_level_device would run for it (the context's tuned variant and schedule at creation
 * time), on its own stream forked from and joined back into the context's stream; the levels
 * overlap (graph branches when captured). Same results. */
#define TDAG_BATCH_GRAPH 1u
#define TDAG_BATCH_CONCURRENT 2u

/* Plans the batch (device scratch owned by the batch) and, with TDAG_BATCH_GRAPH, captures the
 * whole call into a hipGraph on the context's stream. Persistent mode (default): all levels'
 * preparation, one persistent search kernel over a device work queue, all levels'
 * post-processing. The jobs' device buffers must stay valid until tdag_batch_destroy. */
tdag_status tdag_batch_create(tdag_context* ctx, const tdag_level_job* jobs, uint32_t count, const tdag_config* cfg, int use_graph, tdag_batch** out);
/* Runs the batch on the context's stream (graph replay or eager launches) and waits. out_stats
 * (count entries, job order, may be NULL) receives per-level results; ms_total and ms_search
 * are the whole call's time. Fails with TDAG_EINTERNAL if a grid barrier timed out. */
tdag_status tdag_batch_run(tdag_batch* b, tdag_match_stats* out_stats, float* out_ms);
/* Tunes the persistent kernel's work queue by running the batch (1 warm-up + `repeats` timed
 * runs per candidate) and keeps the fastest: first the tiles a block takes per pop (1, 2, 4, 8,
 * 16), then the share of blocks that take the lighter levels' tiles first, beside the heaviest
 * level's (one block in 0 = none, 2, 3, 4 or 6). The captured graph stays valid (it reads the
 * descriptor at run time). A concurrent batch has no queue: *out_grab = 0. */
tdag_status tdag_batch_autotune(tdag_batch* b, uint32_t repeats, uint32_t* out_grab);
/* Sets or reads the persistent kernel's queue tuning: tiles per pop (1..64) and split_every
 * (every split_every-th block takes the lighter levels' tiles first; 0 = off, else 2..64).
 * No reference counterpart. */
tdag_status tdag_batch_set_tuning(tdag_batch* b, uint32_t grab, uint32_t split_every);
tdag_status tdag_batch_get_tuning(const tdag_batch* b, uint32_t* out_grab, uint32_t* out_split_every);
tdag_status tdag_batch_destroy(tdag_batch* b);
/* The batched form with the learned proposal at level 2 (the combined form): the batch's level-2
 * job runs the learned matcher (network `mlp` of the same context, k <= 64 proposals per query,
 * bitwise verification, exhaustive fallback for the queries without a verified proposal, parent
 * chase), the other jobs run as in tdag_batch_create with the same flags. Persistent mode: the
 * learned matcher runs ahead of the persistent kernel on the call's stream (never beside it: the
 * persistent grid needs every CU for its grid barriers); concurrent mode: on the level-2 branch. With TDAG_BATCH_GRAPH the whole call, learned stage included, is one
 * captured hipGraph. Every accepted match is exact; the level-2 parents can differ from the
 * exhaustive minimum where it was not proposed. No reference counterpart. */
tdag_status tdag_batch_create_learned(tdag_context* ctx, const tdag_level_job* jobs, uint32_t count, const tdag_config* cfg, uint32_t flags, tdag_mlp* mlp, uint32_t k,
    tdag_batch** out);
/* Proposal hits and fallback queries of the learned job in the last tdag_batch_run. */
tdag_status tdag_batch_learned_stats(const tdag_batch* b, uint64_t* out_proposal_hits, uint64_t* out_fallback_queries);
/* One-shot batched call without a graph: create, run, destroy. */
tdag_status tdag_match_levels_batched_device(tdag_context* ctx, const tdag_level_job* jobs, uint32_t count, const tdag_config* cfg, tdag_match_stats* out_stats);

/* ------------------------------------------------------------------------------------------
 * Learned proposal (level 2, 4^3 leaves). An MLP embeds each leaf (64 mask bits plus 12
 * transform-invariant descriptor bytes: per-axis occupancy histograms, each trimmed of zero ends
 * and left-aligned, replaced by the smaller of itself and its reverse, the three sorted) into
 * 128 dimensions; 80 -> 256 -> 256 ->
 * 256 -> 128, ReLU, L2-normalised, bf16 MFMA with fp32 accumulate in training and inference.
 * It is trained with InfoNCE on (query, matched parent) pairs from the exact matcher. Matching:
 * the k (<= 64) most similar earlier-ranked leaves are proposed, each proposal is verified
 * bitwise under all 48 transforms and the possible translations, and queries without a
 * verified proposal fall back to the exhaustive search. Every accepted match is exact; the
 * chosen parent can differ from the exhaustive minimum when the minimum is not proposed. No
 * reference counterpart.
 * ---------------------------------------------------------------------------------------- */

typedef struct tdag_learned_stats {
    uint64_t num_items;
    uint64_t num_unique;
    uint64_t num_matched;
    uint64_t proposal_hits; /* queries resolved by a verified proposal */
    uint64_t fallback_queries; /* non-empty queries sent to the exhaustive search */
    float ms_embed, ms_propose, ms_verify, ms_fallback, ms_total;
} tdag_learned_stats;

/* Number of fp32 parameters (W_l row-major in x out, then b_l, for l = 0..3). */
uint64_t tdag_mlp_param_count(void);
/* A network on ctx's device with He-uniform weights from `seed` and zero biases. */
tdag_status tdag_mlp_create(tdag_context* ctx, uint64_t seed, tdag_mlp** out);
tdag_status tdag_mlp_destroy(tdag_mlp* m);
tdag_status tdag_mlp_get_params(tdag_mlp* m, float* out /* host, param_count floats */);
/* Replaces the parameters (host, param_count floats) and resets the optimiser state. */
tdag_status tdag_mlp_set_params(tdag_mlp* m, const float* in);
/* InfoNCE loss over `batch` pairs (d_q[i], d_p[i]) (device 4^3 masks) at temperature tau, and its
 * gradient (host, param_count floats, may be NULL). Parameters are not changed. */
tdag_status tdag_mlp_loss_grad(tdag_mlp* m, const uint64_t* d_q, const uint64_t* d_p, uint32_t batch, float tau, float* out_loss, float* out_grad);
/* `steps` Adam steps (learning rate lr) on batches of `batch` pairs drawn from the n_pairs pairs
 * (a seeded shuffle per epoch); out_losses (steps floats, may be NULL) receives each step's loss. */
tdag_status tdag_mlp_train(tdag_mlp* m, const uint64_t* d_q, const uint64_t* d_p, uint64_t n_pairs, uint32_t steps, uint32_t batch, float lr, float tau, uint64_t seed,
    float* out_losses);
/* L2-normalised embeddings (device, n x 128 fp32) of n device 4^3 masks. */
tdag_status tdag_mlp_embed(tdag_mlp* m, const uint64_t* d_masks, uint64_t n, float* d_out);
/* The bf16 MFMA GEMM of the network, exposed for tests: C (m x n, fp32, row-major, ldc) =
 * op(A) op(B), op(X) = X or X^T of row-major bf16 (raw uint16) device matrices. */
tdag_status tdag_gemm_bf16(tdag_context* ctx, int trans_a, int trans_b, const uint16_t* d_a, int lda, const uint16_t* d_b, int ldb, float* d_c, int ldc, int m, int n, int k);
/* The retrieval step of the learned proposal, exposed for tests: for each of the n rows of bf16
 * embeddings (device, n x 128 raw uint16, row = rank) the k (<= 64) best earlier rows under
 * "higher dot product first (bf16 MFMA, fp32 accumulate), then lower row", best first, into
 * d_cand (device, n x k; 0xFFFFFFFF where fewer than k earlier rows exist). */
tdag_status tdag_mlp_topk(tdag_context* ctx, const uint16_t* d_emb, uint64_t n, uint32_t k, uint32_t* d_cand);
/* How a network's matcher retrieves its proposals. ALL: the k best of all earlier leaves
 * (tdag_mlp_topk). BUCKETED: the k best of the earlier leaves whose transform-invariant
 * descriptor equals the query's (tdag_mlp_topk_bucketed): every match of equal popcount lies in
 * the query's bucket, matches that clip voxels are left to the fallback search, and the work is
 * the sum of the squared bucket sizes instead of n^2 / 2. No reference counterpart. */
#define TDAG_RETRIEVAL_ALL 0u
#define TDAG_RETRIEVAL_BUCKETED 1u
tdag_status tdag_mlp_set_retrieval(tdag_mlp* m, uint32_t mode);
tdag_status tdag_mlp_get_retrieval(const tdag_mlp* m, uint32_t* out_mode);
/* The bucketed retrieval step, exposed for tests: d_masks (device, n 4^3 masks in rank order) and
 * their bf16 embeddings d_emb (as tdag_mlp_topk); for each row the k (<= 64) best earlier rows
 * with the same descriptor under "higher dot product first (fp32 sum of the bf16 products in
 * dimension order), then lower row", best first, into d_cand (0xFFFFFFFF where fewer exist). */
tdag_status tdag_mlp_topk_bucketed(tdag_context* ctx, const uint64_t* d_masks, const uint16_t* d_emb, uint64_t n, uint32_t k, uint32_t* d_cand);
/* Level-2 matcher with proposals from `mlp` (NULL: the k leaves ranked just before each query,
 * i.e. popcount-neighbour proposals without a network). Outputs as tdag_match_level_device. */
tdag_status tdag_match_level_learned_device(tdag_context* ctx, tdag_mlp* mlp, uint32_t level, const uint64_t* d_masks, uint64_t n, const tdag_config* cfg, uint32_t k,
    uint64_t* d_encoded, uint64_t* d_mapping, uint32_t* d_unique, tdag_learned_stats* stats);
/* Host-buffer form of tdag_match_level_learned_device (as tdag_match_level). */
tdag_status tdag_match_level_learned(tdag_context* ctx, tdag_mlp* mlp, uint32_t level, const uint64_t* masks, uint64_t n, const tdag_config* cfg, uint32_t k,
    uint64_t* out_encoded, uint64_t* out_mapping, uint32_t* out_unique, uint64_t* out_num_unique, tdag_learned_stats* stats);

/* ------------------------------------------------------------------------------------------
 * Pipeline around the matcher: scene synthesis, SVDAG producer, flattening, the
 * constructTransformDAG driver and the .tdag writer.
 * ---------------------------------------------------------------------------------------- */

/* An SVDAG: voxcom::EditStructure<void, uint32_t> after toDAG (structure.h:159-211, 703-763). */
typedef struct tdag_svdag tdag_svdag;
/* A Transform-Aware SVDAG: voxcom::EditStructure<void, TransformPointer> plus per-level
 * matcher records. */
typedef struct tdag_tdag tdag_tdag;

/* Per-level record of a tdag_construct_transform_dag run. */
typedef struct tdag_level_record {
    uint32_t level;
    uint32_t has_masks; /* the matcher input masks were kept (TDAG_KEEP_LEVEL_INPUTS) */
    uint64_t items; /* matcher input size: nodes referenced by surviving parents */
    uint64_t unique;
    uint64_t matched;
    float ms_match;
    uint32_t learned; /* level 2 of tdag_construct_transform_dag_learned: 1 */
    uint64_t proposal_hits; /* learned: queries resolved by a verified proposal */
    uint64_t fallback_queries; /* learned: queries sent to the exhaustive search */
} tdag_level_record;

/* Flag of tdag_construct_transform_dag: keep each level's matcher input masks. */
#define TDAG_KEEP_LEVEL_INPUTS 1u
/* Flag of tdag_construct_transform_dag: run the translation stage with every step on the device
 * (referenced-node marking and compaction, flattening, matching, emission of the unique nodes,
 * rewriting of the parents' pointers; levels gated by stream order); only each level's counts come
 * back to the host. Same result. Needs a context; not with TDAG_KEEP_LEVEL_INPUTS. */
#define TDAG_DEVICE_TRANSLATION 2u

/* Seeded synthetic scene as a dense bit grid (bit i = x + R*y + R*R*z, the VoxelGrid layout of
 * voxel_grid.h:70-73). kind 0: 200 sphere shells + 100 box shells; kind 1: kind 0 plus a
 * fractal value-noise heightfield shell; kind 2 (R >= 256): a city of box-shell buildings with
 * window voids on 64-voxel lots, half of its 256-voxel blocks mirrored / axis-swapped copies of
 * others, over a noise ground (the configs directory describes it). out_bytes >= resolution^3 / 8. */
tdag_status tdag_scene_dense(uint32_t kind, uint32_t resolution, uint64_t seed, uint8_t* out_bits, uint64_t out_bytes);

/* EditStructure<void, uint32_t>(VoxelGrid) followed by toDAG() (structure.h:423-429, 658-763). */
tdag_status tdag_svdag_from_dense(const uint8_t* bits, uint32_t resolution, tdag_svdag** out);
/* The same SVDAG, node for node, built on the device from a dense grid already in device memory
 * (d_bits: resolution^3 / 8 bytes, bit i = x + R*y + R*R*z). Leaves are the non-empty 4^3 blocks in
 * Morton order and every level keeps the first occurrence of each node, which is the index order
 * of the reference's DFS construction and toDAG. Work runs on ctx's stream; out_ms (optional)
 * receives the wall time of the call, including the copy of the result to host memory. */
tdag_status tdag_svdag_from_dense_device(tdag_context* ctx, const uint8_t* d_bits, uint32_t resolution, tdag_svdag** out, float* out_ms);
/* resolution, root level and the node count of levels [0, cap) (level 2 = leaves). */
tdag_status tdag_svdag_info(const tdag_svdag* svdag, uint32_t* resolution, uint32_t* root_level, uint64_t* level_sizes, uint32_t cap);
/* Copies level `level`: 8 uint32 children per node (0xFFFFFFFF = empty), or for level 2 the
 * uint64 Morton leaf masks. */
tdag_status tdag_svdag_copy_level(const tdag_svdag* svdag, uint32_t level, void* out);
/* FlatGridGenerator / createLargeSubGrid (transform_dag.cpp:44-116, large_sub_grid.h:127-141):
 * linear LargeSubGrid masks (W words each) of the given level-`level` nodes. */
tdag_status tdag_svdag_flatten(const tdag_svdag* svdag, uint32_t level, const uint32_t* nodes, uint64_t n, uint64_t* out_masks);
/* The same masks written on the device (d_nodes, d_out_masks in device memory; n * W words out),
 * the input of tdag_match_level_device. The SVDAG's levels 2..5 are copied to the device on first
 * use and kept with the handle. A node index outside the level gives TDAG_EINVAL. Returns after
 * the work on ctx's stream is done. tdag_construct_transform_dag flattens its exhaustive levels
 * this way. */
tdag_status tdag_svdag_flatten_device(tdag_context* ctx, const tdag_svdag* svdag, uint32_t level, const uint32_t* d_nodes, uint64_t n, uint64_t* d_out_masks);
tdag_status tdag_svdag_free(tdag_svdag* svdag);

/* voxcom::EditStructure<void, uint32_t> / EditStructureOOC<void, uint32_t> (structure.h:133-171), the
 * input type of constructTransformDAG (transform_dag.h:161-162), as borrowed arrays. For level l in
 * [3, root_level], nodes_per_level[l] holds node_counts[l] EditNode<uint32_t> (8 uint32 child
 * indices into level l - 1, 0xFFFFFFFF = EmptyChild; leaf indices at level 3); entries 0..2 are
 * not read. subgrids: num_subgrids EditSubGrid<void> bitmasks (4x4x4, Morton bit order,
 * structure.h:116-125). structure_type: voxcom::StructureType (0 Tree, 1 DAG). */
typedef struct tdag_svdag_view {
    uint32_t resolution;
    uint32_t root_level;
    const uint32_t* const* nodes_per_level;
    const uint64_t* node_counts;
    const uint64_t* subgrids;
    uint64_t num_subgrids;
    int32_t structure_type;
    uint32_t reserved;
} tdag_svdag_view;

/* Copies an EditStructure view into a tdag_svdag (the handle tdag_construct_transform_dag takes).
 * Validated: resolution == 2^root_level with root_level in [3, 12], structure_type 0 or 1, every
 * child index EmptyChild or inside the level below, no null array with a non-zero count;
 * TDAG_EINVAL otherwise. The driver then requires structure_type == 1 (DAG), as the reference
 * asserts (transform_dag.cpp:141). */
tdag_status tdag_svdag_from_view(const tdag_svdag_view* view, tdag_svdag** out);

/* constructTransformDAG (transform_dag.h:161-162, transform_dag.cpp:138-240): hierarchical
 * sym+perm dedup above max_translation_level, then one GPU matcher call per level from
 * max_translation_level down to 2. max_translation_level must lie in [2, 5]. ctx may be NULL
 * when cfg->translation is 0 (no matcher call is made). */
tdag_status tdag_construct_transform_dag(tdag_context* ctx, const tdag_svdag* in, const tdag_config* cfg, uint32_t flags, tdag_tdag** out);
/* constructTransformDAG with the learned proposal at level 2 (mlp; k proposals per query) and the
 * exhaustive matcher at levels 3..maxT. Every match is exact; the output can differ from
 * tdag_construct_transform_dag where a query's minimum-key parent was not proposed. */
tdag_status tdag_construct_transform_dag_learned(tdag_context* ctx, const tdag_svdag* in, const tdag_config* cfg, uint32_t flags, tdag_mlp* mlp, uint32_t k,
    tdag_tdag** out);
/* A build's translation stage as one captured hipGraph. create runs the device hierarchical pass,
 * then the device translation stage once (learning each level's sizes), then captures the whole
 * stage (every level's marking, compaction, flattening, matching, emission and pointer rewrites;
 * no host synchronisation inside) into a hipGraph on ctx's stream. run replays it (out_ms: device
 * time of the replay); result returns the structure of the last run, the same as
 * tdag_construct_transform_dag's. The graph keeps its own device copy of the SVDAG's levels (the
 * handle may be freed after create); the context must outlive the graph. No reference
 * counterpart: the reference rebuilds its matcher inputs on the host between levels
 * (transform_dag.cpp:165-233). */
typedef struct tdag_build_graph tdag_build_graph;
tdag_status tdag_build_graph_create(tdag_context* ctx, const tdag_svdag* in, const tdag_config* cfg, tdag_build_graph** out);
tdag_status tdag_build_graph_run(tdag_build_graph* g, float* out_ms);
tdag_status tdag_build_graph_result(tdag_build_graph* g, tdag_tdag** out);
tdag_status tdag_build_graph_destroy(tdag_build_graph* g);

/* The .tdag byte stream (voxelizer/src/main.cpp:144-151): pass buf = NULL to query the size; cfg
 * must be the config the structure was built with (TDAG_EINVAL otherwise). The
 * header's structureType is 0 (Tree), as in the reference, whose output structure is
 * default-constructed (transform_dag_hierarchical.cpp:307, structure.h:171). */
tdag_status tdag_tdag_serialize(const tdag_tdag* tdag, const tdag_config* cfg, uint8_t* buf, uint64_t cap, uint64_t* out_size);
tdag_status tdag_tdag_info(const tdag_tdag* tdag, uint32_t* root_level, uint64_t* level_sizes, uint32_t cap);
tdag_status tdag_tdag_level_record(const tdag_tdag* tdag, uint32_t level, tdag_level_record* out);
/* The nodes of `level` of a constructTransformDAG result, the EditStructure<void, TransformPointer>
 * the reference returns: for level >= 3, 8 TransformPointer bit patterns per node (sentinel
 * 0x82082400FFFFFFFF = empty child); for level 2, the uint64 leaf masks. cap in uint64 words;
 * *out_count receives the word count (out may be NULL to query it). */
tdag_status tdag_tdag_copy_level(const tdag_tdag* tdag, uint32_t level, uint64_t* out, uint64_t cap, uint64_t* out_count);
/* Copies the kept matcher input masks of `level` (items * W words). */
tdag_status tdag_tdag_level_masks(const tdag_tdag* tdag, uint32_t level, uint64_t* out, uint64_t cap_words);
tdag_status tdag_tdag_free(tdag_tdag* tdag);

/* ------------------------------------------------------------------------------------------
 * .tdag16 compact encoding (tools/svdag_encoding: TransformDAG16, transform_dag_encoding.h:138-153).
 * ------------------------------------------------------------------------------------------ */
typedef struct tdag_tdag16 tdag_tdag16;
#define TDAG16_POINTER_TABLES 1u /* TransformEncodingConfig::pointerTables */
#define TDAG16_HUFFMAN 2u /* TransformEncodingConfig::huffmanCode */
/* constructTransformDAG16(EditStructure<void, TransformPointer>, ...) (transform_dag_encoding.cpp:
 * 241-478): nodes ordered by reference count with unreferenced nodes pruned, 16-bit node words,
 * optional pointer tables and per-level Huffman code of the pointer types. Pointers of equal
 * reference count keep the order of their first use (the reference takes an unordered_map's
 * order there), so the bytes can differ from the reference's while the volume is the same.
 * Takes the result of tdag_construct_transform_dag (not of the static driver) and the config it
 * was built with (another config gives TDAG_EINVAL: the pointer types would not describe the
 * pointers). */
tdag_status tdag_tdag16_encode(const tdag_tdag* tdag, const tdag_config* cfg, uint32_t flags, tdag_tdag16** out);
/* TransformDAG16::readFrom of a .tdag16 byte stream (transform_dag_encoding.cpp:1029-1036). */
tdag_status tdag_tdag16_from_bytes(const uint8_t* bytes, uint64_t len, tdag_tdag16** out);
/* TransformDAG16::writeTo (transform_dag_encoding.cpp:1020-1027): pass buf = NULL to query the size. */
tdag_status tdag_tdag16_serialize(const tdag_tdag16* h, uint8_t* buf, uint64_t cap, uint64_t* out_size);
/* Resolution, root level, and for levels [0, cap): 16-bit words per level (level 2: 4 per leaf)
 * and pointer-table entries per child level. Any output pointer may be NULL. */
tdag_status tdag_tdag16_info(const tdag_tdag16* h, uint32_t* resolution, uint32_t* root_level, uint64_t* level_words, uint64_t* table_entries, uint32_t cap);
/* The voxel lookup TranslateTransformBaseDAG::get (transform_dag_encoding.cpp:897-942) for every
 * voxel, on the device: writes the dense grid (bit i = x + R*y + R*R*z; resolution^3 / 32 uint32
 * words, d_out_words in device memory). The encoded DAG is copied to the device on first use and
 * kept with the handle. Malformed data (an index outside its level, table or leaf array) gives
 * TDAG_EINVAL. out_ms (optional): device time of the decode kernel. Returns after the work on
 * ctx's stream is done. */
tdag_status tdag_tdag16_decode_device(tdag_context* ctx, const tdag_tdag16* h, uint32_t* d_out_words, uint64_t out_words, float* out_ms);
/* Ray casting through the encoded DAG on the device, the renderer's use of the format
 * (renderer/src/dags/transform_dag/transform_dag.cu:5-41, tracer_impl.h:874-1010): for each ray the
 * first voxel it enters, found with the same voxel lookup as tdag_tdag16_decode_device.
 *   d_rays  device, n x 8 floats: origin x y z, tmin, direction x y z, tmax (voxel units; the
 *           volume is [0, resolution)^3)
 *   d_hits  device, n x 4 int32: the hit voxel x y z (x = -1: no voxel in [tmin, tmax]), then the
 *           number of steps the ray took (-1 if it exceeded its budget of 4 * resolution + 64)
 *   d_t     device, n floats: the distance at which the ray enters the hit voxel
 * Malformed data gives TDAG_EINVAL. out_ms (optional): device time of the call. Synchronises. */
tdag_status tdag_tdag16_trace_device(tdag_context* ctx, const tdag_tdag16* h, const float* d_rays, uint64_t n, int32_t* d_hits, float* d_t, float* out_ms);
tdag_status tdag_tdag16_free(tdag_tdag16* h);
/* createLengthLimitedHuffmanCodeLengths (huffman.cpp:113-183): package-merge code lengths of at
 * most max_len bits for n symbols. */
tdag_status tdag_huffman_code_lengths(const uint64_t* freq, uint32_t n, uint32_t max_len, uint32_t* out_len);
/* assignCodeWords (huffman.cpp:93-111), then convertHuffmanToLSB (192-202) when lsb_first. */
tdag_status tdag_huffman_code_words(const uint32_t* len, uint32_t n, int lsb_first, uint64_t* out_code);

/* ------------------------------------------------------------------------------------------
 * StaticStructure form: constructStaticTransformDAG (transform_dag.h:163).
 * ---------------------------------------------------------------------------------------- */

/* voxcom::StaticStructure<void, uint32_t> (structure.h:299-320) as borrowed arrays. For level l in
 * [3, root_level], level_words[l] holds level_word_counts[l] words: per node its child mask
 * (bits 0..7), then one word per present child in child-index order, the child's word offset in
 * level l - 1 (a leaf index at level 3). Entries 0..2 of both arrays are not read. Leaves are
 * 4x4x4 masks in Morton bit order. resolution = 2^root_level. */
typedef struct tdag_static_svdag_view {
    uint32_t resolution;
    uint32_t root_level;
    const uint32_t* const* level_words;
    const uint64_t* level_word_counts;
    const uint64_t* leaves;
    uint64_t num_leaves;
} tdag_static_svdag_view;

/* StaticStructure(const EditStructure&) (structure.h:766-798): the packed words of `level` of an
 * SVDAG as uint32 (level 2: its leaf masks as uint64). *out_count receives the count; out may be
 * NULL to query it. */
tdag_status tdag_svdag_static_words(const tdag_svdag* svdag, uint32_t level, void* out, uint64_t cap, uint64_t* out_count);

/* constructStaticTransformDAG (transform_dag.h:163, transform_dag.cpp:242-390;
 * constructStaticTransformDAGHierarchical at transform_dag_hierarchical.cpp:531-743). Same
 * matcher calls as tdag_construct_transform_dag on the same SVDAG; the hierarchical levels above
 * max_translation_level + 1 come out in the static pass's own order. The input is validated
 * (masks within 8 bits, nodes inside their level, every child word a node offset or leaf index of
 * the level below: TDAG_EINVAL otherwise). The result is read with tdag_tdag_info,
 * tdag_tdag_level_record, tdag_tdag_static_words and tdag_tdag_static_serialize. */
tdag_status tdag_construct_static_transform_dag(tdag_context* ctx, const tdag_static_svdag_view* in, const tdag_config* cfg, uint32_t flags, tdag_tdag** out);
/* StaticStructure<void, TransformPointer> words of `level` (level 2: the leaf masks); children are
 * TransformPointer bit patterns whose ptr is a word offset in the level below (a leaf index at
 * level 3). out may be NULL to query the count. */
tdag_status tdag_tdag_static_words(const tdag_tdag* tdag, uint32_t level, uint64_t* out, uint64_t cap, uint64_t* out_count);
/* StaticStructure<void, TransformPointer>::writeTo bytes (structure.h:888-895, binary_writer.h:48-78):
 * per level u64 word count + words, u64 leaf count + leaves, u32 resolution, u32 rootLevel,
 * i32 structureType (0, Tree: the reference's output is default-constructed,
 * transform_dag_hierarchical.cpp:533). buf = NULL queries the size. */
tdag_status tdag_tdag_static_serialize(const tdag_tdag* tdag, uint8_t* buf, uint64_t cap, uint64_t* out_size);

#ifdef __cplusplus
}
#endif

#endif /* TDAG_H_ */
