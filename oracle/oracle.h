// TEST INFRASTRUCTURE ONLY: C entry points of the CPU oracle (restatement of the reference's
// Transform-Aware SVDAG path). Used by tests/, the smoke check and bench.py's cpu_baseline.
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

// Byte layout of voxcom::TransformDAGConfig (transform_dag.h:157-160) under GCC x86-64:
// three bools, one padding byte, then uint32 maxTranslationLevel.
typedef struct orc_config {
    uint8_t symmetry;
    uint8_t axisPermutation;
    uint8_t translation;
    uint8_t pad;
    uint32_t maxTranslationLevel;
} orc_config;

uint64_t orc_tp_create(uint64_t ptr, uint32_t symBits, uint32_t permCode, int tx, int ty, int tz);
uint64_t orc_tp_sentinel(void);
uint64_t orc_encode_fixed64(uint64_t tpBits, uint32_t level);
uint64_t orc_decode_fixed64(uint64_t key, uint32_t level);
int orc_transform(uint32_t level, const uint64_t* in, uint32_t permCode, uint32_t symBits, int tx, int ty, int tz, uint64_t* out);
int orc_corner_hash(uint32_t level, const uint64_t* in, uint32_t dirBits, uint64_t* outHash);
int orc_match_encoded(uint32_t level, const uint64_t* masks, uint64_t n, const orc_config* cfg, uint64_t* outEncoded, int nthreads, int brute);
int orc_match_timed(uint32_t level, const uint64_t* masks, uint64_t n, const orc_config* cfg, int nthreads, uint32_t queryStride, double* outSetupS,
    double* outSearchS, uint64_t* outQueries);
int orc_postprocess(uint32_t level, const uint64_t* masks, uint64_t n, const uint64_t* encoded, uint64_t* outMapping, uint32_t* outUnique, uint64_t* outNumUnique);

/* .tdag16 compact encoding (oracle_tdag16.cpp). */
int orc_huffman_code_lengths(const uint64_t* freq, uint32_t n, uint32_t maxLen, uint32_t* outLen);
int orc_huffman_code_words(const uint32_t* len, uint32_t n, int lsbFirst, uint64_t* outCode);
/* First-hit ray casting through a dense grid (oracle_rays.cpp): rays n x 8 floats (origin, tmin,
 * direction, tmax), hits n x 3 int32 (-1: none), t the entry distance (double). */
int orc_ray_march(const uint8_t* bits, uint32_t R, const float* rays, uint64_t n, int32_t* outHit, double* outT);
int64_t orc_tdag16_decode(const uint8_t* bytes, uint64_t len, uint8_t* outBits, uint64_t outCap);

#ifdef __cplusplus
}
#endif
