// Learned proposal for the level-2 matcher on gfx950 (no reference counterpart).
//
//  * Features: 64 mask bits and 12 transform-invariant descriptor bytes (per-axis occupancy
//    histograms, each trimmed of zero ends and left-aligned, replaced by the smaller of itself
//    and its reverse, the three sorted: invariant under the 48 mirror/permutation transforms and
//    unclipped translations), as bf16 rows of 80 (zero-padded).
//  * MLP 80 -> 256 -> 256 -> 256 -> 128 (ReLU), L2-normalised. Every dense layer, forward and
//    backward, is one bf16 MFMA GEMM (v_mfma_f32_16x16x32_bf16, fp32 accumulate, 64x64 tiles).
//  * Training: InfoNCE over (query, matched parent) pairs emitted by the exact matcher, Adam on
//    fp32 master weights; the similarity/softmax part runs in fp32 on the vector units.
//  * Matching: proposals = the k most similar earlier-ranked leaves (bf16 MFMA scores, per-query
//    top-k heaps); every proposal is verified bitwise against all 48 transforms and the possible
//    translations; queries with no verified proposal go to the exhaustive search with only
//    them in the child table. Every accepted match is therefore exact.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <random>
#include <vector>

#include "common.hpp"
#include "learned.hpp"
#include "primitives.hpp"
#include "search_device.hpp"

namespace tdag {
namespace learn {

namespace {

using namespace dev;

typedef short bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 mfma_bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// gfx950's double-K bf16 MFMA: D (16 x 16, fp32) += A (16 x 32) B (32 x 16).
__device__ __forceinline__ f32x4 mfma32(bf16x8 a, bf16x8 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(mfma_bf16x8, a), __builtin_bit_cast(mfma_bf16x8, b), c, 0, 0, 0);
}

#define LEARN_TRY(expr)                  \
    do {                                 \
        const hipError_t e_ = (expr);    \
        if (e_ != hipSuccess)            \
            return e_;                   \
    } while (0)

inline size_t cdiv(size_t a, size_t b) { return (a + b - 1) / b; }
inline size_t align256(size_t v) { return (v + 255) / 256 * 256; }

// float -> bf16 bits, round to nearest, ties to even (finite inputs).
__host__ __device__ __forceinline__ uint16_t f2bf(float f)
{
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// ---------------------------------------------------------------------------------------
// Features
// ---------------------------------------------------------------------------------------

// Per-axis occupancy histogram of a 4^3 mask along axis a (0: x, 1: y, 2: z).
__device__ __forceinline__ void axisHist(uint64_t m, int a, uint32_t h[4])
{
    for (int v = 0; v < 4; ++v) {
        uint64_t plane;
        if (a == 0)
            plane = rangeMaskX4(v, v + 1);
        else if (a == 1)
            plane = rangeMaskY4(v, v + 1);
        else
            plane = rangeMaskZ4(v, v + 1);
        h[v] = (uint32_t)__popcll(m & plane);
    }
}

// Lexicographic comparison of two 4-vectors.
__device__ __forceinline__ bool lexLess(const uint32_t* a, const uint32_t* b)
{
    for (int i = 0; i < 4; ++i)
        if (a[i] != b[i])
            return a[i] < b[i];
    return false;
}

// The transform-invariant descriptor of a 4^3 mask: per axis, the occupancy histogram trimmed of
// zero ends, left-aligned, and the smaller (lexicographically) of it and its reverse; the three
// histograms sorted. Invariant under the 48 permutations and mirrors, and under translations that
// lose no voxel.
__device__ __forceinline__ void canonicalHists(uint64_t m, uint32_t h[3][4])
{
    for (int a = 0; a < 3; ++a) {
        uint32_t raw[4];
        axisHist(m, a, raw);
        int lo = 4, hi = -1;
        for (int v = 0; v < 4; ++v)
            if (raw[v]) {
                lo = lo < v ? lo : v;
                hi = v;
            }
        uint32_t fwd[4] = { 0, 0, 0, 0 }, rev[4] = { 0, 0, 0, 0 };
        for (int v = lo; v <= hi; ++v) {
            fwd[v - lo] = raw[v];
            rev[hi - v] = raw[v];
        }
        const bool useRev = lexLess(rev, fwd);
        for (int i = 0; i < 4; ++i)
            h[a][i] = useRev ? rev[i] : fwd[i];
    }
    // sort the three canonical histograms (3-element sorting network)
    auto swapIf = [&](int i, int j) {
        if (lexLess(h[j], h[i]))
            for (int k = 0; k < 4; ++k) {
                const uint32_t t = h[i][k];
                h[i][k] = h[j][k];
                h[j][k] = t;
            }
    };
    swapIf(0, 1);
    swapIf(1, 2);
    swapIf(0, 1);
}

__global__ __launch_bounds__(256) void kFeatures(const uint64_t* __restrict__ masks, uint32_t n, uint16_t* __restrict__ X)
{
    const uint32_t r = blockIdx.x * 256u + threadIdx.x;
    if (r >= n)
        return;
    const uint64_t m = masks[r];
    uint16_t* row = X + (size_t)r * kIn;
    for (int i = 0; i < 64; ++i)
        row[i] = ((m >> i) & 1ull) ? 0x3F80u : 0u; // 1.0 or 0.0
    uint32_t h[3][4];
    canonicalHists(m, h);
    for (int a = 0; a < 3; ++a)
        for (int i = 0; i < 4; ++i)
            row[64 + 4 * a + i] = f2bf((float)h[a][i] * (1.0f / 16.0f));
    for (int i = 76; i < kIn; ++i)
        row[i] = 0u;
}

__global__ __launch_bounds__(256) void kGatherPairs(const uint64_t* __restrict__ q, const uint64_t* __restrict__ p, const uint32_t* __restrict__ idx, uint32_t B,
    uint64_t* __restrict__ out)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i < B) {
        out[i] = q[idx[i]];
        out[B + i] = p[idx[i]];
    }
}

// ---------------------------------------------------------------------------------------
// bf16 MFMA GEMM: C[M x N] (fp32) = opA[M x K] * opB[K x N].
//   opA[m][k] = TA ? A[k * lda + m] : A[m * lda + k];  opB[k][n] = TB ? B[n * ldb + k] : B[k * ldb + n].
// 256 threads per 64x64 tile of C; waves in 2x2, each 32x32 = 2x2 v_mfma_f32_16x16x32_bf16 tiles.
// Each 32-deep K slice of both operands is staged in LDS (B transposed so a lane's eight k values
// are contiguous); out-of-range elements read as 0 and out-of-range C is not written.
// MFMA 16x16x32 operand layout: lane l holds A[l % 16][8 (l / 16) + i] and B[8 (l / 16) + i][l % 16]
// for i < 8, and receives D[4 (l / 16) + i][l % 16] for i < 4.
// ---------------------------------------------------------------------------------------
constexpr int kTile = 64;
constexpr int kKs = 32;
constexpr int kLdsPitch = kKs + 4; // bf16 elements per staged row (padding against bank conflicts)

// Stages one operand's 64 x 32 slice (rows r0.., k0..) into s[row][k]. RowFast: the row index
// is contiguous in memory (element (row, k) at p[k * ld + row]); thread t takes row t % 64 and
// k = 8 (t / 64) + i, so a wave reads 64 consecutive elements. Otherwise k is contiguous
// (p[row * ld + k]) and thread t takes row t / 4 and k = 8 (t % 4) + i.
template <bool RowFast>
__device__ __forceinline__ void stageSlice(const uint16_t* __restrict__ p, int ld, int r0, int rows, int k0, int kEnd, uint16_t (*sm)[kLdsPitch], int t)
{
    const int row = RowFast ? (t & 63) : (t >> 2);
    const int kb = RowFast ? 8 * (t >> 6) : 8 * (t & 3);
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const int r = r0 + row, k = k0 + kb + i;
        uint16_t v = 0;
        if (r < rows && k < kEnd)
            v = RowFast ? p[(size_t)k * ld + r] : p[(size_t)r * ld + k];
        sm[row][kb + i] = v;
    }
}

// C tile (or, with split K, the partial product of K range [kBegin, kEnd) of slice blockIdx.z,
// written to C + blockIdx.z * sliceStride) of the 64 x 64 output tile (blockIdx.y, blockIdx.x).
template <bool TA, bool TB>
__global__ __launch_bounds__(256) void kGemm(const uint16_t* __restrict__ A, int lda, const uint16_t* __restrict__ B, int ldb, float* __restrict__ C, int ldc, int M, int N,
    int K, int kChunk, size_t sliceStride)
{
    __shared__ __align__(16) uint16_t sA[kTile][kLdsPitch];
    __shared__ __align__(16) uint16_t sB[kTile][kLdsPitch]; // sB[n][k]
    const int t = (int)threadIdx.x;
    const int lane = t & 63, wave = t >> 6;
    const int m0 = (int)blockIdx.y * kTile, n0 = (int)blockIdx.x * kTile;
    const int wr = wave >> 1, wc = wave & 1;
    const int kBegin = (int)blockIdx.z * kChunk, kEnd = min(K, kBegin + kChunk);
    C += (size_t)blockIdx.z * sliceStride;
    f32x4 acc[2][2];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            acc[i][j] = f32x4 { 0.f, 0.f, 0.f, 0.f };
    for (int k0 = kBegin; k0 < kEnd; k0 += kKs) {
        stageSlice<TA>(A, lda, m0, M, k0, kEnd, sA, t);
        stageSlice<!TB>(B, ldb, n0, N, k0, kEnd, sB, t);
        __syncthreads();
        const int kq = 8 * (lane >> 4);
        bf16x8 a[2], b[2];
        for (int i = 0; i < 2; ++i) {
            const uint16_t* pa = &sA[wr * 32 + i * 16 + (lane & 15)][kq];
            const uint16_t* pb = &sB[wc * 32 + i * 16 + (lane & 15)][kq];
            for (int e = 0; e < 8; ++e) {
                a[i][e] = (short)pa[e];
                b[i][e] = (short)pb[e];
            }
        }
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j)
                acc[i][j] = mfma32(a[i], b[j], acc[i][j]);
        __syncthreads();
    }
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            for (int r = 0; r < 4; ++r) {
                const int m = m0 + wr * 32 + i * 16 + 4 * (lane >> 4) + r;
                const int nn = n0 + wc * 32 + j * 16 + (lane & 15);
                if (m < M && nn < N)
                    C[(size_t)m * ldc + nn] = acc[i][j][r];
            }
}

// C[m][n] = sum of the `slices` partial products in slice order (deterministic split K).
__global__ __launch_bounds__(256) void kSumSlices(const float* __restrict__ P, int slices, int M, int N, float* __restrict__ C, int ldc)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    const size_t mn = (size_t)M * N;
    if (i >= mn)
        return;
    float acc = 0.f;
    for (int z = 0; z < slices; ++z)
        acc += P[(size_t)z * mn + i];
    C[(i / N) * (size_t)ldc + i % N] = acc;
}

template <bool TA, bool TB>
void launchGemm(dim3 grid, const uint16_t* A, int lda, const uint16_t* B, int ldb, float* C, int ldc, int M, int N, int K, int kChunk, size_t sliceStride, hipStream_t s)
{
    hipLaunchKernelGGL((kGemm<TA, TB>), grid, dim3(256), 0, s, A, lda, B, ldb, C, ldc, M, N, K, kChunk, sliceStride);
}

// Split K applies when the output has few tiles and K is long (the weight gradients, K = batch
// rows): kSplitSlices slices of K, partial products in `splitWork` (kSplitSlices x M x N floats).
constexpr int kSplitSlices = 8;
constexpr size_t kSplitWorkFloats = (size_t)kSplitSlices * 256 * 256;

hipError_t gemm(bool ta, bool tb, const uint16_t* A, int lda, const uint16_t* B, int ldb, float* C, int ldc, int M, int N, int K, hipStream_t s,
    float* splitWork = nullptr)
{
    if (M <= 0 || N <= 0)
        return hipSuccess;
    const int tiles = (int)(cdiv(N, kTile) * cdiv(M, kTile));
    // Split K when few output tiles would leave the GPU idle and each slice keeps >= 128 of K.
    const bool split = splitWork && tiles < 64 && K >= 128 * kSplitSlices && (size_t)M * N * kSplitSlices <= kSplitWorkFloats;
    const int slices = split ? kSplitSlices : 1;
    const int kChunk = split ? (int)(cdiv(cdiv(K, kSplitSlices), kKs) * kKs) : std::max(K, 1);
    float* out = split ? splitWork : C;
    const int ldo = split ? N : ldc;
    const size_t stride = split ? (size_t)M * N : 0;
    const dim3 grid((unsigned)cdiv(N, kTile), (unsigned)cdiv(M, kTile), (unsigned)slices);
    if (!ta && !tb)
        launchGemm<false, false>(grid, A, lda, B, ldb, out, ldo, M, N, K, kChunk, stride, s);
    else if (ta && !tb)
        launchGemm<true, false>(grid, A, lda, B, ldb, out, ldo, M, N, K, kChunk, stride, s);
    else if (!ta && tb)
        launchGemm<false, true>(grid, A, lda, B, ldb, out, ldo, M, N, K, kChunk, stride, s);
    else
        launchGemm<true, true>(grid, A, lda, B, ldb, out, ldo, M, N, K, kChunk, stride, s);
    LEARN_TRY(hipGetLastError());
    if (split) {
        hipLaunchKernelGGL(kSumSlices, dim3((unsigned)cdiv((size_t)M * N, 256)), dim3(256), 0, s, splitWork, slices, M, N, C, ldc);
        LEARN_TRY(hipGetLastError());
    }
    return hipSuccess;
}

// ---------------------------------------------------------------------------------------
// Elementwise parts of the forward and backward passes
// ---------------------------------------------------------------------------------------

// out = C + bias, optionally ReLU; written as bf16 (hidden layers) and/or fp32 (last layer).
__global__ __launch_bounds__(256) void kBiasAct(const float* __restrict__ C, const float* __restrict__ bias, uint32_t M, uint32_t N, int relu, uint16_t* __restrict__ outBf,
    float* __restrict__ outF)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (size_t)M * N)
        return;
    float v = C[i] + bias[i % N];
    if (relu)
        v = v > 0.f ? v : 0.f;
    if (outBf)
        outBf[i] = f2bf(v);
    if (outF)
        outF[i] = v;
}

// e = z / |z| per row of 128; invNorm[r] = 1 / |z|. One wave per row.
__global__ __launch_bounds__(256) void kNormalize(const float* __restrict__ Z, uint32_t rows, float* __restrict__ E, float* __restrict__ invNorm)
{
    const uint32_t r = blockIdx.x * 4u + (threadIdx.x >> 6);
    const uint32_t lane = threadIdx.x & 63u;
    if (r >= rows)
        return;
    const float a = Z[(size_t)r * kEmb + lane], b = Z[(size_t)r * kEmb + 64 + lane];
    float ss = a * a + b * b;
    for (int off = 32; off > 0; off >>= 1)
        ss += __shfl_xor(ss, off, 64);
    const float inv = 1.0f / fmaxf(sqrtf(ss), 1e-20f);
    E[(size_t)r * kEmb + lane] = a * inv;
    E[(size_t)r * kEmb + 64 + lane] = b * inv;
    if (lane == 0)
        invNorm[r] = inv;
}

// InfoNCE, rows: one block per query i. S_ij = <e_q_i, e_p_j> / tau; loss_i = logsumexp_j S_ij -
// S_ii; G_ij = (softmax_j S_ij - [i == j]) / (B tau) is stored, and dE_q[i] = sum_j G_ij e_p_j.
__global__ __launch_bounds__(256) void kInfoNceRows(const float* __restrict__ E, uint32_t B, float invTau, float* __restrict__ G, float* __restrict__ dE, float* __restrict__ loss)
{
    extern __shared__ float s_dyn[];
    float* S = s_dyn; // B scores
    __shared__ float s_red[256];
    const uint32_t i = blockIdx.x;
    const float* q = E + (size_t)i * kEmb;
    for (uint32_t j = threadIdx.x; j < B; j += 256) {
        const float* p = E + (size_t)(B + j) * kEmb;
        float d = 0.f;
        for (int k = 0; k < kEmb; ++k)
            d += q[k] * p[k];
        S[j] = d * invTau;
    }
    __syncthreads();
    float mx = -INFINITY;
    for (uint32_t j = threadIdx.x; j < B; j += 256)
        mx = fmaxf(mx, S[j]);
    s_red[threadIdx.x] = mx;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off)
            s_red[threadIdx.x] = fmaxf(s_red[threadIdx.x], s_red[threadIdx.x + off]);
        __syncthreads();
    }
    mx = s_red[0];
    __syncthreads();
    float sum = 0.f;
    for (uint32_t j = threadIdx.x; j < B; j += 256)
        sum += expf(S[j] - mx);
    s_red[threadIdx.x] = sum;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off)
            s_red[threadIdx.x] += s_red[threadIdx.x + off];
        __syncthreads();
    }
    sum = s_red[0];
    const float scale = invTau / (float)B;
    for (uint32_t j = threadIdx.x; j < B; j += 256) {
        const float pj = expf(S[j] - mx) / sum;
        G[(size_t)i * B + j] = (pj - (j == i ? 1.f : 0.f)) * scale;
    }
    if (threadIdx.x == 0)
        loss[i] = logf(sum) + mx - S[i];
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < kEmb; k += 256) {
        float d = 0.f;
        for (uint32_t j = 0; j < B; ++j)
            d += G[(size_t)i * B + j] * E[(size_t)(B + j) * kEmb + k];
        dE[(size_t)i * kEmb + k] = d;
    }
}

// InfoNCE, columns: dE_p[j] = sum_i G_ij e_q_i. One block per positive j.
__global__ __launch_bounds__(128) void kInfoNceCols(const float* __restrict__ E, uint32_t B, const float* __restrict__ G, float* __restrict__ dE)
{
    const uint32_t j = blockIdx.x;
    const uint32_t k = threadIdx.x;
    float d = 0.f;
    for (uint32_t i = 0; i < B; ++i)
        d += G[(size_t)i * B + j] * E[(size_t)i * kEmb + k];
    dE[(size_t)(B + j) * kEmb + k] = d;
}

// Through the normalisation: dz = (de - e (e . de)) / |z|; fp32 and bf16 copies. One wave per row.
__global__ __launch_bounds__(256) void kNormBackward(const float* __restrict__ dE, const float* __restrict__ E, const float* __restrict__ invNorm, uint32_t rows,
    float* __restrict__ dZ, uint16_t* __restrict__ dZbf)
{
    const uint32_t r = blockIdx.x * 4u + (threadIdx.x >> 6);
    const uint32_t lane = threadIdx.x & 63u;
    if (r >= rows)
        return;
    const size_t o = (size_t)r * kEmb;
    const float e0 = E[o + lane], e1 = E[o + 64 + lane], d0 = dE[o + lane], d1 = dE[o + 64 + lane];
    float dot = e0 * d0 + e1 * d1;
    for (int off = 32; off > 0; off >>= 1)
        dot += __shfl_xor(dot, off, 64);
    const float inv = invNorm[r];
    const float z0 = (d0 - e0 * dot) * inv, z1 = (d1 - e1 * dot) * inv;
    dZ[o + lane] = z0;
    dZ[o + 64 + lane] = z1;
    dZbf[o + lane] = f2bf(z0);
    dZbf[o + 64 + lane] = f2bf(z1);
}

// ReLU backward: g = dH * (H > 0) (H = the layer's bf16 ReLU output); fp32 in place and bf16.
__global__ __launch_bounds__(256) void kReluBack(float* __restrict__ dH, const uint16_t* __restrict__ H, size_t count, uint16_t* __restrict__ outBf)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= count)
        return;
    const float v = (H[i] != 0u && (H[i] & 0x8000u) == 0u) ? dH[i] : 0.f;
    dH[i] = v;
    outBf[i] = f2bf(v);
}

// out[c] = sum over rows of X[r][c] (bias gradient). One block of 16 x 64 threads per 64 columns:
// thread (g, c) sums rows g, g + 16, ... (coalesced along c), then the 16 partial sums are added
// in a fixed order, so the result does not depend on scheduling.
__global__ __launch_bounds__(1024) void kColSum(const float* __restrict__ X, uint32_t rows, uint32_t cols, float* __restrict__ out)
{
    __shared__ float s_part[16][64];
    const uint32_t cl = threadIdx.x & 63u, g = threadIdx.x >> 6;
    const uint32_t c = blockIdx.x * 64u + cl;
    float acc = 0.f;
    if (c < cols)
        for (uint32_t r = g; r < rows; r += 16)
            acc += X[(size_t)r * cols + c];
    s_part[g][cl] = acc;
    __syncthreads();
    if (g == 0 && c < cols) {
        float sum = 0.f;
#pragma unroll
        for (int i = 0; i < 16; ++i)
            sum += s_part[i][cl];
        out[c] = sum;
    }
}

__global__ __launch_bounds__(256) void kCastBf(const float* __restrict__ x, size_t n, uint16_t* __restrict__ out)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n)
        out[i] = f2bf(x[i]);
}

__global__ __launch_bounds__(256) void kAdam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v, uint16_t* __restrict__ wbf, size_t n,
    float lr, float b1, float b2, float eps, float c1, float c2)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n)
        return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wi;
    wbf[i] = f2bf(wi);
}

// ---------------------------------------------------------------------------------------
// Workspace of one forward/backward pass over `rows` feature rows.
// ---------------------------------------------------------------------------------------
struct Work {
    uint64_t* masks;
    uint16_t* act[kLayers]; // layer inputs (bf16): features, h1, h2, h3
    float* Z;
    float* E;
    float* invNorm;
    float* C; // GEMM output scratch (rows x 256)
    float* dE;
    float* G; // B x B
    float* loss;
    uint16_t* gbf; // back-propagated signal (bf16)
    uint32_t* idx;
    float* split; // split-K partial products (kSplitWorkFloats)
};

size_t workBytes(size_t rows, size_t B)
{
    size_t s = 0;
    s += align256(rows * 8);
    s += align256(rows * kIn * 2) + 3 * align256(rows * 256 * 2);
    s += 3 * align256(rows * kEmb * 4) + align256(rows * 4);
    s += align256(rows * 256 * 4);
    s += align256(B * B * 4) + align256(B * 4);
    s += align256(rows * 256 * 2);
    s += align256(B * 4);
    s += align256(kSplitWorkFloats * 4);
    return s;
}

Work carve(void* base, size_t rows, size_t B)
{
    char* p = (char*)base;
    auto take = [&](size_t bytes) {
        char* r = p;
        p += align256(bytes);
        return (void*)r;
    };
    Work w {};
    w.masks = (uint64_t*)take(rows * 8);
    w.act[0] = (uint16_t*)take(rows * kIn * 2);
    for (int l = 1; l < kLayers; ++l)
        w.act[l] = (uint16_t*)take(rows * 256 * 2);
    w.Z = (float*)take(rows * kEmb * 4);
    w.E = (float*)take(rows * kEmb * 4);
    w.dE = (float*)take(rows * kEmb * 4);
    w.invNorm = (float*)take(rows * 4);
    w.C = (float*)take(rows * 256 * 4);
    w.G = (float*)take(B * B * 4);
    w.loss = (float*)take(B * 4);
    w.gbf = (uint16_t*)take(rows * 256 * 2);
    w.idx = (uint32_t*)take(B * 4);
    w.split = (float*)take(kSplitWorkFloats * 4);
    return w;
}

hipError_t ensureWork(Mlp& mlp, size_t rows, size_t B)
{
    const size_t need = workBytes(rows, B);
    if (mlp.wsBytes >= need)
        return hipSuccess;
    if (mlp.ws)
        (void)hipFree(mlp.ws);
    mlp.ws = nullptr;
    mlp.wsBytes = 0;
    LEARN_TRY(hipMalloc(&mlp.ws, need));
    mlp.wsBytes = need;
    return hipSuccess;
}

// Forward pass of `rows` rows whose features are in w.act[0]; leaves Z, E, invNorm.
hipError_t forward(const Mlp& mlp, Work& w, uint32_t rows, hipStream_t s)
{
    for (int l = 0; l < kLayers; ++l) {
        const int in = kLayerIn[l], out = kLayerOut[l];
        LEARN_TRY(gemm(false, false, w.act[l], in, mlp.wbf + paramOffsetW(l), out, w.C, out, (int)rows, out, in, s));
        const size_t cnt = (size_t)rows * out;
        if (l + 1 < kLayers)
            hipLaunchKernelGGL(kBiasAct, dim3((unsigned)cdiv(cnt, 256)), dim3(256), 0, s, w.C, mlp.w + paramOffsetB(l), rows, (uint32_t)out, 1, w.act[l + 1], nullptr);
        else
            hipLaunchKernelGGL(kBiasAct, dim3((unsigned)cdiv(cnt, 256)), dim3(256), 0, s, w.C, mlp.w + paramOffsetB(l), rows, (uint32_t)out, 0, nullptr, w.Z);
        LEARN_TRY(hipGetLastError());
    }
    hipLaunchKernelGGL(kNormalize, dim3((unsigned)cdiv(rows, 4)), dim3(256), 0, s, w.Z, rows, w.E, w.invNorm);
    return hipGetLastError();
}

// Loss and gradients (into mlp.g) for 2B rows of features in w.act[0].
hipError_t lossGrad(Mlp& mlp, Work& w, uint32_t B, float tau, hipStream_t s)
{
    const uint32_t rows = 2 * B;
    LEARN_TRY(forward(mlp, w, rows, s));
    const size_t shm = (size_t)B * 4;
    hipLaunchKernelGGL(kInfoNceRows, dim3(B), dim3(256), shm, s, w.E, B, 1.f / tau, w.G, w.dE, w.loss);
    LEARN_TRY(hipGetLastError());
    hipLaunchKernelGGL(kInfoNceCols, dim3(B), dim3(128), 0, s, w.E, B, w.G, w.dE);
    LEARN_TRY(hipGetLastError());
    // g (fp32) lives in w.C, its bf16 copy in w.gbf
    hipLaunchKernelGGL(kNormBackward, dim3((unsigned)cdiv(rows, 4)), dim3(256), 0, s, w.dE, w.E, w.invNorm, rows, w.C, w.gbf);
    LEARN_TRY(hipGetLastError());
    for (int l = kLayers - 1; l >= 0; --l) {
        const int in = kLayerIn[l], out = kLayerOut[l];
        // dW_l = act_l^T g ; db_l = column sums of g
        LEARN_TRY(gemm(true, false, w.act[l], in, w.gbf, out, mlp.g + paramOffsetW(l), out, in, out, (int)rows, s, w.split));
        hipLaunchKernelGGL(kColSum, dim3((unsigned)cdiv(out, 64)), dim3(1024), 0, s, w.C, rows, (uint32_t)out, mlp.g + paramOffsetB(l));
        LEARN_TRY(hipGetLastError());
        if (l > 0) {
            // dH = g W_l^T, masked by the ReLU of act_l; becomes the next g
            LEARN_TRY(gemm(false, true, w.gbf, out, mlp.wbf + paramOffsetW(l), out, w.C, in, (int)rows, in, out, s));
            const size_t cnt = (size_t)rows * in;
            hipLaunchKernelGGL(kReluBack, dim3((unsigned)cdiv(cnt, 256)), dim3(256), 0, s, w.C, w.act[l], cnt, w.gbf);
            LEARN_TRY(hipGetLastError());
        }
    }
    return hipSuccess;
}

// ---------------------------------------------------------------------------------------
// Proposals and verification (level 2, ranks in popcount-descending order)
// ---------------------------------------------------------------------------------------

// Retrieval: for query q (rank order) the k best earlier leaves c < q under the strict order
// "higher score first, then lower rank", where score = <E_q, E_c> of the bf16 embeddings
// (MFMA, fp32 accumulate). Query block qb holds queries [64 qb, 64 qb + 64) and scores candidate
// tiles 0..qb (64 leaves each); those tiles are cut into chunks of `ct` tiles and every
// (query block, chunk) pair is one block of kTopKPartial, so the triangular total is spread
// evenly over the GPU. Each query's 64 best so far are a wave queue: one 64-bit key per lane,
// sorted descending across the lanes; a tile's candidates that beat the queue's last key are
// merged in with a bitonic sort and merge across the lanes. kTopKMerge merges a query's chunk
// queues the same way. Keys order (score, rank) exactly, so the proposal set is the top k under
// that order, independent of the chunking and of thread timing.
constexpr int kQ = 64;
constexpr int kChasePasses = 3; // parent-chase passes after the fallback search

// Blocks of query blocks 0..qb-1: sum over m = 1..qb of ceil(m / ct).
__host__ __device__ __forceinline__ uint64_t chunksBefore(uint32_t qb, uint32_t ct)
{
    const uint64_t g = qb / ct, rem = qb % ct;
    return (uint64_t)ct * g * (g + 1) / 2 + rem * (g + 1);
}

// Chunk length: at least 32 tiles, and at most 16 chunks per query block (bounds the merge).
constexpr uint32_t kMaxChunks = 16;
__host__ __device__ __forceinline__ uint32_t chunkTiles(uint32_t qblocks) { return qblocks / kMaxChunks + 32u; }

// Key of candidate c with score s: the score's bits mapped to an unsigned integer of the same
// order, then the complemented rank, so a larger key is a better candidate. 0 (below every real
// key) marks an empty entry and rank 0xFFFFFFFF; NaN scores get 0 and never enter a queue.
__device__ __forceinline__ uint64_t scoreKey(float s, uint32_t c)
{
    if (s != s)
        return 0ull;
    const uint32_t b = __float_as_uint(s);
    const uint32_t o = (b & 0x80000000u) ? ~b : (b | 0x80000000u);
    return ((uint64_t)o << 32) | (uint64_t)(0xFFFFFFFFu - c);
}

// Score of a key (inverse of scoreKey's order map); -inf for the empty key 0.
__device__ __forceinline__ float keyScore(uint64_t key)
{
    if (key == 0ull)
        return -INFINITY;
    const uint32_t o = (uint32_t)(key >> 32);
    return __uint_as_float((o & 0x80000000u) ? (o & 0x7FFFFFFFu) : ~o);
}

__device__ __forceinline__ uint64_t shflXor64(uint64_t v, int m)
{
    const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)v, m, 64);
    const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(v >> 32), m, 64);
    return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint64_t shfl64(uint64_t v, uint32_t src)
{
    const uint32_t lo = (uint32_t)__shfl((int)(uint32_t)v, (int)src, 64);
    const uint32_t hi = (uint32_t)__shfl((int)(uint32_t)(v >> 32), (int)src, 64);
    return ((uint64_t)hi << 32) | lo;
}

// Lane keys forming a bitonic sequence -> sorted descending across the 64 lanes.
__device__ __forceinline__ uint64_t bitonicMergeDesc(uint64_t v, uint32_t lane)
{
#pragma unroll
    for (int stride = 32; stride > 0; stride >>= 1) {
        const uint64_t o = shflXor64(v, stride);
        v = (lane & (uint32_t)stride) == 0u ? (v > o ? v : o) : (v < o ? v : o);
    }
    return v;
}

// Arbitrary lane keys -> sorted descending across the 64 lanes.
__device__ __forceinline__ uint64_t bitonicSortDesc(uint64_t v, uint32_t lane)
{
#pragma unroll
    for (int size = 2; size <= 64; size <<= 1)
#pragma unroll
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            const uint64_t o = shflXor64(v, stride);
            const bool keepMax = ((lane & (uint32_t)stride) == 0u) == ((lane & (uint32_t)size) == 0u);
            v = keepMax ? (v > o ? v : o) : (v < o ? v : o);
        }
    return v;
}

// Queue q (sorted descending) and a second sorted-descending list l given reversed (lane i holds
// l[63 - i]): the 64 best of both, sorted descending. The lane-wise maximum of a descending and
// an ascending sequence is bitonic and holds the 64 largest of the union.
__device__ __forceinline__ uint64_t mergeReversed(uint64_t q, uint64_t lRev, uint32_t lane)
{
    return bitonicMergeDesc(q > lRev ? q : lRev, lane);
}

// Loads a lane's 32 embedding values of one row. The 128 dimensions are split as
// k = 32 g + 8 s + i for lane group g = lane / 16, MFMA step s < 4 and element i < 8: both
// operands use the same split, so the 4 steps of v_mfma_f32_16x16x32_bf16 (lane l supplies
// elements 8 (l / 16) .. + 8 of the 32-deep slice for row or column l % 16) sum the full dot
// product, and each lane reads 64 contiguous bytes per row.
__device__ __forceinline__ void loadFrag(const uint16_t* __restrict__ Ebf, uint32_t row, uint32_t g, bf16x8 (&f)[4])
{
    const bf16x8* p = reinterpret_cast<const bf16x8*>(Ebf + (size_t)row * kEmb + 32u * g);
#pragma unroll
    for (int v = 0; v < 4; ++v)
        f[v] = p[v];
}

// Buffered survivors (lane i < count holds one, others 0) merged into a queue; the rare path of
// kTopKPartial, kept out of line so the unrolled per-query filter stays small.
__device__ __noinline__ uint64_t mergeBuffer(uint64_t queue, uint64_t buffered, uint32_t lane)
{
    const uint64_t sorted = bitonicSortDesc(buffered, lane);
    return mergeReversed(queue, shflXor64(sorted, 63), lane);
}

// Block barrier ordering LDS only: unlike __syncthreads it does not wait for outstanding global
// loads, so a prefetch issued before it stays in flight across it.
__device__ __forceinline__ void ldsBarrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Orders the calling wave's LDS accesses (as waveSync, without waiting for global loads).
__device__ __forceinline__ void ldsWaveSync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
}

__device__ __forceinline__ uint64_t readLane64(uint64_t v, int l)
{
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, l);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), l);
    return ((uint64_t)hi << 32) | lo;
}

// One block per (query block, chunk): 256 threads, waves in 2 x 2 over the 64 x 64 score tile,
// each wave 32 x 32 = 2 x 2 MFMA 16x16x32 tiles. The wave's query fragments stay in registers;
// the next candidate tile's fragments are loaded while the current one is filtered. Scores go to
// LDS; wave w then owns queries 16 w + i, i < 16: their queues live in registers (lane = entry),
// their thresholds (the queue's last key) and buffer fill counts are wave-uniform. Lane j takes
// candidate j of the tile; keys above the threshold are appended to the query's 64-entry buffer
// in LDS, and a full buffer is sorted and merged into the queue (then the threshold rises).
// Writes each query's queue (64 keys, sorted descending) to part[(block * 64 + query) * 64 + entry].
// The chunks of one query block share a threshold per query, shared[q]: after each merge a block
// raises it to its queue's last key with a 64-bit atomicMax and filters against the larger of its
// own and the shared one, refreshed every tile. Any chunk's 64th best key is a lower bound of the
// query's final 64th best, so a key at or below it can never be among the 64 best: the result is
// exact whatever the timing, and the shared bound keeps the later chunks' survivors few.
__global__ __launch_bounds__(256) void kTopKPartial(const uint16_t* __restrict__ Ebf, uint32_t n, uint32_t qblocks, uint32_t ct, uint64_t* __restrict__ part,
    uint64_t* __restrict__ shared)
{
    constexpr int QW = kQ / 4; // queries per wave
    __shared__ float s_score[kQ][kQ + 1]; // [candidate][query]
    __shared__ uint64_t s_buf[kQ][64]; // [query][entry]
    const int t = (int)threadIdx.x;
    const uint32_t lane = (uint32_t)t & 63u;
    const int wave = t >> 6;
    const uint32_t g = lane >> 4;
    const int wr = wave >> 1, wc = wave & 1;
    // (query block, chunk) of this block: the largest qb with chunksBefore(qb) <= blockIdx.x.
    uint32_t lo = 0, hi = qblocks - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (chunksBefore(mid, ct) <= blockIdx.x)
            lo = mid;
        else
            hi = mid - 1;
    }
    const uint32_t qb = lo;
    // Chunks and the tiles inside a chunk are visited from the query block's own tile backwards:
    // matches have popcounts at or just above the query's, i.e. nearby ranks, so the best
    // candidates come early and the thresholds rise fast. The result does not depend on order.
    const uint32_t chunks = qb / ct + 1;
    const uint32_t chunk = chunks - 1u - (blockIdx.x - (uint32_t)chunksBefore(qb, ct));
    const uint32_t tile0 = chunk * ct, tile1 = min(tile0 + ct, qb + 1);
    const uint32_t q0 = qb * kQ;
    uint64_t queue[QW], thr[QW];
    float thrScore[QW]; // keyScore(thr[i]): score >= thrScore is a superset test of key > thr
    uint32_t cnt[QW];
#pragma unroll
    for (int i = 0; i < QW; ++i) {
        queue[i] = 0ull;
        thr[i] = 0ull;
        thrScore[i] = -INFINITY;
        cnt[i] = 0u;
    }
    bf16x8 a[2][4], b[2][4];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        loadFrag(Ebf, min(q0 + wr * 32 + i * 16 + (lane & 15), n - 1), g, a[i]);
        loadFrag(Ebf, min((tile1 - 1u) * kQ + wc * 32 + i * 16 + (lane & 15), n - 1), g, b[i]);
    }
    for (uint32_t step = 0; step < tile1 - tile0; ++step) {
        const uint32_t tile = tile1 - 1u - step;
        const uint32_t c0 = tile * kQ;
        f32x4 acc[2][2];
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j)
                acc[i][j] = f32x4 { 0.f, 0.f, 0.f, 0.f };
#pragma unroll
        for (int st = 0; st < 4; ++st)
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    acc[i][j] = mfma32(a[i][st], b[j][st], acc[i][j]);
        // The shared thresholds (lane i < 16: query 16 w + i; the address is clamped so the load
        // is unconditional) are loaded before the next tile's prefetch: loads complete in order,
        // so waiting for them does not wait for the prefetch.
        const uint64_t sharedThr = __hip_atomic_load(&shared[min(q0 + wave * QW + (lane & (QW - 1)), n - 1)], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        // Prefetch of the next tile's candidate fragments (unconditional, rows clamped: after the
        // last tile it reads rows that are never used; a branch here would make the compiler wait
        // for every outstanding load at the join).
#pragma unroll
        for (int j = 0; j < 2; ++j)
            loadFrag(Ebf, min((tile > 0u ? c0 - kQ : 0u) + wc * 32 + j * 16 + (lane & 15), n - 1), g, b[j]);
        // D[row][col]: lane holds rows 4 (lane / 16) + r, column lane % 16 (row = query, col = candidate).
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j)
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    s_score[wc * 32 + j * 16 + (lane & 15)][wr * 32 + i * 16 + 4 * (lane >> 4) + r] = acc[i][j][r];
        ldsBarrier();
        const uint32_t c = c0 + lane;
        const bool refresh = (step & 3u) == 0u; // shared bounds every 4 tiles
#pragma unroll
        for (int i = 0; i < QW; ++i) {
            const int qi = wave * QW + i;
            const uint32_t q = q0 + (uint32_t)qi;
            if (refresh) {
                // For q >= n the clamped load read query n - 1's bound: harmless, as no key passes.
                const uint64_t sh = readLane64(sharedThr, i);
                if (sh > thr[i]) {
                    thr[i] = sh;
                    thrScore[i] = keyScore(sh);
                }
            }
            // One float compare per candidate; the exact key (score, then rank) only where a
            // score reaches the threshold's. NaN fails both tests.
            const float sc = s_score[lane][qi];
            if (__ballot(sc >= thrScore[i]) == 0ull)
                continue;
            const uint64_t key = (c < q && q < n) ? scoreKey(sc, c) : 0ull;
            uint64_t m = __ballot(key > thr[i]);
            if (m == 0ull)
                continue;
            if (cnt[i] + (uint32_t)__popcll(m) > 64u) { // buffer full: merge it, raise the threshold
                ldsWaveSync();
                queue[i] = mergeBuffer(queue[i], lane < cnt[i] ? s_buf[qi][lane] : 0ull, lane);
                const uint64_t last = readLane64(queue[i], 63);
                if (last > thr[i]) {
                    thr[i] = last;
                    thrScore[i] = keyScore(last);
                    if (lane == 0u && q < n)
                        __hip_atomic_fetch_max(&shared[q], last, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
                cnt[i] = 0u;
                ldsWaveSync();
                m = __ballot(key > thr[i]);
            }
            if (key > thr[i])
                s_buf[qi][cnt[i] + __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u))] = key;
            cnt[i] += (uint32_t)__popcll(m);
        }
        ldsBarrier();
    }
#pragma unroll
    for (int i = 0; i < QW; ++i) {
        const int qi = wave * QW + i;
        if (cnt[i] > 0u) {
            ldsWaveSync();
            queue[i] = mergeBuffer(queue[i], lane < cnt[i] ? s_buf[qi][lane] : 0ull, lane);
        }
        part[((size_t)blockIdx.x * kQ + (uint32_t)qi) * 64 + lane] = queue[i];
    }
}

// One wave per query: the 64 best keys of its chunk queues (merged pairwise), then the first k
// ranks, best first, into cand[q * k ..]; 0xFFFFFFFF where fewer than k earlier leaves exist.
__global__ __launch_bounds__(256) void kTopKMerge(const uint64_t* __restrict__ part, uint32_t n, uint32_t k, uint32_t ct, uint32_t* __restrict__ cand)
{
    const uint32_t q = blockIdx.x * 4u + (threadIdx.x >> 6);
    if (q >= n)
        return;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t qb = q / kQ, t = q % kQ;
    const uint32_t b0 = (uint32_t)chunksBefore(qb, ct);
    const uint32_t chunks = qb / ct + 1;
    const uint64_t* base = part + ((size_t)b0 * kQ + t) * 64;
    uint64_t v = base[lane];
    for (uint32_t ch = 1; ch < chunks; ++ch)
        v = mergeReversed(v, base[(size_t)ch * kQ * 64 + (63u - lane)], lane);
    if (lane < k)
        cand[(size_t)q * k + lane] = 0xFFFFFFFFu - (uint32_t)v;
}

// Baseline proposals without a network: the k leaves ranked just before the query (the nearest
// popcounts at or above the query's).
__global__ __launch_bounds__(256) void kNeighbourProposals(uint32_t n, uint32_t k, uint32_t* __restrict__ cand)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (size_t)n * k)
        return;
    const uint32_t q = (uint32_t)(i / k), s = (uint32_t)(i % k);
    cand[i] = s < q ? q - 1 - s : 0xFFFFFFFFu;
}

// Bucketed retrieval: the candidates of query q are the earlier leaves with q's exact descriptor
// (canonicalHists, 12 entries of at most 16, packed 5 bits each: 60 bits), and among them the k
// best under kTopKPartial's order (higher score, then lower rank). A match that loses no voxel to
// the clipping translation has the query's descriptor, so a bucket holds every match of equal
// popcount; matches that clip voxels are left to the fallback search. The work is the sum of the
// squared bucket sizes instead of n^2 / 2.
__device__ __forceinline__ uint64_t descriptorKey(uint64_t m)
{
    uint32_t h[3][4];
    canonicalHists(m, h);
    uint64_t key = 0;
    for (int a = 0; a < 3; ++a)
        for (int i = 0; i < 4; ++i)
            key = (key << 5) | h[a][i];
    return key;
}

__global__ __launch_bounds__(256) void kDescriptorKeys(const uint64_t* __restrict__ sortedMasks, uint32_t n, uint64_t* __restrict__ keys, uint32_t* __restrict__ vals)
{
    const uint32_t r = blockIdx.x * 256u + threadIdx.x;
    if (r >= n)
        return;
    keys[r] = descriptorKey(sortedMasks[r]);
    vals[r] = r;
}

// One wave per sorted position i (keys sorted, ties in rank order, so the bucket's earlier
// leaves are positions [start, i) with start the bucket's first position). Lane j scores
// candidate j of each chunk of 64 (fp32 dot product of the bf16 rows; the query row is in LDS),
// and the chunk is sorted and merged into the wave's queue (lane = entry, descending).
__global__ __launch_bounds__(256) void kTopKBucket(const uint16_t* __restrict__ Ebf, const uint64_t* __restrict__ keys, const uint32_t* __restrict__ vals, uint32_t n,
    uint32_t k, uint32_t* __restrict__ cand)
{
    __shared__ float s_q[4][kEmb];
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    const uint32_t i = blockIdx.x * 4u + wave;
    if (i >= n)
        return; // the whole wave: the kernel has no block barrier
    const uint32_t q = vals[i];
    const uint64_t key = keys[i];
    uint32_t lo = 0, hi = i; // first position of the bucket
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (keys[mid] < key)
            lo = mid + 1;
        else
            hi = mid;
    }
    for (uint32_t d = lane; d < (uint32_t)kEmb; d += 64)
        s_q[wave][d] = __uint_as_float((uint32_t)Ebf[(size_t)q * kEmb + d] << 16);
    ldsWaveSync();
    uint64_t queue = 0;
    for (uint32_t j0 = lo; j0 < i; j0 += 64) {
        const uint32_t j = j0 + lane;
        uint64_t kk = 0;
        if (j < i) {
            const uint32_t c = vals[j];
            const uint4* row = reinterpret_cast<const uint4*>(Ebf + (size_t)c * kEmb);
            float acc = 0.f;
#pragma unroll 4
            for (int v = 0; v < kEmb / 8; ++v) {
                const uint4 x = row[v];
                const uint32_t w[4] = { x.x, x.y, x.z, x.w };
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    acc = fmaf(__uint_as_float(w[e] << 16), s_q[wave][8 * v + 2 * e], acc);
                    acc = fmaf(__uint_as_float(w[e] & 0xFFFF0000u), s_q[wave][8 * v + 2 * e + 1], acc);
                }
            }
            kk = scoreKey(acc, c);
        }
        const uint64_t sorted = bitonicSortDesc(kk, lane);
        queue = mergeReversed(queue, shflXor64(sorted, 63), lane);
    }
    if (lane < k)
        cand[(size_t)q * k + lane] = 0xFFFFFFFFu - (uint32_t)queue;
}

// Lowest / highest occupied coordinate of a non-empty 4^3 mask along each axis.
__device__ __forceinline__ void bbox4(uint64_t m, int lo[3], int hi[3])
{
    for (int a = 0; a < 3; ++a) {
        lo[a] = 4;
        hi[a] = -1;
        for (int v = 0; v < 4; ++v) {
            const uint64_t plane = a == 0 ? rangeMaskX4(v, v + 1) : (a == 1 ? rangeMaskY4(v, v + 1) : rangeMaskZ4(v, v + 1));
            if (m & plane) {
                lo[a] = min(lo[a], v);
                hi[a] = v;
            }
        }
    }
}

// permute (find_translation.cpp:109-125): out(o) = P(v) with v[p[a]] = o[a].
__device__ __forceinline__ uint64_t permute4(uint64_t P, uint32_t pc)
{
    uint64_t out = 0;
    const uint32_t a0 = permAxis(pc, 0), a1 = permAxis(pc, 1), a2 = permAxis(pc, 2);
    for (uint32_t o = 0; o < 64; ++o) {
        int v[3];
        v[a0] = (int)(o & 3u);
        v[a1] = (int)((o >> 2) & 3u);
        v[a2] = (int)(o >> 4);
        out |= ((P >> (v[0] + 4 * v[1] + 16 * v[2])) & 1ull) << o;
    }
    return out;
}

// Packs the lowest / highest occupied coordinate per axis (2 bits each) of a non-empty mask.
__device__ __forceinline__ uint32_t packBox(uint64_t m)
{
    int lo[3], hi[3];
    bbox4(m, lo, hi);
    return (uint32_t)(lo[0] | lo[1] << 2 | lo[2] << 4 | hi[0] << 6 | hi[1] << 8 | hi[2] << 10);
}

// The 48 images M = mirror(permute(P_r)) of every leaf (index pc * 8 + sym) and their boxes.
__global__ __launch_bounds__(256) void kTransforms48(const uint64_t* __restrict__ sortedMasks, uint32_t n, uint64_t* __restrict__ images, uint16_t* __restrict__ boxes)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (size_t)n * 48)
        return;
    const uint32_t r = (uint32_t)(i / 48), tf = (uint32_t)(i % 48);
    const uint64_t P = sortedMasks[r];
    const uint64_t M = P ? mirror4(permute4(P, tf >> 3), tf & 7u) : 0ull;
    images[i] = M;
    boxes[i] = (uint16_t)(M ? packBox(M) : 0u);
}

// One thread per (query, proposal): the smallest key of a transform T with C == T(P_r) (perm,
// then mirror, then the clipping translation), merged into best[q]. For each axis one side of
// the translated grid is not clipped, so t_a is C.lo_a - M.lo_a or C.hi_a - M.hi_a.
// The smallest key of the clipping translations t with C == translate(M, t), M = the image of
// P_r under permutation pc then mirror sym (box mb of M, clo/chi of C), or ~0. For each axis one
// side of the translated grid is not clipped, so t_a is C.lo_a - M.lo_a or C.hi_a - M.hi_a.
__device__ __forceinline__ uint64_t keyForImage(uint64_t C, const int (&clo)[3], const int (&chi)[3], uint64_t M, uint32_t mb, uint32_t r, uint32_t pc,
    uint32_t sym, uint32_t cfgBits)
{
    int cands[3][2];
    int nc[3];
    for (int a = 0; a < 3; ++a) {
        cands[a][0] = clo[a] - (int)((mb >> (2 * a)) & 3u);
        cands[a][1] = chi[a] - (int)((mb >> (6 + 2 * a)) & 3u);
        nc[a] = cands[a][1] != cands[a][0] ? 2 : 1;
    }
    uint64_t bestKey = ~0ull;
    for (int ix = 0; ix < nc[0]; ++ix)
        for (int iy = 0; iy < nc[1]; ++iy)
            for (int iz = 0; iz < nc[2]; ++iz) {
                const int tx = cands[0][ix], ty = cands[1][iy], tz = cands[2][iz];
                if (!configAllows(cfgBits, pc, sym, tx, ty, tz))
                    continue;
                if (translate4(M, tx, ty, tz) != C)
                    continue;
                const uint64_t key = encodeKey(r, pc, sym, tx, ty, tz, 2);
                bestKey = key < bestKey ? key : bestKey;
            }
    return bestKey;
}

__device__ __forceinline__ void boxOf(uint32_t cb, int (&lo)[3], int (&hi)[3])
{
    for (int a = 0; a < 3; ++a) {
        lo[a] = (int)((cb >> (2 * a)) & 3u);
        hi[a] = (int)((cb >> (6 + 2 * a)) & 3u);
    }
}

// One wave per query q. The usable proposals (earlier ranks with at least q's voxel count) are
// sorted by rank, ascending, across the lanes; a key's leading field is the parent rank, so the
// smallest verified key comes from the first proposal in that order that verifies. They are
// checked four at a time: lane l handles proposal l / 16 of the group and transforms
// 3 (l % 16) .. 3 (l % 16) + 2 (index pc * 8 + sym) of it, so a proposal's 48 images are one
// coalesced read, and the loop stops after the first group with a verified key. That key is
// merged into best[q]; `hits` counts the queries whose key moved off the self-match here.
__global__ __launch_bounds__(256) void kVerifyProposals(const uint64_t* __restrict__ sortedMasks, const uint64_t* __restrict__ images, const uint16_t* __restrict__ boxes,
    uint32_t n, uint32_t k, const uint32_t* __restrict__ cand, uint32_t cfgBits, uint64_t* __restrict__ best, unsigned long long* __restrict__ hits)
{
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t q = blockIdx.x * 4u + (threadIdx.x >> 6);
    if (q >= n)
        return;
    const uint64_t C = sortedMasks[q];
    if (C == 0)
        return;
    int clo[3], chi[3];
    boxOf(boxes[(size_t)q * 48], clo, chi);
    const int popC = __popcll(C);
    // Lane s < k loads proposal s; usable ones enter the sort as 0xFFFFFFFF - rank (others as 0),
    // so after a descending sort lane i holds the i-th smallest usable rank.
    uint64_t sortKey = 0ull;
    if (lane < k) {
        const uint32_t r = cand[(size_t)q * k + lane];
        if (r != 0xFFFFFFFFu && r < q) {
            const uint64_t P = sortedMasks[r];
            if (P != 0 && __popcll(P) >= popC)
                sortKey = 0xFFFFFFFFull - r;
        }
    }
    sortKey = bitonicSortDesc(sortKey, lane);
    const uint32_t usable = (uint32_t)__popcll(__ballot(sortKey != 0ull));
    const uint32_t numPerms = (cfgBits & kCfgAxis) ? 6u : 1u;
    const uint32_t numSyms = (cfgBits & kCfgSymmetry) ? 8u : 1u;
    const uint32_t slot = lane >> 4, sub = lane & 15u;
    uint64_t bestKey = ~0ull;
    for (uint32_t g = 0; 4 * g < usable; ++g) {
        const uint32_t s = 4 * g + slot;
        const uint64_t sk = shfl64(sortKey, s & 63u); // lane s's key
        if (s < usable) {
            const uint32_t r = 0xFFFFFFFFu - (uint32_t)sk;
#pragma unroll
            for (uint32_t i = 0; i < 3; ++i) {
                const uint32_t t = sub * 3 + i, pc = t >> 3, sym = t & 7u;
                if (pc >= numPerms || sym >= numSyms)
                    continue;
                const uint64_t key = keyForImage(C, clo, chi, images[(size_t)r * 48 + t], boxes[(size_t)r * 48 + t], r, pc, sym, cfgBits);
                bestKey = key < bestKey ? key : bestKey;
            }
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const uint64_t o = shflXor64(bestKey, off);
            bestKey = o < bestKey ? o : bestKey;
        }
        if (bestKey != ~0ull)
            break; // later groups hold higher ranks only
    }
    if (lane == 0 && bestKey != ~0ull) {
        const unsigned long long prev = atomicMin((unsigned long long*)&best[q], (unsigned long long)bestKey);
        if (prev == (unsigned long long)selfKey(q, 2))
            atomicAdd(hits, 1ull); // the query's first verified proposal
    }
}

// Parent chase: a query q matched to p, where p is itself matched to a lower rank p', tries p'
// (for unclipped transforms the 48-element group composes, so p' is often a parent of q too).
// Accepting the smaller verified key moves q's parent down the chain as the exhaustive minimum
// would, so fewer leaves stay parents. Keys are read from `prev`, a snapshot of best[] taken
// before the pass, so the outcome does not depend on thread order; every accepted key is
// verified bitwise. Four queries per wave: lane l takes query l / 16 of the wave and transforms
// 3 (l % 16) .. 3 (l % 16) + 2 of its candidate p'.
__global__ __launch_bounds__(256) void kChaseParents(const uint64_t* __restrict__ sortedMasks, const uint64_t* __restrict__ images, const uint16_t* __restrict__ boxes,
    uint32_t n, uint32_t cfgBits, const uint64_t* __restrict__ prev, uint64_t* __restrict__ best)
{
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t q = (blockIdx.x * 4u + (threadIdx.x >> 6)) * 4u + (lane >> 4);
    const uint32_t sub = lane & 15u;
    uint64_t kq = ~0ull, key = ~0ull;
    uint32_t pp = 0xFFFFFFFFu;
    if (q < n) {
        kq = prev[q];
        if (kq != sentinelKey(2) && kq != selfKey(q, 2)) {
            const uint32_t p = (uint32_t)keyPtr(kq, 2);
            const uint64_t kp = prev[p];
            if (kp != sentinelKey(2)) {
                const uint32_t cand = (uint32_t)keyPtr(kp, 2);
                if (cand < p)
                    pp = cand;
            }
        }
    }
    if (pp != 0xFFFFFFFFu) {
        const uint64_t C = sortedMasks[q], P = sortedMasks[pp];
        if (C != 0 && P != 0 && __popcll(P) >= __popcll(C)) {
            int clo[3], chi[3];
            boxOf(boxes[(size_t)q * 48], clo, chi);
            const uint32_t numPerms = (cfgBits & kCfgAxis) ? 6u : 1u;
            const uint32_t numSyms = (cfgBits & kCfgSymmetry) ? 8u : 1u;
#pragma unroll
            for (uint32_t i = 0; i < 3; ++i) {
                const uint32_t t = sub * 3 + i, pc = t >> 3, sym = t & 7u;
                if (pc >= numPerms || sym >= numSyms)
                    continue;
                const uint64_t kk = keyForImage(C, clo, chi, images[(size_t)pp * 48 + t], boxes[(size_t)pp * 48 + t], pp, pc, sym, cfgBits);
                key = kk < key ? kk : key;
            }
        }
    }
#pragma unroll
    for (int off = 8; off > 0; off >>= 1) { // minimum over the query's 16 lanes
        const uint64_t o = shflXor64(key, off);
        key = o < key ? o : key;
    }
    if (sub == 0 && q < n && key < kq)
        best[q] = key; // only this lane writes best[q] during the pass
}

// Fallback flags: non-empty queries whose best key is still the self-match.
__global__ __launch_bounds__(256) void kFallbackFlags(const uint64_t* __restrict__ sortedMasks, const uint64_t* __restrict__ best, uint32_t n, uint32_t* __restrict__ flags,
    unsigned long long* __restrict__ count)
{
    const uint32_t q = blockIdx.x * 256u + threadIdx.x;
    if (q >= n)
        return;
    const bool f = sortedMasks[q] != 0 && best[q] == selfKey(q, 2);
    flags[q] = f ? 1u : 0u;
    if (f)
        atomicAdd(count, 1ull);
}

} // namespace

size_t topkWorkBytes(uint64_t n)
{
    if (n == 0)
        return 256;
    const uint32_t qblocks = (uint32_t)cdiv(n, kQ);
    const uint64_t blocks = chunksBefore(qblocks, chunkTiles(qblocks));
    return align256((size_t)blocks * kQ * 64 * 8) + align256((size_t)n * 8);
}

hipError_t topkPropose(const uint16_t* Ebf, uint32_t n, uint32_t k, uint32_t* cand, void* work, hipStream_t s)
{
    if (n == 0)
        return hipSuccess;
    if (k == 0 || k > kMaxProposals)
        return hipErrorInvalidValue;
    const uint32_t qblocks = (uint32_t)cdiv(n, kQ);
    const uint32_t ct = chunkTiles(qblocks);
    const uint64_t blocks = chunksBefore(qblocks, ct);
    if (blocks > 0x7FFFFFFFull)
        return hipErrorInvalidValue;
    uint64_t* shared = (uint64_t*)((char*)work + align256((size_t)blocks * kQ * 64 * 8));
    LEARN_TRY(fillU64(shared, n, 0ull, s));
    hipLaunchKernelGGL(kTopKPartial, dim3((unsigned)blocks), dim3(256), 0, s, Ebf, n, qblocks, ct, (uint64_t*)work, shared);
    LEARN_TRY(hipGetLastError());
    hipLaunchKernelGGL(kTopKMerge, dim3((unsigned)cdiv(n, 4)), dim3(256), 0, s, (const uint64_t*)work, n, k, ct, cand);
    return hipGetLastError();
}

size_t bucketWorkBytes(uint64_t n)
{
    return align256(n * 8) + align256(n * 4) + align256(radixScratchBytes(std::max<uint64_t>(n, 1), 8));
}

hipError_t bucketPropose(const uint64_t* sortedMasks, const uint16_t* Ebf, uint32_t n, uint32_t k, uint32_t* cand, void* work, hipStream_t s)
{
    if (n == 0)
        return hipSuccess;
    if (k == 0 || k > kMaxProposals)
        return hipErrorInvalidValue;
    uint64_t* keys = (uint64_t*)work;
    uint32_t* vals = (uint32_t*)((char*)work + align256((size_t)n * 8));
    void* sortScratch = (char*)vals + align256((size_t)n * 4);
    hipLaunchKernelGGL(kDescriptorKeys, dim3((unsigned)cdiv(n, 256)), dim3(256), 0, s, sortedMasks, n, keys, vals);
    LEARN_TRY(hipGetLastError());
    LEARN_TRY(radixSortPairsU64(keys, vals, n, 0, 60, sortScratch, s));
    hipLaunchKernelGGL(kTopKBucket, dim3((unsigned)cdiv(n, 4)), dim3(256), 0, s, Ebf, keys, vals, n, k, cand);
    return hipGetLastError();
}

size_t learnedWorkBytes(uint64_t n)
{
    // proposals, flags, bf16 embeddings, counters, 48 images and boxes per leaf, the retrieval's
    // workspace (per-chunk lists, or the bucket sort) (the embedding workspace itself lives in
    // the Mlp)
    return align256(n * kMaxProposals * 4) + align256(n * 4) + align256((n + kQ) * kEmb * 2) + 256 + align256(n * 48 * 8) + align256(n * 48 * 2) +
        align256(n * 8) + align256(n * kEmb * 4) + std::max(topkWorkBytes(n), bucketWorkBytes(n));
}

hipError_t mlpPrepareEmbed(Mlp& mlp, uint64_t n)
{
    return n == 0 ? hipSuccess : ensureWork(mlp, std::min<uint64_t>(n, 65536), 2);
}

hipError_t gemmBf16(bool ta, bool tb, const uint16_t* A, int lda, const uint16_t* B, int ldb, float* C, int ldc, int M, int N, int K, hipStream_t stream)
{
    // With a split-K workspace, as the network's weight gradients use it (applies to few output
    // tiles and long K).
    float* split = nullptr;
    LEARN_TRY(hipMallocAsync((void**)&split, kSplitWorkFloats * 4, stream));
    const hipError_t e = gemm(ta, tb, A, lda, B, ldb, C, ldc, M, N, K, stream, split);
    const hipError_t f = hipFreeAsync(split, stream);
    return e != hipSuccess ? e : f;
}

hipError_t mlpAlloc(Mlp& mlp)
{
    LEARN_TRY(hipMalloc(&mlp.w, kParams * 4));
    LEARN_TRY(hipMalloc(&mlp.wbf, kParams * 2));
    LEARN_TRY(hipMalloc(&mlp.g, kParams * 4));
    LEARN_TRY(hipMalloc(&mlp.m, kParams * 4));
    LEARN_TRY(hipMalloc(&mlp.v, kParams * 4));
    return hipSuccess;
}

void mlpFree(Mlp& mlp)
{
    for (void* p : { (void*)mlp.w, (void*)mlp.wbf, (void*)mlp.g, (void*)mlp.m, (void*)mlp.v, mlp.ws })
        if (p)
            (void)hipFree(p);
    mlp = Mlp {};
}

hipError_t mlpSetParams(Mlp& mlp, const float* host, hipStream_t s)
{
    LEARN_TRY(hipMemcpyAsync(mlp.w, host, kParams * 4, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(kCastBf, dim3((unsigned)cdiv(kParams, 256)), dim3(256), 0, s, mlp.w, kParams, mlp.wbf);
    LEARN_TRY(hipGetLastError());
    LEARN_TRY(fillU32((uint32_t*)mlp.m, kParams, 0u, s));
    LEARN_TRY(fillU32((uint32_t*)mlp.v, kParams, 0u, s));
    mlp.step = 0;
    return hipStreamSynchronize(s);
}

hipError_t mlpGetParams(const Mlp& mlp, float* host, hipStream_t s)
{
    LEARN_TRY(hipMemcpyAsync(host, mlp.w, kParams * 4, hipMemcpyDeviceToHost, s));
    return hipStreamSynchronize(s);
}

hipError_t mlpInit(Mlp& mlp, uint64_t seed, hipStream_t s)
{
    std::vector<float> h(kParams, 0.f);
    std::mt19937_64 rng(seed);
    for (int l = 0; l < kLayers; ++l) {
        const float a = std::sqrt(6.0f / (float)kLayerIn[l]);
        std::uniform_real_distribution<float> U(-a, a);
        float* W = h.data() + paramOffsetW(l);
        for (size_t i = 0; i < (size_t)kLayerIn[l] * kLayerOut[l]; ++i)
            W[i] = U(rng);
    }
    return mlpSetParams(mlp, h.data(), s);
}

hipError_t mlpLossGrad(Mlp& mlp, const uint64_t* d_q, const uint64_t* d_p, uint32_t B, float tau, float* out_loss, hipStream_t s)
{
    if (B == 0)
        return hipErrorInvalidValue;
    LEARN_TRY(ensureWork(mlp, 2 * (size_t)B, B));
    Work w = carve(mlp.ws, 2 * (size_t)B, B);
    LEARN_TRY(hipMemcpyAsync(w.masks, d_q, (size_t)B * 8, hipMemcpyDeviceToDevice, s));
    LEARN_TRY(hipMemcpyAsync(w.masks + B, d_p, (size_t)B * 8, hipMemcpyDeviceToDevice, s));
    hipLaunchKernelGGL(kFeatures, dim3((unsigned)cdiv(2 * B, 256)), dim3(256), 0, s, w.masks, 2 * B, w.act[0]);
    LEARN_TRY(hipGetLastError());
    LEARN_TRY(lossGrad(mlp, w, B, tau, s));
    if (out_loss) {
        std::vector<float> l(B);
        LEARN_TRY(hipMemcpyAsync(l.data(), w.loss, (size_t)B * 4, hipMemcpyDeviceToHost, s));
        LEARN_TRY(hipStreamSynchronize(s));
        double sum = 0;
        for (float v : l)
            sum += v;
        *out_loss = (float)(sum / B);
    }
    return hipSuccess;
}

hipError_t mlpTrain(Mlp& mlp, const uint64_t* d_q, const uint64_t* d_p, uint64_t n, uint32_t steps, uint32_t batch, float lr, float tau, uint64_t seed, float* out_losses,
    hipStream_t s)
{
    if (n < 2 || batch < 2)
        return hipErrorInvalidValue;
    const uint32_t B = (uint32_t)std::min<uint64_t>(batch, n);
    LEARN_TRY(ensureWork(mlp, 2 * (size_t)B, B));
    Work w = carve(mlp.ws, 2 * (size_t)B, B);
    std::vector<uint32_t> perm(n);
    std::iota(perm.begin(), perm.end(), 0u);
    std::mt19937_64 rng(seed);
    uint64_t pos = n;
    std::vector<uint32_t> idx(B);
    std::vector<float> lossHost(B);
    uint32_t* pinnedIdx = nullptr;
    LEARN_TRY(hipHostMalloc((void**)&pinnedIdx, (size_t)B * 4 * 2));
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    hipError_t err = hipSuccess;
    for (uint32_t st = 0; st < steps && err == hipSuccess; ++st) {
        if (pos + B > n) { // new epoch
            std::shuffle(perm.begin(), perm.end(), rng);
            pos = 0;
        }
        uint32_t* slot = pinnedIdx + (st & 1u) * B; // double-buffered: the previous copy may be in flight
        if (st >= 2)
            err = hipStreamSynchronize(s);
        std::copy(perm.begin() + pos, perm.begin() + pos + B, slot);
        pos += B;
        if (err == hipSuccess)
            err = hipMemcpyAsync(w.idx, slot, (size_t)B * 4, hipMemcpyHostToDevice, s);
        if (err != hipSuccess)
            break;
        hipLaunchKernelGGL(kGatherPairs, dim3((unsigned)cdiv(B, 256)), dim3(256), 0, s, d_q, d_p, w.idx, B, w.masks);
        hipLaunchKernelGGL(kFeatures, dim3((unsigned)cdiv(2 * B, 256)), dim3(256), 0, s, w.masks, 2 * B, w.act[0]);
        if ((err = hipGetLastError()) != hipSuccess)
            break;
        if ((err = lossGrad(mlp, w, B, tau, s)) != hipSuccess)
            break;
        ++mlp.step;
        const float c1 = 1.f - std::pow(b1, (float)mlp.step), c2 = 1.f - std::pow(b2, (float)mlp.step);
        hipLaunchKernelGGL(kAdam, dim3((unsigned)cdiv(kParams, 256)), dim3(256), 0, s, mlp.w, mlp.g, mlp.m, mlp.v, mlp.wbf, kParams, lr, b1, b2, eps, c1, c2);
        if ((err = hipGetLastError()) != hipSuccess)
            break;
        if (out_losses) {
            if ((err = hipMemcpyAsync(lossHost.data(), w.loss, (size_t)B * 4, hipMemcpyDeviceToHost, s)) != hipSuccess)
                break;
            if ((err = hipStreamSynchronize(s)) != hipSuccess)
                break;
            double sum = 0;
            for (float v : lossHost)
                sum += v;
            out_losses[st] = (float)(sum / B);
        }
    }
    const hipError_t e2 = hipStreamSynchronize(s);
    (void)hipHostFree(pinnedIdx);
    return err != hipSuccess ? err : e2;
}

hipError_t mlpEmbed(Mlp& mlp, const uint64_t* d_masks, uint64_t n, float* d_out, hipStream_t s)
{
    if (n == 0)
        return hipSuccess;
    // Chunks of at most 65536 rows keep the workspace bounded.
    const uint64_t chunk = std::min<uint64_t>(n, 65536);
    LEARN_TRY(ensureWork(mlp, chunk, 2));
    Work w = carve(mlp.ws, chunk, 2);
    for (uint64_t r0 = 0; r0 < n; r0 += chunk) {
        const uint32_t rows = (uint32_t)std::min<uint64_t>(chunk, n - r0);
        hipLaunchKernelGGL(kFeatures, dim3((unsigned)cdiv(rows, 256)), dim3(256), 0, s, d_masks + r0, rows, w.act[0]);
        LEARN_TRY(hipGetLastError());
        LEARN_TRY(forward(mlp, w, rows, s));
        LEARN_TRY(hipMemcpyAsync(d_out + r0 * kEmb, w.E, (size_t)rows * kEmb * 4, hipMemcpyDeviceToDevice, s));
    }
    return hipSuccess;
}

hipError_t enqueueLearned2(const MatchPlan& p, Mlp* mlp, uint32_t k, const uint64_t* d_masks, uint32_t cfgBits, void* d_scratch, void* d_work, hipStream_t s,
    uint64_t* d_encoded, uint64_t* d_mapping, uint32_t* d_unique, hipEvent_t* ev)
{
    if (p.level != 2 || k == 0 || k > kMaxProposals)
        return hipErrorInvalidValue;
    const uint32_t n = (uint32_t)p.n;
    const LevelView v = levelView(p, d_scratch);
    char* wb = (char*)d_work;
    uint32_t* cand = (uint32_t*)wb;
    uint32_t* flags = (uint32_t*)(wb + align256((size_t)n * kMaxProposals * 4));
    uint16_t* Ebf = (uint16_t*)((char*)flags + align256((size_t)n * 4));
    unsigned long long* cnt = (unsigned long long*)((char*)Ebf + align256(((size_t)n + kQ) * kEmb * 2));
    uint64_t* images = (uint64_t*)((char*)cnt + 256);
    uint16_t* boxes = (uint16_t*)((char*)images + align256((size_t)n * 48 * 8));
    uint64_t* prevBest = (uint64_t*)((char*)boxes + align256((size_t)n * 48 * 2));
    float* E = (float*)((char*)prevBest + align256((size_t)n * 8));
    void* topkWork = (char*)E + align256((size_t)n * kEmb * 4);
    auto mark = [&](int i) {
        if (ev)
            (void)hipEventRecord(ev[i], s);
    };
    LEARN_TRY(fillU64((uint64_t*)cnt, 2, 0ull, s));
    LEARN_TRY(enqueuePrepareOrder(p, d_masks, d_scratch, s));
    mark(0);
    if (n > 0) {
        // 1. Proposals.
        if (mlp) {
            if (mlp->wsBytes < workBytes(std::min<uint64_t>(n, 65536), 2))
                return hipErrorInvalidValue; // mlpPrepareEmbed was not called for this size
            LEARN_TRY(mlpEmbed(*mlp, v.sortedMasks, n, E, s));
            hipLaunchKernelGGL(kCastBf, dim3((unsigned)cdiv((size_t)n * kEmb, 256)), dim3(256), 0, s, E, (size_t)n * kEmb, Ebf);
            LEARN_TRY(hipGetLastError());
            mark(1);
            if (mlp->retrieval == kRetrieveBucketed)
                LEARN_TRY(bucketPropose(v.sortedMasks, Ebf, n, k, cand, topkWork, s));
            else
                LEARN_TRY(topkPropose(Ebf, n, k, cand, topkWork, s));
        } else {
            mark(1);
            hipLaunchKernelGGL(kNeighbourProposals, dim3((unsigned)cdiv((size_t)n * k, 256)), dim3(256), 0, s, n, k, cand);
        }
        LEARN_TRY(hipGetLastError());
        mark(2);
        // 2. Bitwise verification of every proposal against the 48 precomputed images.
        hipLaunchKernelGGL(kTransforms48, dim3((unsigned)cdiv((size_t)n * 48, 256)), dim3(256), 0, s, v.sortedMasks, n, images, boxes);
        hipLaunchKernelGGL(kVerifyProposals, dim3((unsigned)cdiv(n, 4)), dim3(256), 0, s, v.sortedMasks, images, boxes, n, k, cand, cfgBits, v.best, cnt);
        hipLaunchKernelGGL(kFallbackFlags, dim3((unsigned)cdiv(n, 256)), dim3(256), 0, s, v.sortedMasks, v.best, n, flags, cnt + 1);
        LEARN_TRY(hipGetLastError());
        mark(3);
        // 3. Exhaustive search for the queries without a verified proposal.
        LEARN_TRY(enqueueBuildTable(p, d_scratch, s, flags));
        LEARN_TRY(enqueueSearch(p, cfgBits, d_scratch, s));
        // 4. Parent chase (pointer jumping along verified parent chains).
        for (int pass = 0; pass < kChasePasses; ++pass) {
            LEARN_TRY(hipMemcpyAsync(prevBest, v.best, (size_t)n * 8, hipMemcpyDeviceToDevice, s));
            hipLaunchKernelGGL(kChaseParents, dim3((unsigned)cdiv(n, 16)), dim3(256), 0, s, v.sortedMasks, images, boxes, n, cfgBits, prevBest, v.best);
            LEARN_TRY(hipGetLastError());
        }
        mark(4);
    }
    return enqueueFinish(p, d_scratch, s, d_encoded, d_mapping, d_unique);
}

const unsigned long long* learnedCounters(const MatchPlan& p, void* d_work)
{
    const size_t n = p.n;
    const size_t off = align256(n * kMaxProposals * 4) + align256(n * 4) + align256((n + kQ) * kEmb * 2);
    return (const unsigned long long*)((char*)d_work + off);
}

hipError_t matchLearned2(const MatchPlan& p, Mlp* mlp, uint32_t k, const uint64_t* d_masks, uint32_t cfgBits, void* d_scratch, void* d_work, hipStream_t s,
    uint64_t* d_encoded, uint64_t* d_mapping, uint32_t* d_unique, LearnedStats* stats)
{
    if (mlp)
        LEARN_TRY(mlpPrepareEmbed(*mlp, p.n));
    hipEvent_t ev[5] = {};
    for (auto& e : ev)
        LEARN_TRY(hipEventCreate(&e));
    auto cleanup = [&](hipError_t e) {
        for (auto& x : ev)
            (void)hipEventDestroy(x);
        return e;
    };
    hipError_t e = enqueueLearned2(p, mlp, k, d_masks, cfgBits, d_scratch, d_work, s, d_encoded, d_mapping, d_unique, ev);
    if (e == hipSuccess)
        e = hipStreamSynchronize(s);
    if (e != hipSuccess)
        return cleanup(e);
    if (stats) {
        unsigned long long c[2] = { 0, 0 };
        if ((e = hipMemcpy(c, learnedCounters(p, d_work), 16, hipMemcpyDeviceToHost)) != hipSuccess)
            return cleanup(e);
        stats->proposalHits = c[0];
        stats->fallback = c[1];
        if (p.n > 0) {
            (void)hipEventElapsedTime(&stats->msEmbed, ev[0], ev[1]);
            (void)hipEventElapsedTime(&stats->msPropose, ev[1], ev[2]);
            (void)hipEventElapsedTime(&stats->msVerify, ev[2], ev[3]);
            (void)hipEventElapsedTime(&stats->msFallback, ev[3], ev[4]);
        }
    }
    return cleanup(hipSuccess);
}

} // namespace learn
} // namespace tdag
