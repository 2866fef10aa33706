// Shared host/device definitions of the Transform-Aware SVDAG matching path on gfx950.
//
// Encodings follow the reference bit for bit where they are part of the output:
//   TransformPointer bit layout          transform_dag.h:19-54 (GCC x86-64 bit-fields)
//   encodeFixed64 tie-break key          transform_dag.h:375-397
//   axis-permutation table and its code  transform_dag.h:136-147, 326-331
// Everything else (hashes, table layout, kernel structure) is this library's own design.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tdag {

#define TDAG_HD __host__ __device__ __forceinline__

// transform_dag.h:138-145 — axis permutation (p0, p1, p2) for each code 0..5.
TDAG_HD uint32_t permAxis(uint32_t code, uint32_t a)
{
    // Packed table: code -> 3 two-bit axis indices (p0 | p1 << 2 | p2 << 4).
    constexpr uint32_t packed[6] = { 0x24u, 0x18u, 0x21u, 0x09u, 0x12u, 0x06u };
    return (packed[code] >> (2 * a)) & 3u;
}

// Code of the permutation (p0, p1, p2), the inverse of permAxis (transform_dag.h:326-331).
TDAG_HD uint32_t permCode(uint32_t p0, uint32_t p1, uint32_t p2) { return 2u * p0 + (p2 < p1 ? 1u : 0u); }

// Key field widths: translation fields are level+1 bits each, biased by 2^level.
TDAG_HD uint32_t keyTBits(uint32_t level) { return level + 1; }
TDAG_HD uint32_t keyLowBits(uint32_t level) { return 6 + 3 * (level + 1); }

// encodeFixed64 of (ptr, permCode, symBits, t) at `level` (transform_dag.h:375-397).
TDAG_HD uint64_t encodeKey(uint64_t ptr, uint32_t permCode, uint32_t symBits, int tx, int ty, int tz, uint32_t level)
{
    const uint32_t b = level + 1;
    const int bias = 1 << level;
    uint64_t out = (ptr << 6) | ((uint64_t)permCode << 3) | (uint64_t)symBits;
    out = (out << b) | (uint64_t)(uint32_t)(tz + bias);
    out = (out << b) | (uint64_t)(uint32_t)(ty + bias);
    out = (out << b) | (uint64_t)(uint32_t)(tx + bias);
    return out;
}

// TransformPointer::create(ptr) with identity transform: tx=ty=tz=32, p=(0,1,2).
TDAG_HD uint64_t tpIdentity(uint64_t ptr) { return 0x8208240000000000ull | (ptr & ((1ull << 37) - 1)); }
TDAG_HD uint64_t tpSentinel() { return tpIdentity(0xFFFFFFFFull); }

// Packs (ptr, sym, permCode, t) into the TransformPointer bit layout.
TDAG_HD uint64_t tpPack(uint64_t ptr, uint32_t symBits, uint32_t permCode, int tx, int ty, int tz)
{
    uint64_t v = ptr & ((1ull << 37) - 1);
    v |= (uint64_t)(symBits & 7u) << 37;
    v |= (uint64_t)permAxis(permCode, 0) << 40;
    v |= (uint64_t)permAxis(permCode, 1) << 42;
    v |= (uint64_t)permAxis(permCode, 2) << 44;
    v |= (uint64_t)((uint32_t)(tx + 32) & 63u) << 46;
    v |= (uint64_t)((uint32_t)(ty + 32) & 63u) << 52;
    v |= (uint64_t)((uint32_t)(tz + 32) & 63u) << 58;
    return v;
}

// decodeFixed64 into TransformPointer bits, replacing ptr by `newPtr` (transform_dag.h:349-373).
TDAG_HD uint64_t keyToTp(uint64_t key, uint32_t level, uint64_t newPtr)
{
    const uint32_t b = level + 1;
    const uint64_t m = (1ull << b) - 1ull;
    const int bias = 1 << level;
    const int tx = (int)(key & m) - bias;
    const int ty = (int)((key >> b) & m) - bias;
    const int tz = (int)((key >> (2 * b)) & m) - bias;
    const uint32_t sym = (uint32_t)((key >> (3 * b)) & 7u);
    const uint32_t perm = (uint32_t)((key >> (3 * b + 3)) & 7u);
    return tpPack(newPtr, sym, perm, tx, ty, tz);
}

TDAG_HD uint64_t keyPtr(uint64_t key, uint32_t level) { return key >> keyLowBits(level); }

// Key of the identity self-match of rank r (every non-empty item matches itself).
TDAG_HD uint64_t selfKey(uint64_t rank, uint32_t level) { return encodeKey(rank, 0, 0, 0, 0, 0, level); }
// Key of TransformPointer::sentinel() (find_translation.cpp:652).
TDAG_HD uint64_t sentinelKey(uint32_t level) { return encodeKey(0xFFFFFFFFull, 0, 0, 0, 0, 0, level); }

// ---------------------------------------------------------------------------------------
// 4^3 masks in linear layout (bit = x + 4y + 16z): mirrors, clipping translation.
// ---------------------------------------------------------------------------------------
TDAG_HD uint64_t mirrorX4(uint64_t m)
{
    m = ((m & 0x5555555555555555ull) << 1) | ((m >> 1) & 0x5555555555555555ull);
    m = ((m & 0x3333333333333333ull) << 2) | ((m >> 2) & 0x3333333333333333ull);
    return m;
}
TDAG_HD uint64_t mirrorY4(uint64_t m)
{
    m = ((m & 0x0F0F0F0F0F0F0F0Full) << 4) | ((m >> 4) & 0x0F0F0F0F0F0F0F0Full);
    m = ((m & 0x00FF00FF00FF00FFull) << 8) | ((m >> 8) & 0x00FF00FF00FF00FFull);
    return m;
}
TDAG_HD uint64_t mirrorZ4(uint64_t m)
{
    m = ((m & 0x0000FFFF0000FFFFull) << 16) | ((m >> 16) & 0x0000FFFF0000FFFFull);
    m = (m << 32) | (m >> 32);
    return m;
}
TDAG_HD uint64_t mirror4(uint64_t m, uint32_t symBits)
{
    if (symBits & 1u)
        m = mirrorX4(m);
    if (symBits & 2u)
        m = mirrorY4(m);
    if (symBits & 4u)
        m = mirrorZ4(m);
    return m;
}

// Bits whose coordinate along one axis lies in [lo, hi) — for x, y and z respectively.
TDAG_HD uint64_t rangeMaskX4(int lo, int hi)
{
    uint64_t row = 0;
    for (int x = lo; x < hi; ++x)
        row |= 1ull << x;
    return row * 0x1111111111111111ull;
}
TDAG_HD uint64_t rangeMaskY4(int lo, int hi)
{
    uint64_t slab = 0;
    for (int y = lo; y < hi; ++y)
        slab |= 0xFull << (4 * y);
    return slab * 0x0001000100010001ull;
}
TDAG_HD uint64_t rangeMaskZ4(int lo, int hi)
{
    uint64_t m = 0;
    for (int z = lo; z < hi; ++z)
        m |= 0xFFFFull << (16 * z);
    return m;
}

// translateSubGrid semantics (find_translation.cpp:66-84): out[o] = in[o - t] if in range.
TDAG_HD uint64_t translate4(uint64_t m, int tx, int ty, int tz)
{
    const int s = tx + 4 * ty + 16 * tz;
    const uint64_t shifted = s >= 0 ? (s >= 64 ? 0ull : (m << s)) : (-s >= 64 ? 0ull : (m >> (-s)));
    const int lox = tx > 0 ? tx : 0, hix = tx < 0 ? 4 + tx : 4;
    const int loy = ty > 0 ? ty : 0, hiy = ty < 0 ? 4 + ty : 4;
    const int loz = tz > 0 ? tz : 0, hiz = tz < 0 ? 4 + tz : 4;
    return shifted & rangeMaskX4(lox, hix) & rangeMaskY4(loy, hiy) & rangeMaskZ4(loz, hiz);
}

// ---------------------------------------------------------------------------------------
// 32-bit nested sub-box hash for levels >= 3 (this library's own function). Same "started"
// rule as the reference's passes (find_translation.cpp:213-217): the state stays 0 while only
// zeros were seen and is never 0 afterwards, so a hash is 0 exactly for an empty sub-box.
// Per step: rotate, xor, odd multiply (a bijection of the rotated state for a fixed input),
// then max(x, min(s, 1)): 0 before anything was seen, >= 1 afterwards (only the product 0 is
// merged into 1). The product is 0 exactly when v == rotl(s, 7), and then s and v are both 0 or
// both non-zero, so min(s, 1) says whether anything was seen. The final value of the last pass
// goes through finalHash before it is used as a table or filter key.
// ---------------------------------------------------------------------------------------
TDAG_HD uint32_t rotl32(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }

TDAG_HD uint32_t hashStep(uint32_t s, uint32_t v)
{
    const uint32_t x = (rotl32(s, 7) ^ v) * 0x9E3779B1u;
#if defined(__HIP_DEVICE_COMPILE__)
    // min(s, 1) as one v_min_u32: written as a compare and a select, it costs two instructions
    // and a VCC hazard in the unrolled sweeps.
    uint32_t seen;
    asm("v_min_u32 %0, 1, %1" : "=v"(seen) : "v"(s));
#else
    const uint32_t seen = s < 1u ? s : 1u;
#endif
    return x > seen ? x : seen;
}

// xorshift: a bijection with finalHash(0) == 0 that folds the well-mixed high half of the last
// product into the low bits (the Bloom filter's bit positions use the low ten bits).
TDAG_HD uint32_t finalHash(uint32_t h) { return h ^ (h >> 16); }

TDAG_HD uint32_t homeSlot64(uint64_t key, uint32_t logCap)
{
    return (uint32_t)((key * 0x9E3779B97F4A7C15ull) >> (64 - logCap));
}
// Home slot of a 32-bit nested hash: its top bits, as the filter word (bloomWordNested) is, so
// the table's slots run in the order of the filter words (the prune rebuilds the filter from a
// sweep of the table with few atomics). finalHash leaves the well-mixed top half of the last
// product in the top bits.
TDAG_HD uint32_t homeSlot32(uint32_t h, uint32_t logCap) { return h >> (32 - logCap); }

// Blocked Bloom filter over the child table keys: one 32-bit word per key, two bits set in it.
// `f` is a well-mixed 32-bit value of the key; the word index uses its top bits and the two bit
// positions its low ten bits (disjoint for filters up to 2^22 words).
TDAG_HD uint32_t bloomWord(uint32_t f, uint32_t logWords) { return f >> (32 - logWords); }
TDAG_HD uint32_t bloomBits(uint32_t f) { return (1u << (f & 31u)) | (1u << ((f >> 5) & 31u)); }
// 1 when both bits of f are set in the filter word bw. On the device: two bit-field extracts
// (v_bfe_u32 reads its offset from the low five bits of the operand), the second of width 0 or 1
// (the first bit), so a clear first bit yields 0 at once.
TDAG_HD uint32_t bloomHit(uint32_t bw, uint32_t f)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_ubfe(bw, f, __builtin_amdgcn_ubfe(bw, f >> 5, 1u));
#else
    return (bw >> (f & 31u)) & (bw >> ((f >> 5) & 31u)) & 1u;
#endif
}
// Word of a level-3..5 nested hash: its top bits. (The search loads no word for the hash 0 of an
// empty region: its buffer offset is out of range.)
TDAG_HD uint32_t bloomWordNested(uint32_t f, uint32_t logWords) { return f >> (32 - logWords); }
TDAG_HD uint32_t bloomMix64(uint64_t key) { return (uint32_t)((key * 0xD6E8FEB86659FD93ull) >> 32); }

} // namespace tdag
