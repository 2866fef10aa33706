"""Seeded test sets of LargeSubGrid masks (n x W uint64, linear layout x + R*y + R*R*z).

Random sparse masks plus copies planted under random (permutation, mirror, clipping
translation), and the edge cases the matcher must handle: empty masks, full masks, single
voxels at corners, exact duplicates and equal-popcount pairs that match each other.
"""
from __future__ import annotations

import numpy as np


def words(level: int) -> int:
    return max(1, (1 << (3 * level)) // 64)


def bits_to_mask(level: int, bits: np.ndarray) -> np.ndarray:
    """bits: bool array of R^3 in linear order -> W uint64 words."""
    W = words(level)
    b = np.zeros(W * 64, dtype=np.uint8)
    b[: bits.size] = bits.astype(np.uint8)
    return np.packbits(b.reshape(W, 64)[:, ::-1], axis=1).view(">u8").astype(np.uint64).reshape(W)


def mask_to_bits(level: int, mask: np.ndarray) -> np.ndarray:
    R3 = 1 << (3 * level)
    m = np.ascontiguousarray(mask, dtype=np.uint64).reshape(-1)
    by = m.astype(">u8").view(np.uint8).reshape(-1, 8)
    bits = np.unpackbits(by, axis=1)[:, ::-1].reshape(-1)
    return bits[:R3].astype(bool)


def random_masks(rng, level: int, n: int, density: float) -> np.ndarray:
    R3 = 1 << (3 * level)
    out = np.zeros((n, words(level)), dtype=np.uint64)
    for i in range(n):
        out[i] = bits_to_mask(level, rng.random(R3) < density)
    return out


def blob_masks(rng, level: int, n: int) -> np.ndarray:
    """Small compact blobs placed at random offsets (many clip-edge translations)."""
    R = 1 << level
    out = np.zeros((n, words(level)), dtype=np.uint64)
    for i in range(n):
        g = np.zeros((R, R, R), dtype=bool)  # indexed [z, y, x]
        size = rng.integers(1, max(2, R // 2) + 1, 3)
        lo = rng.integers(0, R, 3)
        sub = rng.random(tuple(size)) < 0.6
        for z in range(size[2]):
            for y in range(size[1]):
                for x in range(size[0]):
                    X, Y, Z = lo[0] + x, lo[1] + y, lo[2] + z
                    if X < R and Y < R and Z < R and sub[x, y, z]:
                        g[Z, Y, X] = True
        out[i] = bits_to_mask(level, g.reshape(-1))
    return out


def planted_set(rng, level: int, n_base: int, n_planted: int, transform, density=0.12, edge_cases=True) -> np.ndarray:
    """Base masks + transformed copies. `transform(level, mask, perm, sym, t)` applies
    translate(mirror(permute(mask))) (the oracle's restatement)."""
    R = 1 << level
    base = np.concatenate([random_masks(rng, level, n_base // 2, density), blob_masks(rng, level, n_base - n_base // 2)])
    items = [base]
    planted = []
    for _ in range(n_planted):
        src = base[rng.integers(len(base))]
        perm = int(rng.integers(6))
        sym = int(rng.integers(8))
        mode = rng.integers(3)
        if mode == 0:
            t = (0, 0, 0)
        elif mode == 1:
            t = tuple(int(v) for v in rng.integers(-(R - 1), R, 3))
        else:
            t = tuple(int(v) for v in rng.integers(-1, 2, 3))
        planted.append(transform(level, src, perm, sym, t))
    if planted:
        items.append(np.array(planted, dtype=np.uint64).reshape(-1, words(level)))
    if edge_cases:
        W = words(level)
        empty = np.zeros((1, W), dtype=np.uint64)
        full = np.full((1, W), np.uint64(0xFFFFFFFFFFFFFFFF) if R >= 4 else np.uint64((1 << (R ** 3)) - 1), dtype=np.uint64)
        corners = []
        for c in range(8):
            g = np.zeros(R ** 3, dtype=bool)
            x = 0 if c & 1 else R - 1
            y = 0 if c & 2 else R - 1
            z = 0 if c & 4 else R - 1
            g[x + R * y + R * R * z] = True
            corners.append(bits_to_mask(level, g))
        pair = []
        g1 = np.zeros(R ** 3, dtype=bool)
        g1[0] = True
        g2 = np.zeros(R ** 3, dtype=bool)
        g2[1] = True
        pair = [bits_to_mask(level, g1), bits_to_mask(level, g2)]
        dup = base[:2]
        items += [empty, full, np.array(corners), np.array(pair), dup, empty]
    masks = np.concatenate(items).astype(np.uint64)
    perm = rng.permutation(len(masks))
    return np.ascontiguousarray(masks[perm])
