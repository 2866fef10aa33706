"""Independent numpy statements of the reference's grid transforms, used to pin the oracle.

The transforms are defined in find_translation.cpp:66-125 (translate with clipping, mirror,
axis permutation). Here they are expressed with array operations on an [x, y, z]-indexed
boolean cube (np.transpose / np.flip / slicing) rather than per-voxel index loops.
"""
import numpy as np

from sets import bits_to_mask, mask_to_bits

PERMS = ((0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0))


def to_cube(level, mask):
    R = 1 << level
    # linear index = x + R*y + R*R*z -> reshape gives [z, y, x]; transpose to [x, y, z]
    return mask_to_bits(level, mask).reshape(R, R, R).transpose(2, 1, 0)


def from_cube(level, cube):
    return bits_to_mask(level, np.ascontiguousarray(cube.transpose(2, 1, 0)).reshape(-1))


def permute(cube, perm):
    # out[(v[p0], v[p1], v[p2])] = in[v]  <=>  result axis k is input axis perm[k]
    return np.transpose(cube, perm)


def mirror(cube, sym):
    for a in range(3):
        if (sym >> a) & 1:
            cube = np.flip(cube, axis=a)
    return cube


def translate(cube, t):
    R = cube.shape[0]
    out = np.zeros_like(cube)
    src = [slice(max(0, -ta), min(R, R - ta)) for ta in t]
    dst = [slice(max(0, ta), min(R, R + ta)) for ta in t]
    out[tuple(dst)] = cube[tuple(src)]
    return out


def apply(level, mask, perm_code, sym, t):
    c = to_cube(level, mask)
    return from_cube(level, translate(mirror(permute(c, PERMS[perm_code]), sym), t))
