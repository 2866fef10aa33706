"""Independent .tdag reader and expander (numpy), the analogue of verifyComplexTransformDAG
(transform_dag_encoding.h:82-117): parse the byte stream written by the voxelizer
(voxelizer/src/main.cpp:147-150 -> structure.h:493-501 -> binary_writer.h:48-78), expand every
TransformPointer top-down as child = translate(mirror(permute(C, perm), mirror), t) and rebuild
the dense voxel grid.
"""
import struct

import numpy as np

PERM_CODES = {(0, 1, 2): 0, (0, 2, 1): 1, (1, 0, 2): 2, (1, 2, 0): 3, (2, 0, 1): 4, (2, 1, 0): 5}
SENTINEL = 0x82082400FFFFFFFF


def parse(b: bytes):
    off = 0

    def u64():
        nonlocal off
        v = struct.unpack_from("<Q", b, off)[0]
        off += 8
        return v

    nlevels = u64()
    levels = []
    for _ in range(nlevels):
        n = u64()
        arr = np.frombuffer(b, dtype="<u8", count=n * 8, offset=off).reshape(n, 8) if n else np.zeros((0, 8), dtype=np.uint64)
        off += n * 64
        levels.append(arr)
    ns = u64()
    leaves = np.frombuffer(b, dtype="<u8", count=ns, offset=off) if ns else np.zeros(0, dtype=np.uint64)
    off += ns * 8
    resolution, root_level, structure_type = struct.unpack_from("<IIi", b, off)
    off += 12
    sym, axis, trans, _pad, maxT = struct.unpack_from("<BBBBI", b, off)
    off += 8
    assert off == len(b)
    return dict(levels=levels, leaves=leaves, resolution=resolution, root_level=root_level, structure_type=structure_type,
                config=(sym, axis, trans, maxT))


def parse_static(b: bytes):
    """StaticStructure<void, TransformPointer>::writeTo bytes (structure.h:888-895 through
    binary_writer.h:48-78): per level u64 word count + u64 words, u64 leaf count + leaf masks,
    then u32 resolution, u32 rootLevel, i32 structureType. No config follows."""
    off = 0

    def u64():
        nonlocal off
        v = struct.unpack_from("<Q", b, off)[0]
        off += 8
        return v

    nlevels = u64()
    words = []
    for _ in range(nlevels):
        n = u64()
        words.append(np.frombuffer(b, dtype="<u8", count=n, offset=off).copy() if n else np.zeros(0, dtype=np.uint64))
        off += n * 8
    ns = u64()
    leaves = np.frombuffer(b, dtype="<u8", count=ns, offset=off) if ns else np.zeros(0, dtype=np.uint64)
    off += ns * 8
    resolution, root_level, structure_type = struct.unpack_from("<IIi", b, off)
    off += 12
    assert off == len(b)
    return dict(words=words, leaves=leaves, resolution=resolution, root_level=root_level, structure_type=structure_type)


def node_starts(words) -> list:
    """getLevelNodeStarts (structure.h:859-874): a node is its child-mask word plus one word per child."""
    out, cur = [], 0
    while cur < len(words):
        out.append(cur)
        cur += 1 + bin(int(words[cur]) & 0xFFFFFFFF).count("1")
    return out


def static_to_edit(d):
    """Packed levels -> 8 slots per node (SENTINEL for an absent child), pointers as node indices:
    the EditStructure view of the same structure, which `expand` walks."""
    levels = [np.zeros((0, 8), dtype=np.uint64) for _ in d["words"]]
    index_of = None  # start -> index of the level below (None at level 3: leaf indices)
    for level in range(3, len(d["words"])):
        w = d["words"][level]
        starts = node_starts(w)
        arr = np.full((len(starts), 8), SENTINEL, dtype=np.uint64)
        for i, st in enumerate(starts):
            mask = int(w[st]) & 0xFFFFFFFF
            assert mask < 256
            o = 1
            for c in range(8):
                if (mask >> c) & 1:
                    tp = int(w[st + o])
                    o += 1
                    ptr = tp & ((1 << 37) - 1)
                    if index_of is not None:
                        ptr = index_of[ptr]
                    arr[i, c] = (tp & ~((1 << 37) - 1)) | ptr
        levels[level] = arr
        index_of = {st: i for i, st in enumerate(starts)}
    return dict(levels=levels, leaves=d["leaves"], resolution=d["resolution"], root_level=d["root_level"],
                structure_type=d["structure_type"])


def _leaf_cube(mask):
    cube = np.zeros((4, 4, 4), dtype=bool)  # [x, y, z]
    m = int(mask)
    for i in range(64):
        if (m >> i) & 1:
            x = (i & 1) | ((i >> 3) & 1) << 1
            y = ((i >> 1) & 1) | ((i >> 4) & 1) << 1
            z = ((i >> 2) & 1) | ((i >> 5) & 1) << 1
            cube[x, y, z] = True
    return cube


def _apply(cube, tp):
    perm = ((tp >> 40) & 3, (tp >> 42) & 3, (tp >> 44) & 3)
    sym = (tp >> 37) & 7
    t = tuple(((tp >> s) & 63) - 32 for s in (46, 52, 58))
    out = np.transpose(cube, perm)
    for a in range(3):
        if (sym >> a) & 1:
            out = np.flip(out, axis=a)
    if t != (0, 0, 0):
        R = out.shape[0]
        res = np.zeros_like(out)
        src = tuple(slice(max(0, -ta), min(R, R - ta)) for ta in t)
        dst = tuple(slice(max(0, ta), min(R, R + ta)) for ta in t)
        res[dst] = out[src]
        out = res
    return out


def expand(d):
    """Dense [x, y, z] boolean grid of the whole structure."""
    cache = {}

    def node(level, idx):
        key = (level, idx)
        if key in cache:
            return cache[key]
        if level == 2:
            c = _leaf_cube(d["leaves"][idx])
        else:
            R = 1 << level
            h = R // 2
            c = np.zeros((R, R, R), dtype=bool)
            for i, tp in enumerate(d["levels"][level][idx]):
                tp = int(tp)
                if tp == SENTINEL:
                    continue
                child = _apply(node(level - 1, tp & ((1 << 37) - 1)), tp)
                ox, oy, oz = (i & 1) * h, ((i >> 1) & 1) * h, ((i >> 2) & 1) * h
                c[ox:ox + h, oy:oy + h, oz:oz + h] = child
        cache[key] = c
        return c

    root = d["root_level"]
    if len(d["levels"][root]) == 0:
        R = d["resolution"]
        return np.zeros((R, R, R), dtype=bool)
    return node(root, 0)


def dense_to_cube(bits: np.ndarray, R: int) -> np.ndarray:
    """Dense bit grid (bit i = x + R*y + R*R*z) -> [x, y, z] boolean cube."""
    flat = np.unpackbits(bits, bitorder="little")[: R ** 3].astype(bool)
    return flat.reshape(R, R, R).transpose(2, 1, 0)
