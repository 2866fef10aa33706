"""ctypes bindings of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg import this
module, and only as the checker. The product library never loads it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

_u64p = ctypes.POINTER(ctypes.c_uint64)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u8p = ctypes.POINTER(ctypes.c_uint8)


class OrcConfig(ctypes.Structure):
    """Byte layout of voxcom::TransformDAGConfig (transform_dag.h:157-160)."""

    _fields_ = [
        ("symmetry", ctypes.c_uint8),
        ("axisPermutation", ctypes.c_uint8),
        ("translation", ctypes.c_uint8),
        ("pad", ctypes.c_uint8),
        ("maxTranslationLevel", ctypes.c_uint32),
    ]


def build() -> str:
    """Compile the oracle with its Makefile (g++)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_tp_create.restype = ctypes.c_uint64
        L.orc_tp_create.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_tp_sentinel.restype = ctypes.c_uint64
        L.orc_encode_fixed64.restype = ctypes.c_uint64
        L.orc_encode_fixed64.argtypes = [ctypes.c_uint64, ctypes.c_uint32]
        L.orc_decode_fixed64.restype = ctypes.c_uint64
        L.orc_decode_fixed64.argtypes = [ctypes.c_uint64, ctypes.c_uint32]
        L.orc_transform.argtypes = [ctypes.c_uint32, _u64p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, ctypes.c_int, ctypes.c_int, _u64p]
        L.orc_corner_hash.argtypes = [ctypes.c_uint32, _u64p, ctypes.c_uint32, _u64p]
        L.orc_match_encoded.argtypes = [ctypes.c_uint32, _u64p, ctypes.c_uint64, ctypes.POINTER(OrcConfig), _u64p, ctypes.c_int, ctypes.c_int]
        L.orc_match_timed.argtypes = [ctypes.c_uint32, _u64p, ctypes.c_uint64, ctypes.POINTER(OrcConfig), ctypes.c_int, ctypes.c_uint32,
                                      ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double), _u64p]
        L.orc_postprocess.argtypes = [ctypes.c_uint32, _u64p, ctypes.c_uint64, _u64p, _u64p, _u32p, _u64p]
        if hasattr(L, "orc_svdag_build"):
            _bind_svdag(L)
        L.orc_huffman_code_lengths.argtypes = [_u64p, ctypes.c_uint32, ctypes.c_uint32, _u32p]
        L.orc_huffman_code_words.argtypes = [_u32p, ctypes.c_uint32, ctypes.c_int, _u64p]
        L.orc_ray_march.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p, ctypes.c_void_p]
        L.orc_tdag16_decode.argtypes = [ctypes.c_char_p, ctypes.c_uint64, ctypes.c_void_p, ctypes.c_uint64]
        L.orc_tdag16_decode.restype = ctypes.c_int64
        _lib = L
    return _lib


def _bind_svdag(L):
    L.orc_svdag_build.argtypes = [_u8p, ctypes.c_uint32, ctypes.c_void_p]
    L.orc_svdag_build.restype = ctypes.c_int
    L.orc_svdag_free.argtypes = [ctypes.c_void_p]
    L.orc_svdag_level_size.argtypes = [ctypes.c_void_p, ctypes.c_uint32]
    L.orc_svdag_level_size.restype = ctypes.c_uint64
    L.orc_svdag_copy_level.argtypes = [ctypes.c_void_p, ctypes.c_uint32, _u32p]
    L.orc_svdag_copy_leaves.argtypes = [ctypes.c_void_p, _u64p]
    L.orc_svdag_root_level.argtypes = [ctypes.c_void_p]
    L.orc_svdag_root_level.restype = ctypes.c_uint32
    L.orc_flatten.argtypes = [ctypes.c_void_p, ctypes.c_uint32, _u32p, ctypes.c_uint64, _u64p]
    L.orc_flatten.restype = ctypes.c_int
    L.orc_construct_tdag.argtypes = [ctypes.c_void_p, ctypes.POINTER(OrcConfig), ctypes.c_int, ctypes.c_char_p]
    L.orc_construct_tdag.restype = ctypes.c_int64
    L.orc_last_level_inputs.argtypes = [ctypes.c_uint32, _u64p, ctypes.c_uint64]
    L.orc_last_level_inputs.restype = ctypes.c_int64
    L.orc_last_level_encoded.argtypes = [ctypes.c_uint32, _u64p, ctypes.c_uint64]
    L.orc_last_level_encoded.restype = ctypes.c_int64
    L.orc_svdag_static_level.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_uint64]
    L.orc_svdag_static_level.restype = ctypes.c_int64
    L.orc_construct_static_tdag.argtypes = [ctypes.c_void_p, ctypes.POINTER(OrcConfig), ctypes.c_int, ctypes.c_char_p]
    L.orc_construct_static_tdag.restype = ctypes.c_int64


def huffman_code_lengths(freq, max_len: int) -> np.ndarray:
    """Length-limited Huffman code lengths by package-merge (huffman.cpp:113-183)."""
    f = np.ascontiguousarray(freq, dtype=np.uint64)
    out = np.zeros(len(f), dtype=np.uint32)
    if lib().orc_huffman_code_lengths(_p64(f), len(f), max_len, out.ctypes.data_as(_u32p)) != 0:
        raise ValueError("package-merge failed (too few symbols for the length limit?)")
    return out


def huffman_code_words(lengths, lsb_first: bool = False) -> np.ndarray:
    """Canonical code words for the given lengths (huffman.cpp:93-111), optionally bit-reversed
    to LSB-first order (huffman.cpp:192-202)."""
    ln = np.ascontiguousarray(lengths, dtype=np.uint32)
    out = np.zeros(len(ln), dtype=np.uint64)
    lib().orc_huffman_code_words(ln.ctypes.data_as(_u32p), len(ln), int(lsb_first), _p64(out))
    return out


def tdag16_decode(data: bytes) -> np.ndarray:
    """Decodes a .tdag16 stream voxel by voxel into a dense (R, R, R) bool grid indexed [z, y, x]."""
    L = lib()
    R = int(L.orc_tdag16_decode(data, len(data), None, 0))
    if R < 0:
        raise ValueError("malformed .tdag16 stream")
    bits = np.zeros(R * R * R // 8, dtype=np.uint8)
    if L.orc_tdag16_decode(data, len(data), bits.ctypes.data, bits.nbytes) < 0:
        raise ValueError("malformed .tdag16 stream")
    return np.unpackbits(bits, bitorder="little").reshape(R, R, R).astype(bool)


def _p64(a):
    return a.ctypes.data_as(_u64p)


def config(symmetry=True, axis=True, translation=True, max_level=4) -> OrcConfig:
    return OrcConfig(int(symmetry), int(axis), int(translation), 0, int(max_level))


def tp_create(ptr, sym=0, perm=0, tx=0, ty=0, tz=0) -> int:
    return int(lib().orc_tp_create(ptr, sym, perm, tx, ty, tz))


def tp_sentinel() -> int:
    return int(lib().orc_tp_sentinel())


def encode_fixed64(tp_bits: int, level: int) -> int:
    return int(lib().orc_encode_fixed64(tp_bits, level))


def decode_fixed64(key: int, level: int) -> int:
    return int(lib().orc_decode_fixed64(key, level))


def words_per_mask(level: int) -> int:
    return max(1, (1 << (3 * level)) // 64)


def transform(level: int, mask: np.ndarray, perm: int, sym: int, t) -> np.ndarray:
    W = words_per_mask(level)
    src = np.ascontiguousarray(mask, dtype=np.uint64).reshape(W)
    out = np.zeros(W, dtype=np.uint64)
    rc = lib().orc_transform(level, _p64(src), perm, sym, int(t[0]), int(t[1]), int(t[2]), _p64(out))
    assert rc == 0
    return out


def corner_hash(level: int, mask: np.ndarray, dir_bits: int) -> int:
    W = words_per_mask(level)
    src = np.ascontiguousarray(mask, dtype=np.uint64).reshape(W)
    out = np.zeros(1, dtype=np.uint64)
    assert lib().orc_corner_hash(level, _p64(src), dir_bits, _p64(out)) == 0
    return int(out[0])


def match_encoded(level: int, masks: np.ndarray, cfg: OrcConfig | None = None, nthreads: int | None = None, brute=False) -> np.ndarray:
    """encodedParents of findDuplicatesAmongFlattenedItems_cpp2, indexed by original item."""
    cfg = cfg or config()
    W = words_per_mask(level)
    m = np.ascontiguousarray(masks, dtype=np.uint64).reshape(-1, W)
    n = m.shape[0]
    out = np.zeros(max(n, 1), dtype=np.uint64)
    nthreads = nthreads or os.cpu_count() or 1
    rc = lib().orc_match_encoded(level, _p64(m), n, ctypes.byref(cfg), _p64(out), int(nthreads), int(brute))
    assert rc == 0
    return out[:n]


def match_timed(level: int, masks: np.ndarray, cfg: OrcConfig | None = None, nthreads: int | None = None, query_stride: int = 1):
    """Wall times of the hash search with the full table and every query_stride-th query rank:
    (table build seconds, search seconds, queries searched)."""
    cfg = cfg or config()
    W = words_per_mask(level)
    m = np.ascontiguousarray(masks, dtype=np.uint64).reshape(-1, W)
    setup, search = ctypes.c_double(0.0), ctypes.c_double(0.0)
    q = np.zeros(1, dtype=np.uint64)
    rc = lib().orc_match_timed(level, _p64(m), m.shape[0], ctypes.byref(cfg), int(nthreads or os.cpu_count() or 1), int(query_stride),
                               ctypes.byref(setup), ctypes.byref(search), _p64(q))
    assert rc == 0
    return setup.value, search.value, int(q[0])


def postprocess(level: int, masks: np.ndarray, encoded: np.ndarray):
    """(mapping, unique_indices) exactly as find_translation.cpp:923-997 produces them."""
    W = words_per_mask(level)
    m = np.ascontiguousarray(masks, dtype=np.uint64).reshape(-1, W)
    n = m.shape[0]
    enc = np.ascontiguousarray(encoded, dtype=np.uint64)
    mapping = np.zeros(max(n, 1), dtype=np.uint64)
    uniq = np.zeros(max(n, 1), dtype=np.uint32)
    nu = np.zeros(1, dtype=np.uint64)
    rc = lib().orc_postprocess(level, _p64(m), n, _p64(enc), _p64(mapping), uniq.ctypes.data_as(_u32p), _p64(nu))
    assert rc == 0
    return mapping[:n], uniq[: int(nu[0])]


def ray_march(bits: np.ndarray, R: int, rays: np.ndarray):
    """First occupied voxel along each ray (n x 8 float32: origin, tmin, direction, tmax) through
    the dense grid bits (bit i = x + R*y + R*R*z): (hits n x 3 int32, -1 for none; t float64)."""
    b = np.ascontiguousarray(bits, dtype=np.uint8)
    r = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
    hits = np.zeros((max(1, len(r)), 3), dtype=np.int32)
    t = np.zeros(max(1, len(r)), dtype=np.float64)
    assert lib().orc_ray_march(b.ctypes.data, R, r.ctypes.data, len(r), hits.ctypes.data, t.ctypes.data) == 0
    return hits[: len(r)], t[: len(r)]
