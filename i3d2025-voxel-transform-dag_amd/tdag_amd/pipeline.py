"""Pipeline entry points of libtdag.so: scenes, SVDAG producer, driver and .tdag writer.

Mirrors the reference's construction API:

* :class:`Svdag` ~ ``voxcom::EditStructure<void, uint32_t>`` after ``toDAG`` (structure.h:159-211)
* :func:`construct_transform_dag` ~ ``constructTransformDAG`` (transform_dag.h:161-162)
* :meth:`TDag.tdag_bytes` ~ the voxelizer's ``.tdag`` output (voxelizer/src/main.cpp:144-151)
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import TdagConfig, TransformDAGConfig, check, lib

_u64p = ctypes.POINTER(ctypes.c_uint64)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u8p = ctypes.POINTER(ctypes.c_uint8)

KEEP_LEVEL_INPUTS = 1


class LevelRecord(ctypes.Structure):
    _fields_ = [
        ("level", ctypes.c_uint32),
        ("has_masks", ctypes.c_uint32),
        ("items", ctypes.c_uint64),
        ("unique", ctypes.c_uint64),
        ("matched", ctypes.c_uint64),
        ("ms_match", ctypes.c_float),
        ("learned", ctypes.c_uint32),
        ("proposal_hits", ctypes.c_uint64),
        ("fallback_queries", ctypes.c_uint64),
    ]


def bind(L):
    vp = ctypes.c_void_p
    L.tdag_scene_dense.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64, _u8p, ctypes.c_uint64]
    L.tdag_svdag_from_dense.argtypes = [_u8p, ctypes.c_uint32, ctypes.POINTER(vp)]
    L.tdag_svdag_info.argtypes = [vp, _u32p, _u32p, _u64p, ctypes.c_uint32]
    L.tdag_svdag_copy_level.argtypes = [vp, ctypes.c_uint32, vp]
    L.tdag_svdag_flatten.argtypes = [vp, ctypes.c_uint32, _u32p, ctypes.c_uint64, _u64p]
    L.tdag_svdag_free.argtypes = [vp]
    L.tdag_construct_transform_dag.argtypes = [vp, vp, ctypes.POINTER(TdagConfig), ctypes.c_uint32, ctypes.POINTER(vp)]
    L.tdag_construct_transform_dag_learned.argtypes = [vp, vp, ctypes.POINTER(TdagConfig), ctypes.c_uint32, vp, ctypes.c_uint32, ctypes.POINTER(vp)]
    L.tdag_construct_transform_dag_learned.restype = ctypes.c_int
    L.tdag_tdag_serialize.argtypes = [vp, ctypes.POINTER(TdagConfig), _u8p, ctypes.c_uint64, _u64p]
    L.tdag_tdag_info.argtypes = [vp, _u32p, _u64p, ctypes.c_uint32]
    L.tdag_tdag_level_record.argtypes = [vp, ctypes.c_uint32, ctypes.POINTER(LevelRecord)]
    L.tdag_tdag_level_masks.argtypes = [vp, ctypes.c_uint32, _u64p, ctypes.c_uint64]
    L.tdag_tdag_free.argtypes = [vp]
    for name in ("tdag_scene_dense", "tdag_svdag_from_dense", "tdag_svdag_info", "tdag_svdag_copy_level", "tdag_svdag_flatten",
                 "tdag_svdag_free", "tdag_construct_transform_dag", "tdag_tdag_serialize", "tdag_tdag_info",
                 "tdag_tdag_level_record", "tdag_tdag_level_masks", "tdag_tdag_free"):
        getattr(L, name).restype = ctypes.c_int


def scene_dense(kind: int, resolution: int, seed: int) -> np.ndarray:
    """Dense bit grid (uint8, bit i = x + R*y + R*R*z) of a seeded synthetic scene."""
    nbytes = resolution ** 3 // 8
    bits = np.zeros(nbytes, dtype=np.uint8)
    check(lib().tdag_scene_dense(kind, resolution, seed, bits.ctypes.data_as(_u8p), nbytes), "tdag_scene_dense")
    return bits


class Svdag:
    """An SVDAG built from a dense grid by the library (EditStructure + toDAG)."""

    def __init__(self, bits: np.ndarray, resolution: int):
        b = np.ascontiguousarray(bits, dtype=np.uint8)
        self._h = ctypes.c_void_p()
        check(lib().tdag_svdag_from_dense(b.ctypes.data_as(_u8p), resolution, ctypes.byref(self._h)), "tdag_svdag_from_dense")
        r = ctypes.c_uint32()
        root = ctypes.c_uint32()
        sizes = np.zeros(16, dtype=np.uint64)
        check(lib().tdag_svdag_info(self._h, ctypes.byref(r), ctypes.byref(root), sizes.ctypes.data_as(_u64p), 16), "tdag_svdag_info")
        self.resolution = int(r.value)
        self.root_level = int(root.value)
        self.level_sizes = [int(v) for v in sizes[: self.root_level + 1]]

    @property
    def handle(self):
        return self._h

    def leaves(self) -> np.ndarray:
        out = np.zeros(max(1, self.level_sizes[2]), dtype=np.uint64)
        check(lib().tdag_svdag_copy_level(self._h, 2, out.ctypes.data_as(ctypes.c_void_p)), "tdag_svdag_copy_level")
        return out[: self.level_sizes[2]]

    def nodes(self, level: int) -> np.ndarray:
        n = self.level_sizes[level]
        out = np.zeros((max(1, n), 8), dtype=np.uint32)
        check(lib().tdag_svdag_copy_level(self._h, level, out.ctypes.data_as(ctypes.c_void_p)), "tdag_svdag_copy_level")
        return out[:n]

    def flatten(self, level: int, nodes) -> np.ndarray:
        idx = np.ascontiguousarray(nodes, dtype=np.uint32)
        W = max(1, (1 << (3 * level)) // 64)
        out = np.zeros((max(1, idx.size), W), dtype=np.uint64)
        check(lib().tdag_svdag_flatten(self._h, level, idx.ctypes.data_as(_u32p), idx.size, out.ctypes.data_as(_u64p)), "tdag_svdag_flatten")
        return out[: idx.size]

    def close(self):
        if self._h:
            lib().tdag_svdag_free(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class TDag:
    """Result of constructTransformDAG: EditStructure<void, TransformPointer> + level records."""

    def __init__(self, handle, config: TransformDAGConfig):
        self._h = handle
        self.config = config
        root = ctypes.c_uint32()
        sizes = np.zeros(16, dtype=np.uint64)
        check(lib().tdag_tdag_info(self._h, ctypes.byref(root), sizes.ctypes.data_as(_u64p), 16), "tdag_tdag_info")
        self.root_level = int(root.value)
        self.level_sizes = [int(v) for v in sizes[: self.root_level + 1]]

    def tdag_bytes(self) -> bytes:
        cfg = self.config.to_c()
        size = ctypes.c_uint64()
        check(lib().tdag_tdag_serialize(self._h, ctypes.byref(cfg), None, 0, ctypes.byref(size)), "tdag_tdag_serialize")
        buf = np.zeros(int(size.value), dtype=np.uint8)
        check(lib().tdag_tdag_serialize(self._h, ctypes.byref(cfg), buf.ctypes.data_as(_u8p), buf.size, ctypes.byref(size)), "tdag_tdag_serialize")
        return buf.tobytes()

    def record(self, level: int) -> LevelRecord | None:
        rec = LevelRecord()
        if lib().tdag_tdag_level_record(self._h, level, ctypes.byref(rec)) != 0:
            return None
        return rec

    def level_masks(self, level: int) -> np.ndarray:
        rec = self.record(level)
        W = max(1, (1 << (3 * level)) // 64)
        out = np.zeros((max(1, rec.items), W), dtype=np.uint64)
        check(lib().tdag_tdag_level_masks(self._h, level, out.ctypes.data_as(_u64p), out.size), "tdag_tdag_level_masks")
        return out[: rec.items]

    def close(self):
        if self._h:
            lib().tdag_tdag_free(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def construct_transform_dag_learned(ctx, svdag: Svdag, config: TransformDAGConfig, mlp, k: int = 64, keep_level_inputs: bool = False) -> TDag:
    """constructTransformDAG with the learned proposal (a tdag_amd.learned.MLP) at level 2."""
    cfg = config.to_c()
    h = ctypes.c_void_p()
    check(lib().tdag_construct_transform_dag_learned(ctx._ptr, svdag.handle, ctypes.byref(cfg), KEEP_LEVEL_INPUTS if keep_level_inputs else 0, mlp._ptr,
                                                     int(k), ctypes.byref(h)), "tdag_construct_transform_dag_learned")
    return TDag(h, config)


def construct_transform_dag(ctx, svdag: Svdag, config: TransformDAGConfig, keep_level_inputs: bool = False) -> TDag:
    """constructTransformDAG on the GPU matcher. ``ctx`` may be None when translation is off
    (the hierarchical pass runs on the host and no matcher call is made)."""
    cfg = config.to_c()
    h = ctypes.c_void_p()
    check(lib().tdag_construct_transform_dag(ctx._ptr if ctx is not None else None, svdag.handle, ctypes.byref(cfg),
                                             KEEP_LEVEL_INPUTS if keep_level_inputs else 0, ctypes.byref(h)),
          "tdag_construct_transform_dag")
    return TDag(h, config)
