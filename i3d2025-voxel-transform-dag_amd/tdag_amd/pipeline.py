"""Pipeline entry points of libtdag.so: scenes, SVDAG producer, driver and .tdag writer.

Mirrors the reference's construction API:

* :class:`Svdag` ~ ``voxcom::EditStructure<void, uint32_t>`` after ``toDAG`` (structure.h:159-211)
* :func:`construct_transform_dag` ~ ``constructTransformDAG`` (transform_dag.h:161-162)
* :meth:`TDag.tdag_bytes` ~ the voxelizer's ``.tdag`` output (voxelizer/src/main.cpp:144-151)
* :class:`StaticSvdag` ~ ``voxcom::StaticStructure<void, uint32_t>`` (structure.h:299-320)
* :func:`construct_static_transform_dag` ~ ``constructStaticTransformDAG`` (transform_dag.h:163)
* :class:`TDag16` ~ ``TransformDAG16`` and ``constructTransformDAG16``
  (tools/svdag_encoding/src/transform_dag_encoding.h:138-153, transform_dag_encoding.cpp:241-478)
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import TdagConfig, TransformDAGConfig, check, lib

_u64p = ctypes.POINTER(ctypes.c_uint64)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u8p = ctypes.POINTER(ctypes.c_uint8)

KEEP_LEVEL_INPUTS = 1
DEVICE_TRANSLATION = 2  # TDAG_DEVICE_TRANSLATION
POINTER_TABLES = 1  # TDAG16_POINTER_TABLES
HUFFMAN = 2  # TDAG16_HUFFMAN


class StaticSvdagView(ctypes.Structure):
    _fields_ = [
        ("resolution", ctypes.c_uint32),
        ("root_level", ctypes.c_uint32),
        ("level_words", ctypes.POINTER(_u32p)),
        ("level_word_counts", _u64p),
        ("leaves", _u64p),
        ("num_leaves", ctypes.c_uint64),
    ]


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


class SvdagView(ctypes.Structure):
    """tdag_svdag_view: EditStructure<void, uint32_t> as borrowed arrays (include/tdag.h)."""

    _fields_ = [
        ("resolution", ctypes.c_uint32),
        ("root_level", ctypes.c_uint32),
        ("nodes_per_level", ctypes.POINTER(ctypes.POINTER(ctypes.c_uint32))),
        ("node_counts", ctypes.POINTER(ctypes.c_uint64)),
        ("subgrids", ctypes.POINTER(ctypes.c_uint64)),
        ("num_subgrids", ctypes.c_uint64),
        ("structure_type", ctypes.c_int32),
        ("reserved", ctypes.c_uint32),
    ]


def bind(L):
    vp = ctypes.c_void_p
    L.tdag_scene_dense.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64, _u8p, ctypes.c_uint64]
    L.tdag_svdag_from_dense.argtypes = [_u8p, ctypes.c_uint32, ctypes.POINTER(vp)]
    L.tdag_svdag_info.argtypes = [vp, _u32p, _u32p, _u64p, ctypes.c_uint32]
    L.tdag_svdag_from_dense_device.argtypes = [vp, vp, ctypes.c_uint32, ctypes.POINTER(vp), ctypes.POINTER(ctypes.c_float)]
    L.tdag_svdag_from_dense_device.restype = ctypes.c_int
    L.tdag_svdag_copy_level.argtypes = [vp, ctypes.c_uint32, vp]
    L.tdag_svdag_flatten.argtypes = [vp, ctypes.c_uint32, _u32p, ctypes.c_uint64, _u64p]
    L.tdag_svdag_flatten_device.argtypes = [vp, vp, ctypes.c_uint32, vp, ctypes.c_uint64, vp]
    L.tdag_svdag_flatten_device.restype = ctypes.c_int
    L.tdag_svdag_free.argtypes = [vp]
    L.tdag_construct_transform_dag.argtypes = [vp, vp, ctypes.POINTER(TdagConfig), ctypes.c_uint32, ctypes.POINTER(vp)]
    L.tdag_construct_transform_dag_learned.argtypes = [vp, vp, ctypes.POINTER(TdagConfig), ctypes.c_uint32, vp, ctypes.c_uint32, ctypes.POINTER(vp)]
    L.tdag_construct_transform_dag_learned.restype = ctypes.c_int
    L.tdag_tdag_serialize.argtypes = [vp, ctypes.POINTER(TdagConfig), _u8p, ctypes.c_uint64, _u64p]
    L.tdag_tdag_info.argtypes = [vp, _u32p, _u64p, ctypes.c_uint32]
    L.tdag_tdag_level_record.argtypes = [vp, ctypes.c_uint32, ctypes.POINTER(LevelRecord)]
    L.tdag_tdag_level_masks.argtypes = [vp, ctypes.c_uint32, _u64p, ctypes.c_uint64]
    L.tdag_tdag_free.argtypes = [vp]
    L.tdag_svdag_static_words.argtypes = [vp, ctypes.c_uint32, vp, ctypes.c_uint64, _u64p]
    L.tdag_construct_static_transform_dag.argtypes = [vp, ctypes.POINTER(StaticSvdagView), ctypes.POINTER(TdagConfig), ctypes.c_uint32, ctypes.POINTER(vp)]
    L.tdag_tdag_static_words.argtypes = [vp, ctypes.c_uint32, _u64p, ctypes.c_uint64, _u64p]
    L.tdag_tdag_static_serialize.argtypes = [vp, _u8p, ctypes.c_uint64, _u64p]
    L.tdag_tdag16_encode.argtypes = [vp, ctypes.POINTER(TdagConfig), ctypes.c_uint32, ctypes.POINTER(vp)]
    L.tdag_tdag16_from_bytes.argtypes = [ctypes.c_char_p, ctypes.c_uint64, ctypes.POINTER(vp)]
    L.tdag_tdag16_serialize.argtypes = [vp, _u8p, ctypes.c_uint64, _u64p]
    L.tdag_tdag16_info.argtypes = [vp, _u32p, _u32p, _u64p, _u64p, ctypes.c_uint32]
    L.tdag_tdag16_decode_device.argtypes = [vp, vp, vp, ctypes.c_uint64, ctypes.POINTER(ctypes.c_float)]
    L.tdag_tdag16_free.argtypes = [vp]
    L.tdag_tdag16_trace_device.argtypes = [vp, vp, vp, ctypes.c_uint64, vp, vp, ctypes.POINTER(ctypes.c_float)]
    L.tdag_tdag16_trace_device.restype = ctypes.c_int
    L.tdag_svdag_from_view.argtypes = [ctypes.POINTER(SvdagView), ctypes.POINTER(vp)]
    L.tdag_svdag_from_view.restype = ctypes.c_int
    L.tdag_tdag_copy_level.argtypes = [vp, ctypes.c_uint32, _u64p, ctypes.c_uint64, _u64p]
    L.tdag_tdag_copy_level.restype = ctypes.c_int
    L.tdag_build_graph_create.argtypes = [vp, vp, ctypes.POINTER(TdagConfig), ctypes.POINTER(vp)]
    L.tdag_build_graph_run.argtypes = [vp, ctypes.POINTER(ctypes.c_float)]
    L.tdag_build_graph_result.argtypes = [vp, ctypes.POINTER(vp)]
    L.tdag_build_graph_destroy.argtypes = [vp]
    for name in ("tdag_build_graph_create", "tdag_build_graph_run", "tdag_build_graph_result", "tdag_build_graph_destroy"):
        getattr(L, name).restype = ctypes.c_int
    L.tdag_huffman_code_lengths.argtypes = [_u64p, ctypes.c_uint32, ctypes.c_uint32, _u32p]
    L.tdag_huffman_code_words.argtypes = [_u32p, ctypes.c_uint32, ctypes.c_int, _u64p]
    for name in ("tdag_scene_dense", "tdag_svdag_from_dense", "tdag_svdag_info", "tdag_svdag_copy_level", "tdag_svdag_flatten",
                 "tdag_svdag_free", "tdag_construct_transform_dag", "tdag_tdag_serialize", "tdag_tdag_info",
                 "tdag_tdag_level_record", "tdag_tdag_level_masks", "tdag_tdag_free", "tdag_svdag_static_words",
                 "tdag_construct_static_transform_dag", "tdag_tdag_static_words", "tdag_tdag_static_serialize"):
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
        self._load_info()

    @classmethod
    def from_device(cls, ctx, d_bits, resolution: int):
        """Built on the GPU from a dense grid in device memory (a torch uint8 tensor of R^3/8
        bytes, bit i = x + R*y + R*R*z). Sets ``build_ms`` to the wall time of the call."""
        if not d_bits.is_cuda or not d_bits.is_contiguous() or d_bits.numel() * d_bits.element_size() < resolution ** 3 // 8:
            raise ValueError("d_bits must be a contiguous device tensor of at least resolution^3 / 8 bytes")
        import torch

        torch.cuda.synchronize()  # the library works on its own stream
        self = cls.__new__(cls)
        self._h = ctypes.c_void_p()
        ms = ctypes.c_float()
        check(lib().tdag_svdag_from_dense_device(ctx._ptr, ctypes.c_void_p(d_bits.data_ptr()), resolution, ctypes.byref(self._h), ctypes.byref(ms)),
              "tdag_svdag_from_dense_device")
        self.build_ms = float(ms.value)
        self._load_info()
        return self

    @classmethod
    def from_view(cls, resolution: int, root_level: int, nodes: dict, subgrids, structure_type: int = 1):
        """tdag_svdag_from_view of an EditStructure<void, uint32_t> given as arrays: nodes[level]
        (level 3..root) of shape (count, 8) uint32 child indices, subgrids the uint64 leaf masks."""
        keep = {}
        ptrs = (ctypes.POINTER(ctypes.c_uint32) * (root_level + 1))()
        counts = (ctypes.c_uint64 * (root_level + 1))()
        for level in range(3, root_level + 1):
            a = np.ascontiguousarray(nodes.get(level, np.zeros((0, 8), dtype=np.uint32)), dtype=np.uint32).reshape(-1, 8)
            keep[level] = a
            ptrs[level] = a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))
            counts[level] = a.shape[0]
        sg = np.ascontiguousarray(subgrids, dtype=np.uint64).reshape(-1)
        view = SvdagView(resolution, root_level, ptrs, counts, sg.ctypes.data_as(_u64p), sg.size, structure_type, 0)
        self = cls.__new__(cls)
        self._h = ctypes.c_void_p()
        check(lib().tdag_svdag_from_view(ctypes.byref(view), ctypes.byref(self._h)), "tdag_svdag_from_view")
        self._load_info()
        return self

    def _load_info(self):
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

    def flatten_device(self, ctx, level: int, d_nodes, d_out):
        """createLargeSubGrid on the GPU: d_nodes (int32/uint32 device tensor) -> d_out (int64
        device tensor of len(d_nodes) * W words)."""
        n = d_nodes.numel()
        W = max(1, (1 << (3 * level)) // 64)
        if not (d_nodes.is_cuda and d_out.is_cuda and d_nodes.element_size() == 4 and d_out.element_size() == 8 and d_out.numel() >= n * W):
            raise ValueError("d_nodes must be a 4-byte device tensor and d_out an 8-byte device tensor of n * W words")
        import torch

        torch.cuda.synchronize()
        check(lib().tdag_svdag_flatten_device(ctx._ptr, self._h, level, ctypes.c_void_p(d_nodes.data_ptr()), n, ctypes.c_void_p(d_out.data_ptr())),
              "tdag_svdag_flatten_device")

    def static_words(self, level: int) -> np.ndarray:
        """StaticStructure(EditStructure) packed words of `level` (uint32; level 2: uint64 leaves)."""
        n = ctypes.c_uint64()
        check(lib().tdag_svdag_static_words(self._h, level, None, 0, ctypes.byref(n)), "tdag_svdag_static_words")
        out = np.zeros(max(1, n.value), dtype=np.uint64 if level == 2 else np.uint32)
        check(lib().tdag_svdag_static_words(self._h, level, out.ctypes.data_as(ctypes.c_void_p), out.size, ctypes.byref(n)), "tdag_svdag_static_words")
        return out[: n.value]

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

    def static_words(self, level: int) -> np.ndarray:
        """Packed StaticStructure<void, TransformPointer> words of `level` (level 2: leaves)."""
        n = ctypes.c_uint64()
        check(lib().tdag_tdag_static_words(self._h, level, None, 0, ctypes.byref(n)), "tdag_tdag_static_words")
        out = np.zeros(max(1, n.value), dtype=np.uint64)
        check(lib().tdag_tdag_static_words(self._h, level, out.ctypes.data_as(_u64p), out.size, ctypes.byref(n)), "tdag_tdag_static_words")
        return out[: n.value]

    def static_bytes(self) -> bytes:
        """StaticStructure<void, TransformPointer>::writeTo bytes (structure.h:888-895)."""
        size = ctypes.c_uint64()
        check(lib().tdag_tdag_static_serialize(self._h, None, 0, ctypes.byref(size)), "tdag_tdag_static_serialize")
        buf = np.zeros(max(1, int(size.value)), dtype=np.uint8)
        check(lib().tdag_tdag_static_serialize(self._h, buf.ctypes.data_as(_u8p), buf.size, ctypes.byref(size)), "tdag_tdag_static_serialize")
        return buf[: int(size.value)].tobytes()

    def record(self, level: int) -> LevelRecord | None:
        rec = LevelRecord()
        if lib().tdag_tdag_level_record(self._h, level, ctypes.byref(rec)) != 0:
            return None
        return rec

    def nodes(self, level: int) -> np.ndarray:
        """Level `level` of the EditStructure<void, TransformPointer>: (count, 8) uint64 TransformPointer
        bit patterns (level >= 3) or the uint64 leaf masks (level 2)."""
        n = ctypes.c_uint64()
        check(lib().tdag_tdag_copy_level(self._h, level, None, 0, ctypes.byref(n)), "tdag_tdag_copy_level")
        out = np.zeros(max(1, n.value), dtype=np.uint64)
        check(lib().tdag_tdag_copy_level(self._h, level, out.ctypes.data_as(_u64p), out.size, ctypes.byref(n)), "tdag_tdag_copy_level")
        out = out[: n.value]
        return out if level == 2 else out.reshape(-1, 8)

    def level_masks(self, level: int) -> np.ndarray:
        rec = self.record(level)
        W = max(1, (1 << (3 * level)) // 64)
        out = np.zeros((max(1, rec.items), W), dtype=np.uint64)
        check(lib().tdag_tdag_level_masks(self._h, level, out.ctypes.data_as(_u64p), out.size), "tdag_tdag_level_masks")
        return out[: rec.items]

    def encode16(self, pointer_tables: bool = True, huffman: bool = True) -> "TDag16":
        """constructTransformDAG16 of this result (transform_dag_encoding.cpp:241-478)."""
        cfg = self.config.to_c()
        h = ctypes.c_void_p()
        flags = (POINTER_TABLES if pointer_tables else 0) | (HUFFMAN if huffman else 0)
        check(lib().tdag_tdag16_encode(self._h, ctypes.byref(cfg), flags, ctypes.byref(h)), "tdag_tdag16_encode")
        return TDag16(h)

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


def construct_transform_dag(ctx, svdag: Svdag, config: TransformDAGConfig, keep_level_inputs: bool = False, device_translation: bool = False) -> TDag:
    """constructTransformDAG on the GPU matcher. ``ctx`` may be None when translation is off
    (the hierarchical pass runs on the host and no matcher call is made). With
    ``device_translation`` every step of the translation stage runs on the device
    (TDAG_DEVICE_TRANSLATION)."""
    cfg = config.to_c()
    h = ctypes.c_void_p()
    flags = (KEEP_LEVEL_INPUTS if keep_level_inputs else 0) | (DEVICE_TRANSLATION if device_translation else 0)
    check(lib().tdag_construct_transform_dag(ctx._ptr if ctx is not None else None, svdag.handle, ctypes.byref(cfg), flags, ctypes.byref(h)),
          "tdag_construct_transform_dag")
    return TDag(h, config)


class BuildGraph:
    """A build's translation stage captured as one hipGraph (tdag_build_graph_*)."""

    def __init__(self, ctx, svdag: Svdag, config: TransformDAGConfig):
        self._ctx = ctx  # the graph runs on the context's stream (it keeps its own SVDAG copy)
        self.config = config
        cfg = config.to_c()
        self._h = ctypes.c_void_p()
        check(lib().tdag_build_graph_create(ctx._ptr, svdag.handle, ctypes.byref(cfg), ctypes.byref(self._h)), "tdag_build_graph_create")

    def run(self) -> float:
        """Replays the stage; returns its device time in milliseconds."""
        ms = ctypes.c_float()
        check(lib().tdag_build_graph_run(self._h, ctypes.byref(ms)), "tdag_build_graph_run")
        return float(ms.value)

    def result(self) -> TDag:
        h = ctypes.c_void_p()
        check(lib().tdag_build_graph_result(self._h, ctypes.byref(h)), "tdag_build_graph_result")
        return TDag(h, self.config)

    def close(self):
        if self._h:
            lib().tdag_build_graph_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class StaticSvdag:
    """StaticStructure<void, uint32_t>: packed words per level (3..root) and the leaf masks."""

    def __init__(self, resolution: int, root_level: int, level_words: dict, leaves):
        self.resolution = int(resolution)
        self.root_level = int(root_level)
        self.level_words = {l: np.ascontiguousarray(level_words.get(l, np.zeros(0, np.uint32)), dtype=np.uint32) for l in range(3, self.root_level + 1)}
        self.leaves = np.ascontiguousarray(leaves, dtype=np.uint64)

    @classmethod
    def from_svdag(cls, svdag: Svdag) -> "StaticSvdag":
        """StaticStructure(const EditStructure&) (structure.h:766-798), done by the library."""
        return cls(svdag.resolution, svdag.root_level, {l: svdag.static_words(l) for l in range(3, svdag.root_level + 1)}, svdag.static_words(2))

    def view(self):
        n = self.root_level + 1
        ptrs = (_u32p * n)()
        counts = np.zeros(n, dtype=np.uint64)
        for l, w in self.level_words.items():
            ptrs[l] = w.ctypes.data_as(_u32p)
            counts[l] = w.size
        v = StaticSvdagView(self.resolution, self.root_level, ctypes.cast(ptrs, ctypes.POINTER(_u32p)), counts.ctypes.data_as(_u64p),
                            self.leaves.ctypes.data_as(_u64p), self.leaves.size)
        return v, (ptrs, counts)  # keep the arrays alive while the view is used


def construct_static_transform_dag(ctx, static: StaticSvdag, config: TransformDAGConfig, keep_level_inputs: bool = False) -> TDag:
    """constructStaticTransformDAG on the GPU matcher (``ctx`` may be None when translation is
    off). Read the result with :meth:`TDag.static_words` / :meth:`TDag.static_bytes`."""
    cfg = config.to_c()
    v, keep = static.view()
    h = ctypes.c_void_p()
    check(lib().tdag_construct_static_transform_dag(ctx._ptr if ctx is not None else None, ctypes.byref(v), ctypes.byref(cfg),
                                                    KEEP_LEVEL_INPUTS if keep_level_inputs else 0, ctypes.byref(h)),
          "tdag_construct_static_transform_dag")
    del keep
    return TDag(h, config)


class TDag16:
    """TransformDAG16: the .tdag16 compact encoding (transform_dag_encoding.h:138-153)."""

    def __init__(self, handle):
        self._h = handle
        res, root = ctypes.c_uint32(), ctypes.c_uint32()
        words = np.zeros(32, dtype=np.uint64)
        tables = np.zeros(32, dtype=np.uint64)
        check(lib().tdag_tdag16_info(self._h, ctypes.byref(res), ctypes.byref(root), words.ctypes.data_as(_u64p),
                                     tables.ctypes.data_as(_u64p), 32), "tdag_tdag16_info")
        self.resolution = int(res.value)
        self.root_level = int(root.value)
        self.level_words = [int(v) for v in words[: self.root_level + 1]]
        self.table_entries = [int(v) for v in tables[: self.root_level + 1]]

    @classmethod
    def from_bytes(cls, data: bytes) -> "TDag16":
        """TransformDAG16::readFrom (transform_dag_encoding.cpp:1029-1036)."""
        h = ctypes.c_void_p()
        check(lib().tdag_tdag16_from_bytes(data, len(data), ctypes.byref(h)), "tdag_tdag16_from_bytes")
        return cls(h)

    def to_bytes(self) -> bytes:
        """TransformDAG16::writeTo (transform_dag_encoding.cpp:1020-1027)."""
        size = ctypes.c_uint64()
        check(lib().tdag_tdag16_serialize(self._h, None, 0, ctypes.byref(size)), "tdag_tdag16_serialize")
        buf = np.zeros(max(1, int(size.value)), dtype=np.uint8)
        check(lib().tdag_tdag16_serialize(self._h, buf.ctypes.data_as(_u8p), buf.size, ctypes.byref(size)), "tdag_tdag16_serialize")
        return buf[: int(size.value)].tobytes()

    def decode_device(self, ctx, d_out=None):
        """Every voxel through TranslateTransformBaseDAG::get on the GPU (transform_dag_encoding.cpp:
        897-942). Returns (d_out, ms): d_out an int32 device tensor of R^3/32 words, bit i = x + R*y
        + R*R*z; ms the device time of the decode kernel."""
        import torch

        R = self.resolution
        n = max(1, R * R * R // 32)
        if d_out is None:
            d_out = torch.empty(n, dtype=torch.int32, device=f"cuda:{ctx.device}")
        if not (d_out.is_cuda and d_out.element_size() == 4 and d_out.is_contiguous() and d_out.numel() >= n):
            raise ValueError("d_out must be a contiguous 4-byte device tensor of resolution^3 / 32 words")
        torch.cuda.synchronize()
        ms = ctypes.c_float()
        check(lib().tdag_tdag16_decode_device(ctx._ptr, self._h, ctypes.c_void_p(d_out.data_ptr()), d_out.numel(), ctypes.byref(ms)),
              "tdag_tdag16_decode_device")
        return d_out, float(ms.value)

    def trace_device(self, ctx, d_rays, n: int, d_hits=None, d_t=None):
        """First-hit ray casting on the device (tdag_tdag16_trace_device). d_rays: float32 device
        tensor of n x 8 (origin, tmin, direction, tmax). Returns (hits int32 n x 4: voxel x y z or
        -1, steps; t float32 n; device ms)."""
        import torch

        if d_hits is None:
            d_hits = torch.empty((max(1, n), 4), dtype=torch.int32, device=d_rays.device)
        if d_t is None:
            d_t = torch.empty(max(1, n), dtype=torch.float32, device=d_rays.device)
        torch.cuda.synchronize()
        ms = ctypes.c_float()
        check(lib().tdag_tdag16_trace_device(ctx._ptr, self._h, ctypes.c_void_p(d_rays.data_ptr()), int(n), ctypes.c_void_p(d_hits.data_ptr()),
                                             ctypes.c_void_p(d_t.data_ptr()), ctypes.byref(ms)), "tdag_tdag16_trace_device")
        return d_hits, d_t, float(ms.value)

    def close(self):
        if self._h:
            lib().tdag_tdag16_free(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def huffman_code_lengths(freq, max_len: int) -> np.ndarray:
    """createLengthLimitedHuffmanCodeLengths (huffman.cpp:113-183)."""
    f = np.ascontiguousarray(freq, dtype=np.uint64)
    out = np.zeros(len(f), dtype=np.uint32)
    check(lib().tdag_huffman_code_lengths(f.ctypes.data_as(_u64p), len(f), max_len, out.ctypes.data_as(_u32p)), "tdag_huffman_code_lengths")
    return out


def huffman_code_words(lengths, lsb_first: bool = False) -> np.ndarray:
    """assignCodeWords (huffman.cpp:93-111), optionally convertHuffmanToLSB (192-202)."""
    ln = np.ascontiguousarray(lengths, dtype=np.uint32)
    out = np.zeros(len(ln), dtype=np.uint64)
    check(lib().tdag_huffman_code_words(ln.ctypes.data_as(_u32p), len(ln), int(lsb_first), out.ctypes.data_as(_u64p)), "tdag_huffman_code_words")
    return out
