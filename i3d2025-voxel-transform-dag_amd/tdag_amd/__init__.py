"""Python host mirror of the MI355X-native Transform-Aware SVDAG matching path.

The compute lives in ``libtdag.so`` (hand-written HIP for gfx950 behind the C-ABI declared in
``include/tdag.h``). This module only binds that ABI with ctypes, moves numpy/torch buffers
across it, and mirrors the reference's entry points:

* :func:`find_duplicates_among_flattened_items_gpu` mirrors
  ``findDuplicatesAmongFlattenedItems_gpu<Level, Item>`` (transform_dag.h:310-311): it returns
  the unique items and ``outPrevLevelMapping`` as TransformPointer bit patterns.
* :meth:`Context.match_level_device` is the device-resident hot path used by the benchmark.

There is no CPU fallback: if ``libtdag.so`` is missing or no GPU is present, the calls raise.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtdag.so")

_u64p = ctypes.POINTER(ctypes.c_uint64)
_u32p = ctypes.POINTER(ctypes.c_uint32)

TDAG_OK = 0
VARIANT_SLOTS = 32  # TDAG_VARIANT_SLOTS in include/tdag.h
_STATUS = {0: "OK", 1: "EINVAL", 2: "ENOMEM", 3: "EDEVICE", 4: "EINTERNAL"}


class TdagConfig(ctypes.Structure):
    """Byte layout of voxcom::TransformDAGConfig (transform_dag.h:157-160)."""

    _fields_ = [
        ("symmetry", ctypes.c_uint8),
        ("axis_permutation", ctypes.c_uint8),
        ("translation", ctypes.c_uint8),
        ("pad", ctypes.c_uint8),
        ("max_translation_level", ctypes.c_uint32),
    ]


class TdagMatchStats(ctypes.Structure):
    _fields_ = [
        ("num_items", ctypes.c_uint64),
        ("num_unique", ctypes.c_uint64),
        ("num_matched", ctypes.c_uint64),
        ("table_entries", ctypes.c_uint64),
        ("ms_total", ctypes.c_float),
        ("ms_search", ctypes.c_float),
        ("probes", ctypes.c_uint64),
        ("filter_pass", ctypes.c_uint64),
        ("key_hits", ctypes.c_uint64),
        ("candidates", ctypes.c_uint64),
        ("verifications", ctypes.c_uint64),
        ("verified", ctypes.c_uint64),
    ]

    def counters(self) -> dict:
        return {k: int(getattr(self, k)) for k in ("probes", "filter_pass", "key_hits", "candidates", "verifications", "verified")}


class TdagLevelJob(ctypes.Structure):
    """One level of a batched matcher call (tdag_level_job in include/tdag.h)."""

    _fields_ = [
        ("level", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
        ("n", ctypes.c_uint64),
        ("d_masks", ctypes.c_void_p),
        ("d_encoded", ctypes.c_void_p),
        ("d_mapping", ctypes.c_void_p),
        ("d_unique", ctypes.c_void_p),
    ]


@dataclass
class TransformDAGConfig:
    """Mirror of voxcom::TransformDAGConfig with the reference defaults."""

    symmetry: bool = True
    axis_permutation: bool = True
    translation: bool = True
    max_translation_level: int = 4

    def to_c(self) -> TdagConfig:
        return TdagConfig(int(self.symmetry), int(self.axis_permutation), int(self.translation), 0, int(self.max_translation_level))


class TdagError(RuntimeError):
    """Raised when a C-ABI call returns a non-OK status (mirrors assert_always's throw)."""


_lib = None


def lib() -> ctypes.CDLL:
    """Loads libtdag.so. Raises if it was not built: there is no fallback path."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise TdagError(f"{LIB_PATH} is missing: build it with `python build.py` (hipcc, gfx950)")
    L = ctypes.CDLL(LIB_PATH)
    L.tdag_last_error.argtypes = [ctypes.c_char_p, ctypes.c_size_t]
    L.tdag_last_error.restype = ctypes.c_size_t
    L.tdag_version.restype = ctypes.c_char_p
    L.tdag_context_create.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
    L.tdag_context_destroy.argtypes = [ctypes.c_void_p]
    L.tdag_match_level_device.argtypes = [
        ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_uint64, ctypes.POINTER(TdagConfig),
        ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, _u64p, ctypes.c_void_p, ctypes.POINTER(TdagMatchStats),
    ]
    L.tdag_match_level.argtypes = [
        ctypes.c_void_p, ctypes.c_uint32, _u64p, ctypes.c_uint64, ctypes.POINTER(TdagConfig),
        _u64p, _u64p, _u32p, _u64p, ctypes.POINTER(TdagMatchStats),
    ]
    L.tdag_match_level_device_noalloc.argtypes = [
        ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_uint64, ctypes.POINTER(TdagConfig),
        ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
    ]
    L.tdag_match_level_subset_device.argtypes = [
        ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_uint64, ctypes.POINTER(TdagConfig), ctypes.c_void_p, ctypes.c_void_p,
        ctypes.POINTER(TdagMatchStats),
    ]
    L.tdag_last_stats.argtypes = [ctypes.c_void_p, ctypes.POINTER(TdagMatchStats)]
    L.tdag_context_set_variant.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int32]
    L.tdag_context_set_counters.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.tdag_context_get_variant.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.POINTER(ctypes.c_int32)]
    L.tdag_autotune_level.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_uint64, ctypes.POINTER(TdagConfig), ctypes.c_uint32,
                                      ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_uint32)]
    L.tdag_context_set_batching.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32]
    L.tdag_context_set_filter_bits.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32]
    L.tdag_context_get_filter_bits.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32)]
    L.tdag_context_set_child_frames.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32]
    L.tdag_context_get_child_frames.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32)]
    L.tdag_context_get_batching.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32)]
    L.tdag_batch_create.argtypes = [ctypes.c_void_p, ctypes.POINTER(TdagLevelJob), ctypes.c_uint32, ctypes.POINTER(TdagConfig), ctypes.c_int,
                                    ctypes.POINTER(ctypes.c_void_p)]
    L.tdag_batch_run.argtypes = [ctypes.c_void_p, ctypes.POINTER(TdagMatchStats), ctypes.POINTER(ctypes.c_float)]
    L.tdag_batch_destroy.argtypes = [ctypes.c_void_p]
    L.tdag_batch_autotune.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32)]
    L.tdag_match_levels_batched_device.argtypes = [ctypes.c_void_p, ctypes.POINTER(TdagLevelJob), ctypes.c_uint32, ctypes.POINTER(TdagConfig),
                                                   ctypes.POINTER(TdagMatchStats)]
    L.tdag_batch_create_learned.argtypes = [ctypes.c_void_p, ctypes.POINTER(TdagLevelJob), ctypes.c_uint32, ctypes.POINTER(TdagConfig), ctypes.c_uint32,
                                            ctypes.c_void_p, ctypes.c_uint32, ctypes.POINTER(ctypes.c_void_p)]
    L.tdag_batch_learned_stats.argtypes = [ctypes.c_void_p, _u64p, _u64p]
    L.tdag_batch_set_tuning.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32]
    L.tdag_batch_get_tuning.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32)]
    for name in ("tdag_batch_create", "tdag_batch_run", "tdag_batch_destroy", "tdag_batch_autotune", "tdag_match_levels_batched_device",
                 "tdag_batch_create_learned", "tdag_batch_learned_stats", "tdag_batch_set_tuning", "tdag_batch_get_tuning"):
        getattr(L, name).restype = ctypes.c_int
    for name in ("tdag_context_create", "tdag_context_destroy", "tdag_match_level_device", "tdag_match_level_device_noalloc", "tdag_match_level_subset_device", "tdag_match_level", "tdag_last_stats",
                 "tdag_context_set_variant", "tdag_context_get_variant", "tdag_autotune_level", "tdag_context_set_counters",
                 "tdag_context_set_batching", "tdag_context_get_batching", "tdag_context_set_filter_bits", "tdag_context_get_filter_bits",
                 "tdag_context_set_child_frames", "tdag_context_get_child_frames"):
        getattr(L, name).restype = ctypes.c_int
    _extra_bindings(L)
    _lib = L
    return L


def _extra_bindings(L):
    """Bindings of the pipeline entry points (SVDAG build, driver, .tdag writer)."""
    if hasattr(L, "tdag_scene_dense"):
        from . import pipeline  # noqa: F401  (binds its own symbols)

        pipeline.bind(L)


def last_error() -> str:
    buf = ctypes.create_string_buffer(4096)
    lib().tdag_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


def check(status: int, what: str) -> None:
    if status != TDAG_OK:
        raise TdagError(f"{what} failed with {_STATUS.get(status, status)}: {last_error()}")


def _dev_ptr(x):
    if x is None:
        return None
    if isinstance(x, int):
        return x
    return x.data_ptr()


def _jobs_array(jobs):
    arr = (TdagLevelJob * len(jobs))()
    keep = []
    for i, j in enumerate(jobs):
        arr[i].level = int(j["level"])
        arr[i].n = int(j["n"])
        for k in ("d_masks", "d_encoded", "d_mapping", "d_unique"):
            setattr(arr[i], k, _dev_ptr(j.get(k)))
            keep.append(j.get(k))
    return arr, keep


class Batch:
    """A planned (and, with use_graph, captured) batched matcher call; see Context.batch."""

    GRAPH = 1  # TDAG_BATCH_GRAPH
    CONCURRENT = 2  # TDAG_BATCH_CONCURRENT

    def __init__(self, ctx: "Context", jobs, config, use_graph: bool, concurrent: bool = False, mlp=None, k: int = 64):
        self._ptr = ctypes.c_void_p()
        self._ctx = ctx
        self._jobs, self._keep = _jobs_array(jobs)
        self.count = len(jobs)
        self.concurrent = bool(concurrent)
        self.learned = mlp is not None
        cfg = (config or TransformDAGConfig()).to_c()
        flags = (self.GRAPH if use_graph else 0) | (self.CONCURRENT if concurrent else 0)
        if mlp is None:
            check(lib().tdag_batch_create(ctx._ptr, self._jobs, self.count, ctypes.byref(cfg), flags, ctypes.byref(self._ptr)), "tdag_batch_create")
        else:
            self._mlp = mlp  # keeps the network alive as long as the batch
            check(lib().tdag_batch_create_learned(ctx._ptr, self._jobs, self.count, ctypes.byref(cfg), flags, mlp._ptr, int(k), ctypes.byref(self._ptr)),
                  "tdag_batch_create_learned")

    def learned_stats(self):
        """(proposal hits, fallback queries) of the learned level-2 job in the last run."""
        hits, fb = ctypes.c_uint64(), ctypes.c_uint64()
        check(lib().tdag_batch_learned_stats(self._ptr, ctypes.byref(hits), ctypes.byref(fb)), "tdag_batch_learned_stats")
        return int(hits.value), int(fb.value)

    def run(self):
        """Runs the call and waits; returns (per-job stats, milliseconds of the whole call)."""
        stats = (TdagMatchStats * self.count)()
        ms = ctypes.c_float()
        check(lib().tdag_batch_run(self._ptr, stats, ctypes.byref(ms)), "tdag_batch_run")
        return list(stats), float(ms.value)

    def autotune(self, repeats: int = 1) -> int:
        """Tunes the work queue (tiles per pop, then the blocks that start on the lighter levels'
        queue); returns the tiles per pop."""
        g = ctypes.c_uint32()
        check(lib().tdag_batch_autotune(self._ptr, int(repeats), ctypes.byref(g)), "tdag_batch_autotune")
        return int(g.value)

    def set_tuning(self, grab: int, split_every: int):
        check(lib().tdag_batch_set_tuning(self._ptr, int(grab), int(split_every)), "tdag_batch_set_tuning")

    def tuning(self):
        """(tiles per pop, split_every) of the persistent kernel's queue."""
        g, sp = ctypes.c_uint32(), ctypes.c_uint32()
        check(lib().tdag_batch_get_tuning(self._ptr, ctypes.byref(g), ctypes.byref(sp)), "tdag_batch_get_tuning")
        return int(g.value), int(sp.value)

    def close(self):
        if self._ptr:
            lib().tdag_batch_destroy(self._ptr)
            self._ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def version() -> str:
    return lib().tdag_version().decode()


def words_per_mask(level: int) -> int:
    return max(1, (1 << (3 * level)) // 64)


class Context:
    """Owns a tdag_context (HIP stream + cached device scratch) on one device."""

    def __init__(self, device: int = 0):
        self._ptr = ctypes.c_void_p()
        check(lib().tdag_context_create(int(device), ctypes.byref(self._ptr)), "tdag_context_create")
        self.device = device

    def close(self):
        if self._ptr:
            lib().tdag_context_destroy(self._ptr)
            self._ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- device-resident hot path -----------------------------------------------------
    def match_level_device(self, level: int, d_masks, n: int, config: TransformDAGConfig | None = None,
                           d_encoded=None, d_mapping=None, d_unique=None, stream=None, want_stats=True):
        """Runs the matcher on device buffers (torch tensors or raw device pointers).

        Returns (num_unique, stats) when want_stats, else (None, None) without synchronising.
        """
        cfg = (config or TransformDAGConfig()).to_c()

        def ptr(x):
            if x is None:
                return None
            if isinstance(x, int):
                return ctypes.c_void_p(x)
            return ctypes.c_void_p(x.data_ptr())

        nu = ctypes.c_uint64(0)
        st = TdagMatchStats()
        s = ptr(stream) if stream is not None and not hasattr(stream, "cuda_stream") else (
            ctypes.c_void_p(stream.cuda_stream) if stream is not None else None)
        check(lib().tdag_match_level_device(
            self._ptr, int(level), ptr(d_masks), int(n), ctypes.byref(cfg), ptr(d_encoded), ptr(d_mapping), ptr(d_unique),
            ctypes.byref(nu) if want_stats else None, s, ctypes.byref(st) if want_stats else None), "tdag_match_level_device")
        if want_stats:
            return int(nu.value), st
        return None, None

    def match_level_device_noalloc(self, level: int, d_masks, n: int, config: TransformDAGConfig | None = None,
                                   d_encoded=None, d_mapping=None, d_unique=None, d_counts=None, stream=None):
        """The allocation-free matcher (tdag_match_level_device_noalloc): same results as
        match_level_device, scratch in the library's fixed per-GPU array, no synchronisation."""
        cfg = (config or TransformDAGConfig()).to_c()
        s = None if stream is None else ctypes.c_void_p(stream.cuda_stream if hasattr(stream, "cuda_stream") else int(stream))
        check(lib().tdag_match_level_device_noalloc(
            self._ptr, int(level), _dev_ptr(d_masks), int(n), ctypes.byref(cfg), _dev_ptr(d_encoded), _dev_ptr(d_mapping), _dev_ptr(d_unique),
            _dev_ptr(d_counts), s), "tdag_match_level_device_noalloc")

    def match_level_subset_device(self, level: int, d_masks, n: int, d_child_flags, config: TransformDAGConfig | None = None, d_encoded=None):
        """The exhaustive search with only the flagged children (uint32 device tensor, original
        order) in the table (tdag_match_level_subset_device); returns the stats."""
        cfg = (config or TransformDAGConfig()).to_c()
        st = TdagMatchStats()
        check(lib().tdag_match_level_subset_device(self._ptr, int(level), _dev_ptr(d_masks), int(n), ctypes.byref(cfg), _dev_ptr(d_child_flags),
                                                   _dev_ptr(d_encoded), ctypes.byref(st)), "tdag_match_level_subset_device")
        return st

    # -- batched form ------------------------------------------------------------------------
    def batch(self, jobs, config: TransformDAGConfig | None = None, use_graph: bool = True, concurrent: bool = False, mlp=None, k: int = 64) -> "Batch":
        """Batched matcher over several levels (one job per level). By default every level's
        parent search runs in one persistent kernel; with concurrent=True each level runs its own
        tuned matcher on its own stream and the levels overlap. With use_graph the whole call is
        captured once as a hipGraph. `jobs` holds dicts with keys level, n, d_masks and optional
        d_encoded, d_mapping, d_unique (torch tensors or raw device pointers). With `mlp` (a
        learned.MLP of this context) the level-2 job runs the learned proposal matcher (k
        proposals per query, exhaustive fallback) inside the same call."""
        return Batch(self, jobs, config, use_graph, concurrent, mlp, k)

    def match_levels_batched_device(self, jobs, config: TransformDAGConfig | None = None):
        """One-shot batched call (no graph); returns per-job stats."""
        arr, _ = _jobs_array(jobs)
        cfg = (config or TransformDAGConfig()).to_c()
        stats = (TdagMatchStats * len(jobs))()
        check(lib().tdag_match_levels_batched_device(self._ptr, arr, len(jobs), ctypes.byref(cfg), stats), "tdag_match_levels_batched_device")
        return list(stats)

    # -- tuning ---------------------------------------------------------------------------

    def set_counters(self, enable: bool) -> None:
        """Per-stage probe counters in the returned stats (diagnostics)."""
        check(lib().tdag_context_set_counters(self._ptr, int(bool(enable))), "tdag_context_set_counters")

    def set_variant(self, level: int, variant: int) -> None:
        """Search-kernel variant for `level` (-1 = library default)."""
        check(lib().tdag_context_set_variant(self._ptr, int(level), int(variant)), "tdag_context_set_variant")

    def get_variant(self, level: int) -> int:
        v = ctypes.c_int32()
        check(lib().tdag_context_get_variant(self._ptr, int(level), ctypes.byref(v)), "tdag_context_get_variant")
        return int(v.value)

    def autotune(self, level: int, d_masks, n: int, config: TransformDAGConfig | None = None, repeats: int = 2):
        """Times the whole matcher call for every valid search variant (levels 3..5: with 1 and 6
        child frames), then four first-batch sizes with each of the three fastest variants, then
        four filter sizes, then batch growths 3 to 12 around the first batch found, on device
        masks, and keeps the fastest for this context; returns
        (best_variant, first_batch, {variant: call ms}, tuned call ms)."""
        cfg = (config or TransformDAGConfig()).to_c()
        ms = (ctypes.c_float * (VARIANT_SLOTS + 1))()
        best = ctypes.c_int32()
        first = ctypes.c_uint32()
        ptr = ctypes.c_void_p(d_masks if isinstance(d_masks, int) else d_masks.data_ptr())
        check(lib().tdag_autotune_level(self._ptr, int(level), ptr, int(n), ctypes.byref(cfg), int(repeats), ms, ctypes.byref(best),
                                        ctypes.byref(first)), "tdag_autotune_level")
        return int(best.value), int(first.value), {v: round(float(ms[v]), 4) for v in range(VARIANT_SLOTS) if ms[v] >= 0}, round(float(ms[VARIANT_SLOTS]), 4)

    def set_filter_bits(self, level: int, bits_per_entry: int) -> None:
        """Bloom filter bits per table entry (0 = library default: 16 at level 3, 4 elsewhere)."""
        check(lib().tdag_context_set_filter_bits(self._ptr, int(level), int(bits_per_entry)), "tdag_context_set_filter_bits")

    def get_filter_bits(self, level: int) -> int:
        v = ctypes.c_uint32()
        check(lib().tdag_context_get_filter_bits(self._ptr, int(level), ctypes.byref(v)), "tdag_context_get_filter_bits")
        return int(v.value)

    def set_child_frames(self, level: int, frames: int) -> None:
        """Levels 3..5: 1 = parents run all 48 (permutation, direction) combos against 8 table
        entries per child; 6 = parents run 8 directions against 48 entries per child (the
        children in all six permutation frames); 0 = library default."""
        check(lib().tdag_context_set_child_frames(self._ptr, int(level), int(frames)), "tdag_context_set_child_frames")

    def get_child_frames(self, level: int) -> int:
        v = ctypes.c_uint32()
        check(lib().tdag_context_get_child_frames(self._ptr, int(level), ctypes.byref(v)), "tdag_context_get_child_frames")
        return int(v.value)

    def get_batching(self, level: int):
        f, g = ctypes.c_uint32(), ctypes.c_uint32()
        check(lib().tdag_context_get_batching(self._ptr, int(level), ctypes.byref(f), ctypes.byref(g)), "tdag_context_get_batching")
        return int(f.value), int(g.value)

    def set_batching(self, level: int, first_batch: int, growth: int = 2) -> None:
        """Rank-batched search schedule (0, 0 = defaults)."""
        check(lib().tdag_context_set_batching(self._ptr, int(level), int(first_batch), int(growth)), "tdag_context_set_batching")

    # -- host-buffer drop-in ------------------------------------------------------------
    def match_level(self, level: int, masks: np.ndarray, config: TransformDAGConfig | None = None, want_encoded=True):
        """Host arrays in, host arrays out: (encoded, mapping, unique_indices, stats)."""
        W = words_per_mask(level)
        m = np.ascontiguousarray(masks, dtype=np.uint64).reshape(-1, W)
        n = m.shape[0]
        cfg = (config or TransformDAGConfig()).to_c()
        enc = np.zeros(max(n, 1), dtype=np.uint64)
        mapping = np.zeros(max(n, 1), dtype=np.uint64)
        uniq = np.zeros(max(n, 1), dtype=np.uint32)
        nu = ctypes.c_uint64(0)
        st = TdagMatchStats()
        check(lib().tdag_match_level(
            self._ptr, int(level), m.ctypes.data_as(_u64p), n, ctypes.byref(cfg),
            enc.ctypes.data_as(_u64p) if want_encoded else None, mapping.ctypes.data_as(_u64p), uniq.ctypes.data_as(_u32p),
            ctypes.byref(nu), ctypes.byref(st)), "tdag_match_level")
        return (enc[:n] if want_encoded else None), mapping[:n], uniq[: int(nu.value)], st


_default_ctx: Context | None = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


def find_duplicates_among_flattened_items_gpu(level: int, items, grids: np.ndarray, config: TransformDAGConfig | None = None):
    """Mirror of findDuplicatesAmongFlattenedItems_gpu (transform_dag.h:310-311).

    ``items`` is any sequence of per-item payloads (indices, nodes...), ``grids`` the matching
    LargeSubGrid masks (n x W uint64). Returns (unique_items, outPrevLevelMapping) where the
    mapping holds TransformPointer bit patterns (uint64).
    """
    _, mapping, uniq, _ = default_context().match_level(level, grids, config, want_encoded=False)
    items_arr = list(items) if not isinstance(items, np.ndarray) else items
    if isinstance(items_arr, np.ndarray):
        return items_arr[uniq], mapping
    return [items_arr[i] for i in uniq], mapping


# ---------------------------------------------------------------------------------------
# TransformPointer helpers (transform_dag.h:19-54, 349-397), pure bit manipulation.
# ---------------------------------------------------------------------------------------
PERMS = ((0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0))


def tp_unpack(bits: int) -> dict:
    bits = int(bits)
    return {
        "ptr": bits & ((1 << 37) - 1),
        "sym": (bits >> 37) & 7,
        "perm": ((bits >> 40) & 3, (bits >> 42) & 3, (bits >> 44) & 3),
        "t": (((bits >> 46) & 63) - 32, ((bits >> 52) & 63) - 32, ((bits >> 58) & 63) - 32),
    }


def tp_pack(ptr: int, sym: int = 0, perm=(0, 1, 2), t=(0, 0, 0)) -> int:
    v = ptr & ((1 << 37) - 1)
    v |= (sym & 7) << 37
    v |= perm[0] << 40 | perm[1] << 42 | perm[2] << 44
    v |= ((t[0] + 32) & 63) << 46 | ((t[1] + 32) & 63) << 52 | ((t[2] + 32) & 63) << 58
    return v


TP_SENTINEL = tp_pack(0xFFFFFFFF)
