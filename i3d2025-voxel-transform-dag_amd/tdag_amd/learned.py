"""Learned proposal for the level-2 matcher (bindings of the tdag_mlp_* and
tdag_match_level_learned_device entry points of include/tdag.h).

The network, its training and the proposal/verification kernels all run on the GPU; this module
only moves arguments. Training pairs come from the exact matcher: every leaf whose minimum key
names another leaf forms a (leaf, parent) pair.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import TdagError, TransformDAGConfig, TdagConfig, check, lib

PARAMS = None


class TdagLearnedStats(ctypes.Structure):
    _fields_ = [
        ("num_items", ctypes.c_uint64),
        ("num_unique", ctypes.c_uint64),
        ("num_matched", ctypes.c_uint64),
        ("proposal_hits", ctypes.c_uint64),
        ("fallback_queries", ctypes.c_uint64),
        ("ms_embed", ctypes.c_float),
        ("ms_propose", ctypes.c_float),
        ("ms_verify", ctypes.c_float),
        ("ms_fallback", ctypes.c_float),
        ("ms_total", ctypes.c_float),
    ]

    def as_dict(self) -> dict:
        return {k: (float(getattr(self, k)) if k.startswith("ms_") else int(getattr(self, k))) for k, _ in self._fields_}


_bound = False


RETRIEVAL_ALL = 0  # TDAG_RETRIEVAL_ALL
RETRIEVAL_BUCKETED = 1  # TDAG_RETRIEVAL_BUCKETED


def _bind():
    global _bound, PARAMS
    if _bound:
        return lib()
    L = lib()
    vp, u64, u32, f32 = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_float
    L.tdag_mlp_param_count.restype = ctypes.c_uint64
    L.tdag_mlp_create.argtypes = [vp, u64, ctypes.POINTER(vp)]
    L.tdag_mlp_destroy.argtypes = [vp]
    L.tdag_mlp_get_params.argtypes = [vp, vp]
    L.tdag_mlp_set_params.argtypes = [vp, vp]
    L.tdag_mlp_loss_grad.argtypes = [vp, vp, vp, u32, f32, ctypes.POINTER(f32), vp]
    L.tdag_mlp_train.argtypes = [vp, vp, vp, u64, u32, u32, f32, f32, u64, vp]
    L.tdag_mlp_embed.argtypes = [vp, vp, u64, vp]
    L.tdag_gemm_bf16.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp, ctypes.c_int, vp, ctypes.c_int, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.tdag_mlp_topk.argtypes = [vp, vp, u64, u32, vp]
    L.tdag_mlp_topk_bucketed.argtypes = [vp, vp, vp, u64, u32, vp]
    L.tdag_mlp_set_retrieval.argtypes = [vp, u32]
    L.tdag_mlp_get_retrieval.argtypes = [vp, ctypes.POINTER(u32)]
    L.tdag_match_level_learned_device.argtypes = [vp, vp, u32, vp, u64, ctypes.POINTER(TdagConfig), u32, vp, vp, vp, ctypes.POINTER(TdagLearnedStats)]
    for name in ("tdag_mlp_create", "tdag_mlp_destroy", "tdag_mlp_get_params", "tdag_mlp_set_params", "tdag_mlp_loss_grad", "tdag_mlp_train", "tdag_mlp_embed",
                 "tdag_gemm_bf16", "tdag_mlp_topk", "tdag_mlp_topk_bucketed", "tdag_mlp_set_retrieval", "tdag_mlp_get_retrieval",
                 "tdag_match_level_learned_device"):
        getattr(L, name).restype = ctypes.c_int
    PARAMS = int(L.tdag_mlp_param_count())
    _bound = True
    return L


def _ptr(x):
    if x is None:
        return None
    if isinstance(x, int):
        return ctypes.c_void_p(x)
    return ctypes.c_void_p(x.data_ptr())


def param_count() -> int:
    _bind()
    return PARAMS


class MLP:
    """A network on the device of `ctx` (a tdag_amd.Context)."""

    def __init__(self, ctx, seed: int = 0):
        L = _bind()
        self._ctx = ctx
        self._ptr = ctypes.c_void_p()
        check(L.tdag_mlp_create(ctx._ptr, int(seed), ctypes.byref(self._ptr)), "tdag_mlp_create")

    def close(self):
        if self._ptr:
            lib().tdag_mlp_destroy(self._ptr)
            self._ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self) -> np.ndarray:
        out = np.empty(PARAMS, dtype=np.float32)
        check(lib().tdag_mlp_get_params(self._ptr, out.ctypes.data_as(ctypes.c_void_p)), "tdag_mlp_get_params")
        return out

    def set_params(self, params: np.ndarray) -> None:
        p = np.ascontiguousarray(params, dtype=np.float32)
        assert p.size == PARAMS
        check(lib().tdag_mlp_set_params(self._ptr, p.ctypes.data_as(ctypes.c_void_p)), "tdag_mlp_set_params")

    def loss_grad(self, d_q, d_p, batch: int, tau: float):
        """(loss, gradient as float32[PARAMS]) of InfoNCE over device mask pairs."""
        loss = ctypes.c_float()
        g = np.empty(PARAMS, dtype=np.float32)
        check(lib().tdag_mlp_loss_grad(self._ptr, _ptr(d_q), _ptr(d_p), int(batch), float(tau), ctypes.byref(loss), g.ctypes.data_as(ctypes.c_void_p)),
              "tdag_mlp_loss_grad")
        return float(loss.value), g

    def train(self, d_q, d_p, n_pairs: int, steps: int, batch: int = 512, lr: float = 1e-3, tau: float = 0.1, seed: int = 0) -> np.ndarray:
        """Adam steps on the device pairs; returns the per-step losses."""
        losses = np.zeros(max(1, steps), dtype=np.float32)
        check(lib().tdag_mlp_train(self._ptr, _ptr(d_q), _ptr(d_p), int(n_pairs), int(steps), int(batch), float(lr), float(tau), int(seed),
                                   losses.ctypes.data_as(ctypes.c_void_p)), "tdag_mlp_train")
        return losses[:steps]

    def embed(self, d_masks, n: int, d_out) -> None:
        check(lib().tdag_mlp_embed(self._ptr, _ptr(d_masks), int(n), _ptr(d_out)), "tdag_mlp_embed")

    @property
    def retrieval(self) -> int:
        """RETRIEVAL_ALL (the k best of all earlier leaves) or RETRIEVAL_BUCKETED (the k best of the
        earlier leaves with the query's transform-invariant descriptor)."""
        mode = ctypes.c_uint32()
        check(lib().tdag_mlp_get_retrieval(self._ptr, ctypes.byref(mode)), "tdag_mlp_get_retrieval")
        return int(mode.value)

    @retrieval.setter
    def retrieval(self, mode: int) -> None:
        check(lib().tdag_mlp_set_retrieval(self._ptr, int(mode)), "tdag_mlp_set_retrieval")


def gemm_bf16(ctx, ta: bool, tb: bool, d_a, lda: int, d_b, ldb: int, d_c, ldc: int, m: int, n: int, k: int) -> None:
    L = _bind()
    check(L.tdag_gemm_bf16(ctx._ptr, int(ta), int(tb), _ptr(d_a), int(lda), _ptr(d_b), int(ldb), _ptr(d_c), int(ldc), int(m), int(n), int(k)), "tdag_gemm_bf16")


def topk(ctx, d_emb, n: int, k: int, d_cand) -> None:
    """Retrieval step: d_cand (n x k uint32) = the k best earlier rows of d_emb (n x 128 bf16 bits)
    by dot product, ties to the lower row, best first (0xFFFFFFFF where fewer exist)."""
    L = _bind()
    check(L.tdag_mlp_topk(ctx._ptr, _ptr(d_emb), int(n), int(k), _ptr(d_cand)), "tdag_mlp_topk")


def topk_bucketed(ctx, d_masks, d_emb, n: int, k: int, d_cand) -> None:
    """Bucketed retrieval step: as topk, restricted to the earlier rows whose masks (d_masks, n
    4^3 masks in rank order) have the query's transform-invariant descriptor; scores are fp32 sums
    of the bf16 products in dimension order."""
    L = _bind()
    check(L.tdag_mlp_topk_bucketed(ctx._ptr, _ptr(d_masks), _ptr(d_emb), int(n), int(k), _ptr(d_cand)), "tdag_mlp_topk_bucketed")


def match_level_learned_device(ctx, mlp: MLP | None, d_masks, n: int, config: TransformDAGConfig | None = None, k: int = 64, d_encoded=None,
                               d_mapping=None, d_unique=None) -> TdagLearnedStats:
    """Level-2 matcher with proposals (mlp None: popcount-neighbour proposals)."""
    L = _bind()
    cfg = (config or TransformDAGConfig()).to_c()
    st = TdagLearnedStats()
    check(L.tdag_match_level_learned_device(ctx._ptr, mlp._ptr if mlp is not None else None, 2, _ptr(d_masks), int(n), ctypes.byref(cfg), int(k),
                                            _ptr(d_encoded), _ptr(d_mapping), _ptr(d_unique), ctypes.byref(st)), "tdag_match_level_learned_device")
    return st


def rank_order(masks: np.ndarray) -> np.ndarray:
    """Stable popcount-descending order of 4^3 masks (the matcher's rank order,
    find_translation.cpp:631-643)."""
    m = np.asarray(masks, dtype=np.uint64).reshape(-1)
    pc = np.array([bin(int(v)).count("1") for v in m], dtype=np.int64)
    return np.argsort(-pc, kind="stable")


def pairs_from_encoded(masks: np.ndarray, encoded: np.ndarray):
    """(query masks, parent masks) for every leaf whose minimum key names another leaf."""
    m = np.asarray(masks, dtype=np.uint64).reshape(-1)
    order = rank_order(m)
    rank_of = np.empty_like(order)
    rank_of[order] = np.arange(len(order))
    ptr = (np.asarray(encoded, dtype=np.uint64) >> np.uint64(6 + 3 * 3)).astype(np.int64)
    sel = (ptr != rank_of) & (ptr < len(m)) & (m != 0)
    q = m[sel]
    p = m[order[ptr[sel]]]
    return q, p


def training_pairs(ctx, config: TransformDAGConfig, kind: int, resolution: int, seeds):
    """Training data of the network: the (query, parent) 4^3 pairs the exact GPU matcher emits at
    level 2 of constructTransformDAG on seeded scenes. Returns (d_q, d_p, n) device tensors."""
    import torch

    from . import pipeline as pl

    qs, ps = [], []
    for seed in seeds:
        bits = pl.scene_dense(kind, resolution, seed)
        t = pl.construct_transform_dag(ctx, pl.Svdag(bits, resolution), config, keep_level_inputs=True)
        m2 = np.ascontiguousarray(t.level_masks(2).reshape(-1))
        t.close()
        d_m = torch.from_numpy(m2.view(np.int64)).cuda()
        d_enc = torch.empty(max(1, len(m2)), dtype=torch.int64, device="cuda")
        torch.cuda.synchronize()
        ctx.match_level_device(2, d_m, len(m2), config, d_encoded=d_enc)
        q, p = pairs_from_encoded(m2, d_enc.cpu().numpy().view(np.uint64)[: len(m2)])
        qs.append(q)
        ps.append(p)
    q, p = np.concatenate(qs), np.concatenate(ps)
    return torch.from_numpy(q.view(np.int64).copy()).cuda(), torch.from_numpy(p.view(np.int64).copy()).cuda(), len(q)
