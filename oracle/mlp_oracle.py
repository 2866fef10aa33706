"""CPU oracle of the learned proposal (test infrastructure only; nothing in the library calls it).

The learned part has no counterpart in the reference. It is pinned here by mathematics:
  * descriptors: closed form, invariant under the 48 mirror/axis-permutation transforms and
    under translations that clip nothing (checked exhaustively in tests against the transform
    restatement of find_translation.cpp:66-125);
  * MLP + InfoNCE loss: float64 forward and analytic backward, whose gradients are checked
    against central finite differences in tests;
  * a bf16 emulation mode that rounds exactly where the GPU kernels round (weights, layer
    inputs, back-propagated signals), so the GPU gradients can be compared with a tight
    tolerance.

Network: 76 inputs (64 mask bits of a 4^3 grid, bit i = x + 4y + 16z, plus the 12 descriptor
bytes / 16) padded to 80, three hidden ReLU layers of 256, a 128-d output, L2-normalised.
Loss over a batch of B (query, positive) pairs: mean_i [logsumexp_j S_ij - S_ii] with
S_ij = <e_q_i, e_p_j> / tau.
"""
from __future__ import annotations

import numpy as np

IN = 80
FEAT = 76
DIMS = [(80, 256), (256, 256), (256, 256), (256, 128)]
EMB = 128


def param_count() -> int:
    return sum(i * o + o for i, o in DIMS)


def split_params(flat: np.ndarray):
    """Flat parameter vector -> [(W (in x out, row-major), b (out))] in layer order."""
    out, off = [], 0
    for i, o in DIMS:
        W = flat[off:off + i * o].reshape(i, o)
        off += i * o
        b = flat[off:off + o]
        off += o
        out.append((W, b))
    return out


def histograms(mask: int) -> np.ndarray:
    """3 x 4 per-axis occupancy counts of a 4^3 mask."""
    h = np.zeros((3, 4), dtype=np.int64)
    for i in range(64):
        if (mask >> i) & 1:
            x, y, z = i & 3, (i >> 2) & 3, i >> 4
            h[0, x] += 1
            h[1, y] += 1
            h[2, z] += 1
    return h


def _span(r) -> tuple:
    """Histogram with its zero ends trimmed, left-aligned in 4 slots (invariant under any
    translation that clips nothing)."""
    nz = [i for i, v in enumerate(r) if v]
    if not nz:
        return (0, 0, 0, 0)
    core = list(r[nz[0]:nz[-1] + 1])
    return tuple(core + [0] * (4 - len(core)))


def descriptor(mask: int) -> np.ndarray:
    """12 bytes invariant under mirror, axis permutation and unclipped translation: each axis
    histogram is trimmed and left-aligned, replaced by the lexicographically smaller of itself
    and its reversed span (mirror), and the three are sorted (permutation)."""
    h = histograms(int(mask))
    canon = []
    for r in h:
        a = _span(r)
        b = _span(list(r)[::-1])
        canon.append(min(a, b))
    canon.sort()
    return np.array([v for r in canon for v in r], dtype=np.uint8)


def features(masks: np.ndarray) -> np.ndarray:
    """(n, 80) float64 inputs: bits, descriptor / 16, zero padding."""
    masks = np.asarray(masks, dtype=np.uint64).reshape(-1)
    X = np.zeros((len(masks), IN), dtype=np.float64)
    for r, m in enumerate(masks):
        m = int(m)
        X[r, :64] = [(m >> i) & 1 for i in range(64)]
        X[r, 64:76] = descriptor(m) / 16.0
    return X


def bf16(a: np.ndarray) -> np.ndarray:
    """Round float values to bfloat16 (round to nearest, ties to even), returned as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def forward(params: np.ndarray, X: np.ndarray, emulate_bf16: bool = False):
    """Returns (e, cache). With emulate_bf16, weights and each layer's input are rounded to bf16
    (the GPU's MFMA operands) and products are accumulated in float32-like precision."""
    layers = split_params(params)
    rnd = bf16 if emulate_bf16 else (lambda v: v)
    h = rnd(X)
    acts = [h]
    for li, (W, b) in enumerate(layers):
        pre = h @ rnd(W) + b
        if li < len(layers) - 1:
            h = rnd(np.maximum(pre, 0.0))
            acts.append(h)
        else:
            z = pre
    norm = np.sqrt(np.sum(z * z, axis=1, keepdims=True))
    e = z / norm
    return e, {"acts": acts, "z": z, "norm": norm, "e": e}


def infonce(e: np.ndarray, B: int, tau: float):
    """Loss and d loss / d e for rows [0, B) = queries, [B, 2B) = positives."""
    q, p = e[:B], e[B:2 * B]
    S = q @ p.T / tau
    S = S - S.max(axis=1, keepdims=True)
    P = np.exp(S)
    P /= P.sum(axis=1, keepdims=True)
    loss = float(np.mean(-np.log(np.diag(P))))
    G = (P - np.eye(B)) / B / tau
    de = np.zeros_like(e)
    de[:B] = G @ p
    de[B:2 * B] = G.T @ q
    return loss, de


def loss_and_grad(params: np.ndarray, X: np.ndarray, B: int, tau: float, emulate_bf16: bool = False):
    """Loss and the flat parameter gradient for the 2B feature rows X."""
    layers = split_params(params)
    rnd = bf16 if emulate_bf16 else (lambda v: v)
    e, c = forward(params, X, emulate_bf16)
    loss, de = infonce(e, B, tau)
    z, norm = c["z"], c["norm"]
    dz = (de - e * np.sum(e * de, axis=1, keepdims=True)) / norm
    grads = [None] * len(layers)
    g = dz
    for li in range(len(layers) - 1, -1, -1):
        W, _ = layers[li]
        h = c["acts"][li]
        gr = rnd(g)
        grads[li] = (h.T @ gr, g.sum(axis=0))
        if li > 0:
            g = (gr @ rnd(W).T) * (c["acts"][li] > 0)
    flat = np.concatenate([np.concatenate([dW.reshape(-1), db]) for dW, db in grads])
    return loss, flat


def init_params(seed: int) -> np.ndarray:
    """He-uniform weights, zero biases (any initialisation works for the tests)."""
    rng = np.random.default_rng(seed)
    parts = []
    for i, o in DIMS:
        a = np.sqrt(6.0 / i)
        parts.append(rng.uniform(-a, a, size=i * o))
        parts.append(np.zeros(o))
    return np.concatenate(parts)
