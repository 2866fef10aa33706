"""Pins the learned part's CPU oracle (oracle/mlp_oracle.py) by mathematics.

* Descriptors are invariant under all 48 mirror/axis-permutation transforms and unclipped
  translations (applied with the transform restatement of find_translation.cpp:66-125) and are
  NOT invariant under a generic change of content; histograms sum to the popcount on every axis.
* The float64 analytic gradient of the InfoNCE loss through the MLP equals central finite
  differences (a dropped term, a wrong sign or a transposed operand in the backward fails this).
* The bf16 emulation stays close to float64 and its rounding is round-to-nearest-even.
"""
import numpy as np
import pytest

import mlp_oracle as mo
import sets


def test_histograms_closed_form():
    rng = np.random.default_rng(0)
    for m in sets.random_masks(rng, 2, 50, 0.4).reshape(-1):
        h = mo.histograms(int(m))
        pc = bin(int(m)).count("1")
        assert (h.sum(axis=1) == pc).all()
    assert (mo.histograms(0xFFFFFFFFFFFFFFFF) == 16).all()
    # a single voxel at (1, 2, 3)
    h = mo.histograms(1 << (1 + 4 * 2 + 16 * 3))
    assert h[0].tolist() == [0, 1, 0, 0] and h[1].tolist() == [0, 0, 1, 0] and h[2].tolist() == [0, 0, 0, 1]


def test_descriptor_invariant_under_48_transforms(oracle):
    rng = np.random.default_rng(1)
    masks = sets.random_masks(rng, 2, 40, 0.3).reshape(-1)
    for m in masks:
        d0 = mo.descriptor(int(m))
        for perm in range(6):
            for sym in range(8):
                t = oracle.transform(2, np.array([m], dtype=np.uint64), perm, sym, (0, 0, 0))
                assert np.array_equal(mo.descriptor(int(t[0])), d0), (hex(int(m)), perm, sym)


def test_descriptor_invariant_under_unclipped_translation(oracle):
    # a 2x2x1 blob with a tail, moved around inside the grid without clipping
    m = (1 << 0) | (1 << 1) | (1 << 4) | (1 << 5) | (1 << 16)
    d0 = mo.descriptor(m)
    for t in [(1, 0, 0), (2, 2, 0), (0, 1, 2), (2, 0, 1)]:
        mt = int(oracle.transform(2, np.array([m], dtype=np.uint64), 0, 0, t)[0])
        assert bin(mt).count("1") == bin(m).count("1")
        assert np.array_equal(mo.descriptor(mt), d0)


def test_descriptor_separates_contents():
    # a 1x1x4 rod and a 2x2x1 square have the same popcount but different descriptors
    rod = sum(1 << (16 * z) for z in range(4))
    square = sum(1 << (x + 4 * y) for x in range(2) for y in range(2))
    assert not np.array_equal(mo.descriptor(rod), mo.descriptor(square))


def test_features_layout():
    X = mo.features(np.array([0b101, 0], dtype=np.uint64))
    assert X.shape == (2, 80)
    assert X[0, 0] == 1 and X[0, 1] == 0 and X[0, 2] == 1 and X[0, 3:64].sum() == 0
    assert np.all(X[:, 76:] == 0) and np.all(X[1] == 0)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(2)
    B = 4
    masks = sets.random_masks(rng, 2, 2 * B, 0.35).reshape(-1)
    X = mo.features(masks)
    params = mo.init_params(3) * 0.5
    params[rng.integers(0, params.size, 200)] += rng.normal(0, 0.05, 200)  # non-zero biases too
    loss, g = mo.loss_and_grad(params, X, B, tau=0.5)
    assert np.isfinite(loss)
    idx = rng.choice(params.size, 300, replace=False)
    # include bias entries of every layer explicitly
    off = 0
    for i, o in mo.DIMS:
        idx = np.concatenate([idx, [off + i * o, off + i * o + o - 1]])
        off += i * o + o
    eps = 1e-5
    num = np.empty(idx.size)
    for k, j in enumerate(idx):
        p1 = params.copy()
        p1[j] += eps
        p2 = params.copy()
        p2[j] -= eps
        num[k] = (mo.loss_and_grad(p1, X, B, 0.5)[0] - mo.loss_and_grad(p2, X, B, 0.5)[0]) / (2 * eps)
    err = np.abs(num - g[idx])
    assert np.max(err) < 1e-6 + 1e-4 * np.max(np.abs(num)), (np.max(err), np.max(np.abs(num)))


def test_bf16_rounding():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5, 0.0])
    r = mo.bf16(v)
    assert r[0] == 1.0
    assert r[1] == 1.0  # tie -> even mantissa
    assert r[2] == 1.0 + 4 * 2 ** -8  # tie -> even (rounds up)
    assert r[3] == 1.0
    assert r[4] == -2.5 and r[5] == 0.0


def test_bf16_emulation_close_to_float64():
    rng = np.random.default_rng(4)
    B = 8
    X = mo.features(sets.random_masks(rng, 2, 2 * B, 0.3).reshape(-1))
    params = mo.init_params(5)
    l64, g64 = mo.loss_and_grad(params, X, B, 0.2)
    l16, g16 = mo.loss_and_grad(params, X, B, 0.2, emulate_bf16=True)
    assert abs(l64 - l16) < 0.05 * max(1.0, abs(l64))
    cos = g64 @ g16 / np.linalg.norm(g64) / np.linalg.norm(g16)
    assert cos > 0.99
