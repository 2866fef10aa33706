"""Pins the oracle's TransformPointer codec and grid transforms (no GPU).

Closed forms: bit patterns of the sentinel and identity pointers (GCC x86-64 bit-field
layout of transform_dag.h:24-33), sentinel keys per level, key order = lexicographic order of
(ptr, permCode, mirror, tz, ty, tx) (transform_dag.h:375-397), round trips, and the 48-element
transform group checked against independent numpy array operations.
"""
import itertools

import numpy as np
import pytest

import np_ref
import sets


def test_pointer_bit_layout(oracle):
    assert oracle.tp_sentinel() == 0x82082400FFFFFFFF
    for k in (0, 1, 12345, (1 << 32) - 2):
        assert oracle.tp_create(k) == 0x8208240000000000 | k
    # mirror bits 37..39, permutation axes 40..45, translations (biased 32) 46..63
    assert oracle.tp_create(0, 0b101, 0, 0, 0, 0) == 0x8208240000000000 | (0b101 << 37)
    assert oracle.tp_create(7, 0, 5, -1, 2, 31) == 7 | (2 << 40) | (1 << 42) | (0 << 44) | (31 << 46) | (34 << 52) | (63 << 58)


@pytest.mark.parametrize("level,expected", [(2, 0x7FFFFFFF8124), (3, 0x3FFFFFFFC0888), (4, 0x1FFFFFFFE04210), (5, 0xFFFFFFFF020820)])
def test_sentinel_keys(oracle, level, expected):
    assert oracle.encode_fixed64(oracle.tp_sentinel(), level) == expected


@pytest.mark.parametrize("level", [2, 3, 4, 5])
def test_key_order_and_roundtrip(oracle, level):
    R = 1 << level
    rng = np.random.default_rng(level)
    tuples, keys = [], []
    for _ in range(3000):
        ptr = int(rng.integers(0, 1 << 20))
        perm = int(rng.integers(6))
        sym = int(rng.integers(8))
        t = [int(v) for v in rng.integers(-(R - 1), R, 3)]
        tp = oracle.tp_create(ptr, sym, perm, *t)
        key = oracle.encode_fixed64(tp, level)
        assert oracle.decode_fixed64(key, level) == tp
        tuples.append((ptr, perm, sym, t[2], t[1], t[0]))
        keys.append(key)
    order_keys = sorted(range(len(keys)), key=lambda i: keys[i])
    order_tuples = sorted(range(len(keys)), key=lambda i: tuples[i])
    assert [tuples[i] for i in order_keys] == [tuples[i] for i in order_tuples]


@pytest.mark.parametrize("level", [2, 3])
def test_transforms_match_numpy(oracle, level):
    R = 1 << level
    rng = np.random.default_rng(100 + level)
    for _ in range(60):
        m = sets.random_masks(rng, level, 1, 0.3)[0]
        perm, sym = int(rng.integers(6)), int(rng.integers(8))
        t = tuple(int(v) for v in rng.integers(-(R - 1), R, 3))
        assert np.array_equal(oracle.transform(level, m, perm, sym, t), np_ref.apply(level, m, perm, sym, t))


def test_group_of_48(oracle):
    level = 2
    rng = np.random.default_rng(7)
    m = sets.random_masks(rng, level, 1, 0.4)[0]
    images = {tuple(oracle.transform(level, m, p, s, (0, 0, 0))) for p in range(6) for s in range(8)}
    assert len(images) == 48  # a generic grid has 48 distinct images
    for p, s in itertools.product(range(6), range(8)):
        img = oracle.transform(level, m, p, s, (0, 0, 0))
        # some (p', s') maps the image back to the original (closure + inverse)
        back = [(q, u) for q in range(6) for u in range(8) if np.array_equal(oracle.transform(level, img, q, u, (0, 0, 0)), m)]
        assert len(back) == 1


@pytest.mark.parametrize("level", [2, 3, 4])
def test_translation_clips(oracle, level):
    R = 1 << level
    rng = np.random.default_rng(11 + level)
    m = sets.random_masks(rng, level, 1, 0.5)[0]
    full_pop = int(sum(bin(int(w)).count("1") for w in m))
    for t in [(R - 1, 0, 0), (0, -(R - 1), 0), (1, 1, -1)]:
        out = oracle.transform(level, m, 0, 0, t)
        pop = int(sum(bin(int(w)).count("1") for w in out))
        assert pop <= full_pop
        back = oracle.transform(level, out, 0, 0, tuple(-v for v in t))
        assert np.array_equal(back & m, back)  # translating back lands inside the original
