"""Pins the oracle's matcher (restatement of findDuplicatesAmongFlattenedItems_cpp2) (no GPU).

* The hash-search restatement equals an independent brute force over the full candidate set
  (all refs x 48 transforms x t in [-(R-1), R-1]^3, first hit in key order) on seeded sets with
  planted transformed copies and edge cases, at every level and for every config switch.
* Every chosen parent satisfies the exact equation child == T(parent) under an independent
  numpy statement of the transforms.
* Post-processing: the unique set is exactly the set of parents, uniques map to themselves,
  and expanding every mapping from the unique list reproduces every item (decode invariant).
"""
import numpy as np
import pytest

import np_ref
import sets


def _parents(oracle, level, masks, enc):
    """(parent original index, perm, sym, t) per item from encodedParents."""
    n = len(masks)
    pops = [int(sum(bin(int(w)).count("1") for w in m)) for m in masks]
    order = sorted(range(n), key=lambda i: -pops[i])  # Python's sort is stable
    sentinel_key = oracle.encode_fixed64(oracle.tp_sentinel(), level)
    out = []
    for i in range(n):
        if int(enc[i]) == sentinel_key:
            out.append(None)
            continue
        tp = oracle.decode_fixed64(int(enc[i]), level)
        ptr = tp & ((1 << 37) - 1)
        sym = (tp >> 37) & 7
        perm = ((tp >> 40) & 3, (tp >> 42) & 3, (tp >> 44) & 3)
        t = tuple(((tp >> s) & 63) - 32 for s in (46, 52, 58))
        out.append((order[ptr], np_ref.PERMS.index(perm), sym, t))
    return out


CASES = [(2, 90, 90, 0), (2, 60, 60, 1), (3, 24, 24, 2), (4, 6, 6, 3), (5, 3, 3, 4)]


@pytest.mark.parametrize("level,n_base,n_planted,seed", CASES)
def test_hash_search_equals_brute_force(oracle, level, n_base, n_planted, seed):
    rng = np.random.default_rng(seed)
    masks = sets.planted_set(rng, level, n_base, n_planted, oracle.transform, edge_cases=(level <= 3))
    enc_hash = oracle.match_encoded(level, masks)
    enc_brute = oracle.match_encoded(level, masks, brute=True)
    assert np.array_equal(enc_hash, enc_brute)
    parents = _parents(oracle, level, masks, enc_hash)
    foreign = 0
    for i, p in enumerate(parents):
        if p is None:
            assert not masks[i].any()  # only empty items keep the sentinel
            continue
        src, perm, sym, t = p
        assert np.array_equal(np_ref.apply(level, masks[src], perm, sym, t), masks[i])
        foreign += src != i
    assert foreign >= n_planted // 2  # planted copies really are found


@pytest.mark.parametrize("sym,axis,trans", [(False, True, True), (True, False, True), (True, True, False), (False, False, True), (False, False, False)])
def test_config_switches(oracle, sym, axis, trans):
    level = 2
    rng = np.random.default_rng(21)
    masks = sets.planted_set(rng, level, 70, 70, oracle.transform)
    cfg = oracle.config(sym, axis, trans)
    enc_hash = oracle.match_encoded(level, masks, cfg)
    enc_brute = oracle.match_encoded(level, masks, cfg, brute=True)
    assert np.array_equal(enc_hash, enc_brute)
    for p in _parents(oracle, level, masks, enc_hash):
        if p is None:
            continue
        _, perm, s, t = p
        assert axis or perm == 0
        assert sym or s == 0
        assert trans or t == (0, 0, 0)


@pytest.mark.parametrize("level", [2, 3])
def test_postprocess_decode_invariant(oracle, level):
    rng = np.random.default_rng(31 + level)
    masks = sets.planted_set(rng, level, 40, 40, oracle.transform)
    enc = oracle.match_encoded(level, masks)
    mapping, uniq = oracle.postprocess(level, masks, enc)
    parents = _parents(oracle, level, masks, enc)
    parent_set = {i if p is None else p[0] for i, p in enumerate(parents)}
    assert set(int(u) for u in uniq) == parent_set
    assert list(uniq) == sorted(uniq)
    for new_idx, u in enumerate(uniq):
        assert int(mapping[u]) == oracle.tp_create(new_idx)
    for i in range(len(masks)):
        tp = int(mapping[i])
        ptr = tp & ((1 << 37) - 1)
        sym = (tp >> 37) & 7
        perm = ((tp >> 40) & 3, (tp >> 42) & 3, (tp >> 44) & 3)
        t = tuple(((tp >> s) & 63) - 32 for s in (46, 52, 58))
        rebuilt = np_ref.apply(level, masks[uniq[ptr]], np_ref.PERMS.index(perm), sym, t)
        assert np.array_equal(rebuilt, masks[i])


def test_empty_and_single(oracle):
    for level in (2, 3):
        W = sets.words(level)
        assert oracle.match_encoded(level, np.zeros((0, W), dtype=np.uint64)).size == 0
        one = np.zeros((1, W), dtype=np.uint64)
        one[0, 0] = 1
        assert int(oracle.match_encoded(level, one)[0]) == oracle.encode_fixed64(oracle.tp_create(0), level)


def test_timed_sampling_counts_queries(oracle):
    """The cpu_baseline timing entry searches exactly the ranks 0, s, 2s, ... and leaves the
    keys of a full (stride 1) run unchanged."""
    rng = np.random.default_rng(5)
    masks = sets.planted_set(rng, 3, 150, 150, oracle.transform)
    n = len(masks)
    for stride in (1, 3, 7, n + 5):
        setup, search, q = oracle.match_timed(3, masks, oracle.config(), nthreads=2, query_stride=stride)
        assert q == (n + stride - 1) // stride
        assert setup >= 0.0 and search >= 0.0
    assert np.array_equal(oracle.match_encoded(3, masks), oracle.match_encoded(3, masks, brute=True))
