"""Pipeline rows on the CPU (no GPU): SVDAG producer, flattening, hierarchical dedup, .tdag.

* The library's bottom-up SVDAG builder equals the oracle's restatement of the reference's
  recursive constructor + toDAG, node for node (same indices, hence the same tie-breaks).
* Flattened LargeSubGrid masks equal the oracle's structure.get-based createLargeSubGrid.
* With translation off, constructTransformDAG is the hierarchical pass alone: its .tdag bytes
  equal the oracle's, expand back to the original voxels, and the leaf table has the 22
  canonical 2x2x2 classes under the 48 transforms.
"""
import ctypes
import os
import tempfile

import numpy as np
import pytest

import tdag_decode

SCENES = [(0, 64, 1), (0, 128, 2), (1, 128, 3), (0, 256, 0), (2, 256, 4)]


def _oracle_svdag(oracle, bits, R):
    L = oracle.lib()
    h = ctypes.c_void_p()
    assert L.orc_svdag_build(bits.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), R, ctypes.byref(h)) == 0
    return h


@pytest.fixture(scope="module")
def pl():
    from tdag_amd import pipeline

    import tdag_amd

    tdag_amd.lib()
    return pipeline


@pytest.mark.parametrize("kind,R,seed", SCENES)
def test_svdag_matches_oracle(oracle, pl, kind, R, seed):
    bits = pl.scene_dense(kind, R, seed)
    s = pl.Svdag(bits, R)
    h = _oracle_svdag(oracle, bits, R)
    L = oracle.lib()
    assert L.orc_svdag_root_level(h) == s.root_level
    for level in range(2, s.root_level + 1):
        n = L.orc_svdag_level_size(h, level)
        assert n == s.level_sizes[level]
        if level == 2:
            ref = np.zeros(max(1, n), dtype=np.uint64)
            L.orc_svdag_copy_leaves(h, ref.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)))
            assert np.array_equal(ref[:n], s.leaves())
        else:
            ref = np.zeros((max(1, n), 8), dtype=np.uint32)
            L.orc_svdag_copy_level(h, level, ref.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)))
            assert np.array_equal(ref[:n], s.nodes(level))
    L.orc_svdag_free(h)


@pytest.mark.parametrize("level", [2, 3, 4, 5])
def test_flatten_matches_oracle(oracle, pl, level):
    bits = pl.scene_dense(1, 128, 5)
    s = pl.Svdag(bits, 128)
    h = _oracle_svdag(oracle, bits, 128)
    rng = np.random.default_rng(level)
    idx = rng.choice(s.level_sizes[level], size=min(200, s.level_sizes[level]), replace=False).astype(np.uint32)
    mine = s.flatten(level, idx)
    W = mine.shape[1]
    ref = np.zeros((len(idx), W), dtype=np.uint64)
    L = oracle.lib()
    assert L.orc_flatten(h, level, idx.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), len(idx), ref.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))) == 0
    assert np.array_equal(mine, ref)
    L.orc_svdag_free(h)


def _oracle_tdag_bytes(oracle, bits, R, cfg):
    h = _oracle_svdag(oracle, bits, R)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "o.tdag")
        n = oracle.lib().orc_construct_tdag(h, ctypes.byref(cfg), os.cpu_count() or 1, path.encode())
        assert n > 0
        data = open(path, "rb").read()
    oracle.lib().orc_svdag_free(h)
    return data


@pytest.mark.parametrize("kind,R,seed", SCENES[:3])
@pytest.mark.parametrize("sym,axis", [(True, True), (True, False), (False, True), (False, False)])
def test_hierarchical_tdag_bytes(oracle, pl, kind, R, seed, sym, axis):
    import tdag_amd

    bits = pl.scene_dense(kind, R, seed)
    cfg = tdag_amd.TransformDAGConfig(sym, axis, False, 3)
    t = pl.construct_transform_dag(None, pl.Svdag(bits, R), cfg)
    mine = t.tdag_bytes()
    ref = _oracle_tdag_bytes(oracle, bits, R, oracle.config(sym, axis, False, 3))
    assert mine == ref
    d = tdag_decode.parse(mine)
    # The reference's output EditStructure is default-constructed (transform_dag_hierarchical.cpp:307)
    # and never marked a DAG, so the header holds StructureType::Tree = 0 (structure.h:127-131, 171).
    assert d["structure_type"] == 0
    assert np.array_equal(tdag_decode.expand(d), tdag_decode.dense_to_cube(bits, R))
    # Every level comes out in the order of its group key: hash_combine over the node's 8 child
    # pointers sorted by ptr, an empty child counting as the sentinel's ptr 0xFFFFFFFF
    # (transform_dag_hierarchical.cpp:365-378 groups, :417-430 emits in sorted key order).
    for level in range(3, d["root_level"] + 1):
        keys = [_edit_group_key(node) for node in d["levels"][level]]
        assert all(a <= b for a, b in zip(keys, keys[1:])), f"level {level} not in group-key order"


def _edit_group_key(node):
    m64 = (1 << 64) - 1
    seed = 0
    for p in sorted(int(c) & ((1 << 37) - 1) for c in node):
        seed = (seed ^ ((p + 0x9E3779B9 + (seed << 6) + (seed >> 2)) & m64)) & m64
    return seed


def test_canonical_leaf_classes():
    """22 classes of 2x2x2 masks under the 48 mirror/permutation transforms
    (transform_dag_hierarchical.cpp:200 log line, sym+perm on)."""
    import itertools

    def transform(m, perm, sym):
        out = 0
        for i in range(8):
            if (m >> i) & 1:
                v = (i & 1, (i >> 1) & 1, (i >> 2) & 1)
                o = [v[perm[a]] ^ ((sym >> a) & 1) for a in range(3)]
                out |= 1 << (o[0] | o[1] << 1 | o[2] << 2)
        return out

    orbits = set()
    for m in range(256):
        orbits.add(min(transform(m, p, s) for p in itertools.permutations(range(3)) for s in range(8)))
    assert len(orbits) == 22


def test_empty_scene(oracle, pl):
    import tdag_amd

    bits = np.zeros(64 ** 3 // 8, dtype=np.uint8)
    s = pl.Svdag(bits, 64)
    assert all(v == 0 for v in s.level_sizes)
    cfg = tdag_amd.TransformDAGConfig(True, True, False, 3)
    mine = pl.construct_transform_dag(None, s, cfg).tdag_bytes()
    assert mine == _oracle_tdag_bytes(oracle, bits, 64, oracle.config(True, True, False, 3))


def test_city_scene_deterministic_with_transformed_copies(oracle, pl):
    """The city (scene kind 2) is reproducible from its seed, and its mirrored / axis-swapped
    block copies produce matches that need a mirror or a permutation (oracle matcher, level 4)."""
    a = pl.scene_dense(2, 512, 9)
    assert np.array_equal(a, pl.scene_dense(2, 512, 9))
    assert not np.array_equal(a, pl.scene_dense(2, 512, 10))
    s = pl.Svdag(a, 512)
    nodes = np.arange(s.level_sizes[4], dtype=np.uint32)
    masks = s.flatten(4, nodes)
    enc = oracle.match_encoded(4, masks, oracle.config(True, True, True))
    b = 4 + 1
    low = 6 + 3 * b
    order = np.argsort(-np.array([sum(bin(int(w)).count("1") for w in m) for m in masks]), kind="stable")
    rank_of = np.empty_like(order)
    rank_of[order] = np.arange(len(order))
    matched = (enc >> np.uint64(low)).astype(np.int64) != rank_of
    sym = (enc >> np.uint64(3 * b)) & np.uint64(7)
    perm = (enc >> np.uint64(3 * b + 3)) & np.uint64(7)
    assert matched.sum() > 0
    assert np.sum(matched & ((sym != 0) | (perm != 0))) > 0


@pytest.mark.parametrize("kind,R,seed", [(0, 64, 1), (1, 128, 3), (2, 256, 4)])
def test_edit_structure_view_matches_oracle(oracle, pl, kind, R, seed):
    """constructTransformDAG on an EditStructure<void, uint32_t> handed over as arrays (the
    reference's input type, transform_dag.h:161-162), built by the oracle rather than from a
    dense grid: with translation off the .tdag bytes equal the oracle's driver, and the result's
    levels read back through tdag_tdag_copy_level are the nodes those bytes hold."""
    import svdag_arrays
    import tdag_amd

    bits = pl.scene_dense(kind, R, seed)
    root, nodes, leaves = svdag_arrays.oracle_svdag_arrays(oracle, bits, R)
    s = pl.Svdag.from_view(R, root, nodes, leaves)
    assert s.level_sizes[2] == len(leaves) and all(s.level_sizes[l] == len(nodes[l]) for l in nodes)
    for maxT in (2, 4):
        cfg = tdag_amd.TransformDAGConfig(True, True, False, maxT)
        t = pl.construct_transform_dag(None, s, cfg)
        mine = t.tdag_bytes()
        assert mine == _oracle_tdag_bytes(oracle, bits, R, oracle.config(True, True, False, maxT))
        d = tdag_decode.parse(mine)
        assert np.array_equal(t.nodes(2), d["leaves"])
        for level in range(3, root + 1):
            assert np.array_equal(t.nodes(level), d["levels"][level])


def test_edit_structure_view_is_validated(oracle, pl):
    import svdag_arrays
    import tdag_amd

    bits = pl.scene_dense(0, 64, 1)
    root, nodes, leaves = svdag_arrays.oracle_svdag_arrays(oracle, bits, 64)
    bad = {l: a.copy() for l, a in nodes.items()}
    bad[4][0, int(np.argmax(bad[4][0] != 0xFFFFFFFF))] = len(nodes[3])  # one past the end of level 3
    with pytest.raises(tdag_amd.TdagError, match="EINVAL.*outside level 3"):
        pl.Svdag.from_view(64, root, bad, leaves)
    bad = {l: a.copy() for l, a in nodes.items()}
    bad[3][0, int(np.argmax(bad[3][0] != 0xFFFFFFFF))] = len(leaves)
    with pytest.raises(tdag_amd.TdagError, match="EINVAL"):
        pl.Svdag.from_view(64, root, bad, leaves)
    with pytest.raises(tdag_amd.TdagError, match="EINVAL"):
        pl.Svdag.from_view(128, root, nodes, leaves)  # resolution != 2^root_level
    with pytest.raises(tdag_amd.TdagError, match="EINVAL"):
        pl.Svdag.from_view(64, root, nodes, leaves, structure_type=7)
    # A Tree is accepted as a structure but refused by the driver (transform_dag.cpp:141).
    tree = pl.Svdag.from_view(64, root, nodes, leaves, structure_type=0)
    with pytest.raises(tdag_amd.TdagError, match="EINVAL.*DAG"):
        pl.construct_transform_dag(None, tree, tdag_amd.TransformDAGConfig(True, True, False, 3))


def test_device_translation_flag_validation(pl):
    """TDAG_DEVICE_TRANSLATION needs a context (checked before any device work, so on the CPU)."""
    import tdag_amd

    bits = pl.scene_dense(0, 64, 1)
    svdag = pl.Svdag(bits, 64)
    with pytest.raises(tdag_amd.TdagError, match="EINVAL.*context"):
        pl.construct_transform_dag(None, svdag, tdag_amd.TransformDAGConfig(True, True, True, 3), device_translation=True)
    with pytest.raises(tdag_amd.TdagError, match="EINVAL.*context"):
        pl.construct_transform_dag(None, svdag, tdag_amd.TransformDAGConfig(True, True, False, 3), device_translation=True)
    cfg = tdag_amd.TransformDAGConfig(True, True, False, 3).to_c()
    h = ctypes.c_void_p()
    assert tdag_amd.lib().tdag_construct_transform_dag(None, svdag.handle, ctypes.byref(cfg), 4, ctypes.byref(h)) != 0
    assert "unknown flag" in tdag_amd.last_error()
