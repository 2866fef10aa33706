"""constructStaticTransformDAG on the CPU (no GPU): the StaticStructure row.

* StaticStructure(EditStructure) packing (structure.h:766-798) of the library equals the oracle's.
* With translation off the static driver is the static hierarchical pass alone
  (transform_dag_hierarchical.cpp:531-743): the library's writeTo bytes equal the oracle's, expand
  back to the scene's voxels, and each level comes out in the order of its group key, a hash of
  the children's sorted output word offsets (the key of :600-612); the EditStructure pass, keyed by
  node indices, orders levels >= 4 differently while keeping the same nodes.
* The oracle's static driver with translation (oracle matcher) agrees with its separately
  restated EditStructure driver: the same level sizes everywhere and the same nodes, pointer for
  pointer, at levels <= maxTranslationLevel and at the leaves.
* Malformed static inputs are rejected with TDAG_EINVAL; each result type is read only by its own
  serializer.
"""
import ctypes
import os
import tempfile

import numpy as np
import pytest

import tdag_decode as td

SCENES = [(0, 64, 1), (1, 128, 3), (2, 256, 4)]
M64 = (1 << 64) - 1
SENTINEL_PTR = 0xFFFFFFFF


@pytest.fixture(scope="module")
def pl():
    import tdag_amd
    from tdag_amd import pipeline

    tdag_amd.lib()
    return pipeline


def _oracle_handle(oracle, bits, R):
    h = ctypes.c_void_p()
    assert oracle.lib().orc_svdag_build(bits.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), R, ctypes.byref(h)) == 0
    return h


def _oracle_run(oracle, fn, h, cfg):
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "o.bin")
        assert fn(h, ctypes.byref(cfg), os.cpu_count() or 1, path.encode()) > 0
        return open(path, "rb").read()


def _hash_combine(seed, v):
    return (seed ^ ((v + 0x9E3779B9 + (seed << 6) + (seed >> 2)) & M64)) & M64


def _group_keys(words):
    """Group key of every packed node: hash_combine over its 8 child pointers sorted ascending,
    an absent child counting as the sentinel's ptr 0xFFFFFFFF (transform_dag_hierarchical.cpp:600-612)."""
    keys = []
    for st in td.node_starts(words):
        mask = int(words[st]) & 0xFFFFFFFF
        ptrs, o = [], 1
        for c in range(8):
            if (mask >> c) & 1:
                ptrs.append(int(words[st + o]) & ((1 << 37) - 1))
                o += 1
            else:
                ptrs.append(SENTINEL_PTR)
        seed = 0
        for p in sorted(ptrs):
            seed = _hash_combine(seed, p)
        keys.append(seed)
    return keys


@pytest.mark.parametrize("kind,R,seed", SCENES)
def test_static_words_match_oracle(oracle, pl, kind, R, seed):
    bits = pl.scene_dense(kind, R, seed)
    s = pl.Svdag(bits, R)
    h = _oracle_handle(oracle, bits, R)
    L = oracle.lib()
    for level in range(2, s.root_level + 1):
        n = L.orc_svdag_static_level(h, level, None, 0)
        ref = np.zeros(max(1, n), dtype=np.uint64 if level == 2 else np.uint32)
        L.orc_svdag_static_level(h, level, ref.ctypes.data_as(ctypes.c_void_p), ref.size)
        assert np.array_equal(s.static_words(level), ref[:n]), level
        if level >= 3:  # one node per EditStructure node, in the same order
            assert len(td.node_starts(ref[:n])) == s.level_sizes[level]
    L.orc_svdag_free(h)


@pytest.mark.parametrize("kind,R,seed", SCENES)
@pytest.mark.parametrize("sym,axis", [(True, True), (True, False), (False, True), (False, False)])
def test_static_hierarchical_bytes(oracle, pl, kind, R, seed, sym, axis):
    import tdag_amd

    bits = pl.scene_dense(kind, R, seed)
    svdag = pl.Svdag(bits, R)
    t = pl.construct_static_transform_dag(None, pl.StaticSvdag.from_svdag(svdag), tdag_amd.TransformDAGConfig(sym, axis, False, 3))
    mine = t.static_bytes()
    h = _oracle_handle(oracle, bits, R)
    assert mine == _oracle_run(oracle, oracle.lib().orc_construct_static_tdag, h, oracle.config(sym, axis, False, 3))
    edit = td.parse(_oracle_run(oracle, oracle.lib().orc_construct_tdag, h, oracle.config(sym, axis, False, 3)))
    oracle.lib().orc_svdag_free(h)
    d = td.parse_static(mine)
    assert d["structure_type"] == 0  # default-constructed output (transform_dag_hierarchical.cpp:533)
    as_edit = td.static_to_edit(d)
    assert np.array_equal(td.expand(as_edit), td.dense_to_cube(bits, R))
    for level in range(3, d["root_level"] + 1):
        keys = _group_keys(d["words"][level])
        assert all(a <= b for a, b in zip(keys, keys[1:])), f"level {level} not in group-key order"
        assert len(as_edit["levels"][level]) == len(edit["levels"][level])
        if level == 3:  # leaf indices: both passes key level 3 alike
            assert np.array_equal(as_edit["levels"][3], edit["levels"][3])
    reordered = [l for l in range(4, d["root_level"] + 1) if not np.array_equal(as_edit["levels"][l], edit["levels"][l])]
    assert reordered, "offset-keyed groups should reorder some level >= 4"


@pytest.mark.parametrize("kind,R,seed,maxT", [(0, 64, 1, 2), (0, 64, 2, 3), (1, 128, 3, 4), (2, 256, 4, 3)])
def test_oracle_static_driver_agrees_with_edit_driver(oracle, pl, kind, R, seed, maxT):
    bits = pl.scene_dense(kind, R, seed)
    h = _oracle_handle(oracle, bits, R)
    cfg = oracle.config(True, True, True, maxT)
    stat = td.static_to_edit(td.parse_static(_oracle_run(oracle, oracle.lib().orc_construct_static_tdag, h, cfg)))
    edit = td.parse(_oracle_run(oracle, oracle.lib().orc_construct_tdag, h, cfg))
    oracle.lib().orc_svdag_free(h)
    assert np.array_equal(stat["leaves"], edit["leaves"])
    for level in range(3, edit["root_level"] + 1):
        assert len(stat["levels"][level]) == len(edit["levels"][level]), level
        if level <= maxT:
            assert np.array_equal(stat["levels"][level], edit["levels"][level]), level
    assert np.array_equal(td.expand(stat), td.dense_to_cube(bits, R))


def _bad_view(pl, root, words, leaves):
    return pl.StaticSvdag(1 << root, root, {l: np.array(w, dtype=np.uint32) for l, w in words.items()}, np.array(leaves, dtype=np.uint64))


@pytest.mark.parametrize("case,words,leaves,msg", [
    ("mask above 8 bits", {3: [0x100]}, [1], "bits above 8"),
    ("node past the end", {3: [0x3, 0]}, [1], "runs past"),
    ("leaf index out of range", {3: [0x1, 5]}, [1], "leaf index"),
    ("child not a node offset", {3: [0x3, 0, 0, 0x1, 0], 4: [0x1, 1]}, [1], "not a node offset"),
])
def test_static_view_rejects_malformed(pl, case, words, leaves, msg):
    import tdag_amd

    st = _bad_view(pl, 4, words, leaves)
    with pytest.raises(tdag_amd.TdagError) as e:
        pl.construct_static_transform_dag(None, st, tdag_amd.TransformDAGConfig(True, True, False, 2))
    assert "EINVAL" in str(e.value) and msg in str(e.value), case


def test_static_view_rejects_bad_resolution(pl):
    import tdag_amd

    st = pl.StaticSvdag(48, 4, {3: np.array([0x1, 0], np.uint32), 4: np.array([0x1, 0], np.uint32)}, np.array([1], np.uint64))
    with pytest.raises(tdag_amd.TdagError, match="EINVAL"):
        pl.construct_static_transform_dag(None, st, tdag_amd.TransformDAGConfig(True, True, False, 2))


def test_result_types_keep_their_serializers(pl):
    import tdag_amd

    bits = pl.scene_dense(0, 64, 1)
    svdag = pl.Svdag(bits, 64)
    cfg = tdag_amd.TransformDAGConfig(True, True, False, 3)
    stat = pl.construct_static_transform_dag(None, pl.StaticSvdag.from_svdag(svdag), cfg)
    edit = pl.construct_transform_dag(None, svdag, cfg)
    with pytest.raises(tdag_amd.TdagError, match="EINVAL"):
        stat.tdag_bytes()
    with pytest.raises(tdag_amd.TdagError, match="EINVAL"):
        edit.static_bytes()
    assert stat.level_sizes == edit.level_sizes
    # words of one level read back as the bytes hold them
    d = td.parse_static(stat.static_bytes())
    for level in range(3, stat.root_level + 1):
        assert np.array_equal(stat.static_words(level), d["words"][level])
    assert np.array_equal(stat.static_words(2), d["leaves"])
