"""GPU parity of the whole constructTransformDAG pipeline (driver + GPU matcher).

The library's .tdag bytes must equal the oracle's (restated driver + restated CPU matcher)
byte for byte, and expand back to the scene's voxels (decode invariant). The per-level matcher
inputs recorded by the driver are also matched directly against the oracle matcher.
"""
import ctypes
import os
import tempfile

import numpy as np
import pytest

import tdag_decode

pytestmark = pytest.mark.gpu


def _oracle_bytes(oracle, bits, R, cfg):
    L = oracle.lib()
    h = ctypes.c_void_p()
    assert L.orc_svdag_build(bits.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), R, ctypes.byref(h)) == 0
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "o.tdag")
        assert L.orc_construct_tdag(h, ctypes.byref(cfg), os.cpu_count() or 1, path.encode()) > 0
        data = open(path, "rb").read()
    L.orc_svdag_free(h)
    return data


@pytest.mark.parametrize("kind,R,seed,maxT", [(0, 64, 1, 2), (0, 64, 2, 5), (1, 128, 3, 3), (1, 128, 4, 4), (0, 256, 0, 2), (2, 256, 5, 3)])
def test_tdag_bytes_match_oracle(oracle, gpu_ctx, kind, R, seed, maxT):
    import tdag_amd
    from tdag_amd import pipeline as pl

    print(f"scene kind={kind} R={R} seed={seed} maxT={maxT}", flush=True)
    bits = pl.scene_dense(kind, R, seed)
    cfg = tdag_amd.TransformDAGConfig(True, True, True, maxT)
    t = pl.construct_transform_dag(gpu_ctx, pl.Svdag(bits, R), cfg)
    mine = t.tdag_bytes()
    ref = _oracle_bytes(oracle, bits, R, oracle.config(True, True, True, maxT))
    assert len(mine) == len(ref)
    assert mine == ref
    d = tdag_decode.parse(mine)
    assert np.array_equal(tdag_decode.expand(d), tdag_decode.dense_to_cube(bits, R))


@pytest.mark.parametrize("sym,axis", [(False, True), (True, False), (False, False)])
def test_config_switches_pipeline(oracle, gpu_ctx, sym, axis):
    import tdag_amd
    from tdag_amd import pipeline as pl

    bits = pl.scene_dense(1, 128, 9)
    cfg = tdag_amd.TransformDAGConfig(sym, axis, True, 3)
    print(f"pipeline cfg sym={sym} axis={axis}", flush=True)
    mine = pl.construct_transform_dag(gpu_ctx, pl.Svdag(bits, 128), cfg).tdag_bytes()
    assert mine == _oracle_bytes(oracle, bits, 128, oracle.config(sym, axis, True, 3))


def test_level_inputs_match_oracle_matcher(oracle, gpu_ctx):
    import tdag_amd
    from tdag_amd import pipeline as pl

    bits = pl.scene_dense(1, 256, 11)
    cfg = tdag_amd.TransformDAGConfig(True, True, True, 5)
    t = pl.construct_transform_dag(gpu_ctx, pl.Svdag(bits, 256), cfg, keep_level_inputs=True)
    for level in (5, 4, 3, 2):
        masks = t.level_masks(level)
        print(f"level {level}: {len(masks)} items", flush=True)
        enc, _, _, st = gpu_ctx.match_level(level, masks, cfg)
        ref = oracle.match_encoded(level, masks, oracle.config(True, True, True, 5))
        assert np.array_equal(enc, ref)
        rec = t.record(level)
        assert rec.items == len(masks) and rec.matched == st.num_matched


def _oracle_static_bytes(oracle, bits, R, cfg):
    L = oracle.lib()
    h = ctypes.c_void_p()
    assert L.orc_svdag_build(bits.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), R, ctypes.byref(h)) == 0
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "o.static")
        assert L.orc_construct_static_tdag(h, ctypes.byref(cfg), os.cpu_count() or 1, path.encode()) > 0
        data = open(path, "rb").read()
    L.orc_svdag_free(h)
    return data


@pytest.mark.parametrize("kind,R,seed,maxT", [(0, 64, 1, 2), (0, 64, 2, 5), (1, 128, 3, 3), (1, 128, 4, 4), (2, 256, 5, 3)])
def test_static_transform_dag_matches_oracle(oracle, gpu_ctx, kind, R, seed, maxT):
    """constructStaticTransformDAG through the GPU matcher: writeTo bytes equal the oracle's
    restated static driver, the result expands back to the voxels, and every level's matcher call
    equals the EditStructure driver's (same items, uniques and matches)."""
    import tdag_amd
    from tdag_amd import pipeline as pl

    print(f"static scene kind={kind} R={R} seed={seed} maxT={maxT}", flush=True)
    bits = pl.scene_dense(kind, R, seed)
    svdag = pl.Svdag(bits, R)
    cfg = tdag_amd.TransformDAGConfig(True, True, True, maxT)
    t = pl.construct_static_transform_dag(gpu_ctx, pl.StaticSvdag.from_svdag(svdag), cfg)
    mine = t.static_bytes()
    assert mine == _oracle_static_bytes(oracle, bits, R, oracle.config(True, True, True, maxT))
    d = tdag_decode.parse_static(mine)
    assert np.array_equal(tdag_decode.expand(tdag_decode.static_to_edit(d)), tdag_decode.dense_to_cube(bits, R))
    e = pl.construct_transform_dag(gpu_ctx, svdag, cfg)
    assert t.level_sizes == e.level_sizes
    for level in range(2, maxT + 1):
        a, b = t.record(level), e.record(level)
        assert (a.items, a.unique, a.matched) == (b.items, b.unique, b.matched), level


@pytest.mark.parametrize("sym,axis", [(False, True), (True, False)])
def test_static_config_switches(oracle, gpu_ctx, sym, axis):
    import tdag_amd
    from tdag_amd import pipeline as pl

    print(f"static cfg sym={sym} axis={axis}", flush=True)
    bits = pl.scene_dense(1, 128, 9)
    cfg = tdag_amd.TransformDAGConfig(sym, axis, True, 3)
    t = pl.construct_static_transform_dag(gpu_ctx, pl.StaticSvdag.from_svdag(pl.Svdag(bits, 128)), cfg)
    assert t.static_bytes() == _oracle_static_bytes(oracle, bits, 128, oracle.config(sym, axis, True, 3))


@pytest.mark.parametrize("kind,R,seed", [(0, 64, 1), (1, 128, 3), (2, 256, 4), (1, 512, 7)])
@pytest.mark.parametrize("sym,axis", [(True, True), (True, False), (False, True), (False, False)])
def test_device_hierarchical_equals_host(gpu_ctx, kind, R, seed, sym, axis):
    """The hierarchical pass on the device (canonical keys, min-index classes, sorted emission)
    gives the host pass's output byte for byte, in the EditStructure and the StaticStructure form
    (translation off: the driver is the hierarchical pass alone). The host pass is pinned to the
    oracle's literal restatement by the CPU tests."""
    import tdag_amd
    from tdag_amd import pipeline as pl

    print(f"device hierarchical kind={kind} R={R} seed={seed} sym={sym} axis={axis}", flush=True)
    bits = pl.scene_dense(kind, R, seed)
    svdag = pl.Svdag(bits, R)
    cfg = tdag_amd.TransformDAGConfig(sym, axis, False, 3)
    host = pl.construct_transform_dag(None, svdag, cfg).tdag_bytes()
    dev = pl.construct_transform_dag(gpu_ctx, svdag, cfg).tdag_bytes()
    assert dev == host
    st = pl.StaticSvdag.from_svdag(svdag)
    assert pl.construct_static_transform_dag(gpu_ctx, st, cfg).static_bytes() == pl.construct_static_transform_dag(None, st, cfg).static_bytes()


@pytest.mark.parametrize("kind,R,seed,maxT", [(0, 64, 2, 5), (2, 256, 5, 4)])
def test_edit_structure_view_bytes_match_oracle(oracle, gpu_ctx, kind, R, seed, maxT):
    """constructTransformDAG on the oracle's SVDAG handed over as an EditStructure view (the
    reference's input type), not built from a dense grid: the same .tdag bytes as the oracle."""
    import svdag_arrays
    import tdag_amd
    from tdag_amd import pipeline as pl

    print(f"view scene kind={kind} R={R} seed={seed} maxT={maxT}", flush=True)
    bits = pl.scene_dense(kind, R, seed)
    root, nodes, leaves = svdag_arrays.oracle_svdag_arrays(oracle, bits, R)
    s = pl.Svdag.from_view(R, root, nodes, leaves)
    mine = pl.construct_transform_dag(gpu_ctx, s, tdag_amd.TransformDAGConfig(True, True, True, maxT)).tdag_bytes()
    assert mine == _oracle_bytes(oracle, bits, R, oracle.config(True, True, True, maxT))


@pytest.mark.parametrize("kind,R,seed,maxT", [(0, 64, 1, 2), (0, 64, 2, 5), (1, 128, 3, 3), (2, 256, 5, 4), (1, 256, 7, 5)])
def test_device_translation_stage_bytes_match_oracle(oracle, gpu_ctx, kind, R, seed, maxT):
    """TDAG_DEVICE_TRANSLATION (marking, compaction, flattening, matching, emission and pointer
    rewrites all on the device) and the captured build graph, replayed three times with other GPU
    work in between and after the SVDAG handle was freed: the same .tdag bytes as the oracle's
    driver every time."""
    import torch

    import tdag_amd
    from tdag_amd import pipeline as pl

    print(f"device translation kind={kind} R={R} seed={seed} maxT={maxT}", flush=True)
    bits = pl.scene_dense(kind, R, seed)
    svdag = pl.Svdag(bits, R)
    cfg = tdag_amd.TransformDAGConfig(True, True, True, maxT)
    ref = _oracle_bytes(oracle, bits, R, oracle.config(True, True, True, maxT))
    t = pl.construct_transform_dag(gpu_ctx, svdag, cfg, device_translation=True)
    assert t.tdag_bytes() == ref
    host = pl.construct_transform_dag(gpu_ctx, svdag, cfg)
    for level in range(2, maxT + 1):
        a, b = t.record(level), host.record(level)
        assert (a.items, a.unique, a.matched) == (b.items, b.unique, b.matched), level
    g = pl.BuildGraph(gpu_ctx, svdag, cfg)
    svdag.close()  # the graph keeps its own device copy of the input levels
    try:
        assert g.result().tdag_bytes() == ref
        for replay in range(3):
            ms = g.run()
            assert ms > 0
            assert g.result().tdag_bytes() == ref, f"replay {replay}"
            _ = torch.ones(1 << 20, device="cuda").sum().item()
    finally:
        g.close()
