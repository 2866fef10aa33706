"""GPU parity of the matcher through the C-ABI: bit-exact against the CPU oracle.

Compares the raw encodedParents keys (minimum key per item, find_translation.cpp:864), the
outPrevLevelMapping TransformPointer patterns and the unique list, for every level, every
config switch, edge cases (empty input, single item, empty masks, duplicates) and larger
seeded sets.
"""
import numpy as np
import pytest

import sets

pytestmark = pytest.mark.gpu


def _check(oracle, gpu_ctx, level, masks, cfg_tuple=(True, True, True)):
    import tdag_amd

    cfg = tdag_amd.TransformDAGConfig(*cfg_tuple)
    enc, mapping, uniq, stats = gpu_ctx.match_level(level, masks, cfg)
    ref_enc = oracle.match_encoded(level, masks, oracle.config(*cfg_tuple))
    ref_map, ref_uniq = oracle.postprocess(level, masks, ref_enc)
    bad = np.nonzero(enc != ref_enc)[0]
    assert bad.size == 0, f"level {level}: {bad.size} keys differ, first at {bad[:5]}: gpu={enc[bad[:3]]} ref={ref_enc[bad[:3]]}"
    assert np.array_equal(mapping, ref_map)
    assert np.array_equal(uniq, ref_uniq)
    assert stats.num_unique == len(ref_uniq)
    return stats


@pytest.mark.parametrize("level,n_base,n_planted,seed", [(2, 300, 300, 0), (3, 120, 120, 1), (4, 30, 30, 2), (5, 8, 8, 3)])
def test_planted_sets(oracle, gpu_ctx, level, n_base, n_planted, seed):
    rng = np.random.default_rng(seed)
    masks = sets.planted_set(rng, level, n_base, n_planted, oracle.transform)
    print(f"level={level} n={len(masks)} seed={seed}", flush=True)
    st = _check(oracle, gpu_ctx, level, masks)
    assert st.num_matched > 0


@pytest.mark.parametrize("cfg", [(False, True, True), (True, False, True), (True, True, False), (False, False, False)])
@pytest.mark.parametrize("level", [2, 3])
def test_config_switches(oracle, gpu_ctx, level, cfg):
    rng = np.random.default_rng(40 + level)
    masks = sets.planted_set(rng, level, 80, 80, oracle.transform)
    print(f"level={level} cfg={cfg} n={len(masks)}", flush=True)
    _check(oracle, gpu_ctx, level, masks, cfg)


@pytest.mark.parametrize("level", [2, 3, 4, 5])
def test_edge_cases(oracle, gpu_ctx, level):
    W = sets.words(level)
    print(f"level={level} edge cases", flush=True)
    _check(oracle, gpu_ctx, level, np.zeros((0, W), dtype=np.uint64))
    one = np.zeros((1, W), dtype=np.uint64)
    one[0, -1] = np.uint64(1) << np.uint64(63)
    _check(oracle, gpu_ctx, level, one)
    _check(oracle, gpu_ctx, level, np.zeros((3, W), dtype=np.uint64))
    rng = np.random.default_rng(level)
    m = sets.random_masks(rng, level, 3, 0.3)
    _check(oracle, gpu_ctx, level, np.concatenate([m, m, m, m]))


@pytest.mark.parametrize("level,n", [(2, 20000), (3, 3000)])
def test_larger_sets(oracle, gpu_ctx, level, n):
    rng = np.random.default_rng(1000 + level)
    masks = sets.planted_set(rng, level, n // 2, n // 2, oracle.transform, edge_cases=False)
    print(f"level={level} n={len(masks)}", flush=True)
    _check(oracle, gpu_ctx, level, masks)


# Search-kernel variants per level: unit kind (combo / family / shared passes) x plane split x
# block order; variants whose blocks do not fit the LDS are rejected by the library.
VALID_VARIANTS = {3: [0, 1, 2, 9, 10, 11], 4: list(range(18)), 5: [0, 1, 3, 4, 5, 6, 7, 8, 9, 10, 12, 13, 14, 15, 16, 17]}


@pytest.mark.parametrize("level,n_base,variant", [(3, 60, v) for v in VALID_VARIANTS[3]] + [(4, 20, v) for v in VALID_VARIANTS[4]]
                         + [(5, 6, v) for v in VALID_VARIANTS[5]])
def test_search_variants(oracle, gpu_ctx, level, n_base, variant):
    """Every search-kernel variant (tile shape) the autotuner may pick is bit-exact."""
    rng = np.random.default_rng(500 + 10 * level + variant)
    masks = sets.planted_set(rng, level, n_base, n_base, oracle.transform)
    print(f"level={level} variant={variant} n={len(masks)}", flush=True)
    gpu_ctx.set_variant(level, variant)
    try:
        _check(oracle, gpu_ctx, level, masks)
    finally:
        gpu_ctx.set_variant(level, -1)


@pytest.mark.parametrize("level,n_base,frames,variant", [(lv, nb, f, v) for lv, nb in ((3, 60), (4, 20), (5, 6)) for f in (1, 6)
                                                         for v in (0, 1, 2, 16) if v in VALID_VARIANTS[lv]])
def test_child_frames(oracle, gpu_ctx, level, n_base, frames, variant):
    """Both places of the axis permutations (48 parent combos against 8 entries per child, or 8
    parent directions against 48 entries per child) give the oracle's result bit for bit."""
    rng = np.random.default_rng(700 + 10 * level + variant + frames)
    masks = sets.planted_set(rng, level, n_base, n_base, oracle.transform)
    print(f"level={level} frames={frames} variant={variant} n={len(masks)}", flush=True)
    gpu_ctx.set_variant(level, variant)
    gpu_ctx.set_child_frames(level, frames)
    try:
        assert gpu_ctx.get_child_frames(level) == frames
        _check(oracle, gpu_ctx, level, masks)
    finally:
        gpu_ctx.set_variant(level, -1)
        gpu_ctx.set_child_frames(level, 0)


@pytest.mark.parametrize("level", [3, 4, 5])
@pytest.mark.parametrize("frames", [1, 6])
def test_child_frames_config_switches(oracle, gpu_ctx, level, frames):
    """With axis permutation or symmetry switched off, the permuted or mirrored table entries
    must not produce matches the configuration forbids."""
    rng = np.random.default_rng(800 + level + frames)
    n = 30 if level < 5 else 5
    masks = sets.planted_set(rng, level, n, n, oracle.transform)
    gpu_ctx.set_child_frames(level, frames)
    try:
        for cfg in ((False, True, True), (True, False, True), (False, False, True), (True, True, False)):
            print(f"level={level} frames={frames} cfg={cfg} n={len(masks)}", flush=True)
            _check(oracle, gpu_ctx, level, masks, cfg)
    finally:
        gpu_ctx.set_child_frames(level, 0)


@pytest.mark.parametrize("level", [2, 3, 5])
def test_probe_counters_consistent(oracle, gpu_ctx, level):
    """Per-stage counters are ordered (probes >= filter passes, candidates >= verifications >=
    verified >= foreign matches) and turning them on does not change the result."""
    rng = np.random.default_rng(900 + level)
    masks = sets.planted_set(rng, level, 40 if level < 5 else 6, 40 if level < 5 else 6, oracle.transform)
    gpu_ctx.set_counters(True)
    try:
        st = _check(oracle, gpu_ctx, level, masks)
    finally:
        gpu_ctx.set_counters(False)
    c = st.counters()
    print(level, c, st.num_matched, flush=True)
    assert c["probes"] >= c["filter_pass"] >= 0
    assert c["key_hits"] >= c["candidates"] >= c["verifications"] >= c["verified"] >= st.num_matched
    assert c["probes"] > 0


@pytest.mark.parametrize("level,first,growth", [(2, 1, 2), (2, 7, 3), (3, 1, 2), (3, 5, 4), (4, 2, 2), (5, 1, 2)])
def test_rank_batched_search(oracle, gpu_ctx, level, first, growth):
    """Searching parents in rank batches with table pruning in between gives the same result
    as one batch (pruning only drops children that can no longer improve)."""
    rng = np.random.default_rng(1200 + level * 10 + first)
    nb = {2: 150, 3: 50, 4: 16, 5: 5}[level]
    masks = sets.planted_set(rng, level, nb, nb, oracle.transform)
    print(f"level={level} first={first} growth={growth} n={len(masks)}", flush=True)
    gpu_ctx.set_batching(level, first, growth)
    try:
        _check(oracle, gpu_ctx, level, masks)
    finally:
        gpu_ctx.set_batching(level, 0, 0)


@pytest.mark.parametrize("level,bits", [(2, 1), (3, 4), (5, 2)])
def test_small_filters(oracle, gpu_ctx, level, bits):
    """A tiny (mostly saturated) Bloom filter only costs time, never changes the result."""
    rng = np.random.default_rng(1300 + level)
    nb = {2: 150, 3: 50, 5: 5}[level]
    masks = sets.planted_set(rng, level, nb, nb, oracle.transform)
    gpu_ctx.set_filter_bits(level, bits)
    try:
        _check(oracle, gpu_ctx, level, masks)
    finally:
        gpu_ctx.set_filter_bits(level, 0)


@pytest.mark.parametrize("level,n", [(2, 400), (3, 150), (5, 16)])
def test_subset_search_keeps_the_flagged_minima(oracle, gpu_ctx, level, n):
    """tdag_match_level_subset_device: flagged children get exactly the full search's minimum key,
    the others keep their identity key (empty ones the sentinel)."""
    import torch

    import tdag_amd

    rng = np.random.default_rng(90 + level)
    masks = sets.planted_set(rng, level, n // 2, n - n // 2, oracle.transform)
    m = len(masks)
    full = oracle.match_encoded(level, masks, oracle.config(True, True, True))
    flags = (rng.random(m) < 0.4).astype(np.int32)
    d_m = torch.from_numpy(np.ascontiguousarray(masks).view(np.int64).reshape(-1).copy()).cuda()
    d_f = torch.from_numpy(flags).cuda()
    d_e = torch.zeros(max(1, m), dtype=torch.int64, device="cuda")
    torch.cuda.synchronize()
    print(f"subset level={level} n={m}", flush=True)
    gpu_ctx.match_level_subset_device(level, d_m, m, d_f, tdag_amd.TransformDAGConfig(True, True, True), d_encoded=d_e)
    got = d_e.cpu().numpy().view(np.uint64)[:m]
    # identity key of the item's own rank (encodeFixed64 of TransformPointer::create(rank)); empty items the sentinel's
    pop = np.bitwise_count(np.ascontiguousarray(masks, dtype=np.uint64).reshape(m, -1)).sum(axis=1).astype(np.int64)
    rank = np.empty(m, dtype=np.int64)
    rank[np.argsort(-pop, kind="stable")] = np.arange(m)
    b, bias = level + 1, 1 << level

    def ident(ptr):
        k = ptr << 6
        for _ in range(3):
            k = (k << b) | bias
        return k

    expect = np.array([ident(0xFFFFFFFF) if pop[i] == 0 else ident(int(rank[i])) for i in range(m)], dtype=np.uint64)
    assert np.array_equal(got[flags == 1], full[flags == 1])
    assert np.array_equal(got[flags == 0], expect[flags == 0])
