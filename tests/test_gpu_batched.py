"""GPU parity of the batched form: the persistent multi-level kernel with its device work queue,
and the concurrent form (each level's own matcher on its own stream / graph branch), with and
without a hipGraph.

Every level of a batch must produce the same encodedParents keys, mapping and unique list as the
CPU oracle (and hence as the per-level launches), whatever the phase schedule: the tests force
many phases (tiny first batches) so the in-kernel pruning and the grid barriers run, replay one
captured graph several times with other GPU work in between, and cover config switches, empty
levels and level subsets.
"""
import numpy as np
import pytest

import sets

pytestmark = pytest.mark.gpu


def _level_sets(oracle, seed, levels=(2, 3, 4, 5)):
    rng = np.random.default_rng(seed)
    sizes = {2: 300, 3: 120, 4: 30, 5: 8}
    return {lv: sets.planted_set(rng, lv, sizes[lv], sizes[lv], oracle.transform) for lv in levels}


def _device_jobs(level_masks):
    import torch

    jobs, bufs = [], {}
    for lv, m in level_masks.items():
        n = len(m)
        d = {
            "masks": torch.from_numpy(np.ascontiguousarray(m).view(np.int64).reshape(-1)).cuda() if n else torch.zeros(1, dtype=torch.int64, device="cuda"),
            "enc": torch.zeros(max(1, n), dtype=torch.int64, device="cuda"),
            "map": torch.zeros(max(1, n), dtype=torch.int64, device="cuda"),
            "uniq": torch.zeros(max(1, n), dtype=torch.int32, device="cuda"),
        }
        bufs[lv] = d
        jobs.append({"level": lv, "n": n, "d_masks": d["masks"], "d_encoded": d["enc"], "d_mapping": d["map"], "d_unique": d["uniq"]})
    torch.cuda.synchronize()  # torch's stream is not ordered with the context's stream
    return jobs, bufs


def _check_outputs(oracle, level_masks, bufs, stats, cfg_tuple=(True, True, True)):
    import torch

    torch.cuda.synchronize()
    for i, (lv, m) in enumerate(level_masks.items()):
        n = len(m)
        ref_enc = oracle.match_encoded(lv, m, oracle.config(*cfg_tuple))
        ref_map, ref_uniq = oracle.postprocess(lv, m, ref_enc)
        enc = bufs[lv]["enc"].cpu().numpy().view(np.uint64)[:n]
        mapping = bufs[lv]["map"].cpu().numpy().view(np.uint64)[:n]
        nu = int(stats[i].num_unique)
        uniq = bufs[lv]["uniq"].cpu().numpy().view(np.uint32)[:nu]
        bad = np.nonzero(enc != ref_enc)[0]
        assert bad.size == 0, f"level {lv}: {bad.size} keys differ, first at {bad[:5]}"
        assert np.array_equal(mapping, ref_map), f"level {lv}: mapping differs"
        assert nu == len(ref_uniq) and np.array_equal(uniq, ref_uniq), f"level {lv}: unique list differs"


@pytest.mark.parametrize("use_graph,concurrent", [(False, False), (True, False), (False, True), (True, True)])
def test_batched_all_levels(oracle, gpu_ctx, use_graph, concurrent):
    import tdag_amd

    level_masks = _level_sets(oracle, 7)
    jobs, bufs = _device_jobs(level_masks)
    print(f"batched use_graph={use_graph} sizes={ {lv: len(m) for lv, m in level_masks.items()} }", flush=True)
    b = gpu_ctx.batch(jobs, tdag_amd.TransformDAGConfig(True, True, True), use_graph=use_graph, concurrent=concurrent)
    try:
        stats, ms = b.run()
    finally:
        b.close()
    _check_outputs(oracle, level_masks, bufs, stats)
    assert sum(int(s.num_matched) for s in stats) > 0


@pytest.mark.parametrize("first_batch", [1, 3])
def test_batched_many_phases(oracle, gpu_ctx, first_batch):
    """Tiny first batches: many phases, each followed by in-kernel pruning behind grid barriers."""
    import tdag_amd

    level_masks = _level_sets(oracle, 11)
    jobs, bufs = _device_jobs(level_masks)
    for lv in level_masks:
        gpu_ctx.set_batching(lv, first_batch, 2)
    try:
        print(f"batched first_batch={first_batch}", flush=True)
        b = gpu_ctx.batch(jobs, tdag_amd.TransformDAGConfig(True, True, True), use_graph=True)
        try:
            stats, _ = b.run()
        finally:
            b.close()
    finally:
        for lv in level_masks:
            gpu_ctx.set_batching(lv, 0, 0)
    _check_outputs(oracle, level_masks, bufs, stats)


@pytest.mark.parametrize("concurrent", [False, True])
def test_batched_graph_replays(oracle, gpu_ctx, concurrent):
    """One captured graph replayed several times, with other GPU work (a plain per-level call and
    torch kernels) between the replays; every replay equals the oracle. Concurrent mode: the graph
    holds one branch per level."""
    import torch

    import tdag_amd

    level_masks = _level_sets(oracle, 23)
    jobs, bufs = _device_jobs(level_masks)
    b = gpu_ctx.batch(jobs, tdag_amd.TransformDAGConfig(True, True, True), use_graph=True, concurrent=concurrent)
    try:
        grab = b.autotune(1)  # rewrites the descriptor the graph reads; replays must stay exact
        assert grab in ((0,) if concurrent else (1, 2, 4, 8, 16))
        if not concurrent:
            assert b.tuning()[0] == grab and b.tuning()[1] in (0, 2, 3, 4, 6)
        for bad in ((0, 3), (1, 1), (65, 0)):
            with pytest.raises(RuntimeError):
                b.set_tuning(*bad)
        # queue tunings the replays run with: tiles per pop, split_every (the persistent form)
        settings = [None, (1, 0), (4, 2), (2, 6)]
        for rep in range(4):
            if settings[rep] is not None:
                b.set_tuning(*settings[rep])
                assert b.tuning() == settings[rep]
            print(f"replay {rep} concurrent={concurrent} tuning={b.tuning()}", flush=True)
            for d in bufs.values():
                d["enc"].fill_(-1)
                d["map"].fill_(-1)
            torch.cuda.synchronize()
            stats, _ = b.run()
            _check_outputs(oracle, level_masks, bufs, stats)
            # unrelated eager work between replays
            m2 = bufs[2]["masks"]
            gpu_ctx.match_level_device(2, m2, len(level_masks[2]), tdag_amd.TransformDAGConfig(True, True, True))
            torch.cuda.synchronize()
    finally:
        b.close()


@pytest.mark.parametrize("cfg", [(False, True, True), (True, False, True), (True, True, False)])
def test_batched_config_switches(oracle, gpu_ctx, cfg):
    import tdag_amd

    level_masks = _level_sets(oracle, 31, levels=(2, 3, 5))
    jobs, bufs = _device_jobs(level_masks)
    print(f"batched cfg={cfg}", flush=True)
    stats = gpu_ctx.match_levels_batched_device(jobs, tdag_amd.TransformDAGConfig(*cfg))
    _check_outputs(oracle, level_masks, bufs, stats, cfg)


def test_batched_concurrent_subset_with_empty_level(oracle, gpu_ctx):
    import tdag_amd

    rng = np.random.default_rng(43)
    level_masks = {
        5: sets.planted_set(rng, 5, 6, 6, oracle.transform),
        3: np.zeros((0, sets.words(3)), dtype=np.uint64),
        2: sets.random_masks(rng, 2, 40, 0.3),
    }
    jobs, bufs = _device_jobs(level_masks)
    print("concurrent batch with an empty level", flush=True)
    b = gpu_ctx.batch(jobs, tdag_amd.TransformDAGConfig(True, True, True), use_graph=True, concurrent=True)
    try:
        stats, _ = b.run()
    finally:
        b.close()
    _check_outputs(oracle, level_masks, bufs, stats)


def test_batched_subsets_and_empty(oracle, gpu_ctx):
    import tdag_amd

    rng = np.random.default_rng(41)
    level_masks = {
        4: sets.planted_set(rng, 4, 20, 20, oracle.transform),
        3: np.zeros((0, sets.words(3)), dtype=np.uint64),
        2: sets.random_masks(rng, 2, 5, 0.3),
    }
    jobs, bufs = _device_jobs(level_masks)
    print("batched subset with an empty level", flush=True)
    stats = gpu_ctx.match_levels_batched_device(jobs, tdag_amd.TransformDAGConfig(True, True, True))
    _check_outputs(oracle, level_masks, bufs, stats)


def test_batched_rejects_duplicate_levels(gpu_ctx):
    import torch

    import tdag_amd

    m = torch.zeros(4, dtype=torch.int64, device="cuda")
    torch.cuda.synchronize()
    jobs = [{"level": 2, "n": 4, "d_masks": m}, {"level": 2, "n": 4, "d_masks": m}]
    with pytest.raises(tdag_amd.TdagError, match="one job per level"):
        gpu_ctx.match_levels_batched_device(jobs, tdag_amd.TransformDAGConfig())


@pytest.mark.parametrize("concurrent", [False, True])
def test_combined_batch_with_learned_level2(oracle, gpu_ctx, concurrent):
    """The combined form (tdag_batch_create_learned): levels 3..5 equal the oracle, level 2 equals
    the standalone learned matcher (same network, bit for bit), over several graph replays with
    other GPU work in between."""
    import torch

    import tdag_amd
    from tdag_amd import learned

    level_masks = _level_sets(oracle, 31)
    jobs, bufs = _device_jobs(level_masks)
    mlp = learned.MLP(gpu_ctx, seed=5)
    m2 = level_masks[2]
    n2 = len(m2)
    cfg = tdag_amd.TransformDAGConfig(True, True, True, 5)
    ref2 = torch.zeros(max(1, n2), dtype=torch.int64, device="cuda")
    learned.match_level_learned_device(gpu_ctx, mlp, bufs[2]["masks"], n2, cfg, 64, d_encoded=ref2)
    torch.cuda.synchronize()
    ref2 = ref2.cpu().numpy().view(np.uint64)[:n2].copy()
    exact2 = oracle.match_encoded(2, m2, oracle.config(True, True, True))
    print(f"combined batch concurrent={concurrent}", flush=True)
    b = gpu_ctx.batch(jobs, cfg, use_graph=True, concurrent=concurrent, mlp=mlp, k=64)
    try:
        for replay in range(3):
            for d in bufs.values():
                d["enc"].zero_()
            torch.cuda.synchronize()
            stats, _ = b.run()
            hits, fb = b.learned_stats()
            assert hits + fb <= n2
            rest = {lv: m for lv, m in level_masks.items() if lv != 2}
            _check_outputs(oracle, rest, bufs, [s for lv, s in zip(level_masks, stats) if lv != 2])
            got2 = bufs[2]["enc"].cpu().numpy().view(np.uint64)[:n2]
            assert np.array_equal(got2, ref2), f"replay {replay}: level 2 differs from the standalone learned matcher"
            assert np.all(got2 >= exact2)  # never below the exhaustive minimum
            _ = torch.ones(1 << 20, device="cuda").sum().item()  # other GPU work between replays
    finally:
        b.close()
        mlp.close()
