"""The configurations at full size on the GPU.

config 1: the 1024^3 scene; the driver's .tdag and every level's keys equal the CPU oracle's
(through the digests it wrote, tests/golden/config1_digest.json).
config 2: the 1024^3 generator at 2048^3 with the committed autotuner picks; both batched forms
(one persistent kernel for all levels with a device work queue, and concurrent per-level branches;
captured hipGraphs) must equal the per-level launches bit for bit on every level's full input.
config 3: the learned proposal on the 2048^3 scene, trained on three 1024^3 scenes.
config 4: the 4096^3 city: batched forms, the combined form (batched + learned + fallback with a
fine-tuned network) and the .tdag16 decode of the whole volume.
"""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _tuned_context(name: str):
    """A context with the committed autotuner output configs/<name> applied (search variant, batch
    schedule, filter size and child frames per level)."""
    import tdag_amd

    tuning = json.load(open(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "configs", name)))
    ctx = tdag_amd.Context(0)
    for level_s, tt in tuning["levels"].items():
        level = int(level_s)
        ctx.set_variant(level, int(tt["variant"]))
        ctx.set_batching(level, int(tt["first_batch"]), int(tt["growth"]))
        ctx.set_filter_bits(level, int(tt["filter_bits"]))
        if level >= 3:
            ctx.set_child_frames(level, int(tt["child_frames"]))
    return ctx


def test_config2_batched_bit_exact_2048():
    """Config 2 at full size (2048^3 scene, seed 0) with the autotuner's picks: the committed
    per-level tuning of the workload, then the batched call's own autotune of its tile grab and
    phases. Both batched forms (one persistent kernel, and concurrent per-level branches), each
    captured as a hipGraph, equal the plain per-level launches key for key and mapping for
    mapping."""
    import torch

    import tdag_amd
    from tdag_amd import pipeline as pl

    R = 2048
    cfg = tdag_amd.TransformDAGConfig(True, True, True, 5)
    ctx = _tuned_context("tuning_scene_2048_seed0_maxT5.json")
    try:
        print("config 2: building the 2048^3 scene", flush=True)
        bits = pl.scene_dense(1, R, 0)
        t = pl.construct_transform_dag(ctx, pl.Svdag(bits, R), cfg, keep_level_inputs=True)
        del bits
        ref, masks = {}, {}
        for level in (5, 4, 3, 2):
            m = np.ascontiguousarray(t.level_masks(level))
            n = len(m)
            d_m = torch.from_numpy(m.view(np.int64).reshape(-1)).cuda()
            enc = torch.empty(max(1, n), dtype=torch.int64, device="cuda")
            mp = torch.empty(max(1, n), dtype=torch.int64, device="cuda")
            torch.cuda.synchronize()
            _, st = ctx.match_level_device(level, d_m, n, cfg, d_encoded=enc, d_mapping=mp)
            ref[level] = (enc, mp, n, int(st.num_matched))
            masks[level] = d_m
            print(f"level {level}: {n} items, {st.num_matched} matched", flush=True)
        t.close()
        for concurrent in (False, True):
            jobs = []
            for level in (5, 4, 3, 2):
                enc, mp, n, _ = ref[level]
                jobs.append({"level": level, "n": n, "d_masks": masks[level], "d_encoded": torch.empty_like(enc), "d_mapping": torch.empty_like(mp),
                             "d_unique": torch.empty(max(1, n), dtype=torch.int32, device="cuda")})
            torch.cuda.synchronize()
            b = ctx.batch(jobs, cfg, use_graph=True, concurrent=concurrent)
            try:
                if not concurrent:
                    b.autotune(1)
                stats, ms = b.run()
            finally:
                b.close()
            torch.cuda.synchronize()
            print(f"batched call (concurrent={concurrent}) {ms:.2f} ms", flush=True)
            for j, s in zip(jobs, stats):
                enc, mp, n, matched = ref[j["level"]]
                assert torch.equal(j["d_encoded"][:n], enc[:n]), f"level {j['level']}: keys differ (concurrent={concurrent})"
                assert torch.equal(j["d_mapping"][:n], mp[:n]), f"level {j['level']}: mapping differs (concurrent={concurrent})"
                assert int(s.num_matched) == matched
    finally:
        ctx.close()


def _sha(a) -> str:
    import hashlib

    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def test_config1_full_size_equals_the_oracle_digest(gpu_ctx):
    """Config 1 at full size (1024^3, kind 1, seed 0, maxT 5): the HIP driver's .tdag and every
    level's matcher keys equal the CPU oracle's, through the digests the oracle wrote
    (tests/golden/make_config1_digest.py: input grid, level inputs, encodedParents, .tdag)."""
    import json
    import os

    import torch

    import tdag_amd
    from tdag_amd import pipeline as pl

    golden = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "config1_digest.json")))
    sc = golden["scene"]
    print("config 1: 1024^3 scene against the oracle digest", flush=True)
    bits = pl.scene_dense(sc["kind"], sc["resolution"], sc["seed"])
    assert _sha(bits) == sc["bits_sha256"]
    cfg = tdag_amd.TransformDAGConfig(True, True, True, golden["config"]["max_translation_level"])
    t = pl.construct_transform_dag(gpu_ctx, pl.Svdag(bits, sc["resolution"]), cfg, keep_level_inputs=True)
    data = t.tdag_bytes()
    assert len(data) == golden["tdag_bytes"]
    import hashlib

    assert hashlib.sha256(data).hexdigest() == golden["tdag_sha256"]
    for level_s, g in golden["levels"].items():
        level = int(level_s)
        m = np.ascontiguousarray(t.level_masks(level))
        n = len(m)
        assert n == g["items"] and _sha(m.astype("<u8")) == g["masks_sha256"]
        d_m = torch.from_numpy(m.view(np.int64).reshape(-1)).cuda()
        enc = torch.empty(max(1, n), dtype=torch.int64, device="cuda")
        torch.cuda.synchronize()
        nu, st = gpu_ctx.match_level_device(level, d_m, n, cfg, d_encoded=enc)
        keys = enc.cpu().numpy().view(np.uint64)[:n]
        assert _sha(keys.astype("<u8")) == g["encoded_sha256"], f"level {level}: keys differ from the oracle's"
        assert int(st.num_matched) == g["matched"] and nu == g["unique"]
        print(f"level {level}: {n} items, keys equal the oracle's", flush=True)
    t.close()


def _verified_exact(oracle, masks, keys, level=2):
    """Every key whose parent is another item names a transform that maps the parent onto the item."""
    from tdag_amd import learned

    order = learned.rank_order(masks)
    rank_of = np.empty_like(order)
    rank_of[order] = np.arange(len(order))
    low = 6 + 3 * (level + 1)
    b = level + 1
    bias = 1 << level
    checked = 0
    for i in range(len(masks)):
        k = int(keys[i])
        ptr = k >> low
        if ptr == rank_of[i] or ptr >= len(masks):
            continue
        tx = (k & ((1 << b) - 1)) - bias
        ty = ((k >> b) & ((1 << b) - 1)) - bias
        tz = ((k >> (2 * b)) & ((1 << b) - 1)) - bias
        sym = (k >> (3 * b)) & 7
        perm = (k >> (3 * b + 3)) & 7
        t = oracle.transform(level, np.array([masks[order[ptr]]], dtype=np.uint64), perm, sym, (tx, ty, tz))
        assert int(t[0]) == int(masks[i]), f"item {i}: accepted match is not exact"
        checked += 1
    return checked


def _learned_level2(oracle, gpu_ctx, mlp, t, cfg, label):
    import torch

    from tdag_amd import learned

    m2 = np.ascontiguousarray(t.level_masks(2).reshape(-1))
    n = len(m2)
    d_m = torch.from_numpy(m2.view(np.int64)).cuda()
    ex = torch.empty(max(1, n), dtype=torch.int64, device="cuda")
    le = torch.empty(max(1, n), dtype=torch.int64, device="cuda")
    torch.cuda.synchronize()
    _, st_ex = gpu_ctx.match_level_device(2, d_m, n, cfg, d_encoded=ex)
    st = learned.match_level_learned_device(gpu_ctx, mlp, d_m, n, cfg, 64, d_encoded=le)
    ex = ex.cpu().numpy().view(np.uint64)[:n]
    le = le.cpu().numpy().view(np.uint64)[:n]
    assert np.all(le >= ex), "a learned key is below the exhaustive minimum"
    checked = _verified_exact(oracle, m2, le)
    matched = int(st_ex.num_matched)
    print(f"{label}: level 2 {n} items, recall@64 {st.proposal_hits / max(1, matched):.4f}, fallback {st.fallback_queries}, "
          f"{checked} accepted matches re-verified", flush=True)
    return st


def test_config3_learned_2048(oracle, gpu_ctx):
    """Config 3: the learned proposal (trained on level-2 pairs of three seeded 1024^3 scenes) on
    the 2048^3 scene: every accepted match exact and never below the exhaustive key, the learned
    .tdag within 0.5 % of the exhaustive one, recall reported."""
    import tdag_amd
    from tdag_amd import learned
    from tdag_amd import pipeline as pl

    cfg = tdag_amd.TransformDAGConfig(True, True, True, 5)
    print("config 3: training pairs from three 1024^3 scenes", flush=True)
    d_q, d_p, n = learned.training_pairs(gpu_ctx, cfg, 1, 1024, [101, 102, 103])
    mlp = learned.MLP(gpu_ctx, seed=0)
    losses = mlp.train(d_q, d_p, n, steps=1500, batch=512, lr=1e-3, tau=0.1, seed=7)
    assert losses[-20:].mean() < losses[:20].mean()
    bits = pl.scene_dense(1, 2048, 0)
    svdag = pl.Svdag(bits, 2048)
    del bits
    t = pl.construct_transform_dag(gpu_ctx, svdag, cfg, keep_level_inputs=True)
    st = _learned_level2(oracle, gpu_ctx, mlp, t, cfg, "config 3")
    assert st.proposal_hits > 0
    exhaustive = len(t.tdag_bytes())
    tl = pl.construct_transform_dag_learned(gpu_ctx, svdag, cfg, mlp, 64)
    delta = len(tl.tdag_bytes()) / exhaustive - 1.0
    print(f"config 3: .tdag {exhaustive} exhaustive, learned delta {delta:+.5f}", flush=True)
    assert abs(delta) <= 0.005
    tl.close()
    t.close()
    mlp.close()


def test_config4_city_4096(oracle, gpu_ctx):
    """Config 4 at full size (4096^3 city, seed 0, the committed tuning): both batched forms equal
    the per-level launches bit for bit; the combined form (batched + learned level 2, the network
    pre-trained on config-3 pairs and fine-tuned on level-2 pairs of three seeded 1024^3 cities,
    exhaustive fallback) keeps levels 3..5 bit-exact and every accepted level-2 match exact, never
    below the exhaustive key, with the learned .tdag within 0.5 %; the device .tdag16 decode of
    all 4096^3 voxels equals the input grid."""
    import json
    import os

    import torch

    import tdag_amd
    from tdag_amd import learned
    from tdag_amd import pipeline as pl

    R = 4096
    ctx = _tuned_context("tuning_city_4096_seed0_maxT5.json")
    try:
        cfg = tdag_amd.TransformDAGConfig(True, True, True, 5)
        print("config 4: building the 4096^3 city", flush=True)
        bits = pl.scene_dense(2, R, 0)
        svdag = pl.Svdag(bits, R)
        t = pl.construct_transform_dag(ctx, svdag, cfg, keep_level_inputs=True)
        exhaustive_bytes = len(t.tdag_bytes())
        # .tdag16 decode of every voxel on the device, compared with the input grid in HBM.
        e = t.encode16()
        d_out, _ = e.decode_device(ctx)
        d_bits = torch.from_numpy(np.ascontiguousarray(bits).view(np.int32)).cuda()
        assert torch.equal(d_out, d_bits), "the decoded 4096^3 grid differs from the input"
        del d_out, d_bits, bits
        e.close()
        torch.cuda.empty_cache()
        print("config 4: .tdag16 decode of 4096^3 voxels equals the input", flush=True)
        dev, ref = {}, {}
        for level in (5, 4, 3, 2):
            m = np.ascontiguousarray(t.level_masks(level))
            n = len(m)
            d_m = torch.from_numpy(m.view(np.int64).reshape(-1)).cuda()
            enc = torch.empty(max(1, n), dtype=torch.int64, device="cuda")
            mp = torch.empty(max(1, n), dtype=torch.int64, device="cuda")
            torch.cuda.synchronize()
            _, st = ctx.match_level_device(level, d_m, n, cfg, d_encoded=enc, d_mapping=mp)
            dev[level] = (d_m, n, m)
            ref[level] = (enc, mp, int(st.num_matched))
            print(f"level {level}: {n} items, {st.num_matched} matched", flush=True)

        def jobs_for():
            jobs = []
            for level in (5, 4, 3, 2):
                d_m, n, _ = dev[level]
                jobs.append({"level": level, "n": n, "d_masks": d_m, "d_encoded": torch.empty(max(1, n), dtype=torch.int64, device="cuda"),
                             "d_mapping": torch.empty(max(1, n), dtype=torch.int64, device="cuda"),
                             "d_unique": torch.empty(max(1, n), dtype=torch.int32, device="cuda")})
            torch.cuda.synchronize()
            return jobs

        for concurrent in (False, True):
            jobs = jobs_for()
            b = ctx.batch(jobs, cfg, use_graph=True, concurrent=concurrent)
            try:
                stats, ms = b.run()
            finally:
                b.close()
            torch.cuda.synchronize()
            for j, s in zip(jobs, stats):
                enc, mp, matched = ref[j["level"]]
                n = j["n"]
                assert torch.equal(j["d_encoded"][:n], enc[:n]) and torch.equal(j["d_mapping"][:n], mp[:n]), f"level {j['level']} differs (concurrent={concurrent})"
                assert int(s.num_matched) == matched
            print(f"config 4: batched concurrent={concurrent} bit-exact, {ms:.2f} ms", flush=True)

        print("config 4: training the learned proposal", flush=True)
        mlp = learned.MLP(ctx, seed=0)
        d_q, d_p, n = learned.training_pairs(ctx, cfg, 1, 1024, [101, 102, 103])
        mlp.train(d_q, d_p, n, steps=1500, batch=512, lr=1e-3, tau=0.1, seed=7)
        d_q, d_p, n = learned.training_pairs(ctx, cfg, 2, 1024, [101, 102, 103])
        mlp.train(d_q, d_p, n, steps=500, batch=512, lr=3e-4, tau=0.1, seed=8)
        jobs = jobs_for()
        b = ctx.batch(jobs, cfg, use_graph=True, concurrent=False, mlp=mlp, k=64)
        try:
            stats, ms = b.run()
            hits, fb = b.learned_stats()
        finally:
            b.close()
        torch.cuda.synchronize()
        for j in jobs:
            level, n = j["level"], j["n"]
            enc = ref[level][0]
            if level != 2:
                assert torch.equal(j["d_encoded"][:n], enc[:n]), f"combined form: level {level} differs"
        d_m, n2, m2 = dev[2]
        got2 = jobs[3]["d_encoded"][:n2].cpu().numpy().view(np.uint64)
        ex2 = ref[2][0][:n2].cpu().numpy().view(np.uint64)
        assert np.all(got2 >= ex2)
        checked = _verified_exact(oracle, m2.reshape(-1), got2)
        matched2 = ref[2][2]
        print(f"config 4: combined form {ms:.2f} ms, level-2 recall@64 {hits / max(1, matched2):.4f}, fallback {fb}, {checked} matches re-verified",
              flush=True)
        tl = pl.construct_transform_dag_learned(ctx, svdag, cfg, mlp, 64)
        delta = len(tl.tdag_bytes()) / exhaustive_bytes - 1.0
        print(f"config 4: .tdag {exhaustive_bytes} exhaustive, learned delta {delta:+.5f}", flush=True)
        assert abs(delta) <= 0.005
        tl.close()
        mlp.close()
        t.close()
    finally:
        ctx.close()
