"""Larger configurations on the GPU.

config 2: the 1024^3 generator at 2048^3; the batched form (one persistent kernel for all levels,
device work queue, captured hipGraph) must equal the per-level launches bit for bit on every
level's full input.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_config2_batched_bit_exact_2048(gpu_ctx):
    import torch

    import tdag_amd
    from tdag_amd import pipeline as pl

    R = 2048
    cfg = tdag_amd.TransformDAGConfig(True, True, True, 5)
    print("config 2: building the 2048^3 scene", flush=True)
    bits = pl.scene_dense(1, R, 0)
    t = pl.construct_transform_dag(gpu_ctx, pl.Svdag(bits, R), cfg, keep_level_inputs=True)
    del bits
    jobs, ref = [], {}
    for level in (5, 4, 3, 2):
        m = np.ascontiguousarray(t.level_masks(level))
        n = len(m)
        d_m = torch.from_numpy(m.view(np.int64).reshape(-1)).cuda()
        enc = torch.empty(max(1, n), dtype=torch.int64, device="cuda")
        mp = torch.empty(max(1, n), dtype=torch.int64, device="cuda")
        torch.cuda.synchronize()
        _, st = gpu_ctx.match_level_device(level, d_m, n, cfg, d_encoded=enc, d_mapping=mp)
        ref[level] = (enc, mp, n, int(st.num_matched))
        jobs.append({"level": level, "n": n, "d_masks": d_m, "d_encoded": torch.empty_like(enc), "d_mapping": torch.empty_like(mp),
                     "d_unique": torch.empty(max(1, n), dtype=torch.int32, device="cuda")})
        print(f"level {level}: {n} items, {st.num_matched} matched", flush=True)
    torch.cuda.synchronize()
    b = gpu_ctx.batch(jobs, cfg, use_graph=True)
    try:
        stats, ms = b.run()
    finally:
        b.close()
    torch.cuda.synchronize()
    print(f"batched call {ms:.2f} ms", flush=True)
    for j, s in zip(jobs, stats):
        enc, mp, n, matched = ref[j["level"]]
        assert torch.equal(j["d_encoded"][:n], enc[:n]), f"level {j['level']}: keys differ"
        assert torch.equal(j["d_mapping"][:n], mp[:n]), f"level {j['level']}: mapping differs"
        assert int(s.num_matched) == matched
    t.close()
