"""The allocation-free matcher (tdag_match_level_device_noalloc) against tdag_match_level_device.

Both run the same kernels; the allocation-free form keeps its scratch in the library's fixed
per-GPU array, which holds the previous call's contents. Two calls in a row per level, the second
on a smaller input, must each give exactly the keys, mapping, unique list and counts of the
allocating entry point. A plan larger than the array is refused before anything is launched.
"""
import numpy as np
import pytest

import sets

pytestmark = pytest.mark.gpu


def _run_both(gpu_ctx, level, masks, cfg):
    import torch

    n = len(masks)
    d_masks = torch.from_numpy(np.ascontiguousarray(masks).view(np.int64).reshape(-1).copy()).cuda()
    outs = []
    for noalloc in (False, True):
        enc = torch.full((max(1, n),), -1, dtype=torch.int64, device="cuda")
        mp = torch.full((max(1, n),), -1, dtype=torch.int64, device="cuda")
        uq = torch.full((max(1, n),), -1, dtype=torch.int32, device="cuda")
        if noalloc:
            counts = torch.full((2,), -1, dtype=torch.int64, device="cuda")
            gpu_ctx.match_level_device_noalloc(level, d_masks, n, cfg, d_encoded=enc, d_mapping=mp, d_unique=uq, d_counts=counts)
            torch.cuda.synchronize()
            matched, nu = (int(v) for v in counts.cpu().tolist())
        else:
            nu, st = gpu_ctx.match_level_device(level, d_masks, n, cfg, d_encoded=enc, d_mapping=mp, d_unique=uq)
            torch.cuda.synchronize()
            matched = int(st.num_matched)
        outs.append((enc.cpu().numpy()[:n], mp.cpu().numpy()[:n], uq.cpu().numpy()[:nu], matched, nu))
    return outs


@pytest.mark.parametrize("level,n_first,n_second", [(2, 1200, 300), (3, 240, 60), (4, 60, 16), (5, 40, 10)])
def test_noalloc_equals_allocating_entry(oracle, gpu_ctx, level, n_first, n_second):
    import tdag_amd

    cfg = tdag_amd.TransformDAGConfig(True, True, True, 5)
    for call, (n, seed) in enumerate(((n_first, 10 + level), (n_second, 20 + level))):
        rng = np.random.default_rng(seed)
        masks = sets.planted_set(rng, level, n // 2, n - n // 2, oracle.transform, edge_cases=(level <= 3))
        print(f"noalloc level={level} call={call} n={len(masks)} seed={seed}", flush=True)
        ref, got = _run_both(gpu_ctx, level, masks, cfg)
        assert np.array_equal(got[0], ref[0]), f"call {call}: encoded keys differ"
        assert np.array_equal(got[1], ref[1]), f"call {call}: mapping differs"
        assert got[4] == ref[4] and np.array_equal(got[2], ref[2]), f"call {call}: unique list differs"
        assert got[3] == ref[3] > 0
        # and both equal the oracle
        ref_enc = oracle.match_encoded(level, masks, oracle.config(True, True, True))
        assert np.array_equal(got[0].view(np.uint64), ref_enc)


def test_noalloc_refuses_oversized_plans(gpu_ctx):
    import tdag_amd

    cfg = tdag_amd.TransformDAGConfig(True, True, True, 5)
    # 32^3 masks: 20 000 items need far more than the fixed scratch; refused before any launch,
    # so the (unused) device pointer is never read.
    with pytest.raises(tdag_amd.TdagError, match="EINVAL"):
        gpu_ctx.match_level_device_noalloc(5, 1 << 40, 20000, cfg)
