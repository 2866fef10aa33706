"""Ray casting through the .tdag16 on the device (the renderer's use of the format,
renderer/src/dags/transform_dag/transform_dag.cu:5-41): seeded rays through encoded scenes at
128^3-512^3, built with translation, against the CPU oracle's ray march over the dense grid the
scene was built from. Hit voxels must agree except where a ray passes within float32 rounding of
a voxel edge (tolerance in tests/rays_check.py), and no ray may exhaust its step budget."""
import numpy as np
import pytest

import rays_check
from test_rays_cpu import random_rays

pytestmark = pytest.mark.gpu

FLAGS = [(True, True), (False, False)]


@pytest.mark.parametrize("kind,R,seed,maxT,flags", [(0, 128, 2, 3, 0), (1, 128, 5, 4, 1), (1, 256, 11, 4, 0), (2, 256, 5, 5, 0), (1, 512, 0, 5, 0)])
@pytest.mark.parametrize("axis_parallel", [False, True])
def test_trace_matches_ray_march(oracle, gpu_ctx, kind, R, seed, maxT, flags, axis_parallel):
    import torch

    import tdag_amd
    from tdag_amd import pipeline as pl

    print(f"trace kind={kind} R={R} seed={seed} maxT={maxT} flags={FLAGS[flags]} axis_parallel={axis_parallel}", flush=True)
    bits = pl.scene_dense(kind, R, seed)
    t = pl.construct_transform_dag(gpu_ctx, pl.Svdag(bits, R), tdag_amd.TransformDAGConfig(True, True, True, maxT))
    tables, huffman = FLAGS[flags]
    e = t.encode16(pointer_tables=tables, huffman=huffman)
    rng = np.random.default_rng(seed * 7 + R + int(axis_parallel))
    rays = random_rays(rng, R, 20000, axis_parallel)
    d_rays = torch.from_numpy(rays).cuda()
    hits, ts, ms = e.trace_device(gpu_ctx, d_rays, len(rays))
    hits = hits.cpu().numpy()
    ts = ts.cpu().numpy()
    n, same, tolerated = rays_check.compare(oracle, bits, R, rays, hits, ts)
    nhit = int((hits[:, 0] >= 0).sum())
    print(f"{n} rays, {nhit} hits, {same} identical, {tolerated} near-edge, {ms:.3f} ms", flush=True)
    assert nhit > n // 10
    assert tolerated <= n // 200
    e.close()
    t.close()
