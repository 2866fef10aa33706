"""The device SVDAG producer (tdag_svdag_from_dense_device) against the host producer, which the
CPU tests pin node for node to the oracle's restatement of the reference's recursive
EditStructure(VoxelGrid) constructor + toDAG (structure.h:658-763): same leaves, same nodes, same
indices, at every level; edge cases are the empty grid, the full grid, the smallest resolution and
a noise grid whose leaves are nearly all distinct. Grids of 32^3 and more with a 4-byte aligned
pointer take the cube-tiled leaf kernel; 4^3..16^3 and unaligned pointers take the per-block one."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _same(host, dev):
    assert dev.resolution == host.resolution and dev.root_level == host.root_level
    assert dev.level_sizes == host.level_sizes
    assert np.array_equal(dev.leaves(), host.leaves())
    for level in range(3, host.root_level + 1):
        assert np.array_equal(dev.nodes(level), host.nodes(level)), level


def _device_bits(bits):
    import torch

    return torch.from_numpy(np.ascontiguousarray(bits)).cuda()


@pytest.mark.parametrize("kind,R,seed", [(0, 64, 1), (1, 128, 3), (2, 256, 4), (1, 512, 7)])
def test_device_svdag_equals_host(gpu_ctx, kind, R, seed):
    from tdag_amd import pipeline as pl

    print(f"device svdag kind={kind} R={R} seed={seed}", flush=True)
    bits = pl.scene_dense(kind, R, seed)
    _same(pl.Svdag(bits, R), pl.Svdag.from_device(gpu_ctx, _device_bits(bits), R))


@pytest.mark.parametrize("case", ["empty", "full", "R4", "R16", "noise"])
def test_device_svdag_edge_cases(gpu_ctx, case):
    from tdag_amd import pipeline as pl

    print(f"device svdag edge case {case}", flush=True)
    R = {"R4": 4, "R16": 16}.get(case, 64)
    if case == "empty":
        bits = np.zeros(R ** 3 // 8, dtype=np.uint8)
    elif case == "full":
        bits = np.full(R ** 3 // 8, 0xFF, dtype=np.uint8)
    else:
        bits = np.random.default_rng(11).integers(0, 256, R ** 3 // 8, dtype=np.uint8)
        if case == "R16":
            bits &= np.uint8(0x5A)  # sparser, so some blocks stay empty
    host = pl.Svdag(bits, R)
    dev = pl.Svdag.from_device(gpu_ctx, _device_bits(bits), R)
    _same(host, dev)
    if case == "full":
        assert dev.level_sizes[2:] == [1] * (dev.root_level - 1)


def test_device_svdag_feeds_the_driver(oracle, gpu_ctx):
    """The device-built SVDAG gives the same .tdag as the host-built one (full driver)."""
    import tdag_amd
    from tdag_amd import pipeline as pl

    bits = pl.scene_dense(1, 128, 5)
    cfg = tdag_amd.TransformDAGConfig(True, True, True, 3)
    a = pl.construct_transform_dag(gpu_ctx, pl.Svdag(bits, 128), cfg).tdag_bytes()
    b = pl.construct_transform_dag(gpu_ctx, pl.Svdag.from_device(gpu_ctx, _device_bits(bits), 128), cfg).tdag_bytes()
    assert a == b


def test_device_svdag_unaligned_grid(gpu_ctx):
    """A grid pointer that is not 4-byte aligned is read by the per-block kernel."""
    import torch
    from tdag_amd import pipeline as pl

    print("device svdag unaligned grid", flush=True)
    bits = pl.scene_dense(1, 128, 13)
    buf = torch.zeros(bits.size + 1, dtype=torch.uint8, device="cuda")
    buf[1:] = torch.from_numpy(bits).cuda()
    view = buf[1:]
    assert view.data_ptr() % 4 != 0
    _same(pl.Svdag(bits, 128), pl.Svdag.from_device(gpu_ctx, view, 128))


@pytest.mark.parametrize("level", [2, 3, 4, 5])
def test_device_flatten_equals_host(gpu_ctx, level):
    """createLargeSubGrid on the device equals the host flattening (pinned to the oracle's
    structure.get-based createLargeSubGrid by the CPU tests), word for word."""
    import torch
    from tdag_amd import pipeline as pl

    print(f"device flatten level {level}", flush=True)
    bits = pl.scene_dense(1, 256, 21)
    s = pl.Svdag(bits, 256)
    rng = np.random.default_rng(level)
    idx = rng.choice(s.level_sizes[level], size=min(700, s.level_sizes[level]), replace=False).astype(np.uint32)
    host = s.flatten(level, idx)
    d_idx = torch.from_numpy(idx.view(np.int32)).cuda()
    d_out = torch.full((len(idx) * host.shape[1],), -1, dtype=torch.int64, device="cuda")
    s.flatten_device(gpu_ctx, level, d_idx, d_out)
    assert np.array_equal(d_out.cpu().numpy().view(np.uint64).reshape(host.shape), host)


def test_device_flatten_rejects_bad_index(gpu_ctx):
    import torch
    import tdag_amd
    from tdag_amd import pipeline as pl

    print("device flatten bad index", flush=True)
    s = pl.Svdag(pl.scene_dense(0, 64, 2), 64)
    d_idx = torch.tensor([0, s.level_sizes[3]], dtype=torch.int32, device="cuda")
    d_out = torch.zeros(2 * 8, dtype=torch.int64, device="cuda")
    with pytest.raises(tdag_amd.TdagError, match="EINVAL"):
        s.flatten_device(gpu_ctx, 3, d_idx, d_out)


@pytest.mark.parametrize("kind,R,seed", [(0, 64, 2), (1, 128, 5), (2, 256, 6)])
def test_device_svdag_equals_oracle(oracle, gpu_ctx, kind, R, seed):
    """The device build against the oracle's restatement of the recursive constructor + toDAG
    directly (not through the host build): same leaves and nodes, index for index."""
    import svdag_arrays
    from tdag_amd import pipeline as pl

    print(f"device svdag vs oracle kind={kind} R={R} seed={seed}", flush=True)
    bits = pl.scene_dense(kind, R, seed)
    dev = pl.Svdag.from_device(gpu_ctx, _device_bits(bits), R)
    root, nodes, leaves = svdag_arrays.oracle_svdag_arrays(oracle, bits, R)
    assert dev.root_level == root
    assert np.array_equal(dev.leaves(), leaves)
    for level in range(3, root + 1):
        assert np.array_equal(dev.nodes(level), nodes[level]), level
