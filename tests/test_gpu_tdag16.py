""".tdag16 encoder and its device decoder (GPU, through the C-ABI).

The transform-aware DAGs here come from the full driver (GPU matcher, translation on). For every
encoder flag combination, the encoded stream decodes back to the dense grid it was built from:
voxel for voxel on the device (tdag_tdag16_decode_device) and, at sizes the CPU finishes in
seconds, in the oracle's restatement of the reference's lookup as well. At 1024^3 the device
decode is checked against the input grid alone (a size-independent round trip).
"""
import struct

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FLAGS = [(False, False), (True, False), (False, True), (True, True)]


@pytest.fixture(scope="module")
def pl():
    import tdag_amd
    from tdag_amd import pipeline

    tdag_amd.lib()
    return pipeline


def _device_words(e, gpu_ctx):
    d_out, ms = e.decode_device(gpu_ctx)
    return d_out.cpu().numpy().view(np.uint8), ms


@pytest.mark.parametrize("kind,R,seed,maxT", [(0, 16, 1, 3), (1, 16, 4, 2), (0, 32, 3, 4), (1, 32, 6, 3), (0, 128, 2, 3), (1, 128, 5, 4), (0, 256, 0, 5),
                                               (1, 256, 11, 4)])
def test_encode_decode_with_translation(oracle, pl, gpu_ctx, kind, R, seed, maxT):
    import tdag_amd

    print(f"tdag16 encode/decode kind={kind} R={R} seed={seed} maxT={maxT}", flush=True)
    bits = pl.scene_dense(kind, R, seed)
    cfg = tdag_amd.TransformDAGConfig(True, True, True, maxT)
    t = pl.construct_transform_dag(gpu_ctx, pl.Svdag(bits, R), cfg)
    expect = np.unpackbits(bits, bitorder="little").reshape(R, R, R).astype(bool)
    for tables, huffman in FLAGS:
        e = t.encode16(pointer_tables=tables, huffman=huffman)
        data = e.to_bytes()
        got, _ = _device_words(e, gpu_ctx)
        assert np.array_equal(got, bits), (tables, huffman)
        if R <= 128 or (tables and huffman):
            assert np.array_equal(oracle.tdag16_decode(data), expect), (tables, huffman)
        again, _ = _device_words(pl.TDag16.from_bytes(data), gpu_ctx)
        assert np.array_equal(again, bits)


def test_decode_1024_round_trip(pl, gpu_ctx):
    """The 1024^3 sphere+box+heightfield scene of the plain-matcher configuration (kind 1, seed 0)."""
    import tdag_amd

    print("tdag16 decode 1024^3 scene 1 seed 0", flush=True)
    R = 1024
    bits = pl.scene_dense(1, R, 0)
    cfg = tdag_amd.TransformDAGConfig(True, True, True, 5)
    t = pl.construct_transform_dag(gpu_ctx, pl.Svdag(bits, R), cfg)
    e = t.encode16()
    got, ms = _device_words(e, gpu_ctx)
    assert ms > 0
    assert np.array_equal(got, bits)


def _subgrid_count_offset(data: bytes) -> int:
    off = 8 + 2
    (levels,) = struct.unpack_from("<Q", data, off)
    off += 8
    for _ in range(levels):
        (n,) = struct.unpack_from("<Q", data, off)
        off += 8 + 2 * n
    return off


def test_missing_leaf_is_reported(oracle, pl, gpu_ctx):
    import tdag_amd

    bits = pl.scene_dense(0, 128, 2)
    cfg = tdag_amd.TransformDAGConfig(True, True, True, 4)
    data = pl.construct_transform_dag(gpu_ctx, pl.Svdag(bits, 128), cfg).encode16().to_bytes()
    at = _subgrid_count_offset(data)
    (n,) = struct.unpack_from("<Q", data, at)
    # Drop the last leaf: the stream still parses, but some pointer now names a leaf past the end.
    cut = data[:at] + struct.pack("<Q", n - 1) + data[at + 8: at + 8 + 8 * (n - 1)] + data[at + 8 + 8 * n:]
    with pytest.raises(ValueError):
        oracle.tdag16_decode(cut)
    e = pl.TDag16.from_bytes(cut)
    with pytest.raises(tdag_amd.TdagError, match="malformed"):
        e.decode_device(gpu_ctx)
    # The context and the library keep working after the reported error.
    good, _ = _device_words(pl.TDag16.from_bytes(data), gpu_ctx)
    assert np.array_equal(good, bits)
