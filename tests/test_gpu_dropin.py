"""The C++ drop-ins (include/tdag.hpp) run end to end: a C++ program calls
constructTransformDAG and findDuplicatesAmongFlattenedItems_gpu the way the reference calls them
(voxelizer/src/main.cpp:146, transform_dag.cpp:201), on stand-in types with the reference's
layouts, linked against libtdag.so. The input EditStructure is the oracle's SVDAG (not a dense
grid); the program's EditStructure<void, TransformPointer> must be the .tdag the oracle's driver
writes, byte for byte, and its per-level mapping the oracle matcher's."""
import os
import struct
import subprocess

import numpy as np
import pytest

import dropin_build
import sets
import svdag_arrays

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    return dropin_build.build(str(tmp_path_factory.mktemp("dropin")))


def _run(args):
    proc = subprocess.run(args, capture_output=True, text=True, timeout=300)
    assert proc.returncode == 0, proc.stderr


@pytest.mark.parametrize("kind,R,seed,maxT", [(1, 128, 3, 4), (2, 256, 5, 5)])
def test_construct_transform_dag_dropin(oracle, gpu_ctx, exe, tmp_path, kind, R, seed, maxT):
    from tdag_amd import pipeline as pl

    import test_gpu_pipeline

    print(f"drop-in construct kind={kind} R={R} seed={seed} maxT={maxT}", flush=True)
    bits = pl.scene_dense(kind, R, seed)
    root, nodes, leaves = svdag_arrays.oracle_svdag_arrays(oracle, bits, R)
    src, dst = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(src, "wb") as f:
        f.write(struct.pack("<IIQ", R, root, len(leaves)))
        f.write(leaves.astype("<u8").tobytes())
        for level in range(3, root + 1):
            f.write(struct.pack("<Q", len(nodes[level])))
            f.write(nodes[level].astype("<u4").tobytes())
    _run([exe, "construct", str(src), str(dst), "1", "1", "1", str(maxT)])
    b = dst.read_bytes()
    out_root, out_type = struct.unpack_from("<Ii", b, 0)
    off = 8
    levels = {}
    for level in range(2, out_root + 1):
        (n,) = struct.unpack_from("<Q", b, off)
        off += 8
        levels[level] = np.frombuffer(b, dtype="<u8", count=n, offset=off)
        off += 8 * n
    assert off == len(b) and out_root == root and out_type == 0
    # The .tdag stream of that EditStructure (structure.h:493-501 + the config, voxelizer/src/main.cpp:147-150).
    tdag = struct.pack("<Q", root + 1) + struct.pack("<QQQ", 0, 0, 0)
    for level in range(3, root + 1):
        tdag += struct.pack("<Q", len(levels[level]) // 8) + levels[level].tobytes()
    tdag += struct.pack("<Q", len(levels[2])) + levels[2].tobytes()
    tdag += struct.pack("<IIi", R, root, 0) + struct.pack("<BBBBI", 1, 1, 1, 0, maxT)
    assert tdag == test_gpu_pipeline._oracle_bytes(oracle, bits, R, oracle.config(True, True, True, maxT))


@pytest.mark.parametrize("level,n", [(2, 600), (3, 200), (5, 24)])
def test_find_duplicates_dropin(oracle, gpu_ctx, exe, tmp_path, level, n):
    print(f"drop-in findDuplicates level={level} n={n}", flush=True)
    rng = np.random.default_rng(70 + level)
    masks = sets.planted_set(rng, level, n // 2, n - n // 2, oracle.transform)
    src, dst = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(src, "wb") as f:
        f.write(struct.pack("<Q", len(masks)))
        f.write(np.ascontiguousarray(masks, dtype="<u8").tobytes())
    _run([exe, "match", str(level), str(src), str(dst), "1", "1", "1", "5"])
    b = dst.read_bytes()
    m = len(masks)
    mapping = np.frombuffer(b, dtype="<u8", count=m)
    (nu,) = struct.unpack_from("<Q", b, 8 * m)
    uniq = np.frombuffer(b, dtype="<u4", count=nu, offset=8 * m + 8)
    ref_map, ref_uniq = oracle.postprocess(level, masks, oracle.match_encoded(level, masks))
    assert np.array_equal(mapping, ref_map)
    assert np.array_equal(uniq, ref_uniq.astype(np.uint32) + 1000)  # the items are payloads 1000 + index
