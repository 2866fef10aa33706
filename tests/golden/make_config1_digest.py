#!/usr/bin/env python3
"""Writes tests/golden/config1_digest.json: the CPU oracle's constructTransformDAG on the 1024^3
sphere+box+heightfield scene (kind 1, seed 0, maxTranslationLevel 5, all transforms on).

Recorded per level: the matcher input (items, SHA-256 of the LargeSubGrid masks), the oracle
matcher's encodedParents (SHA-256 of the uint64 little-endian keys, find_translation.cpp:864),
verified matches and unique items; and the .tdag stream's length and SHA-256. The input grid's
SHA-256 is recorded too, so a changed scene generator is caught before the keys are compared.

Runs in the build container only (about 10 minutes on 8 cores): the oracle is the checker and
the full-size run is too slow for a test. tests/test_gpu_configs.py compares the HIP driver and
matcher at full size with these digests.
"""
import ctypes
import hashlib
import json
import os
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
for p in (os.path.join(REPO, "i3d2025-voxel-transform-dag_amd"), os.path.join(REPO, "oracle"), os.path.join(REPO, "tests")):
    sys.path.insert(0, p)

import numpy as np  # noqa: E402

import oracle_ctypes as oc  # noqa: E402
from tdag_amd import pipeline as pl  # noqa: E402

KIND, R, SEED, MAXT = 1, 1024, 0, 5


def sha(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main():
    oc.build()
    L = oc.lib()
    t0 = time.time()
    bits = pl.scene_dense(KIND, R, SEED)
    h = ctypes.c_void_p()
    assert L.orc_svdag_build(bits.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), R, ctypes.byref(h)) == 0
    cfg = oc.config(True, True, True, MAXT)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "o.tdag")
        n = L.orc_construct_tdag(h, ctypes.byref(cfg), os.cpu_count() or 1, path.encode())
        assert n > 0
        data = open(path, "rb").read()
    levels = {}
    for level in range(MAXT, 1, -1):
        W = max(1, (1 << (3 * level)) // 64)
        words = L.orc_last_level_inputs(level, None, 0)
        masks = np.zeros(max(1, words), dtype=np.uint64)
        L.orc_last_level_inputs(level, masks.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), masks.size)
        masks = masks[:words].reshape(-1, W)
        items = masks.shape[0]
        enc = np.zeros(max(1, items), dtype=np.uint64)
        assert L.orc_last_level_encoded(level, enc.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), enc.size) == items
        enc = enc[:items]
        mapping, uniq = oc.postprocess(level, masks, enc)
        # verified matches: items whose minimum-key parent (the key's ptr field, a popcount rank,
        # transform_dag.h:375-397) is another item's rank
        pop = np.zeros(items, dtype=np.int64)
        for w in range(W):
            pop += np.array([bin(int(v)).count("1") for v in masks[:, w]], dtype=np.int64)
        order = np.argsort(-pop, kind="stable")
        rank = np.empty(items, dtype=np.int64)
        rank[order] = np.arange(items)
        ptr = (enc >> np.uint64(6 + 3 * (level + 1))).astype(np.int64)
        matched = int(np.sum((ptr != rank) & (ptr != 0xFFFFFFFF)))
        levels[str(level)] = {"items": items, "masks_sha256": sha(masks.astype("<u8")), "encoded_sha256": sha(enc.astype("<u8")),
                              "unique": int(len(uniq)), "matched": matched}
    L.orc_svdag_free(h)
    out = {"scene": {"kind": KIND, "resolution": R, "seed": SEED, "bits_sha256": sha(bits)},
           "config": {"symmetry": True, "axis_permutation": True, "translation": True, "max_translation_level": MAXT},
           "levels": levels, "tdag_bytes": len(data), "tdag_sha256": hashlib.sha256(data).hexdigest(),
           "generated_by": "tests/golden/make_config1_digest.py (the CPU oracle)", "oracle_seconds": round(time.time() - t0, 1)}
    with open(os.path.join(HERE, "config1_digest.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
