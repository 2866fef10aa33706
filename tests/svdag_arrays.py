"""The oracle's SVDAG (its restatement of EditStructure(VoxelGrid) + toDAG) as plain arrays: the
EditStructure<void, uint32_t> a caller of constructTransformDAG holds (structure.h:159-171)."""
import ctypes

import numpy as np


def oracle_svdag_arrays(oracle, bits, R):
    """(root_level, {level: (count, 8) uint32 children}, uint64 leaf masks) of the oracle's SVDAG."""
    L = oracle.lib()
    h = ctypes.c_void_p()
    assert L.orc_svdag_build(np.ascontiguousarray(bits).ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), R, ctypes.byref(h)) == 0
    try:
        root = int(L.orc_svdag_root_level(h))
        nodes = {}
        for level in range(3, root + 1):
            n = int(L.orc_svdag_level_size(h, level))
            a = np.zeros((max(1, n), 8), dtype=np.uint32)
            L.orc_svdag_copy_level(h, level, a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)))
            nodes[level] = a[:n].copy()
        n = int(L.orc_svdag_level_size(h, 2))
        leaves = np.zeros(max(1, n), dtype=np.uint64)
        L.orc_svdag_copy_leaves(h, leaves.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)))
        return root, nodes, leaves[:n].copy()
    finally:
        L.orc_svdag_free(h)
