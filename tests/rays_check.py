"""Comparison of device ray casting (float32) with the CPU oracle's ray march (float64)."""
import numpy as np

# Float32 ray arithmetic at distances up to a few thousand voxels is good to about 1e-3 voxel;
# a ray that passes closer than that to a voxel edge may legitimately pick a neighbouring voxel.
TOL_VOXELS = 2e-3


def crossing(rays, idx, voxels):
    """Length of each ray's crossing of the given voxel boxes inside [tmin, tmax] (float64), and
    the entry distance."""
    r = rays[idx].astype(np.float64)
    o, d = r[:, 0:3], r[:, 4:7]
    lo = voxels.astype(np.float64)
    a0, a1 = r[:, 3].copy(), r[:, 7].copy()
    for a in range(3):
        with np.errstate(divide="ignore", invalid="ignore"):
            x0 = (lo[:, a] - o[:, a]) / d[:, a]
            x1 = (lo[:, a] + 1 - o[:, a]) / d[:, a]
        par = d[:, a] == 0
        inside = (o[:, a] >= lo[:, a]) & (o[:, a] < lo[:, a] + 1)
        a0 = np.where(par, a0, np.maximum(a0, np.minimum(x0, x1)))
        a1 = np.where(par, np.where(inside, a1, -np.inf), np.minimum(a1, np.maximum(x0, x1)))
    return a1 - a0, a0


def compare(oracle, bits, R, rays, hits, t):
    """Counts of (rays, exact voxel agreements, tolerated near-edge disagreements); raises on a
    disagreement that no float32 rounding explains."""
    ch, ct = oracle.ray_march(bits, R, rays)
    g = hits[:, :3]
    assert np.all(hits[:, 3] >= 0), "a ray exhausted its step budget"
    same = np.all(g == ch, axis=1)
    both = same & (ch[:, 0] >= 0)
    tol = TOL_VOXELS * np.maximum(1.0, np.abs(ct))
    assert np.all(np.abs(t[both] - ct[both]) <= tol[both]), "entry distances differ"
    bad = np.nonzero(~same)[0]
    for i in bad:
        gh, chh = g[i], ch[i]
        if gh[0] >= 0:  # the GPU's voxel must be occupied and crossed (at least grazed)
            k = int(gh[0]) + R * (int(gh[1]) + R * int(gh[2]))
            assert (bits[k >> 3] >> (k & 7)) & 1, f"ray {i}: the device hit an empty voxel {gh}"
        if gh[0] >= 0 and chh[0] >= 0:
            assert abs(float(t[i]) - float(ct[i])) <= tol[i], f"ray {i}: {gh} at {t[i]} vs {chh} at {ct[i]}"
        else:  # one side missed: the other's voxel is only grazed
            v = (gh if gh[0] >= 0 else chh).reshape(1, 3)
            length, _ = crossing(rays, np.array([i]), v)
            assert length[0] <= TOL_VOXELS * max(1.0, float(np.abs(ct[i]) + np.abs(t[i]))), f"ray {i}: {gh} vs {chh}"
    return len(rays), int(same.sum()), len(bad)
