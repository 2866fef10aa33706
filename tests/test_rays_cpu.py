"""The CPU ray-casting oracle (oracle/oracle_rays.cpp) against brute force on small grids: the
first occupied voxel a ray crosses is the one whose box it enters first (smallest entry time over
all occupied voxels it crosses with positive length, inside [tmin, tmax])."""
import numpy as np
import pytest


def _brute(bits, R, rays):
    occ = np.unpackbits(bits, bitorder="little").reshape(R, R, R)  # [z, y, x]
    zz, yy, xx = np.nonzero(occ)
    lo = np.stack([xx, yy, zz], axis=1).astype(np.float64)
    out_v, out_t, out_gap = [], [], []
    for r in rays.astype(np.float64):
        o, tmin, d, tmax = r[:3], r[3], r[4:7], r[7]
        a0 = np.full(len(lo), tmin)
        a1 = np.full(len(lo), tmax)
        for a in range(3):
            if d[a] == 0:
                inside = (o[a] >= lo[:, a]) & (o[a] < lo[:, a] + 1)
                a1 = np.where(inside, a1, -np.inf)
                continue
            x0 = (lo[:, a] - o[a]) / d[a]
            x1 = (lo[:, a] + 1 - o[a]) / d[a]
            a0 = np.maximum(a0, np.minimum(x0, x1))
            a1 = np.minimum(a1, np.maximum(x0, x1))
        ok = a0 < a1
        if not ok.any():
            out_v.append((-1, -1, -1))
            out_t.append(0.0)
            out_gap.append(np.inf)
            continue
        idx = np.nonzero(ok)[0]
        order = idx[np.argsort(a0[idx], kind="stable")]
        best = order[0]
        out_v.append(tuple(int(c) for c in lo[best]))
        out_t.append(a0[best])
        out_gap.append(a0[order[1]] - a0[best] if len(order) > 1 else np.inf)
    return np.array(out_v, dtype=np.int32), np.array(out_t), np.array(out_gap)


def random_rays(rng, R, n, axis_parallel=False):
    """Rays through random points of the volume: from outside or inside, any direction (or along
    one axis), with random [tmin, tmax]."""
    target = rng.uniform(0, R, size=(n, 3))
    if axis_parallel:
        d = np.zeros((n, 3))
        d[np.arange(n), rng.integers(0, 3, n)] = rng.choice([-1.0, 1.0], n)
    else:
        d = rng.normal(size=(n, 3))
    o = target - d * rng.uniform(-0.2 * R, 1.5 * R, size=(n, 1))
    rays = np.zeros((n, 8), dtype=np.float32)
    rays[:, 0:3] = o
    rays[:, 3] = rng.uniform(0, 2, n)
    rays[:, 4:7] = d
    rays[:, 7] = rng.uniform(R, 4 * R, n)
    return rays


@pytest.mark.parametrize("R,density,axis_parallel", [(8, 0.1, False), (16, 0.02, False), (16, 0.05, True)])
def test_ray_march_equals_brute_force(oracle, R, density, axis_parallel):
    rng = np.random.default_rng(R + int(axis_parallel))
    bits = np.packbits(rng.random(R ** 3) < density, bitorder="little")
    rays = random_rays(rng, R, 400, axis_parallel)
    hits, t = oracle.ray_march(bits, R, rays)
    bv, bt, gap = _brute(bits, R, rays)
    clear = gap > 1e-9  # the best voxel is not tied with another
    assert clear.sum() > 300
    assert np.array_equal(hits[clear], bv[clear])
    hit = clear & (bv[:, 0] >= 0)
    assert hit.sum() > 50
    assert np.allclose(t[hit], bt[hit], rtol=0, atol=1e-9)


def test_ray_march_closed_forms(oracle):
    R = 8
    bits = np.zeros(R ** 3 // 8, dtype=np.uint8)
    i = 3 + R * (4 + R * 5)  # voxel (3, 4, 5)
    bits[i >> 3] |= 1 << (i & 7)
    rays = np.array([[0.5, 4.5, 5.5, 0, 1, 0, 0, 100],  # along +x: enters (3,4,5) at x = 3 -> t 2.5
                     [7.9, 4.5, 5.5, 0, -1, 0, 0, 100],  # along -x: enters at x = 4 -> t 3.9
                     [0.5, 4.5, 5.5, 0, 1, 0, 0, 2.0],  # tmax before the voxel: miss
                     [-5.0, 4.5, 5.5, 0, 1, 0, 0, 100],  # from outside: enters the volume at t 5, the voxel at t 8
                     [3.5, 4.5, 5.5, 1.0, 0, 0, 1, 100]], dtype=np.float32)  # starts inside the voxel, heading out: t = tmin = 1? no: z from 6.5
    hits, t = oracle.ray_march(bits, R, rays)
    assert hits[0].tolist() == [3, 4, 5] and t[0] == pytest.approx(2.5)
    assert hits[1].tolist() == [3, 4, 5] and t[1] == pytest.approx(3.9, abs=1e-6)
    assert hits[2].tolist() == [-1, -1, -1]
    assert hits[3].tolist() == [3, 4, 5] and t[3] == pytest.approx(8.0)
    assert hits[4].tolist() == [-1, -1, -1]  # at tmin = 1 the ray is at z = 6.5, past the voxel
