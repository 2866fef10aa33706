"""GPU tests of the learned proposal (level 2) through the C-ABI.

* The bf16 MFMA GEMM equals float64 products of the same bf16 values (all transposes, ragged
  sizes).
* Features: with pass-through weights the embedding is the normalised feature row, so the GPU's
  bits and descriptor bytes equal the CPU oracle's exactly.
* Loss and gradients equal the float64 CPU oracle (whose gradients are pinned by finite
  differences in test_oracle_mlp.py) run with the same bf16 rounding points; tolerance stated
  below.
* Training decreases the loss.
* The learned matcher accepts only exact matches (each re-checked with the CPU transform), never
  beats the exhaustive minimum, leaves exhaustively-unmatched leaves unmatched, and a trained
  network proposes more verified parents than popcount-neighbour proposals.
"""
import numpy as np
import pytest

import mlp_oracle as mo
import sets

pytestmark = pytest.mark.gpu

# Gradient tolerance against the bf16-emulating float64 oracle: the GPU accumulates in fp32 and a
# value sitting on a bf16 rounding boundary can round differently, so compare the whole vector.
GRAD_COSINE_MIN = 0.999
GRAD_REL_ERR_MAX = 2e-2


def _bf16_bits(a):
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def _from_bf16_bits(b):
    return (b.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


@pytest.mark.parametrize("ta,tb", [(False, False), (True, False), (False, True), (True, True)])
@pytest.mark.parametrize("shape", [(100, 70, 37), (200, 130, 1500)])
def test_gemm_bf16(gpu_ctx, ta, tb, shape):
    """Ragged sizes; the second shape (few output tiles, long K) takes the split-K path of the
    weight gradients, whose slices are summed in a fixed order."""
    import torch

    from tdag_amd import learned

    rng = np.random.default_rng(10 + 2 * ta + tb)
    M, N, K = shape
    A = rng.normal(size=(K, M) if ta else (M, K))
    B = rng.normal(size=(N, K) if tb else (K, N))
    Ab, Bb = _bf16_bits(A), _bf16_bits(B)
    ref = (_from_bf16_bits(Ab).T if ta else _from_bf16_bits(Ab)) @ (_from_bf16_bits(Bb).T if tb else _from_bf16_bits(Bb))
    dA = torch.from_numpy(Ab.view(np.int16).copy()).cuda()
    dB = torch.from_numpy(Bb.view(np.int16).copy()).cuda()
    dC = torch.zeros((M, N), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    learned.gemm_bf16(gpu_ctx, ta, tb, dA, A.shape[1], dB, B.shape[1], dC, N, M, N, K)
    C = dC.cpu().numpy().astype(np.float64)
    assert np.max(np.abs(C - ref)) < 1e-4 * (1 + np.max(np.abs(ref)))


def _topk_reference(E, k):
    """For each row q, the k rows c < q with the highest <E_q, E_c>, ties to the lower row, best
    first; 0xFFFFFFFF pads rows with fewer than k earlier rows. E holds small integers, so float64
    products are exact."""
    n = E.shape[0]
    out = np.full((n, k), 0xFFFFFFFF, dtype=np.uint32)
    Ef = E.astype(np.float64)
    for q0 in range(0, n, 512):
        q1 = min(n, q0 + 512)
        S = (Ef[q0:q1] @ Ef[:q1].T).astype(np.int64)
        c = np.arange(q1, dtype=np.int64)
        key = S * (1 << 22) + ((1 << 22) - 1 - c)[None, :]
        for qi in range(q1 - q0):
            q = q0 + qi
            if q == 0:
                continue
            kk = key[qi, :q]
            top = np.argsort(-kk, kind="stable")[:k]
            out[q, : len(top)] = top
    return out


@pytest.mark.parametrize("n,k", [(1, 64), (5, 64), (64, 64), (65, 7), (1000, 64), (9000, 64), (9000, 13)])
def test_topk_retrieval_exact(gpu_ctx, n, k):
    """The retrieval returns exactly the k best earlier rows under (higher score, then lower row),
    best first. Sparse {-1, 0, 1} embeddings make every fp32 score exact and ties frequent, so the
    tie-break and the merge of chunk lists (9000 rows span several chunks per query block) are
    checked bit for bit."""
    import torch

    from tdag_amd import learned

    rng = np.random.default_rng(n * 131 + k)
    E = (rng.random((n, 128)) < 0.08) * rng.choice([-1, 1], size=(n, 128))
    Eb = _bf16_bits(E.astype(np.float32))
    d_e = torch.from_numpy(Eb.view(np.int16).copy()).cuda()
    d_c = torch.zeros(n * k, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    learned.topk(gpu_ctx, d_e, n, k, d_c)
    got = d_c.cpu().numpy().view(np.uint32).reshape(n, k)
    ref = _topk_reference(E, k)
    bad = np.nonzero((got != ref).any(axis=1))[0]
    assert bad.size == 0, f"{bad.size} queries differ, first {bad[:5]}: got {got[bad[0]][:8]} want {ref[bad[0]][:8]}"


def _passthrough_params():
    parts = []
    for (i, o) in mo.DIMS:
        W = np.zeros((i, o))
        for d in range(min(i, o)):
            W[d, d] = 1.0
        parts += [W.reshape(-1), np.zeros(o)]
    return np.concatenate(parts).astype(np.float32)


def test_features_exact(gpu_ctx):
    import torch

    from tdag_amd import learned

    rng = np.random.default_rng(3)
    masks = np.concatenate([sets.random_masks(rng, 2, 300, d).reshape(-1) for d in (0.05, 0.3, 0.7)])
    masks = np.concatenate([masks, np.array([0xFFFFFFFFFFFFFFFF, 1, 1 << 63], dtype=np.uint64)])
    masks = masks[masks != 0]
    mlp = learned.MLP(gpu_ctx, seed=1)
    mlp.set_params(_passthrough_params())
    d_m = torch.from_numpy(masks.view(np.int64).copy()).cuda()
    d_e = torch.zeros((len(masks), 128), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    mlp.embed(d_m, len(masks), d_e)
    e = d_e.cpu().numpy()
    X = mo.features(masks)
    ref = X / np.linalg.norm(X, axis=1, keepdims=True)
    assert np.max(np.abs(e[:, :80] - ref)) < 1e-6
    assert np.all(e[:, 80:] == 0)
    mlp.close()


def test_gradients_match_oracle(gpu_ctx):
    import torch

    from tdag_amd import learned

    rng = np.random.default_rng(4)
    B = 32
    q = sets.random_masks(rng, 2, B, 0.3).reshape(-1)
    p = sets.random_masks(rng, 2, B, 0.3).reshape(-1)
    q[q == 0] = 1
    p[p == 0] = 2
    params = mo.init_params(7).astype(np.float32)
    params[rng.integers(0, params.size, 500)] += rng.normal(0, 0.05, 500).astype(np.float32)
    mlp = learned.MLP(gpu_ctx, seed=0)
    mlp.set_params(params)
    assert np.array_equal(mlp.get_params(), params)
    d_q = torch.from_numpy(q.view(np.int64).copy()).cuda()
    d_p = torch.from_numpy(p.view(np.int64).copy()).cuda()
    torch.cuda.synchronize()
    tau = 0.2
    loss, g = mlp.loss_grad(d_q, d_p, B, tau)
    X = mo.features(np.concatenate([q, p]))
    ref_loss, ref_g = mo.loss_and_grad(params.astype(np.float64), X, B, tau, emulate_bf16=True)
    assert abs(loss - ref_loss) < 1e-3 * max(1.0, abs(ref_loss))
    g = g.astype(np.float64)
    cos = g @ ref_g / np.linalg.norm(g) / np.linalg.norm(ref_g)
    rel = np.linalg.norm(g - ref_g) / np.linalg.norm(ref_g)
    print(f"loss gpu={loss:.6f} ref={ref_loss:.6f} grad cos={cos:.6f} rel={rel:.2e}", flush=True)
    assert cos > GRAD_COSINE_MIN and rel < GRAD_REL_ERR_MAX
    # and against the unrounded float64 oracle, more loosely
    _, g64 = mo.loss_and_grad(params.astype(np.float64), X, B, tau)
    assert g @ g64 / np.linalg.norm(g) / np.linalg.norm(g64) > 0.99
    mlp.close()


def _planted_pairs(oracle, seed, n_base, n_planted):
    from tdag_amd import learned

    rng = np.random.default_rng(seed)
    masks = sets.planted_set(rng, 2, n_base, n_planted, oracle.transform, edge_cases=False).reshape(-1)
    enc = oracle.match_encoded(2, masks, oracle.config(True, True, True))
    q, p = learned.pairs_from_encoded(masks, enc)
    return masks, enc, q, p


def test_training_decreases_loss(oracle, gpu_ctx):
    import torch

    from tdag_amd import learned

    _, _, q, p = _planted_pairs(oracle, 5, 600, 1200)
    assert len(q) > 200
    d_q = torch.from_numpy(q.view(np.int64).copy()).cuda()
    d_p = torch.from_numpy(p.view(np.int64).copy()).cuda()
    torch.cuda.synchronize()
    mlp = learned.MLP(gpu_ctx, seed=2)
    losses = mlp.train(d_q, d_p, len(q), steps=300, batch=128, lr=1e-3, tau=0.1, seed=3)
    print(f"loss first={losses[:20].mean():.4f} last={losses[-20:].mean():.4f}", flush=True)
    assert np.all(np.isfinite(losses))
    assert losses[-20:].mean() < 0.7 * losses[:20].mean()
    mlp.close()


def _check_learned_result(oracle, masks, enc_learned, enc_exact):
    """Every accepted match is exact; never below the exhaustive minimum; unmatched stay unmatched."""
    from tdag_amd import learned

    order = learned.rank_order(masks)
    rank_of = np.empty_like(order)
    rank_of[order] = np.arange(len(order))
    assert np.all(enc_learned >= enc_exact)
    low = 15  # 6 + 3 (L + 1) bits below the pointer at level 2
    for i in range(len(masks)):
        kl, ke = int(enc_learned[i]), int(enc_exact[i])
        ptr = kl >> low
        if ptr == rank_of[i] or ptr >= len(masks):
            assert kl == ke
            continue
        assert (ke >> low) != rank_of[i]  # the exhaustive matcher found a match too
        tx = (kl & 7) - 4
        ty = ((kl >> 3) & 7) - 4
        tz = ((kl >> 6) & 7) - 4
        sym = (kl >> 9) & 7
        perm = (kl >> 12) & 7
        parent = masks[order[ptr]]
        t = oracle.transform(2, np.array([parent], dtype=np.uint64), perm, sym, (tx, ty, tz))
        assert int(t[0]) == int(masks[i]), f"item {i}: accepted match is not exact"


def _run_learned(gpu_ctx, mlp, masks, k=64):
    import torch

    import tdag_amd
    from tdag_amd import learned

    n = len(masks)
    d_m = torch.from_numpy(np.ascontiguousarray(masks).view(np.int64).copy()).cuda()
    d_enc = torch.zeros(n, dtype=torch.int64, device="cuda")
    torch.cuda.synchronize()
    st = learned.match_level_learned_device(gpu_ctx, mlp, d_m, n, tdag_amd.TransformDAGConfig(True, True, True), k, d_encoded=d_enc)
    return d_enc.cpu().numpy().view(np.uint64).copy(), st


@pytest.mark.parametrize("use_mlp", [False, True])
def test_learned_matcher_exact(oracle, gpu_ctx, use_mlp):
    from tdag_amd import learned

    masks, enc, q, p = _planted_pairs(oracle, 6, 400, 800)
    mlp = learned.MLP(gpu_ctx, seed=4) if use_mlp else None
    enc_l, st = _run_learned(gpu_ctx, mlp, masks)
    print(f"use_mlp={use_mlp} stats={st.as_dict()}", flush=True)
    _check_learned_result(oracle, masks, enc_l, enc)
    assert st.proposal_hits + st.fallback_queries <= len(masks)
    if mlp:
        mlp.close()


def test_learned_matcher_deterministic(oracle, gpu_ctx):
    """Repeated calls give bit-identical keys: retrieval ties break by rank, verification keeps the
    smallest key of the lowest-ranked verified group, and the parent chase reads a snapshot of
    the keys in each pass, so no result depends on thread timing. Large enough that the retrieval
    splits query blocks into several chunks that share thresholds."""
    from tdag_amd import learned

    masks, enc, q, p = _planted_pairs(oracle, 8, 3000, 6000)
    mlp = learned.MLP(gpu_ctx, seed=9)
    try:
        runs = [_run_learned(gpu_ctx, mlp, masks)[0] for _ in range(3)]
    finally:
        mlp.close()
    assert np.array_equal(runs[0], runs[1]) and np.array_equal(runs[0], runs[2])
    _check_learned_result(oracle, masks, runs[0], enc)


def _fixed_popcount_set(oracle, seed, n_base, n_copies, pop=10):
    """Leaves that all have `pop` voxels: random bases plus copies under a random mirror/axis
    permutation and a translation that clips nothing. Popcount buckets are then useless for
    proposing parents (every leaf is in the same bucket), while a copy and its base share their
    transform-invariant descriptor."""
    rng = np.random.default_rng(seed)
    bases = []
    for _ in range(n_base):
        bits = np.zeros(64, dtype=bool)
        bits[rng.choice(64, pop, replace=False)] = True
        bases.append(sets.bits_to_mask(2, bits)[0])
    bases = np.array(bases, dtype=np.uint64)
    copies = []
    while len(copies) < n_copies:
        src = bases[rng.integers(n_base)]
        t = tuple(int(v) for v in rng.integers(-1, 2, 3))
        c = oracle.transform(2, np.array([src], dtype=np.uint64), int(rng.integers(6)), int(rng.integers(8)), t)
        if bin(int(c[0])).count("1") == pop:  # nothing clipped
            copies.append(c[0])
    masks = np.concatenate([bases, np.array(copies, dtype=np.uint64)])
    return masks[rng.permutation(len(masks))]


def test_trained_mlp_beats_popcount_neighbours(oracle, gpu_ctx):
    """Trained on one seeded set, evaluated on another: more verified proposals than the k leaves
    ranked next to each query (popcount-bucket proposals without a network)."""
    import torch

    from tdag_amd import learned

    train = _fixed_popcount_set(oracle, 7, 3000, 3000)
    enc_tr = oracle.match_encoded(2, train, oracle.config(True, True, True))
    q, p = learned.pairs_from_encoded(train, enc_tr)
    d_q = torch.from_numpy(q.view(np.int64).copy()).cuda()
    d_p = torch.from_numpy(p.view(np.int64).copy()).cuda()
    torch.cuda.synchronize()
    mlp = learned.MLP(gpu_ctx, seed=5)
    mlp.train(d_q, d_p, len(q), steps=600, batch=256, lr=1e-3, tau=0.1, seed=6)
    masks = _fixed_popcount_set(oracle, 8, 3000, 3000)
    enc = oracle.match_encoded(2, masks, oracle.config(True, True, True))
    enc_mlp, st_mlp = _run_learned(gpu_ctx, mlp, masks)
    enc_base, st_base = _run_learned(gpu_ctx, None, masks)
    print(f"mlp hits={st_mlp.proposal_hits} baseline hits={st_base.proposal_hits} matched={st_mlp.num_matched} items={len(masks)}", flush=True)
    _check_learned_result(oracle, masks, enc_mlp, enc)
    _check_learned_result(oracle, masks, enc_base, enc)
    assert st_mlp.num_matched == st_base.num_matched  # the fallback completes both
    assert st_mlp.proposal_hits > 2 * st_base.proposal_hits
    mlp.close()


def test_learned_driver_decodes_exactly(oracle, gpu_ctx):
    """constructTransformDAG with the learned proposal at level 2: the .tdag expands back to the
    scene's voxels exactly (every match exact) and stays close in size to the exhaustive one."""
    import torch

    import tdag_amd
    import tdag_decode
    from tdag_amd import learned
    from tdag_amd import pipeline as pl

    R, maxT = 128, 3
    cfg = tdag_amd.TransformDAGConfig(True, True, True, maxT)
    # training pairs: the exact matcher on another seeded scene's level-2 inputs
    tr = pl.construct_transform_dag(gpu_ctx, pl.Svdag(pl.scene_dense(1, R, 21), R), cfg, keep_level_inputs=True)
    m2 = tr.level_masks(2).reshape(-1)
    enc = oracle.match_encoded(2, m2, oracle.config(True, True, True, maxT))
    q, p = learned.pairs_from_encoded(m2, enc)
    d_q = torch.from_numpy(q.view(np.int64).copy()).cuda()
    d_p = torch.from_numpy(p.view(np.int64).copy()).cuda()
    torch.cuda.synchronize()
    mlp = learned.MLP(gpu_ctx, seed=11)
    mlp.train(d_q, d_p, len(q), steps=300, batch=256, lr=1e-3, tau=0.1, seed=12)
    bits = pl.scene_dense(1, R, 22)
    svdag = pl.Svdag(bits, R)
    exact = pl.construct_transform_dag(gpu_ctx, svdag, cfg).tdag_bytes()
    t = pl.construct_transform_dag_learned(gpu_ctx, svdag, cfg, mlp, 64)
    mine = t.tdag_bytes()
    rec = t.record(2)
    print(f"learned tdag {len(mine)} B vs exhaustive {len(exact)} B; level 2: items={rec.items} matched={rec.matched} "
          f"proposal_hits={rec.proposal_hits} fallback={rec.fallback_queries}", flush=True)
    assert rec.learned == 1
    d = tdag_decode.parse(mine)
    assert np.array_equal(tdag_decode.expand(d), tdag_decode.dense_to_cube(bits, R))
    assert len(mine) <= 1.05 * len(exact)
    mlp.close()


def _bucket_masks(rng, n):
    """4^3 masks with large descriptor buckets: single voxels and voxel pairs at random places
    (a few descriptors shared by hundreds of masks) mixed with random masks."""
    out = np.zeros(n, dtype=np.uint64)
    for i in range(n):
        u = rng.random()
        if u < 0.4:
            out[i] = np.uint64(1) << np.uint64(rng.integers(64))
        elif u < 0.7:
            a, b = rng.choice(64, 2, replace=False)
            out[i] = (np.uint64(1) << np.uint64(a)) | (np.uint64(1) << np.uint64(b))
        else:
            bits = rng.random(64) < 0.3
            out[i] = np.uint64(sum(1 << j for j in range(64) if bits[j]) or 1)
    return out


def _topk_bucketed_reference(masks, E, k):
    """For each row q, the k rows c < q with q's descriptor (the oracle's) and the highest
    <E_q, E_c>, ties to the lower row, best first. E holds small integers: exact products."""
    n = len(masks)
    desc = [mo.descriptor(int(m)).tobytes() for m in masks]
    out = np.full((n, k), 0xFFFFFFFF, dtype=np.uint32)
    Ef = E.astype(np.int64)
    members = {}
    for q in range(n):
        prev = members.setdefault(desc[q], [])
        if prev:
            c = np.array(prev, dtype=np.int64)
            s = Ef[c] @ Ef[q]
            order = np.lexsort((c, -s))[:k]  # higher score first, then lower row
            out[q, : len(order)] = c[order]
        prev.append(q)
    return out


@pytest.mark.parametrize("n,k", [(1, 64), (70, 64), (3000, 64), (3000, 5)])
def test_topk_bucketed_exact(gpu_ctx, n, k):
    """Bucketed retrieval returns exactly the k best earlier rows of the query's descriptor bucket
    (buckets of hundreds of rows span several 64-candidate chunks per query)."""
    import torch

    from tdag_amd import learned

    rng = np.random.default_rng(n * 7 + k)
    masks = _bucket_masks(rng, n)
    E = (rng.random((n, 128)) < 0.08) * rng.choice([-1, 1], size=(n, 128))
    Eb = _bf16_bits(E.astype(np.float32))
    d_m = torch.from_numpy(masks.view(np.int64).copy()).cuda()
    d_e = torch.from_numpy(Eb.view(np.int16).copy()).cuda()
    d_c = torch.zeros(n * k, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    learned.topk_bucketed(gpu_ctx, d_m, d_e, n, k, d_c)
    got = d_c.cpu().numpy().view(np.uint32).reshape(n, k)
    ref = _topk_bucketed_reference(masks, E, k)
    bad = np.nonzero((got != ref).any(axis=1))[0]
    assert bad.size == 0, f"{bad.size} queries differ, first {bad[:5]}: got {got[bad[0]][:8]} want {ref[bad[0]][:8]}"


def test_learned_matcher_bucketed_exact(oracle, gpu_ctx):
    """The learned matcher with bucketed retrieval: every accepted match exact, never below the
    exhaustive key, unmatched leaves unmatched; on leaves whose matches lose no voxel (one
    popcount, copies translated without clipping) the buckets hold the parents, so the
    proposals resolve most queries; and the results are deterministic."""
    from tdag_amd import learned

    masks, enc, q, p = _planted_pairs(oracle, 6, 400, 800)
    mlp = learned.MLP(gpu_ctx, seed=4)
    try:
        mlp.retrieval = learned.RETRIEVAL_BUCKETED
        assert mlp.retrieval == learned.RETRIEVAL_BUCKETED
        enc_l, st = _run_learned(gpu_ctx, mlp, masks)
        print(f"bucketed planted stats={st.as_dict()}", flush=True)
        _check_learned_result(oracle, masks, enc_l, enc)
        same = _fixed_popcount_set(oracle, 8, 1500, 1500)
        enc_s = oracle.match_encoded(2, same, oracle.config(True, True, True))
        runs = [_run_learned(gpu_ctx, mlp, same) for _ in range(2)]
        print(f"bucketed fixed-popcount stats={runs[0][1].as_dict()}", flush=True)
        _check_learned_result(oracle, same, runs[0][0], enc_s)
        assert np.array_equal(runs[0][0], runs[1][0])
        assert runs[0][1].proposal_hits >= 0.9 * runs[0][1].num_matched
    finally:
        mlp.close()
