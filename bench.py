#!/usr/bin/env python3
"""Benchmark: verified matches/s of the exhaustive GPU matcher on one MI355X.

Workload (the largest configuration, the one the metric is quoted on; it fits one GPU): the
seeded 4096^3 city (box-shell buildings with window voids on 64-voxel lots, half of the
256-voxel blocks mirrored / axis-swapped copies, noise ground; seed 0), built into an SVDAG and
run through constructTransformDAG with maxTranslationLevel = 5. The per-level matcher inputs that
build produces (32^3, 16^3, 8^3 and 4^3 subgrids) are uploaded to HBM once; one step is one
exhaustive matcher launch per level over those inputs, in the driver's top-down order. Verified
matches = items whose minimum-key parent is another item, each accepted only after the bitwise
check child == T(parent). The same line reports the batched and learned forms on the same
inputs, the host-buffer (PCIe-inclusive) rate, and the 1024^3 sphere+box+heightfield scene.

Prints one JSON line. `roofline` reports the dominant search kernel against VALU issue (its
binding ceiling) from committed rocprofv3 counters of the same workload and tuning, with the
counter-measured HBM fraction and the reference's own throughput unit (8 directions x 6
permutations x R^3 voxels x 8 bytes per query, find_translation.cpp:879-885) labelled
reference-equivalent; `cpu_baseline` times the CPU oracle (a restatement of the reference's CPU
matcher) on this host.
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "i3d2025-voxel-transform-dag_amd"))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s
METRIC = "verified matches/sec and TSVDAG bytes vs exhaustive GPU matcher, 1 GPU"


def bytes_per_query(level: int) -> int:
    """Reference matcher unit: 8 directions x 6 permutations x R^3 voxels x 8-byte hashes."""
    return 384 * (1 << (3 * level))


SCENE_KINDS = {"scene": 1, "city": 2}


def workload_key(args) -> dict:
    return {"workload": args.workload, "resolution": args.resolution, "seed": args.seed, "max_level": args.max_level}


def default_tuning_path(args) -> str:
    return os.path.join(REPO, "configs", f"tuning_{args.workload}_{args.resolution}_seed{args.seed}_maxT{args.max_level}.json")


TUNING_FIELDS = ("variant", "first_batch", "growth", "filter_bits", "child_frames")


def tuning_key(t: dict) -> dict:
    """The fields of a level's tuning that decide which kernels run and how (timings dropped)."""
    return {k: int(t[k]) for k in TUNING_FIELDS if k in t}


def load_tuning(ctx, path: str, levels) -> dict:
    """Applies a tuning file written by --save-tuning to the context; returns it per level."""
    with open(path) as f:
        data = json.load(f)
    out = {}
    for level in levels:
        t = data["levels"].get(str(level))
        if t is None:
            continue
        ctx.set_variant(level, int(t["variant"]))
        ctx.set_batching(level, int(t["first_batch"]), int(t["growth"]))
        ctx.set_filter_bits(level, int(t["filter_bits"]))
        if level >= 3:
            ctx.set_child_frames(level, int(t["child_frames"]))
        out[level] = t
    return out


def build_workload(ctx, resolution: int, seed: int, max_level: int, kind: int = 1, extras: bool = True):
    """Scene -> SVDAG -> constructTransformDAG (the matcher inputs of every level are kept). With
    extras: the device SVDAG build, a second (warm) driver call and the .tdag16 encode/decode."""
    import tdag_amd
    from tdag_amd import pipeline as pl

    t0 = time.time()
    bits = pl.scene_dense(kind, resolution, seed)
    th = time.perf_counter()
    svdag = pl.Svdag(bits, resolution)
    host_build_ms = (time.perf_counter() - th) * 1e3
    cfg = tdag_amd.TransformDAGConfig(True, True, True, max_level)
    tc = time.perf_counter()
    tdag = pl.construct_transform_dag(ctx, svdag, cfg, keep_level_inputs=True)
    construct_ms = (time.perf_counter() - tc) * 1e3
    levels = {}
    for level in range(max_level, 1, -1):
        masks = tdag.level_masks(level)
        rec = tdag.record(level)
        levels[level] = {"masks": masks, "items": int(rec.items), "unique": int(rec.unique), "matched": int(rec.matched)}
    info = {
        "svdag": svdag,
        "tdag_bytes": len(tdag.tdag_bytes()),
        "svdag_level_sizes": svdag.level_sizes,
        "tdag_level_sizes": tdag.level_sizes,
    }
    if extras:
        info["svdag_build"] = device_svdag_build(ctx, bits, resolution, svdag, host_build_ms)
        tc = time.perf_counter()
        pl.construct_transform_dag(ctx, svdag, cfg).close()  # again, with the context's buffers in place
        # Wall time of one whole constructTransformDAG call (hierarchical pass, flattening and one
        # matcher call per level on the device, pointer rewrites on the host): the first call of
        # the process (it also allocates the context's buffers), then a second one.
        info["construct_ms"] = round(construct_ms, 1)
        info["construct_warm_ms"] = round((time.perf_counter() - tc) * 1e3, 1)
        info["tdag16"] = encode16_and_decode(ctx, tdag, bits, resolution)
        info["device_translation"] = device_translation_stage(ctx, svdag, cfg, tdag.tdag_bytes())
    tdag.close()
    info["setup_s"] = round(time.time() - t0, 2)
    return cfg, levels, info


def device_translation_stage(ctx, svdag, cfg, ref_bytes: bytes, replays: int = 5):
    """The translation stage with every step on the device: one driver call with
    TDAG_DEVICE_TRANSLATION (one count read-back per level), then the stage captured as a hipGraph
    and replayed. Both results are compared with the host-rewrite driver's bytes."""
    import numpy as np
    from tdag_amd import pipeline as pl

    ref = hashlib.sha256(ref_bytes).hexdigest()
    tc = time.perf_counter()
    t = pl.construct_transform_dag(ctx, svdag, cfg, device_translation=True)
    call_ms = (time.perf_counter() - tc) * 1e3
    same = hashlib.sha256(t.tdag_bytes()).hexdigest() == ref
    t.close()
    tc = time.perf_counter()
    g = pl.BuildGraph(ctx, svdag, cfg)
    create_ms = (time.perf_counter() - tc) * 1e3
    try:
        times = [g.run() for _ in range(replays)]
        r = g.result()
        same_graph = hashlib.sha256(r.tdag_bytes()).hexdigest() == ref
        r.close()
    finally:
        g.close()
    return {
        "driver_call_ms": round(call_ms, 1),
        "graph_create_ms": round(create_ms, 1),
        "graph_replay_ms": round(float(np.median(times)), 3),
        "graph_replay_ms_all": [round(x, 3) for x in times],
        "bytes_equal_host_rewrite_driver": bool(same and same_graph),
    }


def encode16_and_decode(ctx, tdag, bits, resolution: int):
    """constructTransformDAG16 of the result on the host (pointer tables + Huffman), then every voxel
    looked up in the encoded DAG on the device and compared with the input grid."""
    import numpy as np
    import torch

    plain = tdag.encode16(pointer_tables=False, huffman=False)
    bytes_plain = len(plain.to_bytes())
    plain.close()
    t0 = time.perf_counter()
    e = tdag.encode16(pointer_tables=True, huffman=True)
    encode_ms = (time.perf_counter() - t0) * 1e3
    data = e.to_bytes()
    d_out, _ = e.decode_device(ctx)  # first call copies the encoded DAG to the device
    ms = [e.decode_device(ctx, d_out)[1] for _ in range(3)]
    expect = torch.from_numpy(np.ascontiguousarray(bits).view(np.int32)).to(d_out.device)
    equal = bool(torch.equal(d_out, expect))
    del d_out, expect
    trace = trace16(ctx, e, resolution)
    e.close()
    avg = sum(ms) / len(ms)
    return {
        "trace": trace,
        "bytes": len(data),
        "bytes_without_tables_or_huffman": bytes_plain,
        "encode_ms_host": round(encode_ms, 1),
        "decode_ms": round(avg, 3),
        "decode_Gvoxels_per_s": round(resolution ** 3 / (avg * 1e-3) / 1e9, 2) if avg > 0 else None,
        "decoded_equal_to_input": equal,
        "note": "decode = the reference's per-voxel lookup (transform_dag_encoding.cpp:897-942) for all "
                "resolution^3 voxels on the device, kernel time from hipEvents",
    }


def camera_rays(resolution: int, width: int = 1024, height: int = 1024):
    """Primary rays of a pinhole camera (60 degree field of view) above a corner of the volume,
    looking at a point near the middle of its ground (y is up): n x 8 float32 (origin, tmin,
    direction, tmax)."""
    import numpy as np

    R = float(resolution)
    eye = np.array([-0.1 * R, 0.8 * R, -0.1 * R])
    look = np.array([0.5 * R, 0.05 * R, 0.5 * R])
    fwd = look - eye
    fwd /= np.linalg.norm(fwd)
    right = np.cross(fwd, [0.0, 1.0, 0.0])
    right /= np.linalg.norm(right)
    up = np.cross(right, fwd)
    th = np.tan(np.radians(30.0))
    u = (np.arange(width) + 0.5) / width * 2 - 1
    v = (np.arange(height) + 0.5) / height * 2 - 1
    uu, vv = np.meshgrid(u * th, v * th * height / width)
    d = fwd[None, None, :] + uu[..., None] * right[None, None, :] + vv[..., None] * up[None, None, :]
    d /= np.linalg.norm(d, axis=2, keepdims=True)
    rays = np.zeros((width * height, 8), dtype=np.float32)
    rays[:, 0:3] = eye
    rays[:, 4:7] = d.reshape(-1, 3)
    rays[:, 7] = 1e9
    return rays


def trace16(ctx, e, resolution: int):
    """First-hit ray casting through the encoded DAG on the device (the renderer's use of the
    format): 1024 x 1024 camera rays, kernel time from hipEvents, best of three."""
    import torch

    rays = camera_rays(resolution)
    d_rays = torch.from_numpy(rays).cuda()
    hits, ts, _ = e.trace_device(ctx, d_rays, len(rays))  # warm-up
    ms = min(e.trace_device(ctx, d_rays, len(rays), hits, ts)[2] for _ in range(3))
    h = hits.cpu().numpy()
    return {"rays": len(rays), "ms": round(ms, 3), "Mrays_per_s": round(len(rays) / (ms * 1e-3) / 1e6, 1) if ms > 0 else None,
            "hit_fraction": round(float((h[:, 0] >= 0).mean()), 4), "mean_steps": round(float(h[:, 3].mean()), 1),
            "camera": "1024 x 1024 pinhole, 60 degrees, from above a corner toward the middle of the ground"}


def device_svdag_build(ctx, bits, resolution: int, host_svdag, host_build_ms: float):
    """The step before the matcher: EditStructure(VoxelGrid) + toDAG built on the GPU from the dense
    grid in HBM (tdag_svdag_from_dense_device), timed against the library's host build of the same
    grid, and checked node for node against it. Skipped above 2048^3 (the dense grid alone is 1 GiB
    there)."""
    import numpy as np
    import torch
    from tdag_amd import pipeline as pl

    if resolution > 2048:
        return None
    d_bits = torch.from_numpy(bits).cuda()
    runs = [pl.Svdag.from_device(ctx, d_bits, resolution) for _ in range(3)]
    dev = runs[-1]
    equal = dev.level_sizes == host_svdag.level_sizes and np.array_equal(dev.leaves(), host_svdag.leaves()) and all(
        np.array_equal(dev.nodes(l), host_svdag.nodes(l)) for l in range(3, host_svdag.root_level + 1))
    if not equal:
        raise RuntimeError("device SVDAG build differs from the host build")
    del d_bits
    return {"device_ms": round(min(r.build_ms for r in runs), 2), "host_ms": round(host_build_ms, 1),
            "host_threads": min(16, os.cpu_count() or 1), "equal_to_host": True,
            "note": "dense grid resident in HBM; device_ms is the wall time of the call including the copy of the SVDAG to the host"}


def host_cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(levels, cfg, l5_queries: int, threads: int):
    """The CPU oracle (a restatement of the reference's findDuplicatesAmongFlattenedItems_cpp2)
    timed on this host: levels below 5 on their full inputs, level 5 with its full table and a
    uniform sample of its query ranks (every s-th rank), after one untimed warm-up call. Per
    level: table-build seconds + search seconds (level 5: search seconds x queries / sampled)."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle_ctypes as oc

    ocfg = oc.config(cfg.symmetry, cfg.axis_permutation, cfg.translation, cfg.max_translation_level)
    lo = min(levels)
    warm = levels[lo]["masks"]
    oc.match_timed(lo, warm, ocfg, nthreads=threads, query_stride=max(1, len(warm) // 64))  # warm-up (threads, pages)
    per_level = {}
    total_s = 0.0
    total_matches = 0
    for level, lv in sorted(levels.items()):
        masks = lv["masks"]
        n = len(masks)
        total_matches += lv["matched"]
        if n == 0:
            continue
        stride = 1 if level < 5 else max(1, n // max(1, l5_queries))
        setup, search, q = oc.match_timed(level, masks, ocfg, nthreads=threads, query_stride=stride)
        est = setup + search * n / max(1, q)
        total_s += est
        per_level[level] = {"items": n, "queries_searched": q, "table_build_s": round(setup, 3), "search_s": round(search, 3),
                            "level_s": round(est, 3), "measured_s": round(setup + search, 3)}
    value = total_matches / total_s if total_s > 0 else None
    return {
        "value": value,
        "unit": "matches/s",
        "cores": threads,
        "kind": "port",
        "cpu_model": host_cpu_model(),
        "host_nproc": os.cpu_count(),
        "levels": per_level,
        "sample": "oracle hash-search matcher (find_translation.cpp:616-921 restated) with std::thread workers over queries: levels "
                  + ", ".join(str(l) for l in sorted(per_level) if l < 5) + " on their full inputs; level 5 on its full table with every "
                  f"s-th query rank searched and its search time scaled by queries/sampled; one untimed warm-up call first; matches from the GPU run",
    }


def bench_batched(ctx, cfg, dev, order, args):
    """The batched forms on the same level inputs, each captured once as a hipGraph and replayed
    per step: (a) every level's parent search in one persistent kernel over a device work queue;
    (b) concurrent: each level's own tuned matcher on its own graph branch. Outputs are compared
    with the per-level launches' on the device (bit-exact)."""
    import torch

    def run_form(concurrent):
        jobs, outs = [], {}
        for level in order:
            d = dev[level]
            n = d["n"]
            outs[level] = {"enc": torch.empty(max(1, n), dtype=torch.int64, device="cuda"),
                           "map": torch.empty(max(1, n), dtype=torch.int64, device="cuda")}
            jobs.append({"level": level, "n": n, "d_masks": d["masks"], "d_encoded": outs[level]["enc"], "d_mapping": outs[level]["map"],
                         "d_unique": torch.empty(max(1, n), dtype=torch.int32, device="cuda")})
        torch.cuda.synchronize()
        b = ctx.batch(jobs, cfg, use_graph=True, concurrent=concurrent)
        try:
            b.autotune(1)
            grab = b.tuning()
            for _ in range(max(1, args.warmup)):
                b.run()
            t0 = time.perf_counter()
            matches = 0
            for _ in range(args.steps):
                stats, _ = b.run()
                matches += sum(int(s.num_matched) for s in stats)
            elapsed = time.perf_counter() - t0
        finally:
            b.close()
        return outs, grab, matches, elapsed

    def exact_vs_per_level(outs):
        exact = True
        for level in order:
            d = dev[level]
            n = d["n"]
            ref = torch.empty(max(1, n), dtype=torch.int64, device="cuda")
            ref_map = torch.empty(max(1, n), dtype=torch.int64, device="cuda")
            ctx.match_level_device(level, d["masks"], n, cfg, d_encoded=ref, d_mapping=ref_map, d_unique=d["unique"])
            torch.cuda.synchronize()
            exact = exact and bool(torch.equal(ref[:n], outs[level]["enc"][:n])) and bool(torch.equal(ref_map[:n], outs[level]["map"][:n]))
        return exact

    outs, grab, matches, elapsed = run_form(False)
    res = {
        "form": "one persistent kernel for all levels (device work queue, phases = popcount buckets of parents), hipGraph replay",
        "ms_per_step": round(elapsed * 1e3 / max(1, args.steps), 4),
        "matches_per_s": matches / elapsed if elapsed > 0 else None,
        "queue_tuning": {"tiles_per_pop": grab[0], "split_every": grab[1],
                         "note": "autotuned per run: tiles per queue pop, then one block in split_every takes the lighter levels' tiles first (0: none)"},
        "bit_exact_vs_per_level": exact_vs_per_level(outs),
    }
    outs, _, matches, elapsed = run_form(True)
    res["concurrent"] = {
        "form": "each level's tuned per-level matcher on its own branch of one hipGraph (levels overlap), replayed per step",
        "ms_per_step": round(elapsed * 1e3 / max(1, args.steps), 4),
        "matches_per_s": matches / elapsed if elapsed > 0 else None,
        "bit_exact_vs_per_level": exact_vs_per_level(outs),
    }
    return res


def ranks_of(masks):
    """Rank of every item in the stable popcount-descending order (find_translation.cpp:631-643)."""
    import numpy as np

    m = np.ascontiguousarray(masks, dtype=np.uint64)
    pop = np.bitwise_count(m.reshape(len(m), -1)).sum(axis=1).astype(np.int64)
    order = np.argsort(-pop, kind="stable")
    rank = np.empty(len(m), dtype=np.int64)
    rank[order] = np.arange(len(m))
    return rank


def bench_combined(ctx, cfg, dev, order, levels, info, args, plain_ms: float):
    """The combined form of the largest configuration: the batched matcher with the learned
    proposal at level 2 (the network pre-trained on the exact matcher's level-2 pairs of three
    seeded 1024^3 sphere+box+heightfield scenes, then fine-tuned on level-2 pairs of three seeded
    scenes of this workload's kind: level-local pairs), proposals verified bitwise, exhaustive
    fallback, the whole call one captured hipGraph. Reported against the plain per-level matcher
    on the same inputs, with recall, the .tdag size against the exhaustive build, and per level
    the time of the exhaustive search restricted to the items that have no match at all: the
    least any proposal scheme with an exhaustive fallback can take there."""
    import numpy as np
    import torch

    from tdag_amd import learned
    from tdag_amd import pipeline as pl

    t0 = time.time()
    train_res = min(args.resolution, 1024)
    seeds = [args.seed + 101, args.seed + 102, args.seed + 103]
    mlp = learned.MLP(ctx, seed=args.seed)
    d_q, d_p, npre = learned.training_pairs(ctx, cfg, SCENE_KINDS["scene"], train_res, seeds)
    tp = time.time()
    losses_pre = mlp.train(d_q, d_p, npre, steps=args.learned_steps, batch=512, lr=1e-3, tau=0.1, seed=args.seed + 7)
    pre_s = time.time() - tp
    d = dev[2]
    n2 = d["n"]
    ref_enc = torch.empty(max(1, n2), dtype=torch.int64, device="cuda")
    ctx.match_level_device(2, d["masks"], n2, cfg, d_encoded=ref_enc)
    base = learned.match_level_learned_device(ctx, None, d["masks"], n2, cfg, 64)
    st_pre = learned.match_level_learned_device(ctx, mlp, d["masks"], n2, cfg, 64)
    fine = {"pairs": 0, "steps": 0}
    losses_ft = None
    if args.workload != "scene":
        d_q, d_p, nft = learned.training_pairs(ctx, cfg, SCENE_KINDS[args.workload], train_res, seeds)
        tf = time.time()
        losses_ft = mlp.train(d_q, d_p, nft, steps=args.finetune_steps, batch=512, lr=3e-4, tau=0.1, seed=args.seed + 8)
        fine = {"pairs": int(nft), "steps": args.finetune_steps, "train_s": round(time.time() - tf, 2),
                "loss_first_last": [round(float(losses_ft[:20].mean()), 4), round(float(losses_ft[-20:].mean()), 4)],
                "scenes": {"workload": args.workload, "resolution": train_res, "seeds": seeds}}
    d_enc = torch.empty(max(1, n2), dtype=torch.int64, device="cuda")
    st = learned.match_level_learned_device(ctx, mlp, d["masks"], n2, cfg, 64, d_encoded=d_enc)  # standalone learned call: the reference
    torch.cuda.synchronize()
    matched2 = int(st.num_matched)
    # The same network with bucketed retrieval (candidates: earlier leaves of the query's exact
    # transform-invariant descriptor), measured beside it.
    mlp.retrieval = learned.RETRIEVAL_BUCKETED
    d_enc_b = torch.empty(max(1, n2), dtype=torch.int64, device="cuda")
    learned.match_level_learned_device(ctx, mlp, d["masks"], n2, cfg, 64, d_encoded=d_enc_b)  # warm-up (workspace)
    st_b = learned.match_level_learned_device(ctx, mlp, d["masks"], n2, cfg, 64, d_encoded=d_enc_b)
    torch.cuda.synchronize()
    mlp.retrieval = learned.RETRIEVAL_ALL

    def run_combined(concurrent: bool):
        jobs, outs = [], {}
        for level in order:
            dl = dev[level]
            outs[level] = {"enc": torch.empty(max(1, dl["n"]), dtype=torch.int64, device="cuda"),
                           "map": torch.empty(max(1, dl["n"]), dtype=torch.int64, device="cuda")}
            jobs.append({"level": level, "n": dl["n"], "d_masks": dl["masks"], "d_encoded": outs[level]["enc"], "d_mapping": outs[level]["map"],
                         "d_unique": torch.empty(max(1, dl["n"]), dtype=torch.int32, device="cuda")})
        torch.cuda.synchronize()
        b = ctx.batch(jobs, cfg, use_graph=True, concurrent=concurrent, mlp=mlp, k=64)
        try:
            if not concurrent:
                b.autotune(1)
            for _ in range(max(1, args.warmup)):
                b.run()
            torch.cuda.synchronize()
            tt = time.perf_counter()
            matches = 0
            for _ in range(args.steps):
                stats, _ = b.run()
                matches += sum(int(x.num_matched) for x in stats)
            elapsed = time.perf_counter() - tt
            hits, fb = b.learned_stats()
        finally:
            b.close()
        exact = True
        for level in order:
            dl = dev[level]
            n = dl["n"]
            if level == 2:
                exact = exact and bool(torch.equal(outs[2]["enc"][:n], d_enc[:n]))
                continue
            ref = torch.empty(max(1, n), dtype=torch.int64, device="cuda")
            ctx.match_level_device(level, dl["masks"], n, cfg, d_encoded=ref)
            torch.cuda.synchronize()
            exact = exact and bool(torch.equal(ref[:n], outs[level]["enc"][:n]))
        ms = elapsed * 1e3 / max(1, args.steps)
        return {"ms_per_step": round(ms, 4), "matches_per_s": matches / elapsed if elapsed > 0 else None,
                "vs_plain": round(plain_ms / ms, 4) if ms > 0 else None, "matches_per_step": matches // max(1, args.steps),
                "levels_3_5_bit_exact_vs_per_level_and_level_2_equal_to_the_learned_matcher": exact,
                "level2_proposal_hits": hits, "level2_fallback_queries": fb}

    persistent = run_combined(False)
    concurrent = run_combined(True)
    # Every accepted level-2 match is exact (verified on the device) and never below the exhaustive key.
    never_below = bool(np.all(d_enc[:n2].cpu().numpy().view(np.uint64) >= ref_enc[:n2].cpu().numpy().view(np.uint64)))
    tdag_l = pl.construct_transform_dag_learned(ctx, info["svdag"], cfg, mlp, 64)
    learned_bytes = len(tdag_l.tdag_bytes())
    tdag_l.close()
    # Least time of any proposal scheme with an exhaustive fallback: the search with only the
    # items that have no match in its table.
    bounds = {}
    for level in order:
        dl = dev[level]
        n = dl["n"]
        enc = torch.empty(max(1, n), dtype=torch.int64, device="cuda")
        _, full = ctx.match_level_device(level, dl["masks"], n, cfg, d_encoded=enc)
        keys = enc[:n].cpu().numpy().view(np.uint64)
        ptr = (keys >> np.uint64(6 + 3 * (level + 1))).astype(np.int64)
        rank = ranks_of(levels[level]["masks"])
        unmatched = (ptr == rank) | (ptr == 0xFFFFFFFF)
        flags = torch.from_numpy(unmatched.astype(np.int32)).cuda()
        sub = ctx.match_level_subset_device(level, dl["masks"], n, flags, cfg)
        bounds[level] = {"unmatched_items": int(unmatched.sum()), "items": n, "search_ms_all_children": round(full.ms_search, 4),
                         "search_ms_unmatched_children_only": round(sub.ms_search, 4)}
    mlp.close()
    return {
        "form": "batched matcher with the learned level-2 proposal (verified, exhaustive fallback, parent chase) in one captured hipGraph",
        "persistent": persistent,
        "concurrent": concurrent,
        "pretrain": {"pairs": int(npre), "steps": args.learned_steps, "train_s": round(pre_s, 2),
                     "loss_first_last": [round(float(losses_pre[:20].mean()), 4), round(float(losses_pre[-20:].mean()), 4)],
                     "scenes": {"workload": "scene", "resolution": train_res, "seeds": seeds}},
        "finetune_level_local": fine,
        "level2": {
            "items": n2, "matched": matched2,
            "recall_at_64": round(st.proposal_hits / matched2, 4) if matched2 else None,
            "recall_at_64_pretrained_only": round(st_pre.proposal_hits / matched2, 4) if matched2 else None,
            "recall_at_64_popcount_neighbours": round(base.proposal_hits / matched2, 4) if matched2 else None,
            "same_parent_as_exhaustive": round(int((d_enc[:n2] == ref_enc[:n2]).sum().item()) / n2, 4) if n2 else None,
            "never_below_exhaustive_key": never_below,
            "learned_matcher_ms": round(st.ms_total, 4),
            "ms_split": {"embed": round(st.ms_embed, 4), "propose": round(st.ms_propose, 4), "verify": round(st.ms_verify, 4), "fallback": round(st.ms_fallback, 4)},
            "bucketed_retrieval": {
                "recall_at_64": round(st_b.proposal_hits / matched2, 4) if matched2 else None,
                "same_parent_as_exhaustive": round(int((d_enc_b[:n2] == ref_enc[:n2]).sum().item()) / n2, 4) if n2 else None,
                "never_below_exhaustive_key": bool(np.all(d_enc_b[:n2].cpu().numpy().view(np.uint64) >= ref_enc[:n2].cpu().numpy().view(np.uint64))),
                "matched": int(st_b.num_matched),
                "learned_matcher_ms": round(st_b.ms_total, 4),
                "ms_split": {"embed": round(st_b.ms_embed, 4), "propose": round(st_b.ms_propose, 4), "verify": round(st_b.ms_verify, 4),
                             "fallback": round(st_b.ms_fallback, 4)},
            },
        },
        "tdag_bytes": learned_bytes,
        "tdag_bytes_exhaustive": info["tdag_bytes"],
        "tdag_size_delta": round(learned_bytes / info["tdag_bytes"] - 1.0, 5),
        "fallback_bound_per_level": bounds,
        "setup_s": round(time.time() - t0, 2),
    }


VALU_PEAK_GINST = 256 * 2.4  # 256 CUs x (4 SIMDs x one wave64 VALU instruction per 4 cycles) x 2.4 GHz, G wave-instructions/s


def counters_path(args, level: int) -> str:
    return os.path.join(REPO, "profiles", f"counters_search_L{level}_{args.workload}_{args.resolution}.json")


def load_counters(args, level: int, tuning: dict):
    """Per-call hardware counters of the level's search (rocprofv3 --pmc on this workload), if they
    were taken with the tuning in use; else None."""
    path = counters_path(args, level)
    if not os.path.exists(path):
        return None
    try:
        prof = json.load(open(path))
    except (OSError, ValueError):
        return None
    if prof.get("match") != workload_key(args) or level not in tuning or prof.get("tuning") != tuning_key(tuning[level]):
        return None
    return prof


def roofline(args, dom: int, dom_ms: float, items: int, tuning: dict, variant: int):
    """The dominant search kernel against the ceiling that binds it. The level-3..5 search is integer
    hashing and table probing: its bound is VALU issue (wave64 instructions/s), its HBM traffic is a
    few percent of peak. achieved = the call's VALU instructions (rocprofv3 SQ_INSTS_VALU, per call,
    same workload and tuning) / the call's search time from hipEvents in this run."""
    kernel = ("kSearch2" if dom == 2 else f"kSearch<{dom}> variant {variant}") + f" (level {dom})"
    ref_equiv = items * bytes_per_query(dom) / (dom_ms * 1e-3) / 1e9 if dom_ms > 0 else None
    out = {"bound": "valu", "kernel": kernel, "achieved": None, "peak": VALU_PEAK_GINST, "unit": "G wave64 VALU instructions/s", "frac": None,
           "traffic": None}
    prof = load_counters(args, dom, tuning)
    if prof and dom_ms > 0:
        pc = prof["per_call"]
        valu = pc["SQ_INSTS_VALU"] / (dom_ms * 1e-3) / 1e9
        traffic = prof["bytes_per_launch"]
        out.update({
            "achieved": round(valu, 2),
            "frac": round(valu / VALU_PEAK_GINST, 4),
            "traffic": traffic,
            "hbm": {"achieved": round(traffic / (dom_ms * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(traffic / (dom_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
            "wait_any_per_wave_cycle": round(pc["SQ_WAIT_ANY"] / pc["SQ_WAVE_CYCLES"], 4) if pc.get("SQ_WAVE_CYCLES") else None,
            "counters": os.path.relpath(counters_path(args, dom), REPO),
            "profiled_search_ms_per_call": prof.get("kernel_ms_per_call"),
        })
    out["reference_equivalent"] = {
        "achieved": round(ref_equiv, 1) if ref_equiv else None, "unit": "GB/s",
        "note": "items x 384 R^3 B (the reference matcher's own throughput unit, find_translation.cpp:879-885) / search time: bytes of "
                "nested hash grids the reference writes per query and this kernel never materialises; can exceed HBM peak"}
    out["note"] = ("peak = 256 CUs x 4 SIMDs x 1/4 wave64 VALU instruction per cycle x 2.4 GHz; traffic = HBM bytes of one search call from "
                   "rocprofv3 FETCH_SIZE/WRITE_SIZE, scaled by a device copy of known size; achieved/frac are null when the committed "
                   "counters were not taken on this workload with this tuning")
    return out


def time_plain(ctx, cfg, dev, order, steps: int, warmup: int):
    """One step = the exhaustive matcher launched once per level (device-resident inputs)."""
    import torch

    def step(record):
        total = 0
        for level in order:
            d = dev[level]
            _, st = ctx.match_level_device(level, d["masks"], d["n"], cfg, d_mapping=d["mapping"], d_unique=d["unique"])
            total += st.num_matched
            if record is not None:
                record[level]["search_ms"].append(st.ms_search)
                record[level]["total_ms"].append(st.ms_total)
        return total

    for _ in range(warmup):
        step(None)
    rec = {level: {"search_ms": [], "total_ms": []} for level in order}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    matches = 0
    for _ in range(steps):
        matches += step(rec)
    torch.cuda.synchronize()
    return matches, time.perf_counter() - t0, rec


def time_host_buffers(ctx, cfg, levels, order, reps: int):
    """The host-buffer drop-in (tdag_match_level: masks host->device, mapping and unique list
    device->host, pageable host memory) per level; whole-build wall time."""
    best = None
    matches = 0
    for _ in range(max(1, reps)):
        t0 = time.perf_counter()
        m = 0
        for level in order:
            _, _, _, st = ctx.match_level(level, levels[level]["masks"], cfg, want_encoded=False)
            m += st.num_matched
        dt = time.perf_counter() - t0
        if best is None or dt < best:
            best, matches = dt, m
    return {"ms_per_build": round(best * 1e3, 3), "matches_per_s": matches / best if best else None,
            "note": "host buffers in and out (tdag_match_level, the findDuplicatesAmongFlattenedItems_gpu drop-in), PCIe included, best of "
                    f"{max(1, reps)}"}


def per_level_summary(levels, order, rec):
    out = {}
    for level in order:
        s = rec[level]["search_ms"]
        avg = sum(s) / len(s) if s else 0.0
        out[level] = {
            "items": levels[level]["items"],
            "matched": levels[level]["matched"],
            "search_ms": round(avg, 4),
            "call_ms": round(sum(rec[level]["total_ms"]) / max(1, len(rec[level]["total_ms"])), 4),
            "reference_equivalent_GBs": round(levels[level]["items"] * bytes_per_query(level) / (avg * 1e-3) / 1e9, 1) if avg > 0 else None,
        }
    return out


def upload_levels(levels):
    import numpy as np
    import torch

    dev = {}
    for level, lv in levels.items():
        m = np.ascontiguousarray(lv["masks"]).view(np.int64)
        n = lv["items"]
        dev[level] = {
            "masks": torch.from_numpy(m).cuda(),
            "mapping": torch.empty(max(1, n), dtype=torch.int64, device="cuda"),
            "unique": torch.empty(max(1, n), dtype=torch.int32, device="cuda"),
            "n": n,
        }
    return dev


def apply_tuning(ctx, args, order_levels):
    """Committed tuning file of the workload (the autotuner's output), else autotune later."""
    tuning_path = args.tuning or default_tuning_path(args)
    if not args.no_autotune and not args.autotune and os.path.exists(tuning_path):
        tuning = load_tuning(ctx, tuning_path, order_levels)
        print(f"[bench] tuning from {tuning_path}: { {l: tuning_key(t) for l, t in tuning.items()} }", file=sys.stderr, flush=True)
        return tuning, f"autotuner output committed in {os.path.relpath(tuning_path, REPO)}"
    return {}, "library defaults"


def workload_name(args) -> str:
    kind = f"{args.resolution}^3 sphere+box+heightfield shells" if args.workload == "scene" else f"{args.resolution}^3 city"
    return (kind + f" seed {args.seed}, constructTransformDAG maxTranslationLevel={args.max_level}, exhaustive GPU matcher one launch per level "
            "(32^3,16^3,8^3,4^3)")


def secondary_config1(args):
    """The 1024^3 sphere+box+heightfield scene (the plain-matcher configuration with one launch per
    level), on its own context with its own committed tuning: per-level steps only."""
    import tdag_amd

    sub = argparse.Namespace(**vars(args))
    sub.workload, sub.resolution, sub.seed, sub.tuning = "scene", 1024, 0, ""
    ctx = tdag_amd.Context(0)
    try:
        tuning, source = apply_tuning(ctx, sub, range(sub.max_level, 1, -1))
        cfg, levels, info = build_workload(ctx, sub.resolution, sub.seed, sub.max_level, SCENE_KINDS[sub.workload], extras=False)
        dev = upload_levels(levels)
        order = sorted(levels, reverse=True)
        matches, elapsed, rec = time_plain(ctx, cfg, dev, order, args.steps, args.warmup)
        per_level = per_level_summary(levels, order, rec)
        dom = max(order, key=lambda l: per_level[l]["search_ms"])
        # Whole builds on this scene: a warm host-rewrite driver call against the device stage.
        from tdag_amd import pipeline as pl

        svdag = info["svdag"]
        pl.construct_transform_dag(ctx, svdag, cfg).close()
        tc = time.perf_counter()
        t = pl.construct_transform_dag(ctx, svdag, cfg)
        warm_ms = (time.perf_counter() - tc) * 1e3
        ref = t.tdag_bytes()
        t.close()
        stage = device_translation_stage(ctx, svdag, cfg, ref)
        return {
            "workload": workload_name(sub),
            "value": matches / elapsed,
            "unit": "matches/s",
            "ms_per_step": elapsed * 1e3 / max(1, args.steps),
            "matches_per_step": matches // max(1, args.steps),
            "tdag_bytes": info["tdag_bytes"],
            "construct_transform_dag_warm_ms": round(warm_ms, 1),
            "device_translation_stage": stage,
            "levels": per_level,
            "tuning_source": source,
            "roofline": roofline(sub, dom, per_level[dom]["search_ms"], levels[dom]["items"], tuning, ctx.get_variant(dom)),
        }
    finally:
        ctx.close()


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--resolution", type=int, default=4096)
    ap.add_argument("--workload", choices=sorted(SCENE_KINDS), default="city",
                    help="city: the seeded 4096^3 city (the default workload); scene: spheres + boxes + heightfield")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--max-level", type=int, default=5)
    ap.add_argument("--l5-cpu-queries", type=int, default=256, help="level-5 query ranks searched by the CPU baseline (uniform sample)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-secondary", action="store_true", help="skip the 1024^3 scene block")
    ap.add_argument("--no-host-buffers", action="store_true", help="skip the PCIe-inclusive host-buffer timing")
    ap.add_argument("--levels", default="", help="comma-separated subset of levels to time (default: all)")
    ap.add_argument("--no-autotune", action="store_true", help="use the library's default search variants instead of tuning them first")
    ap.add_argument("--autotune", action="store_true",
                    help="run the autotuner now even when a committed tuning file exists for the workload")
    ap.add_argument("--tuning", default="",
                    help="apply this tuning file (per level: variant, first batch, growth, filter bits, child frames) instead of autotuning; "
                         "default: configs/tuning_<workload>_<resolution>_seed<seed>_maxT<max level>.json when it exists")
    ap.add_argument("--save-tuning", default="", help="write the tuning in use to this file")
    ap.add_argument("--counters", action="store_true", help="also run each level once with per-stage probe counters (diagnostics, untimed)")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched forms")
    ap.add_argument("--no-learned", action="store_true", help="skip the combined form (batched + learned level-2 proposal + fallback)")
    ap.add_argument("--no-extras", action="store_true", help="skip the .tdag16 encode/decode and the whole-driver timings of the workload build")
    ap.add_argument("--learned-steps", type=int, default=1500, help="pre-training steps of the learned proposal")
    ap.add_argument("--finetune-steps", type=int, default=500, help="level-local fine-tuning steps of the learned proposal")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="only run a device-to-device copy of this many MiB three times (reference kernel for HBM byte counters)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on exactly one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus}).", file=sys.stderr)
        sys.exit(2)

    import torch

    import tdag_amd

    torch.cuda.init()
    if args.calibrate_copy_mib:
        n = args.calibrate_copy_mib * (1 << 20) // 8
        src = torch.arange(n, dtype=torch.int64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_per_launch": n * 8, "read_bytes": n * 8, "write_bytes": n * 8}))
        return
    ctx = tdag_amd.Context(0)
    # A committed tuning file is applied before the workload build, so every matcher call of the
    # run (the build's and the timed ones) uses the same kernels and schedule.
    tuning, tuning_source = apply_tuning(ctx, args, range(args.max_level, 1, -1))
    cfg, levels, info = build_workload(ctx, args.resolution, args.seed, args.max_level, SCENE_KINDS[args.workload], extras=not args.no_extras)
    print(f"[bench] workload ready: { {k: v for k, v in info.items() if k != 'svdag'} } items={ {l: v['items'] for l, v in levels.items()} }", file=sys.stderr,
          flush=True)
    dev = upload_levels(levels)
    order = sorted(levels, reverse=True)
    if args.levels:
        keep = {int(v) for v in args.levels.split(",")}
        order = [l for l in order if l in keep]

    if not tuning and not args.no_autotune:
        for level in order:
            if dev[level]["n"] > 0:
                best, first, ms, tuned = ctx.autotune(level, dev[level]["masks"], dev[level]["n"], cfg, repeats=1)
                tuning[level] = {"variant": best, "first_batch": first, "growth": ctx.get_batching(level)[1], "filter_bits": ctx.get_filter_bits(level),
                                 "child_frames": ctx.get_child_frames(level) if level >= 3 else 1, "call_ms_by_variant": ms, "tuned_call_ms": tuned}
        tuning_source = "autotuned in this run"
        print(f"[bench] autotuned search variants: {tuning}", file=sys.stderr, flush=True)
    if args.save_tuning and tuning:
        with open(args.save_tuning, "w") as f:
            json.dump({"workload": workload_key(args), "levels": {str(l): t for l, t in tuning.items()}}, f, indent=1)
    counters = {}
    if args.counters:
        ctx.set_counters(True)
        for level in order:
            d = dev[level]
            _, st = ctx.match_level_device(level, d["masks"], d["n"], cfg, d_mapping=d["mapping"], d_unique=d["unique"])
            counters[level] = st.counters()
        ctx.set_counters(False)
        print(f"[bench] probe counters: {counters}", file=sys.stderr, flush=True)

    matches, elapsed, rec = time_plain(ctx, cfg, dev, order, args.steps, args.warmup)
    print(f"[bench] timed {args.steps} steps: {elapsed * 1e3 / max(1, args.steps):.3f} ms/step", file=sys.stderr, flush=True)

    host_buffers = None if args.no_host_buffers else time_host_buffers(ctx, cfg, levels, order, 2)
    batched = None if args.no_batched else bench_batched(ctx, cfg, dev, order, args)
    learned_res = None if (args.no_learned or 2 not in dev) else bench_combined(ctx, cfg, dev, order, levels, info, args, elapsed * 1e3 / max(1, args.steps))

    per_step_matches = matches // max(1, args.steps)
    expected = sum(levels[l]["matched"] for l in order)
    if per_step_matches != expected:
        print(f"[bench] WARNING: matches per step {per_step_matches} != driver's {expected}", file=sys.stderr)
    per_level = per_level_summary(levels, order, rec)
    dom = max(order, key=lambda l: per_level[l]["search_ms"])

    out = {
        "metric": METRIC,
        "value": matches / elapsed,
        "unit": "matches/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / max(1, args.steps),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic (seeded scene generator, no external data)",
        "config": {
            "workload": workload_name(args),
            "tdag_bytes": info["tdag_bytes"],
            "matches_per_step": per_step_matches,
            **({"svdag_build": info["svdag_build"]} if info.get("svdag_build") else {}),
            **({"construct_transform_dag_ms": {"first_call": info["construct_ms"], "second_call": info["construct_warm_ms"]}} if "construct_ms" in info else {}),
            **({"tdag16": info["tdag16"]} if info.get("tdag16") else {}),
            **({"device_translation_stage": info["device_translation"]} if info.get("device_translation") else {}),
            "levels": per_level,
            **({"host_buffers": host_buffers} if host_buffers else {}),
            "search_variants": {l: ctx.get_variant(l) for l in order if l >= 3},
            "autotune": tuning,
            "tuning_source": tuning_source,
            **({"probe_counters": counters} if counters else {}),
            **({"batched": batched} if batched else {}),
            **({"combined": learned_res} if learned_res else {}),
            "parallelism": "single GPU",
        },
        "roofline": roofline(args, dom, per_level[dom]["search_ms"], levels[dom]["items"], tuning, ctx.get_variant(dom)),
    }
    if not args.no_secondary and not (args.workload == "scene" and args.resolution == 1024):
        out["config"]["config1_scene_1024"] = secondary_config1(args)
    if not args.no_cpu_baseline:
        threads = min(16, os.cpu_count() or 1)
        out["cpu_baseline"] = cpu_baseline({l: levels[l] for l in order}, cfg, args.l5_cpu_queries, threads)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
