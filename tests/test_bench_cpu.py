"""CPU checks of the benchmark's bookkeeping (no GPU):

* every committed tuning file names a workload and, per level, the fields the library applies;
* each committed HBM-traffic profile was taken with the committed tuning of its workload's
  dominant level, so bench.py reports `traffic` for the default workload;
* load_tuning applies a tuning file to a context through the library's setters;
* bench.py refuses more than one GPU.
"""
import glob
import json
import os
import subprocess
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, REPO)

import bench  # noqa: E402


def _tuning_files():
    return sorted(glob.glob(os.path.join(REPO, "configs", "tuning_*.json")))


def test_tuning_files_are_complete():
    files = _tuning_files()
    assert files, "no committed tuning files"
    for path in files:
        data = json.load(open(path))
        wk = data["workload"]
        assert set(wk) == {"workload", "resolution", "seed", "max_level"}
        assert os.path.basename(path) == os.path.basename(bench.default_tuning_path(type("A", (), wk)))
        for level in range(2, wk["max_level"] + 1):
            t = data["levels"][str(level)]
            key = bench.tuning_key(t)
            assert set(key) == set(bench.TUNING_FIELDS), (path, level)
            assert key["growth"] >= 2 and key["first_batch"] >= 1
            assert key["child_frames"] in ((1,) if level == 2 else (1, 6))
            assert key["filter_bits"] in (4, 8, 16, 32)


def test_traffic_profiles_match_committed_tuning():
    profiles = sorted(glob.glob(os.path.join(REPO, "profiles", "traffic_search_L*.json")))
    assert profiles
    for path in profiles:
        prof = json.load(open(path))
        level = int(os.path.basename(path)[len("traffic_search_L"):-len(".json")])
        wk = prof["match"]
        tuning = json.load(open(bench.default_tuning_path(type("A", (), wk))))
        assert prof["tuning"] == bench.tuning_key(tuning["levels"][str(level)]), path
        assert prof["bytes_per_launch"] > 0
        assert prof["fetch_scale"] > 0 and prof["write_scale"] > 0


def test_counter_profiles_match_committed_tuning_and_give_a_bounded_roofline():
    """Each committed counter profile was taken with the committed tuning of its workload's level,
    and the roofline the bench derives from it, at the profiled kernel time, stays within its
    peaks (VALU issue and HBM bandwidth)."""
    import argparse

    profiles = sorted(glob.glob(os.path.join(REPO, "profiles", "counters_search_L*_*.json")))
    assert profiles
    for path in profiles:
        prof = json.load(open(path))
        level = int(os.path.basename(path)[len("counters_search_L")])
        wk = prof["match"]
        tuning = json.load(open(bench.default_tuning_path(type("A", (), wk))))
        assert prof["tuning"] == bench.tuning_key(tuning["levels"][str(level)]), path
        for k in ("SQ_INSTS_VALU", "SQ_WAVE_CYCLES", "SQ_WAIT_ANY"):
            assert prof["per_call"][k] > 0
        args = argparse.Namespace(workload=wk["workload"], resolution=wk["resolution"], seed=wk["seed"], max_level=wk["max_level"])
        assert bench.counters_path(args, level) == path
        tun = {level: tuning["levels"][str(level)]}
        r = bench.roofline(args, level, prof["kernel_ms_per_call"], 1000, tun, tun[level]["variant"])
        assert r["bound"] == "valu" and r["traffic"] == prof["bytes_per_launch"]
        assert 0 < r["frac"] <= 1 and 0 < r["hbm"]["frac"] <= 1
        # a different tuning disables the figures instead of reporting stale ones
        other = {level: dict(tun[level], first_batch=tun[level]["first_batch"] + 1)}
        assert bench.roofline(args, level, prof["kernel_ms_per_call"], 1000, other, 0)["frac"] is None


class _FakeCtx:
    def __init__(self):
        self.calls = []

    def set_variant(self, level, v):
        self.calls.append(("variant", level, v))

    def set_batching(self, level, first, growth):
        self.calls.append(("batching", level, first, growth))

    def set_filter_bits(self, level, bits):
        self.calls.append(("filter", level, bits))

    def set_child_frames(self, level, frames):
        self.calls.append(("frames", level, frames))


def test_load_tuning_applies_every_field(tmp_path):
    data = {"workload": {"workload": "scene", "resolution": 64, "seed": 0, "max_level": 3},
            "levels": {"3": {"variant": 2, "first_batch": 77, "growth": 2, "filter_bits": 8, "child_frames": 6, "call_ms_by_variant": {}},
                       "2": {"variant": 0, "first_batch": 99, "growth": 4, "filter_bits": 16, "child_frames": 1}}}
    path = tmp_path / "t.json"
    path.write_text(json.dumps(data))
    ctx = _FakeCtx()
    out = bench.load_tuning(ctx, str(path), [3, 2])
    assert set(out) == {3, 2}
    assert ("variant", 3, 2) in ctx.calls and ("batching", 3, 77, 2) in ctx.calls
    assert ("filter", 3, 8) in ctx.calls and ("frames", 3, 6) in ctx.calls
    assert ("batching", 2, 99, 4) in ctx.calls and ("filter", 2, 16) in ctx.calls
    assert not any(c[0] == "frames" and c[1] == 2 for c in ctx.calls)  # level 2 has no frame count


def test_bench_refuses_several_gpus():
    proc = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", "2"], capture_output=True, text=True, timeout=120)
    assert proc.returncode == 2
    assert "one GPU" in proc.stderr
