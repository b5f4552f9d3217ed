"""Autotuner host logic (aurora_amd/tune.py) on the CPU: a stand-in context
and clock stand for the device.  The fastest candidate is chosen and set on
the context, the choice is cached per device and workload key, a cached
choice is reused without measuring, `force` measures again, and an unreadable
cache file is treated as empty."""
from __future__ import annotations

import json

import numpy as np

from aurora_amd import tune


class FakeCtx:
    def __init__(self):
        self.trace_params, self.render_params = None, None
        self.set_calls = 0

    def set_trace_params(self, **p):
        self.trace_params = dict(p)
        self.set_calls += 1

    def set_render_params(self, **p):
        self.render_params = dict(p)


def _trace_cost(p):  # a made-up landscape with one minimum
    return 1.0 + abs(p["refill"] - 24) / 8 + p["stack_lds_rows"] / 10 + (p["block"] != 256) + 3 * (not p["persistent"])


def test_tune_trace_picks_caches_and_reuses(tmp_path, monkeypatch):
    monkeypatch.setattr(tune, "_device_name", lambda: "gfx950-test")
    measured = []

    def fake_time(ctx, n, rays, hits, reps=3, any_hit=False):
        measured.append(dict(ctx.trace_params))
        return _trace_cost(ctx.trace_params)

    monkeypatch.setattr(tune, "time_trace", fake_time)
    cache = str(tmp_path / "tune.json")
    ctx = FakeCtx()
    best, table = tune.tune_trace(ctx, 10, 0, 0, "scene-a", cache=cache)
    assert len(table) == len(tune.TRACE_CANDIDATES) == len(measured)
    assert best == min(tune.TRACE_CANDIDATES, key=_trace_cost)
    assert best == dict(persistent=1, block=256, refill=24, waves_per_cu=0, stack_lds_rows=0)
    assert ctx.trace_params == best
    stored = json.load(open(cache))
    assert stored["gfx950-test|closest|scene-a"]["params"] == best

    # cached: no measurement, same choice set on a new context
    measured.clear()
    ctx2 = FakeCtx()
    best2, table2 = tune.tune_trace(ctx2, 10, 0, 0, "scene-a", cache=cache)
    assert (best2, table2, measured) == (best, [], [])
    assert ctx2.trace_params == best
    # another key, the any-hit variant and force all measure again
    tune.tune_trace(FakeCtx(), 10, 0, 0, "scene-b", cache=cache)
    tune.tune_trace(FakeCtx(), 10, 0, 0, "scene-a", cache=cache, any_hit=True)
    tune.tune_trace(FakeCtx(), 10, 0, 0, "scene-a", cache=cache, force=True)
    assert len(measured) == 3 * len(tune.TRACE_CANDIDATES)
    assert set(json.load(open(cache))) == {"gfx950-test|closest|scene-a", "gfx950-test|closest|scene-b",
                                           "gfx950-test|any|scene-a"}


def test_tune_trace_unreadable_cache(tmp_path, monkeypatch):
    monkeypatch.setattr(tune, "_device_name", lambda: "dev")
    monkeypatch.setattr(tune, "time_trace", lambda ctx, *a, **k: _trace_cost(ctx.trace_params))
    cache = tmp_path / "tune.json"
    cache.write_text("{not json")
    best, table = tune.tune_trace(FakeCtx(), 1, 0, 0, "k", cache=str(cache))
    assert table and json.load(open(cache))["dev|closest|k"]["params"] == best
    # no cache at all
    best_nc, table_nc = tune.tune_trace(FakeCtx(), 1, 0, 0, "k", cache=None)
    assert best_nc == best and table_nc


def test_tune_batch_picks_fastest_on_a_job_sample(tmp_path, monkeypatch):
    monkeypatch.setattr(tune, "_device_name", lambda: "dev")
    clock = [0.0]
    monkeypatch.setattr("time.perf_counter", lambda: clock[0])
    seen = []

    class Ctx(FakeCtx):
        def batch_render_ptr(self, jobs, ptr):
            seen.append(len(jobs))
            p = self.render_params
            clock[0] += 1.0 + abs(p["max_paths"] - (1 << 23)) / (1 << 23) + (0.1 if p["graph"] else 0.0)

    jobs = np.arange(256)
    ctx = Ctx()
    best, table = tune.tune_batch(ctx, jobs, 0, "c4", sample=48, reps=2, cache=str(tmp_path / "t.json"))
    assert best == dict(max_paths=1 << 23, graph=0) and ctx.render_params == best
    assert len(table) == len(tune.BATCH_CANDIDATES)
    assert set(seen) == {48}  # the sample spans the list without repeats
