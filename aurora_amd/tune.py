"""Autotuner for launch shapes (no reference counterpart).

The reference launches fixed shapes (64-thread blocks, 65,536-ray tiles,
src/kernels/monteCarloRaytracer.cu:119-124, src/render/monteCarloRaytracer.cpp:13).
Here launch shapes are picked by measurement on the device at hand:

* tune_trace: persistent vs flat traversal, block size, refill threshold and
  the LDS window of the traversal stack (a smaller window leaves room for
  more resident waves, which pays for deep trees whose nodes miss the
  caches) for aur_trace_closest/any, timed with HIP events on the context's
  stream over the caller's own rays.
* tune_batch: paths in flight per wavefront pass (queue sizes) and direct
  launches vs a replayed HIP graph for aur_batch_render, timed on a sample
  of the caller's jobs.

Results are cached in a JSON file keyed by device name and workload key, so a
later process reuses them without re-measuring.  Results never depend on the
chosen shape (the traversal order per ray is fixed), only speed does.
"""
from __future__ import annotations

import json
import os

DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "aurora_amd", "tune.json")

TRACE_CANDIDATES = (
    [dict(persistent=0, block=64, refill=1, waves_per_cu=0, stack_lds_rows=0)]
    + [dict(persistent=1, block=256, refill=r, waves_per_cu=0, stack_lds_rows=s)
       for r in (16, 24, 32) for s in (0, 10, 12)]
    + [dict(persistent=1, block=128, refill=r, waves_per_cu=0, stack_lds_rows=0) for r in (16, 24)]
)


def _device_name() -> str:
    import torch

    return torch.cuda.get_device_name(0)


def _load(path: str) -> dict:
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(path: str, data: dict) -> None:
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(data, f, indent=1)
        os.replace(tmp, path)
    except OSError:
        pass


def time_trace(ctx, n: int, rays_ptr: int, hits_ptr: int, reps: int = 3, any_hit: bool = False) -> float:
    """Mean milliseconds of one trace launch with the context's current params."""
    import torch

    stream = torch.cuda.ExternalStream(ctx.stream)
    launch = ctx.trace_any_ptr if any_hit else ctx.trace_closest_ptr
    launch(n, rays_ptr, hits_ptr)  # warm-up (code objects, caches)
    ctx.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        launch(n, rays_ptr, hits_ptr)
    e1.record(stream)
    ctx.synchronize()
    return e0.elapsed_time(e1) / reps


def tune_trace(ctx, n: int, rays_ptr: int, hits_ptr: int, key: str, reps: int = 3,
               cache: str | None = DEFAULT_CACHE, candidates=TRACE_CANDIDATES, force: bool = False,
               any_hit: bool = False):
    """Sets the fastest trace params on ctx and returns (params, table).

    table: list of (params, ms) for every candidate measured (empty when the
    cached choice was reused)."""
    full_key = f"{_device_name()}|{'any' if any_hit else 'closest'}|{key}"
    data = _load(cache) if cache else {}
    if not force and full_key in data:
        p = data[full_key]["params"]
        ctx.set_trace_params(**p)
        return p, []
    table = []
    for cand in candidates:
        ctx.set_trace_params(**cand)
        table.append((dict(cand), time_trace(ctx, n, rays_ptr, hits_ptr, reps, any_hit)))
    best = min(table, key=lambda t: t[1])
    ctx.set_trace_params(**best[0])
    if cache:
        data[full_key] = {"params": best[0], "ms": best[1]}
        _store(cache, data)
    return best[0], table


BATCH_CANDIDATES = tuple(dict(max_paths=1 << k, graph=g) for k in (21, 22, 23, 24) for g in (0, 1))


def tune_batch(ctx, jobs, rgba_ptr: int, key: str, sample: int = 48, reps: int = 2,
               cache: str | None = DEFAULT_CACHE, candidates=BATCH_CANDIDATES, force: bool = False):
    """Sets the fastest render params for aur_batch_render on ctx and
    returns (params, table).  Times `sample` jobs spread over the list (wall
    time of the synchronous call, best of reps)."""
    import time

    import numpy as np

    full_key = f"{_device_name()}|batch|{key}"
    data = _load(cache) if cache else {}
    if not force and full_key in data:
        p = data[full_key]["params"]
        ctx.set_render_params(**p)
        return p, []
    idx = np.linspace(0, len(jobs) - 1, min(sample, len(jobs))).astype(int)
    sub = jobs[np.unique(idx)].copy()
    table = []
    for cand in candidates:
        ctx.set_render_params(**cand)
        ctx.batch_render_ptr(sub, rgba_ptr)  # warm-up: allocations, graph path
        best = float("inf")
        for _ in range(reps):
            t0 = time.perf_counter()
            ctx.batch_render_ptr(sub, rgba_ptr)
            best = min(best, time.perf_counter() - t0)
        table.append((dict(cand), best * 1e3))
    best = min(table, key=lambda t: t[1])
    ctx.set_render_params(**best[0])
    if cache:
        data[full_key] = {"params": best[0], "ms": best[1]}
        _store(cache, data)
    return best[0], table
