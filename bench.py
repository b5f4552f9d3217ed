"""Benchmark of the render core's hot path on one MI355X.

Headline (default): incoherent closest-hit traversal at 1M triangles -- the
config of 200 subdivision-4 icospheres (1,024,002 triangles, 100 m box) with
2^24 rays, origins uniform in the box and directions uniform on the sphere.
A step is one launch over the 2^24 rays, already resident in HBM.  value =
rays traced per second (Mrays/s), timed with HIP events on the library's
stream.

Roofline: algorithmic bytes per ray B = 32 + 16 + 24 P + 48 S (ray in, hit
out, two triangles' axis coordinates for every popped in-range node P, their
other twelve floats for every slab-passing node S), with P and S counted by
an instrumented launch over the same rays.  achieved = sum(B) / mean launch
time; peak = 8 TB/s HBM.

cpu_baseline: the CPU restatement of the reference traversal (the oracle)
on a bounded sample of the same rays, all host threads.

    python bench.py [--gpus 1] [--steps 20] [--warmup 3]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse_traffic(kernel: str, summary: str):
    """HBM bytes per launch (calibrated FETCH_SIZE + WRITE_SIZE) from the
    committed rocprofv3 summary's 'traffic_bytes_per_launch <kernel> <B>' line."""
    path = os.path.join(ROOT, "profiles", summary)
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if len(f) == 3 and f[0] == "traffic_bytes_per_launch" and f[1] == kernel:
            return int(f[2])
    return None


def cpu_baseline(tree_vertices: np.ndarray, rays: np.ndarray, budget_s: float = 15.0):
    from tests import oracle_ffi

    threads = os.cpu_count() or 1
    oracle_ffi.lib()
    n = 4096
    t0 = time.perf_counter()
    oracle_ffi.trace(tree_vertices, rays[:n], threads=threads)
    dt = time.perf_counter() - t0
    # grow the sample to roughly budget_s of CPU work
    n2 = int(min(rays.shape[0], max(n, n * budget_s / max(dt, 1e-3) * 0.5)))
    t0 = time.perf_counter()
    oracle_ffi.trace(tree_vertices, rays[:n2], threads=threads)
    dt = time.perf_counter() - t0
    return {
        "value": round(n2 / dt / 1e6, 3),
        "unit": "Mrays/s",
        "cores": threads,
        "kind": "port",
        "sample": f"first {n2} of the 2^24 rays, closest hit, oracle traversal on {threads} threads",
    }


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--rays", type=int, default=1 << 24)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device copy of MIB MiB (read MIB, write MIB) 5 times: the "
                         "reference kernel for scaling rocprofv3 FETCH_SIZE/WRITE_SIZE")
    a = ap.parse_args()
    if a.copy_calibration:
        import torch

        n = a.copy_calibration * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_mib": a.copy_calibration, "launches": 5}))
        return
    if a.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {a.gpus} is not supported "
              "(the render core is single-GPU; see DESIGN.md for how it would shard)",
              file=sys.stderr)
        sys.exit(2)

    import torch

    from aurora_amd import native, scenes

    if not torch.cuda.is_available():
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(1)

    t0 = time.perf_counter()
    scene = scenes.spheres()
    log(f"scene {scene.name}: {scene.count} triangles ({time.perf_counter() - t0:.1f}s)")
    ctx = native.Context(0)
    ctx.upload_scene(scene)
    stats = ctx.rebuild_nmh()
    stats = ctx.rebuild_nmh()  # second build: warm caches/allocations
    log(f"NMH build: {stats.levels} levels, {stats.build_ms:.2f} ms")
    perm = ctx.read_permutation()

    t0 = time.perf_counter()
    rays_np = scenes.incoherent_rays(a.rays, 0.0, 100.0, seed=2024)
    log(f"rays: {a.rays} ({time.perf_counter() - t0:.1f}s)")
    rays = torch.from_numpy(rays_np).to("cuda")
    hits = torch.empty((a.rays, 4), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()

    popped, passed = ctx.trace_closest_stats_ptr(a.rays, rays.data_ptr(), hits.data_ptr())
    bytes_per_launch = 48 * a.rays + 24 * popped + 48 * passed
    hit_rate = float((hits[:, 0].view(torch.int32) >= 0).float().mean().item())
    log(f"P={popped / a.rays:.2f} S={passed / a.rays:.2f} per ray; hit rate {hit_rate:.3f}")

    stream = torch.cuda.ExternalStream(ctx.stream)
    for _ in range(a.warmup):
        ctx.trace_closest_ptr(a.rays, rays.data_ptr(), hits.data_ptr())
    ctx.synchronize()
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(a.steps + 1)]
    wall0 = time.perf_counter()
    ev[0].record(stream)
    for i in range(a.steps):
        ctx.trace_closest_ptr(a.rays, rays.data_ptr(), hits.data_ptr())
        ev[i + 1].record(stream)
    ctx.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    per = [ev[i].elapsed_time(ev[i + 1]) for i in range(a.steps)]
    ms_total = ev[0].elapsed_time(ev[-1])
    ms_step = ms_total / a.steps
    mrays = a.rays * a.steps / (ms_total * 1e-3) / 1e6
    achieved = bytes_per_launch / (ms_step * 1e-3) / 1e9
    log(f"{ms_step:.3f} ms/launch (median {np.median(per):.3f}), {mrays:.1f} Mrays/s, "
        f"wall {wall:.2f}s")

    cpu = None
    if not a.no_cpu:
        log("cpu baseline ...")
        cpu = cpu_baseline(scene.vertices[perm], rays_np)
        log(f"cpu baseline {cpu['value']} Mrays/s on {cpu['cores']} threads")

    traffic = parse_traffic("k_trace", "k_trace_incoherent_1M.md")
    out = {
        "metric": "Mrays/s @1M tri (incoherent closest hit)",
        "value": round(mrays, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded generator: 200 icospheres + ground plane; 2^24 incoherent rays)",
        "config": {
            "workload": "incoherent closest-hit traversal, 1,024,002 triangles, 2^24 rays",
            "triangles": scene.count,
            "rays": a.rays,
            "nmh_levels": int(stats.levels),
            "nmh_build_ms": round(float(stats.build_ms), 3),
            "popped_per_ray": round(popped / a.rays, 3),
            "passed_per_ray": round(passed / a.rays, 3),
            "hit_rate": round(hit_rate, 4),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
