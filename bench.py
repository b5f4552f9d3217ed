"""Benchmark of the render core's hot path on one MI355X.

Headline (default workload "trace1m"): incoherent closest-hit traversal at
1M triangles -- 200 subdivision-4 icospheres + ground plane (1,024,002
triangles, 100 m box) with 2^24 rays, origins uniform in the box and
directions uniform on the sphere.  A step is one launch over the 2^24 rays,
already resident in HBM.  value = rays traced per second (Mrays/s), timed
with HIP events on the library's stream.

Roofline: the traversal kernel is bound by VALU issue, not by HBM (its node
records, 49 MB at 1M triangles, are served from L2 and the Infinity Cache).
roofline.achieved = VALU lane-instructions per second (wave VALU
instructions per launch from the committed rocprofv3 counter summary of the
same command, x 64 lanes, over the launch time measured here); peak = 39.3 T
lane-instructions/s.  Beside it: the HBM model, algorithmic bytes per ray
B = 32 + 16 + 24 P + 48 S (ray in, hit out, two triangles' axis coordinates
for every popped in-range node P, their other twelve floats for every
slab-passing node S; P and S counted by an instrumented launch over the same
rays) against 8 TB/s, and traffic = calibrated FETCH_SIZE + WRITE_SIZE bytes
per launch.

cpu_baseline: the CPU restatement of the reference traversal (the oracle),
built with -O3 -march=native on the host that runs it, on a bounded sample of
the same rays with OpenMP on every host thread; its hits must equal the GPU's.

Other workloads print their own JSON line:
  soup10m          10M-triangle soup: NMH build ms, incoherent and coherent Mrays/s
  render_cornell   PBR seconds to 1% rel-RMSE, Cornell box 256^2
  render_spheres   PBR seconds to 1% rel-RMSE, 1M triangles 1080p
  compat_cornell   compat-mode (reference MC ray tracer) frame, 256^2 16 spp, and
                   the compat estimator's seconds to 1% rel-RMSE
  compat_spheres   compat-mode frame, 1M triangles, 1080p 64 spp, and seconds to 1%
  batch256         256 batched jobs vs per-job frames
  adjoint_cornell  adjoint/forward time, finite differences, Adam fit
  emitters         emitter table (compaction + prefix sums) over 1M triangles

    python bench.py [--gpus 1] [--steps 20] [--warmup 3] [--workload trace1m]
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
# VALU issue peak of MI355X: 256 CUs x 4 SIMDs x 16 lanes per clock x 2.4 GHz
# = 39.3 T lane-instructions/s (the 157.3 TFLOP/s FP32 vector peak counts each
# lane-instruction as a fused multiply-add on a packed pair: x 2 x 2)
VALU_PEAK_T = 256 * 4 * 16 * 2.4e9 / 1e12


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse_profile(key: str, kernel: str, summary: str):
    """A per-launch figure of a kernel from the committed rocprofv3 summary's
    machine-readable '<key> <kernel> <value>' lines: traffic_bytes_per_launch
    (calibrated FETCH_SIZE + WRITE_SIZE) or valu_insts_per_launch
    (SQ_INSTS_VALU, wave instructions)."""
    path = os.path.join(ROOT, "profiles", summary)
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if len(f) == 3 and f[0] == key and f[1] == kernel:
            return int(float(f[2]))
    return None


def parse_traffic(kernel: str, summary: str):
    return parse_profile("traffic_bytes_per_launch", kernel, summary)


def cpu_trace_baseline(tree_vertices: np.ndarray, rays: np.ndarray, gpu_ids: np.ndarray,
                       gpu_tuv: np.ndarray, budget_s: float = 12.0):
    """The oracle's closest-hit traversal (the reference's intersect(),
    restated) built with -O3 -march=native on this host, OpenMP over rays on
    every host thread: one warm-up, then three timed runs over a sample sized
    to about budget_s in total (median).  The sample's hit ids and t/u/v
    must equal the GPU's before the timing counts."""
    from tests import oracle_ffi

    threads = os.cpu_count() or 1
    oracle_ffi.use_native()
    n = 8192
    t0 = time.perf_counter()
    oracle_ffi.trace(tree_vertices, rays[:n], threads=threads)  # warm-up
    dt = time.perf_counter() - t0
    n2 = int(min(rays.shape[0], max(n, n * budget_s / 3 / max(dt, 1e-3))))
    runs = []
    for _ in range(3):
        t0 = time.perf_counter()
        ids, tuv = oracle_ffi.trace(tree_vertices, rays[:n2], threads=threads)
        runs.append(time.perf_counter() - t0)
    same = bool(np.array_equal(ids, gpu_ids[:n2]) and
                np.array_equal(tuv.view(np.uint32), gpu_tuv[:n2].view(np.uint32)))
    if not same:
        raise RuntimeError("cpu baseline: CPU and GPU hits differ on the timed sample")
    return {
        "value": round(n2 / float(np.median(runs)) / 1e6, 3),
        "unit": "Mrays/s",
        "cores": threads,
        "kind": "port",
        "sample": f"first {n2} of the 2^24 rays, closest hit, median of 3 runs after a warm-up; "
                  f"hit ids and t/u/v equal the GPU's",
        "cpu_model": oracle_ffi.cpu_model(),
        "build": "g++ -O3 -march=native -ffp-contract=off -fopenmp",
    }


def time_launches(ctx, launch, steps: int, warmup: int):
    """HIP-event timing on the context's stream; returns (ms per step, list)."""
    import torch

    stream = torch.cuda.ExternalStream(ctx.stream)
    for _ in range(warmup):
        launch()
    ctx.synchronize()
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
    ev[0].record(stream)
    for i in range(steps):
        launch()
        ev[i + 1].record(stream)
    ctx.synchronize()
    torch.cuda.synchronize()
    per = [ev[i].elapsed_time(ev[i + 1]) for i in range(steps)]
    return ev[0].elapsed_time(ev[-1]) / steps, per


def host_inclusive_trace(ctx, rays_np: np.ndarray, rays, hits, launch) -> dict:
    """The rate for a caller whose rays and hits live in host memory (the
    trace entry points take device pointers, so the headline value has no
    transfer in it): pinned host rays copied in, one launch, the hits copied
    back, each step synchronised; median of 3 after a warm-up."""
    import torch

    n = rays_np.shape[0]
    h_rays = torch.from_numpy(rays_np).pin_memory()
    h_hits = torch.empty(tuple(hits.shape), dtype=hits.dtype, pin_memory=True)
    times = []
    for it in range(4):
        t0 = time.perf_counter()
        rays.copy_(h_rays)
        torch.cuda.synchronize()
        launch()
        ctx.synchronize()
        h_hits.copy_(hits)
        torch.cuda.synchronize()
        if it:
            times.append(time.perf_counter() - t0)
    t = float(np.median(times))
    moved = rays_np.nbytes + hits.numel() * hits.element_size()
    log(f"host-inclusive: {t * 1e3:.2f} ms ({n / t / 1e6:.0f} Mrays/s; {moved / 1e6:.0f} MB over PCIe)")
    return {"ms": round(t * 1e3, 3), "mrays": round(n / t / 1e6, 1), "pcie_bytes": int(moved),
            "note": "pinned host rays in, one launch, hits out; synchronous"}


def workload_trace1m(a) -> dict:
    import torch

    from aurora_amd import native, scenes, tune

    t0 = time.perf_counter()
    scene = scenes.spheres()
    log(f"scene {scene.name}: {scene.count} triangles ({time.perf_counter() - t0:.1f}s)")
    ctx = native.Context(0)
    ctx.upload_scene(scene)
    stats = ctx.rebuild_nmh()
    stats = ctx.rebuild_nmh()  # second build: warm code objects and allocations
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
    if a.tune:
        params, _ = tune.tune_trace(ctx, a.rays, rays.data_ptr(), hits.data_ptr(), key="trace1m")
        log(f"tuned trace params: {params}")
    params = ctx.get_trace_params()

    launch = lambda: ctx.trace_closest_ptr(a.rays, rays.data_ptr(), hits.data_ptr())  # noqa: E731
    ms_step, per = time_launches(ctx, launch, a.steps, a.warmup)
    mrays = a.rays / (ms_step * 1e-3) / 1e6
    achieved = bytes_per_launch / (ms_step * 1e-3) / 1e9
    log(f"{ms_step:.3f} ms/launch (median {np.median(per):.3f}), {mrays:.1f} Mrays/s")
    host = host_inclusive_trace(ctx, rays_np, rays, hits, launch)

    cpu = None
    if not a.no_cpu:
        log("cpu baseline ...")
        h = hits.cpu().numpy()
        cpu = cpu_trace_baseline(scene.vertices[perm], rays_np, h[:, 0].view(np.int32), h[:, 1:4])
        log(f"cpu baseline {cpu['value']} Mrays/s on {cpu['cores']} threads")
    ctx.close()
    return {
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
            "trace_params": params,
            "host_inclusive": host,
        },
        "roofline": trace_roofline(bytes_per_launch, ms_step, "k_trace_incoherent_1M.md"),
        "cpu_baseline": cpu,
    }


def trace_roofline(model_bytes: int, ms: float, summary: str) -> dict:
    """The traversal kernel's binding roof is VALU issue: its node records
    (49 MB at 1M triangles) are served from L2 and the Infinity Cache, so
    HBM traffic is a fraction of the algorithmic bytes.  achieved = wave
    VALU instructions per launch (SQ_INSTS_VALU from the committed profile of
    the same command) x 64 lanes / the launch time measured here; peak =
    VALU_PEAK_T.  The HBM model (32 + 16 + 24 P + 48 S bytes per ray) and the
    measured HBM bytes stay beside it."""
    valu = parse_profile("valu_insts_per_launch", "k_trace_pt", summary)
    traffic = parse_traffic("k_trace_pt", summary)
    sec = ms * 1e-3
    model_gbs = model_bytes / sec / 1e9
    out = {"bound": "valu", "achieved": None, "peak": round(VALU_PEAK_T, 2), "unit": "T VALU lane-instr/s",
           "frac": None, "traffic": traffic,
           "valu_insts_per_launch": valu,
           "hbm_model": {"bytes_per_launch": int(model_bytes), "achieved_gbs": round(model_gbs, 1),
                         "frac_of_8tbs": round(model_gbs / HBM_PEAK_GBS, 4)},
           "hbm_measured": None}
    if valu:
        t = valu * 64 / sec / 1e12
        out["achieved"] = round(t, 2)
        out["frac"] = round(t / VALU_PEAK_T, 4)
    if traffic:
        g = traffic / sec / 1e9
        out["hbm_measured"] = {"achieved_gbs": round(g, 1), "frac_of_8tbs": round(g / HBM_PEAK_GBS, 4)}
    return out


def workload_soup10m(a) -> dict:
    import torch

    from aurora_amd import native, scenes, tune

    t0 = time.perf_counter()
    scene = scenes.soup(n=10_000_000)
    log(f"soup: {scene.count} triangles ({time.perf_counter() - t0:.1f}s)")
    ctx = native.Context(0)
    ctx.upload_scene(scene)
    builds = [ctx.rebuild_nmh().build_ms for _ in range(3)]
    log(f"NMH build ms: {builds}")
    out, shapes, trace_roof = {"incoherent": None, "coherent": None}, {}, {}
    sets = {"incoherent": lambda: scenes.incoherent_rays(1 << 24, 0.0, 1000.0, seed=33),
            "coherent": lambda: scenes.coherent_rays(scene.camera, 4096, 4096)}
    for name in ([] if a.only == "build" else [a.only] if a.only else list(sets)):
        rays_np = sets[name]()
        rays = torch.from_numpy(rays_np).to("cuda")
        hits = torch.empty((rays.shape[0], 4), dtype=torch.float32, device="cuda")
        n = rays.shape[0]
        # launch shape per ray set by the autotuner (cached per device)
        shapes[name], _ = tune.tune_trace(ctx, n, rays.data_ptr(), hits.data_ptr(), key=f"soup10m-{name}",
                                          reps=2, force=a.tune)
        ms, _ = time_launches(ctx, lambda: ctx.trace_closest_ptr(n, rays.data_ptr(), hits.data_ptr()),
                              a.steps, a.warmup)
        out[name] = round(n / (ms * 1e-3) / 1e6, 1)
        log(f"{name}: {ms:.3f} ms, {out[name]} Mrays/s (launch shape {shapes[name]})")
        popped, passed = ctx.trace_closest_stats_ptr(n, rays.data_ptr(), hits.data_ptr())
        model = 48 * n + 24 * popped + 48 * passed
        traffic = parse_traffic(f"k_trace_pt:{name}", "soup10m_traversal.md")
        valu = parse_profile("valu_insts_per_launch", f"k_trace_pt:{name}", "soup10m_traversal.md")
        trace_roof[name] = {"ms": round(ms, 4), "popped_per_ray": round(popped / n, 3),
                            "passed_per_ray": round(passed / n, 3), "hbm_model_bytes": int(model),
                            "hbm_model_frac": round(model / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                            "traffic": traffic,
                            "hbm_measured_frac": round(traffic / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4) if traffic else None,
                            "valu_insts_per_launch": valu,
                            "valu_frac": round(valu * 64 / (ms * 1e-3) / 1e12 / VALU_PEAK_T, 4) if valu else None}
        log(f"{name}: P={popped / n:.1f} S={passed / n:.1f}, model {model / 1e9:.1f} GB per launch")
        del rays, hits
    cpu = None
    if not a.no_cpu:
        from tests import oracle_ffi

        # the literal restatement of cudaRebuildNMH (stable sorts per level)
        # over all 10M triangles, one thread and then OpenMP-parallel sorts on
        # every host thread; both permutations must equal the GPU's
        oracle_ffi.use_native()
        gpu_perm = ctx.read_permutation()
        threads = os.cpu_count() or 1
        # one thread: all 10M with --cpu-full, else the first 2^21 triangles
        # (a build of its own, checked against a GPU build of the same subset)
        n1 = scene.count if a.cpu_full else (1 << 21)
        sub_perm = gpu_perm
        if n1 != scene.count:
            ctx.upload_geometry(scene.vertices[:n1], scene.normals[:n1], scene.texcoords[:n1],
                                scene.shader_ids[:n1])
            ctx.rebuild_nmh()
            sub_perm = ctx.read_permutation()
        cpu_ms = {}
        for th, n, ref in ((1, n1, sub_perm), (threads, scene.count, gpu_perm)):
            t0 = time.perf_counter()
            p = oracle_ffi.nmh_build(scene.vertices[:n], threads=th)
            cpu_ms[th] = (time.perf_counter() - t0) * 1e3
            if not np.array_equal(p, ref):
                raise RuntimeError(f"cpu baseline: the {th}-thread CPU tree differs from the GPU tree")
            log(f"cpu NMH build of {n} triangles, {th} thread(s): {cpu_ms[th]:.0f} ms (tree equals the GPU's)")
        cpu = {"value": round(cpu_ms[threads], 1), "unit": "ms", "cores": threads, "kind": "port",
               "single_thread_ms": round(cpu_ms[1], 1), "single_thread_triangles": n1,
               "sample": f"literal NMH build of all 10,000,000 soup triangles on {threads} threads, and of "
                         f"the first {n1} on one thread (one run each); permutations equal the GPU's",
               "cpu_model": oracle_ffi.cpu_model(),
               "build": "g++ -O3 -march=native -ffp-contract=off -fopenmp"}
    ctx.close()
    build_ms = float(np.median(builds[1:]))
    nb = nmh_build_bytes(scene.count)
    achieved = nb / (build_ms * 1e-3) / 1e9
    return {
        "metric": "NMH build ms @10M tri", "value": round(build_ms, 3),
        "unit": "ms", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "higher_is_better": False, "dtype": "f32", "data": "synthetic (seeded 10M-triangle soup)",
        "config": {"workload": "10M-triangle soup build + 2^24 incoherent + 4096^2 coherent rays",
                   "triangles": scene.count, "incoherent_mrays": out["incoherent"],
                   "coherent_mrays": out["coherent"], "trace_params": shapes, "traversal_roofline": trace_roof,
                   "build_ms_all": [round(b, 3) for b in builds]},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": parse_traffic("nmh_build", "nmh_build_10M.md"),
                     "algorithmic_bytes": nb},
        "cpu_baseline": cpu,
    }


def nmh_build_bytes(n: int) -> int:
    """Algorithmic bytes of one NMH build of n triangles:
    60 n (bounds) + sum_{l<L} 12 P_l (one stable pass per level) +
    sum_{l<L-1} 8 P_l (argmax) + 196 n (final gather), P_l = n - 2(2^l - 1)."""
    levels = (n // 2).bit_length()  # log2i(n/2) + 1
    pend = [n - 2 * ((1 << l) - 1) for l in range(levels)]
    return 60 * n + 12 * sum(pend) + 8 * sum(pend[:-1]) + 196 * n


def two_seed_rel_rmse(img_a: np.ndarray, img_b: np.ndarray) -> float:
    """rel-RMSE of one n-spp render against the converged image, estimated
    from two independent n-spp renders: E||I_a - I_b||^2 = 2 E||I - mu||^2, so
    ||I_a - I_b|| / (sqrt(2) ||(I_a + I_b) / 2||).  Needs no reference render
    (a finite-spp reference would add its own variance to the distance)."""
    a = img_a[:, :3].astype(np.float64)
    b = img_b[:, :3].astype(np.float64)
    d = a - b
    m = 0.5 * (a + b)
    return float(np.sqrt((d * d).sum()) / (np.sqrt(2.0) * np.sqrt((m * m).sum())))


def render_convergence(a, scene, label: str, max_spp: int, mode: int | None = None) -> dict:
    """Wall seconds to 1% rel-RMSE for the PBR path loop (or, with mode, the
    compat estimator): renders of 1, 2, 4, ... spp (seed 7, timed from the
    call to the image in host memory), each with a second render at seed 8
    for the two-seed error estimate, until the estimate is <= 1%."""
    from aurora_amd import native

    mode = native.MODE_PBR if mode is None else mode
    ctx = native.Context(0)
    ctx.upload_scene(scene)
    ctx.rebuild_nmh()
    if a.max_paths:
        ctx.set_render_params(a.max_paths)
    w, h = scene.width, scene.height
    # warm-up: code objects and the largest path buffers
    ctx.frame_create(w, h, 64)
    ctx.frame_render(scene.camera, native.Context.opts(mode, 64, scene.max_depth, scene.rr_start, 1))
    ctx.synchronize()
    spp, rows, hit = (a.spp or 1), [], None
    if a.spp:
        max_spp = a.spp  # one render at --spp (profiling)
    while spp <= max_spp:
        ctx.frame_create(w, h, spp)
        t0 = time.perf_counter()
        ctx.frame_render(scene.camera, native.Context.opts(mode, spp, scene.max_depth, scene.rr_start, 7))
        img_a = ctx.frame_read()
        dt = time.perf_counter() - t0
        rays = ctx.rays_traced()
        rays_ext, rays_sh = ctx.ray_counts()
        ctx.frame_render(scene.camera, native.Context.opts(mode, spp, scene.max_depth, scene.rr_start, 8))
        img_b = ctx.frame_read()
        err = two_seed_rel_rmse(img_a, img_b)
        rows.append({"spp": spp, "s": round(dt, 4), "rel_rmse": round(err, 5),
                     "mrays": round(rays / dt / 1e6, 1), "rays": rays, "extension_rays": rays_ext,
                     "shadow_rays": rays_sh})
        log(f"{label}: {spp} spp {dt:.3f}s rel-RMSE {err:.4f} {rays / dt / 1e6:.0f} Mrays/s")
        if err <= 0.01:
            hit = rows[-1]
            break
        spp *= 2
    ref = None
    if hit and not a.spp:
        # against a converged reference: R at 16x the spp; R's own variance
        # (1/16 of the image's) is taken out of the distance.  One render
        # holds fewer than 2^32 pixel samples, so R is the mean of equal
        # renders with seeds 9, 10, ...
        rs = 16 * hit["spp"]
        part = rs
        while part > 1 and w * h * part >= 1 << 32:
            part //= 2
        ctx.frame_create(w, h, part)
        R = np.zeros((w * h, 3))
        for k in range(rs // part):
            ctx.frame_render(scene.camera, native.Context.opts(mode, part, scene.max_depth, scene.rr_start,
                                                               9 + k))
            R += ctx.frame_read()[:, :3]
        R /= rs // part
        d = float(np.linalg.norm(img_a[:, :3] - R) / np.linalg.norm(R))
        ref = {"reference_spp": rs, "rel_rmse_vs_reference": round(d, 5),
               "rel_rmse_vs_reference_corrected": round(d / np.sqrt(1.0 + 1.0 / 16.0), 5)}
        log(f"{label}: {hit['spp']} spp vs {rs}-spp reference: rel-RMSE {d:.4f} "
            f"({d / np.sqrt(1 + 1 / 16):.4f} without the reference's variance)")
    params = ctx.get_render_params()
    ctx.close()
    return {"curve": rows, "to_1pct": hit, "render_params": params, "against_reference": ref,
            "error_estimate": "two-seed: ||I7 - I8|| / (sqrt2 ||(I7 + I8)/2||)"}


def workload_render_cornell(a) -> dict:
    from aurora_amd import scenes

    scene = scenes.cornell()  # 256^2, depth 4, RR from bounce 3
    res = render_convergence(a, scene, "cornell", 1 << 16)
    cpu = None
    if not a.no_cpu and res["to_1pct"]:
        from tests import oracle_ffi

        perm = oracle_ffi.nmh_build(scene.vertices)
        tree = oracle_ffi.tree_arrays(scene, perm)
        spp = 4
        th = min(16, os.cpu_count() or 1)
        t0 = time.perf_counter()
        oracle_ffi.render(tree, scene, 1, spp, scene.max_depth, scene.rr_start, 7, threads=th)
        dt = (time.perf_counter() - t0) * res["to_1pct"]["spp"] / spp
        cpu = {"value": round(dt, 2), "unit": "s", "cores": th, "kind": "port",
               "sample": f"oracle PBR render at {spp} spp, scaled to {res['to_1pct']['spp']} spp"}
    hit = res["to_1pct"]
    return {"metric": "s to 1% rel-RMSE (Cornell box + 2 spheres, 256x256, PBR)", "value": hit["s"] if hit else None,
            "unit": "s", "n_gpus": 1, "higher_is_better": False, "dtype": "f32",
            "data": "synthetic (seeded Cornell box + 2 icospheres)",
            "config": {"workload": "Cornell box + 2 subdivision-3 spheres (2,584 triangles), 256x256, depth 4, RR from bounce 3", **res},
            "cpu_baseline": cpu}


def workload_render_spheres(a) -> dict:
    from aurora_amd import scenes

    scene = scenes.spheres()  # 1920x1080, depth 8
    res = render_convergence(a, scene, "spheres", 4096)
    hit = res["to_1pct"]
    return {"metric": "s to 1% rel-RMSE (200 spheres, 1M triangles, 1920x1080, PBR)", "value": hit["s"] if hit else None,
            "unit": "s", "n_gpus": 1, "higher_is_better": False, "dtype": "f32",
            "data": "synthetic (seeded 200 icospheres + ground, 4 quad lights)",
            "config": {"workload": "200 subdivision-4 spheres + ground (1,024,002 triangles), 4 quad lights, 1920x1080, depth 8, RR from bounce 3", **res}}


def c5_scene():
    from aurora_amd import scenes

    return scenes.c5_scene()


def workload_adjoint_cornell(a) -> dict:
    """Adjoint (gradient of an L2 image loss w.r.t. 64 RGB albedos and 4 RGB
    light radiances) vs forward render at equal spp; central-difference
    check (h = 1e-3, common random numbers) of the largest gradients; 100
    Adam steps from a +-30% perturbation toward a 1024-spp target."""
    import torch

    from aurora_amd import diff

    scene = c5_scene()
    spp, depth = a.spp or 64, scene.max_depth
    log(f"adjoint scene: {scene.count} triangles, {scene.shaders.shape[0]} shaders, {scene.lights.shape[0]} lights")
    r = diff.DiffRenderer(scene, spp=spp, max_depth=depth, seed=7)
    if a.max_paths:
        r.ctx.set_render_params(a.max_paths)
    gt_a = torch.tensor(scene.shaders["diffuse_color"], device="cuda")
    gt_r = torch.tensor(scene.lights["color"], device="cuda")
    t0 = time.perf_counter()
    target = r.render(gt_a, gt_r, spp=1024, seed=999)[:, :3].contiguous()
    log(f"target 1024 spp in {time.perf_counter() - t0:.2f}s")
    g = torch.Generator(device="cpu").manual_seed(5)
    p0a = (gt_a.cpu() * (1 + 0.3 * (2 * torch.rand(gt_a.shape, generator=g) - 1))).cuda()
    p0r = (gt_r.cpu() * (1 + 0.3 * (2 * torch.rand(gt_r.shape, generator=g) - 1))).cuda()
    npix = scene.width * scene.height

    # forward vs adjoint time at equal spp
    img = r.render(p0a, p0r)
    d_img = (2.0 / (npix * 3)) * (img[:, :3] - target)
    r.adjoint(p0a, p0r, d_img)  # warm-up
    fw, bw = [], []
    for _ in range(max(3, a.steps // 4)):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r.ctx.frame_render(scene.camera, r.opts())
        fw.append(time.perf_counter() - t0)
        rays_fwd = r.ctx.rays_traced()
        t0 = time.perf_counter()
        ga, gr = r.adjoint(p0a, p0r, d_img)
        bw.append(time.perf_counter() - t0)
    t_fwd, t_adj = float(np.median(fw)), float(np.median(bw))
    log(f"forward {t_fwd * 1e3:.1f} ms, adjoint {t_adj * 1e3:.1f} ms at {spp} spp "
        f"({rays_fwd / t_fwd / 1e6:.0f} Mrays/s forward)")

    # central differences, h = 1e-3, same seed (common random numbers)
    def loss_at(pa, pr):
        im = r.render(pa, pr)[:, :3]
        return float(((im.double() - target.double()) ** 2).mean())

    h = 1e-3
    fd_rows = []
    order_a = torch.argsort(ga.abs().flatten(), descending=True)[:8].tolist()
    order_r = torch.argsort(gr.abs().flatten(), descending=True)[:4].tolist()
    for kind, idx_list in (("albedo", order_a), ("radiance", order_r)):
        for idx in idx_list:
            m, c = divmod(idx, 3)
            pp_a, pm_a, pp_r, pm_r = p0a.clone(), p0a.clone(), p0r.clone(), p0r.clone()
            if kind == "albedo":
                pp_a[m, c] += h
                pm_a[m, c] -= h
                grad = float(ga[m, c])
            else:
                pp_r[m, c] += h
                pm_r[m, c] -= h
                grad = float(gr[m, c])
            fd = (loss_at(pp_a, pp_r) - loss_at(pm_a, pm_r)) / (2 * h)
            fd_rows.append({"param": f"{kind}[{m}][{c}]", "adjoint": grad, "fd": fd,
                            "rel_err": abs(grad - fd) / max(abs(fd), 1e-30)})
    max_fd_err = max(row["rel_err"] for row in fd_rows)
    log(f"central differences: max rel err {max_fd_err:.2e} over {len(fd_rows)} params")

    # visible parameters: those the image depends on at all (gradient of the
    # image sum); albedos of hidden faces cannot be recovered from the image
    vis_a, vis_r = r.adjoint(p0a, p0r, torch.ones((npix, 3), device="cuda"))
    vis_a = vis_a > 1e-4 * float(vis_a.max())
    vis_r = vis_r > 1e-4 * float(vis_r.max())

    def eval_loss(pa, pr):  # 256-spp loss with a fixed seed, outside the fit
        im = r.render(pa, pr, spp=256, seed=4242)[:, :3]
        return float(((im.double() - target.double()) ** 2).mean())

    # 100 Adam steps, 16 spp per step, fresh seed per step, decorrelated
    # backward replay (unbiased gradient of ||E[I] - T||^2)
    steps_spp = 16
    pa = p0a.clone().requires_grad_(True)
    pr = p0r.clone().requires_grad_(True)
    lr_r = 0.01 * float(p0r.mean())  # same relative step for radiances (~15) as for albedos
    opt = torch.optim.Adam([{"params": [pa], "lr": 0.01}, {"params": [pr], "lr": lr_r}])
    r.spp = steps_spp
    r.decorrelate = True
    r.ctx.frame_create(scene.width, scene.height, steps_spp)
    losses = []
    t0 = time.perf_counter()
    for it in range(100):
        r.seed = 1000 + it
        opt.zero_grad()
        loss = ((diff.render(r, pa, pr) - target) ** 2).mean()
        loss.backward()
        opt.step()
        with torch.no_grad():
            pa.clamp_(0.01, 1.0)
            pr.clamp_(0.0, None)
        losses.append(float(loss.detach()))
    t_adam = time.perf_counter() - t0

    def rel_err(p, gt, vis):
        return float(((p - gt).abs() / gt)[vis].mean())

    err0_a, err1_a = rel_err(p0a, gt_a, vis_a), rel_err(pa.detach(), gt_a, vis_a)
    err0_r, err1_r = rel_err(p0r, gt_r, vis_r), rel_err(pr.detach(), gt_r, vis_r)
    ev0, ev1 = eval_loss(p0a, p0r), eval_loss(pa.detach(), pr.detach())
    ev_gt = eval_loss(gt_a, gt_r)
    log(f"Adam: 16-spp loss {losses[0]:.3e} -> {losses[-1]:.3e} in {t_adam:.1f}s; 256-spp loss "
        f"{ev0:.3e} -> {ev1:.3e} (ground truth {ev_gt:.3e}); visible albedo rel err "
        f"{err0_a:.3f} -> {err1_a:.3f} ({int(vis_a.sum())} channels), radiance {err0_r:.3f} -> {err1_r:.3f}")
    r.close()
    ratio = t_adj / t_fwd
    return {
        "metric": "adjoint/fwd time (Cornell box + 20 spheres, 512x512)", "value": round(ratio, 3), "unit": "x",
        "n_gpus": 1, "steps": len(fw), "higher_is_better": False, "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (seeded Cornell box + 20 icospheres, 64 shaders, 4 lights)",
        "config": {
            "workload": "Cornell box + 20 subdivision-4 spheres (102,424 triangles), 64 albedos + 4 emitters, 512x512, depth 6 (no Russian roulette), L2 loss vs 1024-spp target",
            "triangles": scene.count, "spp": spp, "forward_ms": round(t_fwd * 1e3, 2),
            "adjoint_ms": round(t_adj * 1e3, 2), "forward_mrays": round(rays_fwd / t_fwd / 1e6, 1),
            "fd_h": h, "fd_max_rel_err": max_fd_err, "fd": fd_rows,
            "adam": {"steps": 100, "spp_per_step": steps_spp, "lr_albedo": 0.01, "lr_radiance": round(lr_r, 4), "seconds": round(t_adam, 2),
                     "decorrelated_backward": True,
                     "loss16_first": losses[0], "loss16_last": losses[-1],
                     "loss16_curve_every_10": losses[::10],
                     "loss256_start_end_groundtruth": [ev0, ev1, ev_gt],
                     "visible_albedo_channels": int(vis_a.sum()),
                     "albedo_rel_err_start_end": [round(err0_a, 4), round(err1_a, 4)],
                     "radiance_rel_err_start_end": [round(err0_r, 4), round(err1_r, 4)]},
        },
    }


def c4_jobs(n_jobs: int = 256, seed: int = 44):
    from aurora_amd import scenes

    return scenes.c4_jobs(n_jobs, seed)


def workload_batch256(a) -> dict:
    """256 jobs in shared wavefront passes (aur_batch_render) vs the same
    jobs rendered one by one with aur_frame_render (scenes built beforehand
    in both cases).  Images are compared bit for bit, so both sides reach the
    same rel-RMSE."""
    import torch

    from aurora_amd import native, tune

    t0 = time.perf_counter()
    sc = c4_jobs()
    tri = np.array([s.count for s in sc])
    paths = np.array([s.width * s.height * s.spp for s in sc], dtype=np.int64)
    log(f"batch: {len(sc)} jobs, {tri.min()}..{tri.max()} triangles, {paths.sum() / 1e6:.0f}M paths "
        f"({time.perf_counter() - t0:.1f}s)")
    order = np.argsort(paths, kind="stable")  # both sides run shortest jobs first

    # per-job launches: one frame render per job, (A) with the frame
    # renderer's multi-sample passes, (B) one sample per pass, the
    # reference's launch structure (src/render/monteCarloRaytracer.cpp:81-120)
    if a.batch_only:  # profiling the batched passes alone
        order = order[:0]
    single = native.Context(0)
    t_single = np.zeros((2, len(sc)))
    rays_single = 0
    images = {}
    check = set(range(len(sc)))  # every image is compared
    for k in order:
        s = sc[k]
        single.upload_scene(s)
        single.rebuild_nmh()
        single.frame_create(s.width, s.height, s.spp)
        o = native.Context.opts(native.MODE_PBR, s.spp, s.max_depth, s.rr_start, 7 + int(k))
        for mode, max_paths in ((0, 0),) if a.per_job_only else ((0, 0), (1, 1)):
            single.set_render_params(max_paths)
            if k == order[0]:
                single.frame_render(s.camera, o)  # warm-up
            t1 = time.perf_counter()
            single.frame_render(s.camera, o)
            t_single[mode, k] = time.perf_counter() - t1
        rays_single += single.rays_traced()
        if k in check:
            images[int(k)] = single.frame_read()
    single.close()
    T_single = float(t_single[0].sum()) or float("nan")
    T_plain = float(t_single[1].sum()) or float("nan")
    # all jobs queued at t=0, shortest first
    lat_single = np.cumsum(t_single[0][order]) if len(order) else np.array([np.nan])
    lat_plain = np.cumsum(t_single[1][order]) if len(order) else np.array([np.nan])
    log(f"per-job (A, multi-sample passes): {T_single:.3f}s total, {len(sc) / T_single:.1f} jobs/s, "
        f"{rays_single / T_single / 1e6:.0f} Mrays/s")
    if a.per_job_only:  # profiling the per-job frames alone
        return {"metric": "per-job frames (256 jobs)", "value": round(len(sc) / T_single, 2), "unit": "jobs/s",
                "n_gpus": 1, "higher_is_better": True, "dtype": "f32",
                "data": "synthetic (256 seeded Cornell scenes, 2-40 spheres)",
                "config": {"workload": "the 256 batch jobs as one frame render each (multi-sample passes)",
                           "per_job_multisample_s": round(T_single, 4),
                           "per_job_mrays": round(rays_single / T_single / 1e6, 1)}}
    log(f"per-job (B, one sample per pass): {T_plain:.3f}s total, {len(sc) / T_plain:.1f} jobs/s, "
        f"{rays_single / T_plain / 1e6:.0f} Mrays/s")

    # batched
    ctx = native.Context(0)
    ids = [ctx.scene_add(s) for s in sc]
    jobs = native.Context.make_jobs(
        dict(scene=ids[k], width=s.width, height=s.height, spp=s.spp, max_depth=s.max_depth,
             rr_start=s.rr_start, seed=7 + k, camera=s.camera) for k, s in enumerate(sc))
    total_px = int(sum(s.width * s.height for s in sc))
    out = torch.zeros((total_px, 4), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    params, table = tune.tune_batch(ctx, jobs, out.data_ptr(), key="batch256", force=a.tune)
    log(f"batch params {params}" + (f" (tuned: {[(p['max_paths'], p['graph'], round(ms, 1)) for p, ms in table]})"
                                    if table else " (cached)"))
    ctx.batch_render_ptr(jobs, out.data_ptr())  # warm-up
    runs = []
    for _ in range(3):
        t1 = time.perf_counter()
        job_ms = ctx.batch_render_ptr(jobs, out.data_ptr(), with_times=True)
        runs.append((time.perf_counter() - t1, job_ms))
    T_batch, job_ms = min(runs, key=lambda r: r[0])
    rays_batch = ctx.rays_traced()
    # per-job error: a second batch at independent seeds gives every job's
    # two-seed rel-RMSE at its own spp, and the spp it would need for 1%
    # (error ~ 1 / sqrt(spp))
    img_a = out.cpu().numpy()
    jobs_b = jobs.copy()
    jobs_b["seed"] = jobs["seed"] + 100_000
    ctx.batch_render_ptr(jobs_b, out.data_ptr())
    img_b = out.cpu().numpy()
    job_err = []
    for k, s in enumerate(sc):
        o0, n = int(jobs[k]["out_offset"]), s.width * s.height
        job_err.append(two_seed_rel_rmse(img_a[o0:o0 + n], img_b[o0:o0 + n]))
    job_err = np.array(job_err)
    spp_1pct = np.array([s.spp for s in sc]) * (job_err / 0.01) ** 2
    log(f"per-job two-seed rel-RMSE at the jobs' spp: min {job_err.min():.4f} median {np.median(job_err):.4f} "
        f"max {job_err.max():.4f}; spp for 1%: median {np.median(spp_1pct):.0f}")
    log(f"batched: {T_batch:.3f}s, {len(sc) / T_batch:.1f} jobs/s, {rays_batch / T_batch / 1e6:.0f} Mrays/s")
    at1 = None if a.batch_only else batch_at_1pct(ctx, ids, sc, job_err)
    img = img_a
    same = 0
    for k in check if not a.batch_only else ():
        o0 = int(jobs[k]["out_offset"])
        n = sc[k].width * sc[k].height
        same += int(np.array_equal(img[o0:o0 + n], images[k]))
    if a.batch_only:
        log("identical images: not checked (--batch-only renders no single frames)")
    else:
        log(f"identical images: {same}/{len(check)}")
    ctx.close()
    lat_batch = np.asarray(job_ms, dtype=np.float64) / 1e3
    speedup = T_single / T_batch
    return {
        "metric": "batched jobs/s (256 jobs)", "value": round(len(sc) / T_batch, 2), "unit": "jobs/s",
        "n_gpus": 1, "steps": 3, "higher_is_better": True, "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (256 seeded Cornell scenes, 2-40 spheres)",
        "config": {
            "workload": "256 Cornell-box jobs with 2-40 subdivision-4 spheres (10,264-204,824 triangles), sides {64..512}, spp {4,16,64}, depth 6, RR from bounce 3",
            "triangles_min_max": [int(tri.min()), int(tri.max())], "paths_total": int(paths.sum()),
            "batch_s": round(T_batch, 4),
            "per_job_multisample_s": round(T_single, 4),
            "per_job_one_sample_per_pass_s": round(T_plain, 4),
            "speedup_vs_per_job_multisample": round(speedup, 3),
            "speedup_vs_per_job_one_sample_per_pass": round(T_plain / T_batch, 3),
            "per_job_jobs_per_s": [round(len(sc) / T_single, 2), round(len(sc) / T_plain, 2)],
            "batch_latency_p50_p99_s": [round(float(np.percentile(lat_batch, 50)), 4),
                                        round(float(np.percentile(lat_batch, 99)), 4)],
            "per_job_multisample_latency_p50_p99_s": [round(float(np.percentile(lat_single, 50)), 4),
                                                      round(float(np.percentile(lat_single, 99)), 4)],
            "per_job_one_sample_per_pass_latency_p50_p99_s": [
                round(float(np.percentile(lat_plain, 50)), 4), round(float(np.percentile(lat_plain, 99)), 4)],
            "batch_mrays": round(rays_batch / T_batch / 1e6, 1),
            "per_job_mrays": round(rays_single / T_single / 1e6, 1),
            "identical_images": f"{same}/{len(check)}",
            "render_params": params,
            "per_job_rel_rmse_two_seed": {"min": round(float(job_err.min()), 5),
                                          "median": round(float(np.median(job_err)), 5),
                                          "max": round(float(job_err.max()), 5),
                                          "jobs_at_or_below_1pct": int((job_err <= 0.01).sum())},
            "per_job_spp_for_1pct_median": round(float(np.median(spp_1pct)), 1),
            "at_1pct": at1,
        },
    }


def batch_at_1pct(ctx, ids, sc, job_err, n_jobs: int = 16) -> dict:
    """The n smallest jobs at the spp each needs for 1% rel-RMSE (from its
    two-seed estimate at its own spp, error ~ 1/sqrt(spp), with twice that
    spp as margin for the estimate's own noise): one batch against per-job frame renders of the same jobs and
    seeds, and every batch image against a 4x-spp render of its job (rel-RMSE
    with the reference's own variance, 1/4 of the image's, taken out)."""
    import torch

    from aurora_amd import native

    pix = np.array([s.width * s.height for s in sc])
    sel = [int(k) for k in np.argsort(pix, kind="stable")[:n_jobs]]
    spp = [int(np.ceil(sc[k].spp * (job_err[k] / 0.01) ** 2 * 2.0)) for k in sel]
    jobs = native.Context.make_jobs(
        dict(scene=ids[k], width=sc[k].width, height=sc[k].height, spp=spp[i], max_depth=sc[k].max_depth,
             rr_start=sc[k].rr_start, seed=500 + k, camera=sc[k].camera) for i, k in enumerate(sel))
    out = torch.zeros((int(sum(pix[k] for k in sel)), 4), dtype=torch.float32, device="cuda")
    ctx.batch_render_ptr(jobs, out.data_ptr())  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ctx.batch_render_ptr(jobs, out.data_ptr())
    torch.cuda.synchronize()
    T_b = time.perf_counter() - t0
    img = out.cpu().numpy()
    single = native.Context(0)
    T_a, errs, same = 0.0, [], 0
    for i, k in enumerate(sel):
        s = sc[k]
        single.upload_scene(s)
        single.rebuild_nmh()
        single.frame_create(s.width, s.height, spp[i])
        o = native.Context.opts(native.MODE_PBR, spp[i], s.max_depth, s.rr_start, 500 + k)
        if i == 0:
            single.frame_render(s.camera, o)  # warm-up
        t1 = time.perf_counter()
        single.frame_render(s.camera, o)
        T_a += time.perf_counter() - t1
        o0, n = int(jobs[i]["out_offset"]), int(pix[k])
        mine = img[o0:o0 + n]
        same += int(np.array_equal(single.frame_read(), mine))
        single.frame_create(s.width, s.height, 4 * spp[i])
        single.frame_render(s.camera, native.Context.opts(native.MODE_PBR, 4 * spp[i], s.max_depth, s.rr_start,
                                                          900_000 + k))
        R = single.frame_read()[:, :3].astype(np.float64)
        d = float(np.linalg.norm(mine[:, :3] - R) / np.linalg.norm(R))
        errs.append(d / np.sqrt(1.0 + 1.0 / 4.0))
    single.close()
    errs = np.array(errs)
    log(f"at 1%: {len(sel)} jobs of {int(pix[sel[0]])}..{int(pix[sel[-1]])} pixels, spp {min(spp)}..{max(spp)}: "
        f"batch {T_b:.3f}s, per-job frames {T_a:.3f}s ({T_a / T_b:.2f}x); rel-RMSE against 4x-spp references "
        f"max {errs.max():.4f}, median {np.median(errs):.4f}; identical images {same}/{len(sel)}")
    return {"jobs": len(sel), "job_indices": sel, "spp": spp, "batch_s": round(T_b, 4),
            "per_job_multisample_s": round(T_a, 4), "speedup": round(T_a / T_b, 3),
            "rel_rmse_vs_4x_reference_corrected": [round(float(e), 5) for e in errs],
            "rel_rmse_max": round(float(errs.max()), 5), "identical_images": f"{same}/{len(sel)}"}


def compat_frame(a, scene, label: str) -> dict:
    """Compat mode (the reference's MonteCarloRaytracer::render: 4x4-strata
    camera rays, estimators F1-F3, mirror and refraction rays) at the
    scene's spp: wall time of the render call, rays traced, and the two-seed
    rel-RMSE of the image at that spp."""
    from aurora_amd import native

    ctx = native.Context(0)
    ctx.upload_scene(scene)
    ctx.rebuild_nmh()
    if a.refill:
        ctx.set_render_params(refill=a.refill)
    w, h, spp = scene.width, scene.height, a.spp or scene.spp
    ctx.frame_create(w, h, spp)
    o = lambda seed: native.Context.opts(native.MODE_COMPAT, spp, scene.max_depth, scene.rr_start, seed)  # noqa: E731
    ctx.frame_render(scene.camera, o(1))  # warm-up
    ctx.synchronize()
    times = []
    for _ in range(max(1, a.steps // 4)):
        t0 = time.perf_counter()
        ctx.frame_render(scene.camera, o(7))
        ctx.synchronize()
        times.append(time.perf_counter() - t0)
    img_a = ctx.frame_read()
    rays = ctx.rays_traced()
    ctx.frame_render(scene.camera, o(8))
    img_b = ctx.frame_read()
    err = two_seed_rel_rmse(img_a, img_b)
    t = float(np.median(times))
    params = ctx.get_render_params()
    ctx.close()
    log(f"{label}: compat {w}x{h} {spp} spp: {t * 1e3:.1f} ms, {rays / t / 1e6:.0f} Mrays/s, rel-RMSE {err:.4f}")
    return {"ms": round(t * 1e3, 3), "mrays": round(rays / t / 1e6, 1), "rays": int(rays), "spp": spp,
            "rel_rmse_two_seed": round(err, 5), "runs": len(times), "render_params": params}


def compat_convergence(a, scene, label: str, max_spp: int) -> dict | None:
    """The compat estimator's seconds to 1% rel-RMSE (render_convergence in
    compat mode), unless one render at --spp was asked for."""
    from aurora_amd import native

    if a.spp:
        return None
    return render_convergence(a, scene, f"{label} compat", max_spp, mode=native.MODE_COMPAT)


def workload_compat_cornell(a) -> dict:
    from aurora_amd import scenes

    scene = scenes.cornell()  # 256^2, 16 spp
    r = compat_frame(a, scene, "cornell")
    r["convergence"] = compat_convergence(a, scene, "cornell", 1 << 14)
    return {"metric": "compat render ms (Cornell box + 2 spheres, 256x256, 16 spp)", "value": r["ms"],
            "unit": "ms", "n_gpus": 1, "higher_is_better": False, "dtype": "f32",
            "data": "synthetic (seeded Cornell box + 2 icospheres)",
            "config": {"workload": "compat mode (reference MC ray tracer), Cornell box + 2 subdivision-3 spheres (2,584 triangles), 256x256, 16 spp", **r}}


def workload_compat_spheres(a) -> dict:
    from aurora_amd import scenes

    scene = scenes.spheres()  # 1920x1080, 64 spp, Phong materials with reflectivity 0.2
    r = compat_frame(a, scene, "spheres")
    r["convergence"] = compat_convergence(a, scene, "spheres", 4096)
    return {"metric": "compat render ms (200 spheres, 1M triangles, 1920x1080, 64 spp)", "value": r["ms"],
            "unit": "ms", "n_gpus": 1, "higher_is_better": False, "dtype": "f32",
            "data": "synthetic (seeded 200 icospheres + ground, 4 quad lights)",
            "config": {"workload": "compat mode (reference MC ray tracer), 200 subdivision-4 spheres + ground (1,024,002 triangles), 4 quad lights, 1920x1080, 64 spp", **r}}


def emitter_bytes(n: int, emitters: int) -> int:
    """Algorithmic bytes of one emitter-table build: the shader ids read by
    the counting and the writing pass (2 B each per triangle), and per
    emitter its vertices (36 B) and its record (28 B)."""
    return 4 * n + 64 * emitters


def workload_emitters(a) -> dict:
    """aur_create_emitters on the 1M-triangle sphere scene with every other
    material emissive (cudaCreateEmitters, src/kernels/emitter.cu:79-119)."""
    import torch

    from aurora_amd import native, scenes

    scene = scenes.spheres()
    scene.shaders["emission"][::2] = np.linspace(0.5, 2.0, 16, dtype=np.float32)
    ctx = native.Context(0)
    ctx.upload_scene(scene)
    ctx.rebuild_nmh()
    n = scene.count
    buf = torch.empty(n * 28, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    count = ctx.emitters_ptr(buf.data_ptr(), n)
    times = []
    for i in range(a.warmup + a.steps):
        t0 = time.perf_counter()
        ctx.emitters_ptr(buf.data_ptr(), n)
        if i >= a.warmup:
            times.append(time.perf_counter() - t0)
    t = float(np.median(times))
    nb = emitter_bytes(n, count)
    cdf = torch.empty(max(1, ctx.num_lights), dtype=torch.float32, device="cuda")
    t0 = time.perf_counter()
    ctx.lights_cdf_ptr(cdf.data_ptr())
    t_cdf = time.perf_counter() - t0
    ctx.close()
    log(f"emitters: {count} of {n} triangles, {t * 1e3:.3f} ms per call (host wall), lights cdf {t_cdf * 1e3:.3f} ms")
    return {"metric": "emitter table ms (1M triangles)", "value": round(t * 1e3, 4), "unit": "ms", "n_gpus": 1,
            "steps": a.steps, "warmup": a.warmup, "higher_is_better": False, "dtype": "u32/f64",
            "data": "synthetic (200 icospheres + ground, 16 of 32 materials emissive)",
            "config": {"workload": "emitter table (compaction + prefix sums) over 1,024,002 triangles", "triangles": n,
                       "emitters": count, "algorithmic_bytes": nb, "lights_cdf_ms": round(t_cdf * 1e3, 4),
                       "timing": "host wall time of the synchronous call (3 launches + one count read-back)"}}


WORKLOADS = {"trace1m": workload_trace1m, "soup10m": workload_soup10m,
             "compat_cornell": workload_compat_cornell, "compat_spheres": workload_compat_spheres,
             "emitters": workload_emitters,
             "render_cornell": workload_render_cornell, "render_spheres": workload_render_spheres,
             "adjoint_cornell": workload_adjoint_cornell, "batch256": workload_batch256}


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--rays", type=int, default=1 << 24)
    ap.add_argument("--workload", default="trace1m", choices=sorted(WORKLOADS))
    ap.add_argument("--tune", action="store_true", help="autotune trace launch params first")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-full", action="store_true", help="soup10m: one-thread CPU build of all 10M triangles")
    ap.add_argument("--only", choices=["build", "incoherent", "coherent"], default=None,
                    help="soup10m: run only the builds or one ray set (profiling)")
    ap.add_argument("--max-paths", type=int, default=0, help="render: paths in flight per pass")
    ap.add_argument("--spp", type=int, default=0, help="render: one render at this spp only")
    ap.add_argument("--refill", type=int, default=0, help="render: refill threshold of the persistent launches")
    ap.add_argument("--batch-only", action="store_true", help="batch256: skip the per-job side")
    ap.add_argument("--per-job-only", action="store_true",
                    help="batch256: only the per-job frames with multi-sample passes (profiling)")
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
        print(f"bench.py: measures one GPU; --gpus {a.gpus} is not supported (the render core "
              "runs on a single GPU: rays, pixels and jobs are independent, so N GPUs would each "
              "run this benchmark's work on their own share)", file=sys.stderr)
        sys.exit(2)
    import torch

    if not torch.cuda.is_available():
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(1)
    print(json.dumps(WORKLOADS[a.workload](a)), flush=True)


if __name__ == "__main__":
    main()
