"""Batched rendering (aur_batch_render): every job's image equals the
single-frame render of the same scene, camera and options bit for bit, for
any number of paths per pass (jobs and samples split across passes) and with
the pass captured in a HIP graph and replayed."""
from __future__ import annotations

import numpy as np
import pytest

from aurora_amd import native, scenes

pytestmark = pytest.mark.gpu


def _scenes():
    return [
        scenes.cornell(seed=1, n_spheres=2, subdivisions=1, n_lights=1),
        scenes.cornell(seed=2, n_spheres=5, subdivisions=2, n_lights=4, split_materials=True),
        scenes.cornell(seed=3, n_spheres=1, subdivisions=1, n_lights=2),
    ]


JOBS = [  # scene, w, h, spp, depth, rr, seed
    (0, 16, 16, 1, 2, 2, 5),
    (1, 40, 24, 4, 6, 3, 6),
    (2, 33, 17, 3, 4, 4, 7),
    (0, 64, 48, 9, 4, 3, 8),
    (1, 8, 8, 2, 1, 1, 9),
    (2, 20, 30, 5, 6, 2, 10),
]


def _single(ctx, s, w, h, spp, depth, rr, seed):
    ctx.upload_scene(s)
    ctx.rebuild_nmh()
    ctx.frame_create(w, h, spp)
    ctx.frame_render(s.camera, native.Context.opts(native.MODE_PBR, spp, depth, rr, seed))
    return ctx.frame_read()


@pytest.fixture(scope="module")
def reference_images(gpu_ctx):
    sc = _scenes()
    return [_single(gpu_ctx, sc[j[0]], *j[1:]) for j in JOBS]


@pytest.fixture(scope="module")
def batch_ctx():
    ctx = native.Context(0)
    sc = _scenes()
    ids = [ctx.scene_add(s) for s in sc]
    jobs = native.Context.make_jobs(
        dict(scene=ids[j[0]], width=j[1], height=j[2], spp=j[3], max_depth=j[4], rr_start=j[5],
             seed=j[6], camera=sc[j[0]].camera) for j in JOBS)
    yield ctx, jobs
    ctx.close()


def _run(ctx, jobs, max_paths=0, graph=0, times=False):
    import torch

    total = int(sum(int(j["width"]) * int(j["height"]) for j in jobs))
    out = torch.full((total, 4), float("nan"), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    ctx.set_render_params(max_paths, graph)
    try:
        t = ctx.batch_render_ptr(jobs, out.data_ptr(), with_times=times)
    finally:
        ctx.set_render_params(0, 0)
    img = out.cpu().numpy()
    parts = []
    for j in jobs:
        o, n = int(j["out_offset"]), int(j["width"]) * int(j["height"])
        parts.append(img[o:o + n])
    return parts, t


@pytest.mark.parametrize("max_paths,graph", [(0, 0), (1000, 0), (777, 1), (5000, 1)])
def test_batch_equals_single_frames(batch_ctx, reference_images, max_paths, graph):
    ctx, jobs = batch_ctx
    parts, _ = _run(ctx, jobs, max_paths, graph)
    for k, (got, ref) in enumerate(zip(parts, reference_images)):
        assert np.array_equal(got, ref), f"job {k}"


def test_batch_graph_replays_repeatable(batch_ctx, reference_images):
    ctx, jobs = batch_ctx
    for _ in range(3):
        parts, _ = _run(ctx, jobs, 1500, 1)
        assert all(np.array_equal(g, r) for g, r in zip(parts, reference_images))


def test_batch_job_times(batch_ctx):
    ctx, jobs = batch_ctx
    _, t = _run(ctx, jobs, 2000, 0, times=True)
    paths = np.array([int(j["width"]) * int(j["height"]) * int(j["spp"]) for j in jobs])
    assert np.all(t > 0)
    order = np.argsort(paths, kind="stable")
    assert np.all(np.diff(t[order]) >= 0)  # shortest jobs complete first


def test_frame_graph_equals_direct(gpu_ctx):
    s = scenes.cornell(width=48, height=40)
    img0 = None
    for graph in (0, 1, 1):
        try:
            gpu_ctx.set_render_params(48 * 40 * 2, graph)  # 2 samples per pass: 4 passes
            img = _single(gpu_ctx, s, 48, 40, 8, 4, 3, 12)
        finally:
            gpu_ctx.set_render_params(0, 0)
        if img0 is None:
            img0 = img
        assert np.array_equal(img, img0)


@pytest.mark.parametrize("job", [1, 3])
def test_one_job_batch_clears_its_image(batch_ctx, reference_images, job):
    """A batch of one job renders its image whole: into a NaN-filled buffer,
    twice in a row into the same buffer, with and without graph replay, it
    equals the single-frame render bit for bit."""
    import torch

    ctx, jobs = batch_ctx
    one = jobs[job:job + 1].copy()
    one["out_offset"] = 0
    n = int(one[0]["width"]) * int(one[0]["height"])
    out = torch.full((n, 4), float("nan"), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    for max_paths, graph in ((0, 0), (0, 0), (n, 1), (n, 1)):
        ctx.set_render_params(max_paths, graph)
        try:
            ctx.batch_render_ptr(one, out.data_ptr())
        finally:
            ctx.set_render_params(0, 0)
        assert np.array_equal(out.cpu().numpy(), reference_images[job])
