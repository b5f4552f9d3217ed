"""Trees whose compact node array would exceed 128 MB keep one node record per
128-byte line.  A 600-sphere scene has 3,072,002 triangles (1,536,001 nodes:
147 MB of compact records), so it takes that layout.  Its renders replay the
oracle per pixel in both modes, and a batch that mixes it with a small
(compact-record) scene gives each job its single-frame image bit for bit."""
from __future__ import annotations

import numpy as np
import pytest

from aurora_amd import native, scenes

pytestmark = pytest.mark.gpu

W, H = 32, 24


@pytest.fixture(scope="module")
def big():
    s = scenes.spheres(seed=9, n_spheres=600, width=W, height=H)
    assert s.count == 3_072_002 and (s.count // 2 + 1) * 96 > 128 << 20
    return s


def _render(ctx, s, mode, spp, depth, seed, w=W, h=H):
    ctx.frame_create(w, h, spp)
    ctx.frame_render(s.camera, native.Context.opts(mode, spp, depth, 2, seed))
    return ctx.frame_read()


@pytest.mark.parametrize("mode", [native.MODE_COMPAT, native.MODE_PBR])
def test_wide_records_replay_oracle(gpu_ctx, oracle, big, mode):
    gpu_ctx.upload_scene(big)
    gpu_ctx.rebuild_nmh()
    perm = gpu_ctx.read_permutation()
    img = _render(gpu_ctx, big, mode, 4, 3, 11)
    ref = oracle.render(oracle.tree_arrays(big, perm), big, mode, 4, 3, 2, 11)
    assert ref[:, 3].max() > 0 and ref[:, :3].max() > 0
    den = np.maximum(np.abs(ref), 1e-6)
    assert (np.abs(img - ref) / den).max() <= 1e-4


def test_batch_mixes_record_layouts(gpu_ctx, big):
    import torch

    small = scenes.cornell(seed=4, n_spheres=2, subdivisions=1, n_lights=2)
    specs = [(big, W, H, 4, 4, 21), (small, 20, 16, 3, 5, 22), (big, 16, 16, 2, 3, 23)]
    singles = []
    for s, w, h, spp, depth, seed in specs:
        gpu_ctx.upload_scene(s)
        gpu_ctx.rebuild_nmh()
        singles.append(_render(gpu_ctx, s, native.MODE_PBR, spp, depth, seed, w, h))
    ctx = native.Context(0)
    try:
        ids = {id(big): ctx.scene_add(big), id(small): ctx.scene_add(small)}
        jobs = native.Context.make_jobs(
            dict(scene=ids[id(s)], width=w, height=h, spp=spp, max_depth=depth, rr_start=2, seed=seed,
                 camera=s.camera) for s, w, h, spp, depth, seed in specs)
        total = sum(w * h for _, w, h, *_ in specs)
        out = torch.full((total, 4), float("nan"), dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        ctx.batch_render_ptr(jobs, out.data_ptr())
        img = out.cpu().numpy()
    finally:
        ctx.close()
    for j, ref in zip(jobs, singles):
        o, n = int(j["out_offset"]), int(j["width"]) * int(j["height"])
        assert np.array_equal(img[o:o + n].view(np.uint32), ref.view(np.uint32))
