"""Point, directional and ambient lights, alone and mixed with area lights, on
an open scene (20 spheres over a ground quad, so a directional light reaches
part of it unoccluded).  Per-pixel replay against the oracle in both modes
within 1e-4 relative; the adjoint's albedo and light gradients against the
oracle's within 1e-4 of the largest, and each light's within 1e-3 of its
own largest component; each delta light's radiance gradient against central
differences of the GPU forward render, 1%.  Light gradients are with respect
to radiance (colour * intensity), so the intensities are not 1 here.  A batch
whose scenes take different next-event slot counts gives every job its
single-frame image bit for bit.
"""
from __future__ import annotations

import numpy as np
import pytest

from aurora_amd import native, scenes

pytestmark = pytest.mark.gpu

W, H = 32, 24


def _delta_lights():
    return {
        "point": scenes.point_light((40.0, 35.0, 30.0), (1.0, 0.9, 0.7), 900.0),
        "directional": scenes.directional_light((0.3, -1.0, 0.5), (0.5, 0.45, 0.4)),
        "ambient": scenes.ambient_light((70.0, 80.0, 40.0), (0.9, 0.95, 1.0), 0.3, 1.2),
    }


def _scene(which):
    s = scenes.spheres(seed=5, n_spheres=20, subdivisions=2, width=W, height=H)
    dl = _delta_lights()
    if which == "mixed":
        s.lights = np.concatenate([s.lights[:2], np.array(list(dl.values()), dtype=scenes.LIGHT_DTYPE)])
    else:
        s.lights = np.array([dl[which]], dtype=scenes.LIGHT_DTYPE)
    return s


def _setup(ctx, s, spp):
    ctx.upload_scene(s)
    ctx.rebuild_nmh()
    ctx.frame_create(s.width, s.height, spp)
    return ctx.read_permutation()


@pytest.mark.parametrize("mode", [native.MODE_COMPAT, native.MODE_PBR])
@pytest.mark.parametrize("which", ["point", "directional", "ambient", "mixed"])
def test_light_type_replays_oracle(gpu_ctx, oracle, mode, which):
    s = _scene(which)
    perm = _setup(gpu_ctx, s, 4)
    gpu_ctx.frame_render(s.camera, native.Context.opts(mode, 4, 3, 2, 13))
    img = gpu_ctx.frame_read()
    ref = oracle.render(oracle.tree_arrays(s, perm), s, mode, 4, 3, 2, 13)
    lit = ref[:, :3].max(axis=1)
    assert (lit > 1e-3).mean() > 0.3  # the light reaches a good share of the pixels
    den = np.maximum(np.abs(ref), 1e-6)
    assert (np.abs(img - ref) / den).max() <= 1e-4


def _dimg(npix, seed):
    d = np.random.default_rng(seed).uniform(-1.0, 1.0, (npix, 4)).astype(np.float32)
    d[:, 3] = 0.0
    return d


def _adjoint(ctx, s, spp, depth, rr, seed, d):
    import torch

    dt = torch.from_numpy(d).cuda()
    ga = torch.empty((s.shaders.shape[0], 3), dtype=torch.float32, device="cuda")
    gl = torch.empty((s.lights.shape[0], 3), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    ctx.frame_adjoint_ptr(s.camera, native.Context.opts(native.MODE_PBR, spp, depth, rr, seed),
                          dt.data_ptr(), ga.data_ptr(), gl.data_ptr())
    return ga.cpu().numpy(), gl.cpu().numpy()


def test_mixed_lights_adjoint_matches_oracle(gpu_ctx, oracle):
    s = _scene("mixed")
    perm = _setup(gpu_ctx, s, 8)
    d = _dimg(W * H, 3)
    ga, gl = _adjoint(gpu_ctx, s, 8, 4, 2, 19, d)
    ra, rl = oracle.adjoint(oracle.tree_arrays(s, perm), s, 8, 4, 2, 19, d)
    scale = max(np.abs(ra).max(), np.abs(rl).max())
    assert scale > 1e-2
    assert np.abs(ga - ra).max() <= 1e-4 * scale
    assert np.abs(gl - rl).max() <= 1e-4 * scale
    own = np.abs(rl).max(axis=1, keepdims=True)
    assert (own > 0).all()
    assert (np.abs(gl - rl) <= 1e-3 * own).all()


def test_delta_light_gradients_match_central_differences(gpu_ctx):
    """The image is linear in each light's radiance R = colour * intensity,
    so the central difference of sum <d, I> along R (colour scaled by 1 +- h,
    h = 1e-3, common random numbers) equals <gl, R>."""
    s = _scene("mixed")
    spp, depth, seed, h = 8, 4, 23, 1e-3
    _setup(gpu_ctx, s, spp)
    d = _dimg(W * H, 7)
    _, gl = _adjoint(gpu_ctx, s, spp, depth, depth, seed, d)

    def loss(lights):
        gpu_ctx.upload_lights(lights)
        gpu_ctx.frame_render(s.camera, native.Context.opts(native.MODE_PBR, spp, depth, depth, seed))
        return float((gpu_ctx.frame_read()[:, :3].astype(np.float64) * d[:, :3]).sum())

    try:
        for l in range(2, s.lights.shape[0]):
            lp, lm = s.lights.copy(), s.lights.copy()
            lp["color"][l] *= 1 + h
            lm["color"][l] *= 1 - h
            fd = (loss(lp) - loss(lm)) / (2 * h)
            radiance = s.lights["color"][l].astype(np.float64) * float(s.lights["intensity"][l])
            want = float(np.dot(gl[l].astype(np.float64), radiance))
            assert abs(want) > 1e-3
            assert fd == pytest.approx(want, rel=1e-2), l
    finally:
        gpu_ctx.upload_lights(s.lights)


def test_batch_with_delta_lights_equals_single_renders(gpu_ctx):
    """Jobs whose scenes take different next-event slot counts (mixed lights:
    2 area + point + directional + ambient = 6 slots; ambient alone = 2;
    a Cornell box with 2 area lights = 2) give their single-frame images."""
    import torch

    mixed, amb = _scene("mixed"), _scene("ambient")
    box = scenes.cornell(seed=4, n_spheres=2, subdivisions=1, n_lights=2)
    specs = [(mixed, W, H, 4, 4, 41), (amb, 20, 16, 3, 3, 42), (box, 16, 16, 2, 5, 43), (mixed, 16, 8, 5, 2, 44)]
    singles = []
    for s, w, h, spp, depth, seed in specs:
        gpu_ctx.upload_scene(s)
        gpu_ctx.rebuild_nmh()
        gpu_ctx.frame_create(w, h, spp)
        gpu_ctx.frame_render(s.camera, native.Context.opts(native.MODE_PBR, spp, depth, 2, seed))
        singles.append(gpu_ctx.frame_read())
    ctx = native.Context(0)
    try:
        ids = {}
        for s, *_ in specs:
            if id(s) not in ids:
                ids[id(s)] = ctx.scene_add(s)
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
        assert ref[:, :3].max() > 0
        assert np.array_equal(img[o:o + n].view(np.uint32), ref.view(np.uint32))
