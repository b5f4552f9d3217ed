"""The HIP compat path against closed forms, and per pixel against the oracle.

The analytic scenes of tests/test_oracle_compat_rays.py fix what the
reference's ray stack (src/kernels/monteCarloRaytracer.cu:23-112) gives for a
mirror floor in front of a lit wall, a mirror corridor that runs into the
depth limit, glass panes (refraction into and out of the medium, and total
internal reflection), a mirror-glass pane whose refraction ray is traced
first, and a textured floor (texture coordinates an affine map of position,
sampled from a ramp).  Here the GPU renders them through the C ABI; every
pixel must equal the closed form (relative 2e-6; 2e-5 for the textured floor)
and the oracle's pixel (relative 1e-4), in compat mode, and for the textured
floor in PBR mode too.
"""
from __future__ import annotations

import numpy as np
import pytest

from aurora_amd import native
from tests import test_oracle_compat_rays as cf

pytestmark = pytest.mark.gpu


def _render(ctx, s, mode, spp, textures=None):
    ctx.upload_scene(s)
    if textures is not None:
        ctx.upload_textures(*textures)
    ctx.rebuild_nmh()
    perm = ctx.read_permutation()
    ctx.frame_create(s.width, s.height, spp)
    ctx.frame_render(s.camera, native.Context.opts(mode, spp, 8, 3, 7))
    return ctx.frame_read(), perm


@pytest.mark.parametrize("name", list(cf.CLOSED_FORMS))
def test_compat_ray_stack_gpu_closed_form(gpu_ctx, oracle, name):
    s, expected, _ = cf.CLOSED_FORMS[name]()
    img, perm = _render(gpu_ctx, s, native.MODE_COMPAT, 16)
    assert np.all(img[:, 3] == 1.0)
    np.testing.assert_allclose(img[:, :3], np.broadcast_to(expected, img[:, :3].shape), rtol=cf.RTOL, atol=0)
    ref = oracle.render(oracle.tree_arrays(s, perm), s, native.MODE_COMPAT, 16, 8, 3, 7)
    assert (np.abs(img - ref) / np.maximum(np.abs(ref), 1e-6)).max() <= 1e-4


@pytest.mark.parametrize("mode", [native.MODE_COMPAT, native.MODE_PBR])
def test_textured_floor_gpu(gpu_ctx, oracle, mode):
    s = cf.textured_floor()
    tex = cf._ramp_texture()
    spp = 16 if mode == native.MODE_COMPAT else 64
    img, perm = _render(gpu_ctx, s, mode, spp, textures=tex)
    ref = oracle.render(oracle.tree_arrays(s, perm), s, mode, spp, 8, 3, 7, textures=tex)
    gpu_ctx.upload_textures(np.zeros(0, cf.scenes.TEXTURE_DTYPE), np.zeros(0, np.uint8))
    assert (np.abs(img - ref) / np.maximum(np.abs(ref), 1e-6)).max() <= 1e-4
    if mode == native.MODE_COMPAT:
        exp, _, _ = cf.textured_expected(s)
        np.testing.assert_allclose(img[:, :3], exp, rtol=2e-5)
