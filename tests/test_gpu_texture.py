"""Texture maps on the GPU: per-pixel replay of textured renders (compat and
PBR) against the oracle within 1e-4 relative (bit-exact in practice), and the
entry points that do not take textured shaders say so."""
from __future__ import annotations

import numpy as np
import pytest

from aurora_amd import native, scenes

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", [native.MODE_COMPAT, native.MODE_PBR])
def test_textured_render_replay(gpu_ctx, oracle, mode):
    s, tex, texels = scenes.textured(scenes.cornell(width=48, height=48))
    gpu_ctx.upload_scene(s)
    gpu_ctx.upload_textures(tex, texels)
    gpu_ctx.rebuild_nmh()
    perm = gpu_ctx.read_permutation()
    gpu_ctx.frame_create(48, 48, 4)
    gpu_ctx.frame_render(s.camera, native.Context.opts(mode, 4, 4, 3, 13))
    img = gpu_ctx.frame_read()
    ref = oracle.render(oracle.tree_arrays(s, perm), s, mode, 4, 4, 3, 13, textures=(tex, texels))
    den = np.maximum(np.abs(ref), 1e-6)
    assert (np.abs(img - ref) / den).max() <= 1e-4
    # the textures change the image
    plain = oracle.render(oracle.tree_arrays(s, perm), s, mode, 4, 4, 3, 13)
    assert np.abs(plain - ref).max() > 1e-2
    gpu_ctx.upload_textures(np.zeros(0, scenes.TEXTURE_DTYPE), np.zeros(0, np.uint8))


def test_textured_shaders_rejected_where_unsupported(gpu_ctx):
    import torch

    s, tex, texels = scenes.textured(scenes.cornell(width=16, height=16))
    gpu_ctx.upload_scene(s)
    gpu_ctx.rebuild_nmh()
    gpu_ctx.frame_create(16, 16, 1)
    d = torch.zeros((256, 4), device="cuda")
    ga = torch.zeros((s.shaders.shape[0], 3), device="cuda")
    gl = torch.zeros((s.lights.shape[0], 3), device="cuda")
    with pytest.raises(native.AuroraError) as e:
        gpu_ctx.frame_adjoint_ptr(s.camera, native.Context.opts(native.MODE_PBR, 1, 2, 2, 1), d.data_ptr(),
                                  ga.data_ptr(), gl.data_ptr())
    assert e.value.status == 5  # AUR_ERR_UNSUPPORTED
    ctx = native.Context(0)
    try:
        with pytest.raises(native.AuroraError):
            ctx.scene_add(s)
    finally:
        ctx.close()
