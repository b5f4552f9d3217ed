"""aur_set_alloc_hook: with PyTorch's caching allocator installed, every
buffer the library allocates comes from the hook and goes back to it when
the context is destroyed; renders equal those made with hipMalloc."""
from __future__ import annotations

import numpy as np
import pytest

from aurora_amd import native, scenes

pytestmark = pytest.mark.gpu


def _render(ctx, s):
    ctx.upload_scene(s)
    ctx.rebuild_nmh()
    ctx.frame_create(s.width, s.height, 4)
    ctx.frame_render(s.camera, native.Context.opts(native.MODE_PBR, 4, 4, 3, 11))
    return ctx.frame_read()


def test_alloc_hook_routes_every_buffer():
    import torch

    s = scenes.cornell(width=48, height=40)
    with native.Context(0) as ctx:
        ref = _render(ctx, s)
    talloc, tfree = native.torch_alloc_hook()
    live, log = set(), {"alloc": 0, "free": 0}

    def alloc(n):
        p = talloc(n)
        if p:
            assert p % 256 == 0
            live.add(p)
            log["alloc"] += 1
        return p

    def free(p):
        assert p in live, "freed a pointer the hook did not allocate"
        live.discard(p)
        log["free"] += 1
        tfree(p)

    native.set_alloc_hook(alloc, free)
    try:
        with native.Context(0) as ctx:
            img = _render(ctx, s)
            assert log["alloc"] > 10
        assert np.array_equal(img, ref)
        assert not live and log["free"] == log["alloc"]
    finally:
        native.set_alloc_hook()
    torch.cuda.synchronize()


def test_alloc_hook_out_of_memory_is_a_status():
    native.set_alloc_hook(lambda n: 0, lambda p: None)
    try:
        s = scenes.cornell(width=16, height=16)
        with native.Context(0) as ctx:
            with pytest.raises(native.AuroraError) as e:
                ctx.upload_scene(s)
                ctx.rebuild_nmh()
            assert e.value.status == 1  # AUR_ERR_OOM
    finally:
        native.set_alloc_hook()
