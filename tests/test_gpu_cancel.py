"""aur_cancel: a render running on the context (another thread) stops at a
pass boundary with status CANCELLED; a request made while nothing runs
cancels the next render; the request is consumed, so the render after it
runs whole and equals an uninterrupted render."""
from __future__ import annotations

import threading
import time

import numpy as np
import pytest

from aurora_amd import native, scenes

pytestmark = pytest.mark.gpu


def _setup(ctx, s):
    ctx.upload_scene(s)
    ctx.rebuild_nmh()


@pytest.mark.parametrize("mode", [native.MODE_COMPAT, native.MODE_PBR])
def test_cancel_before_render_is_consumed(gpu_ctx, mode):
    s = scenes.cornell(width=32, height=32)
    _setup(gpu_ctx, s)
    o = native.Context.opts(mode, 4, 4, 3, 9)
    gpu_ctx.frame_create(32, 32, 4)
    gpu_ctx.frame_render(s.camera, o)
    full = gpu_ctx.frame_read()
    gpu_ctx.cancel()
    with pytest.raises(native.AuroraError) as e:
        gpu_ctx.frame_render(s.camera, o)
    assert e.value.status == native.AUR_ERR_CANCELLED
    gpu_ctx.frame_render(s.camera, o)  # the request was consumed
    assert np.array_equal(gpu_ctx.frame_read(), full)


@pytest.mark.parametrize("mode", [native.MODE_COMPAT, native.MODE_PBR])
def test_cancel_running_render(gpu_ctx, mode):
    s = scenes.cornell(width=128, height=128)
    _setup(gpu_ctx, s)
    try:
        gpu_ctx.set_render_params(128 * 128)  # one sample per pass: many pass boundaries
        gpu_ctx.frame_create(128, 128, 64)
        t0 = time.perf_counter()
        gpu_ctx.frame_render(s.camera, native.Context.opts(mode, 64, 4, 3, 5))
        t64 = time.perf_counter() - t0  # time of 64 passes
        spp = int(min(1 << 15, 64 * max(8, int(np.ceil(2.0 / t64)))))  # about 2 s or more
        gpu_ctx.frame_create(128, 128, spp)
        o = native.Context.opts(mode, spp, 4, 3, 5)
        out = {}

        def run():
            t = time.perf_counter()
            try:
                gpu_ctx.frame_render(s.camera, o)
                out["status"] = native.AUR_OK
            except native.AuroraError as err:
                out["status"] = err.status
            out["t"] = time.perf_counter() - t

        th = threading.Thread(target=run)
        th.start()
        time.sleep(0.1)
        gpu_ctx.cancel()
        th.join(timeout=120)
        assert not th.is_alive()
        full = t64 * spp / 64
        print(f"64 passes {t64:.3f}s, {spp} spp expected {full:.2f}s, cancelled after {out['t']:.3f}s",
              flush=True)
        assert out["status"] == native.AUR_ERR_CANCELLED
        assert out["t"] < 0.5 * full  # stopped well before the end
        gpu_ctx.frame_create(128, 128, 4)
        gpu_ctx.frame_render(s.camera, native.Context.opts(mode, 4, 4, 3, 5))  # usable afterwards
    finally:
        gpu_ctx.set_render_params(0)
