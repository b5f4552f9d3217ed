"""The C++ drop-in class (tests/abi/aurora_renderer.hpp: MonteCarloRaytracer
over the C ABI, as the reference's src/render/monteCarloRaytracer.cpp:23-127
would call it) renders through libaurora_amd.so from a compiled program: its
framebuffers equal the Python binding's renders of the same scene bit for
bit, in compat and PBR mode, final and progressive (IPR) render.  The
consumer is built by __graft_entry__.build() (make -C tests/abi)."""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

from aurora_amd import native, scenes

pytestmark = pytest.mark.gpu

ABI_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "abi")


def _scene():
    v = np.array([[-5, -5, 5, 0, 0, 0, -5, 5, 5], [-5, 5, 5, 0, 0, 0, -5, 5, -5]], np.float32)
    n = np.zeros((2, 9), np.float32)
    n[:, 1::3] = 1.0
    sh = np.zeros(1, scenes.SHADER_DTYPE)
    sh[0]["diffuse_color"] = (0.5, 0.4, 0.3)
    sh[0]["diffuse"] = 1.0
    sh[0]["refractive_index"] = 1.0
    L = np.zeros(1, scenes.LIGHT_DTYPE)
    L[0]["type"] = scenes.LIGHT_AREA
    L[0]["intensity"] = 1.0
    L[0]["area"] = 1.0
    L[0]["samples"] = 1
    L[0]["position"] = (-0.5, 2.0, -0.5)
    L[0]["color"] = (10, 10, 10)
    L[0]["direction"] = (0, -1, 0)
    L[0]["e1"] = (1, 0, 0)
    L[0]["e2"] = (0, 0, 1)
    cam = np.zeros((), scenes.CAMERA_DTYPE)
    cam["position"] = (0.0, 3.0, -4.0)
    cam["forward"] = (0.0, -0.6, 0.8)
    cam["right"] = (-0.8, 0.0, 0.0)
    cam["up"] = (0.0, 0.64, 0.48)
    cam["tanfov"] = (0.6, 0.6)
    cam["aspect"] = 1.0
    return scenes.Scene(vertices=v, normals=n, texcoords=np.zeros((2, 6), np.float32),
                        shader_ids=np.ones(2, np.uint16), shaders=sh, lights=L, camera=cam, width=8, height=8)


def test_cpp_renderer_equals_python_binding(gpu_ctx, tmp_path):
    exe = os.path.join(ABI_DIR, "cpp_consumer")
    assert os.path.exists(exe), "build the consumers first: make -C tests/abi"
    out = tmp_path / "frames.f32"
    res = subprocess.run([exe, str(out)], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0 and "rendered" in res.stdout, res.stdout + res.stderr
    frames = np.fromfile(out, np.float32).reshape(4, 64, 4)
    s = _scene()
    gpu_ctx.upload_scene(s)
    gpu_ctx.rebuild_nmh()
    gpu_ctx.frame_create(8, 8, 4)
    k = 0
    for mode in (native.MODE_COMPAT, native.MODE_PBR):
        gpu_ctx.frame_render(s.camera, native.Context.opts(mode, 4, 8, 3, 5))
        img = gpu_ctx.frame_read()
        assert img[:, 3].min() == 1.0 and img[:, :3].max() > 0
        for _ipr in (False, True):
            assert np.array_equal(frames[k].view(np.uint32), img.view(np.uint32)), (mode, _ipr)
            k += 1
