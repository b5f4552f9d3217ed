"""Host-side mirror of the reference's renderer layer over the C ABI.

Same names, argument meaning and status behaviour as
  MonteCarloRaytracer src/render/monteCarloRaytracer.h:14-40, .cpp:23-127
                      (and the Renderer interface, src/core/renderer.h:21-42)
  Scene::update       src/core/scene.cpp:398-449 (mesh/shader/light upload,
                      rebuild, generateTB)
with Maya's MStatus replaced by the Status values below and MDagPath by an
aur_camera record (src/util/camera.h:12-36).
"""
from __future__ import annotations

import dataclasses
import enum

import numpy as np

from . import native, scenes


class Status(enum.IntEnum):
    kSuccess = 0
    kFailure = 1
    kInsufficientMemory = 2


def _status_of(err: native.AuroraError) -> Status:
    return Status.kInsufficientMemory if err.status == 1 else Status.kFailure


@dataclasses.dataclass
class Rect:
    """src/util/math.h:283-297 (inclusive pixel bounds)."""

    left: int
    right: int
    bottom: int
    top: int


@dataclasses.dataclass
class Dim:
    """src/util/math.h:300-313; depth = samples per pixel."""

    width: int
    height: int
    depth: int = 1


class Scene:
    """Scene upload and NMH rebuild (Scene::update / updateMeshes,
    src/core/scene.cpp:85-240, 398-449)."""

    def __init__(self, ctx: native.Context):
        self.ctx = ctx
        self.build_stats = None
        self.source: scenes.Scene | None = None

    def update(self, scene: scenes.Scene, full: bool = True) -> Status:
        try:
            if full:
                self.ctx.upload_geometry(scene.vertices, scene.normals, scene.texcoords,
                                         scene.shader_ids)
            self.ctx.upload_shaders(scene.shaders)
            self.ctx.upload_lights(scene.lights)
            if full:
                self.build_stats = self.ctx.rebuild_nmh()  # rebuild + generateTB
            self.source = scene
        except native.AuroraError as e:
            return _status_of(e)
        return Status.kSuccess


class MonteCarloRaytracer:
    """src/render/monteCarloRaytracer.cpp:23-127 over aur_frame_*, with the
    Renderer interface it implements (src/core/renderer.h:21-42: createFrame,
    destroyFrame, setRegion, update, render, framebuffer, filterPixels).

    mode selects the reference-compatible estimator (native.MODE_COMPAT) or
    the path loop (native.MODE_PBR); seed keys the stateless random numbers
    (compat: Philox counters; PBR: Owen-scrambled Sobol points) that replace
    the tick-count-seeded cuRAND states (monteCarloRaytracer.cpp:54).
    """

    def __init__(self, ctx: native.Context, mode: int = native.MODE_COMPAT, max_depth: int = 4,
                 rr_start: int = 3, seed: int = 1):
        self.ctx = ctx
        self.mode, self.max_depth, self.rr_start, self.seed = mode, max_depth, rr_start, seed
        self.m_size: Dim | None = None
        self.m_region: Rect | None = None
        self.m_camera = None
        self.m_scene = None

    def createFrame(self, width, height, samples, scene, camera) -> Status:  # noqa: N802
        try:
            self.ctx.frame_create(width, height, samples)
        except native.AuroraError as e:
            return _status_of(e)
        self.m_scene = scene
        self.m_camera = camera
        self.m_region = Rect(0, width - 1, 0, height - 1)
        self.m_size = Dim(width, height, samples)
        return Status.kSuccess

    def destroyFrame(self) -> Status:  # noqa: N802
        try:
            self.ctx.frame_destroy()
        except native.AuroraError as e:
            return _status_of(e)
        self.m_size = None
        return Status.kSuccess

    def setRegion(self, region: Rect) -> Status:  # noqa: N802
        if self.m_size is None:
            return Status.kFailure
        try:
            self.ctx.frame_set_region(region.left, region.right, region.bottom, region.top)
        except native.AuroraError as e:
            return _status_of(e)
        self.m_region = region
        return Status.kSuccess

    def filterPixels(self) -> Status:  # noqa: N802
        """Renderer::filterPixels (src/core/renderer.cpp:43-46): 3x3 tent over
        the region, in place."""
        try:
            self.ctx.frame_filter()
        except native.AuroraError as e:
            return _status_of(e)
        return Status.kSuccess

    # -- progressive (IPR) rendering ------------------------------------------
    # The reference's IPR thread calls render(ipr=True) in a loop, restarting
    # from zero samples every time (src/core/engine.cpp:282-305).  Here the
    # accumulation persists: each step adds the next samples of a fixed
    # budget, and a checkpoint (accumulation + samples done) resumes to the
    # same image an uninterrupted render gives.

    def beginProgressive(self, total_samples: int) -> Status:  # noqa: N802
        if self.m_size is None:
            return Status.kFailure
        self.m_total = int(total_samples)
        self.m_done = 0
        try:
            self.ctx.frame_clear()
        except native.AuroraError as e:
            return _status_of(e)
        return Status.kSuccess

    def progressiveStep(self, samples: int = 1) -> Status:  # noqa: N802
        """Adds the next `samples` samples; the image is the running estimate
        weighted for the full budget (scale by total / done to view it)."""
        s0, s1 = self.m_done, min(self.m_done + int(samples), self.m_total)
        if s1 <= s0:
            return Status.kSuccess
        opts = native.Context.opts(self.mode, self.m_total, self.max_depth, self.rr_start, self.seed)
        try:
            self.ctx.frame_render_samples(self.m_camera, opts, s0, s1)
        except native.AuroraError as e:
            return _status_of(e)
        self.m_done = s1
        return Status.kSuccess

    def checkpoint(self) -> dict:
        return {"samples_done": self.m_done, "total": self.m_total, "rgba": self.ctx.frame_read()}

    def resume(self, cp: dict) -> Status:
        try:
            self.ctx.frame_write(cp["rgba"])
        except native.AuroraError as e:
            return _status_of(e)
        self.m_done, self.m_total = int(cp["samples_done"]), int(cp["total"])
        return Status.kSuccess

    def update(self) -> Status:
        """Renderer::update: nothing to refresh (scene data lives in the context)."""
        return Status.kSuccess

    def render(self, engine=None, ipr: bool = False) -> Status:
        if self.m_size is None:
            return Status.kFailure
        opts = native.Context.opts(self.mode, self.m_size.depth, self.max_depth, self.rr_start,
                                   self.seed)
        try:
            self.ctx.frame_render(self.m_camera, opts)
        except native.AuroraError as e:
            return _status_of(e)
        return Status.kSuccess

    def framebuffer(self) -> np.ndarray | None:
        """float RGBA (height, width, 4), row 0 = bottom; None on failure."""
        try:
            px = self.ctx.frame_read()
        except native.AuroraError:
            return None
        return px.reshape(self.m_size.height, self.m_size.width, 4)
