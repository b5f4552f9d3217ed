"""ctypes binding of the C ABI in include/aurora_amd.h (libaurora_amd.so).

The library is the only compute path: if it is missing or a call fails, an
exception is raised; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
import re

import numpy as np

from . import scenes

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libaurora_amd.so")
HEADER = os.path.join(os.path.dirname(PKG), "include", "aurora_amd.h")

AUR_OK = 0
AUR_ERR_CANCELLED = 4
STATUS_NAMES = {0: "OK", 1: "OOM", 2: "INVALID_ARG", 3: "DEVICE", 4: "CANCELLED", 5: "UNSUPPORTED",
                6: "STATE"}
MODE_COMPAT, MODE_PBR = 0, 1
ABI_VERSION = 3

RAY_DTYPE = np.dtype([("org", "<f4", 3), ("tmax", "<f4"), ("dir", "<f4", 3), ("reserved", "<f4")])
HIT_DTYPE = np.dtype([("id", "<i4"), ("t", "<f4"), ("u", "<f4"), ("v", "<f4")])


class AuroraError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"aurora_amd: {STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


class BuildStats(ctypes.Structure):
    _fields_ = [("count", ctypes.c_uint32), ("levels", ctypes.c_uint32), ("build_ms", ctypes.c_float)]


class RenderOpts(ctypes.Structure):
    _fields_ = [
        ("abi_version", ctypes.c_uint32),
        ("mode", ctypes.c_uint32),
        ("spp", ctypes.c_uint32),
        ("max_depth", ctypes.c_uint32),
        ("rr_start", ctypes.c_uint32),
        ("seed", ctypes.c_uint32),
    ]


class TraceParams(ctypes.Structure):
    _fields_ = [("persistent", ctypes.c_uint32), ("block", ctypes.c_uint32),
                ("refill", ctypes.c_uint32), ("waves_per_cu", ctypes.c_uint32),
                ("stack_lds_rows", ctypes.c_uint32)]


class RenderParams(ctypes.Structure):
    _fields_ = [("max_paths", ctypes.c_uint32), ("graph", ctypes.c_uint32),
                ("refill", ctypes.c_uint32), ("stack_lds_rows", ctypes.c_uint32)]


# aur_job (include/aurora_amd.h): 8 x u32, u64 out_offset, aur_camera, padded to 8 B
JOB_DTYPE = np.dtype({
    "names": ["abi_version", "scene", "width", "height", "spp", "max_depth", "rr_start", "seed",
              "out_offset", "camera"],
    "formats": [np.uint32] * 8 + [np.uint64, scenes.CAMERA_DTYPE],
    "offsets": [0, 4, 8, 12, 16, 20, 24, 28, 32, 40],
    "itemsize": 104,
})

# aur_emitter (include/aurora_amd.h): the reference's Emitter, 28 B
EMITTER_DTYPE = np.dtype([("area", "<f4"), ("pdf", "<f4"), ("cdf", "<f4"), ("power", "<f4", 3),
                          ("triangle", "<u4")])
assert EMITTER_DTYPE.itemsize == 28

# aur_transform (include/aurora_amd.h): the reference's Transform, 136 B
TRANSFORM_DTYPE = scenes.TRANSFORM_DTYPE

# aur_alloc_hook (include/aurora_amd.h)
ALLOC_FN = ctypes.CFUNCTYPE(ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p)
FREE_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_void_p)


class AllocHook(ctypes.Structure):
    _fields_ = [("alloc", ALLOC_FN), ("free", FREE_FN), ("user", ctypes.c_void_p)]


_lib = None
_hooks = []  # every installed hook: callbacks must outlive the buffers they made


def header_symbols() -> list[str]:
    """Every function declared by include/aurora_amd.h."""
    text = re.sub(r"/\*.*?\*/", " ", open(HEADER).read(), flags=re.S)
    return sorted(set(re.findall(r"\b(aur_[a-z0-9_]+)\s*\(", text)))


def load():
    """Loads libaurora_amd.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise AuroraError(5, f"{LIB_PATH} is missing: run __graft_entry__.build() or "
                             "python -m aurora_amd.build")
    L = ctypes.CDLL(LIB_PATH)
    P, u32, i32, u64 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int32, ctypes.c_uint64
    sig = {
        "aur_create": ([i32, ctypes.POINTER(P)], i32),
        "aur_destroy": ([P], None),
        "aur_last_error": ([P], ctypes.c_char_p),
        "aur_abi_version": ([], u32),
        "aur_stream": ([P], P),
        "aur_synchronize": ([P], i32),
        "aur_cancel": ([P], i32),
        "aur_set_alloc_hook": ([ctypes.POINTER(AllocHook)], i32),
        "aur_geometry_upload": ([P, u32, P, P, P, P], i32),
        "aur_shaders_upload": ([P, u32, P], i32),
        "aur_lights_upload": ([P, u32, P], i32),
        "aur_textures_upload": ([P, u32, P, P, u64], i32),
        "aur_rebuild_nmh": ([P, ctypes.POINTER(BuildStats)], i32),
        "aur_generate_tb": ([P], i32),
        "aur_read_permutation": ([P, P], i32),
        "aur_read_geometry": ([P, P, P, P, P, P, P], i32),
        "aur_geometry_count": ([P], u32),
        "aur_trace_closest": ([P, u64, P, P], i32),
        "aur_trace_any": ([P, u64, P, P], i32),
        "aur_trace_closest_stats": ([P, u64, P, P, ctypes.POINTER(u64), ctypes.POINTER(u64)], i32),
        "aur_check_division": ([P, u64, u64, ctypes.POINTER(u64), ctypes.POINTER(u64),
                                ctypes.POINTER(u64)], i32),
        "aur_set_trace_params": ([P, ctypes.POINTER(TraceParams)], i32),
        "aur_get_trace_params": ([P, ctypes.POINTER(TraceParams)], i32),
        "aur_set_render_params": ([P, ctypes.POINTER(RenderParams)], i32),
        "aur_get_render_params": ([P, ctypes.POINTER(RenderParams)], i32),
        "aur_frame_create": ([P, u32, u32, u32], i32),
        "aur_frame_destroy": ([P], i32),
        "aur_frame_render": ([P, P, ctypes.POINTER(RenderOpts)], i32),
        "aur_frame_render_samples": ([P, P, ctypes.POINTER(RenderOpts), u32, u32], i32),
        "aur_frame_clear": ([P], i32),
        "aur_frame_read": ([P, P], i32),
        "aur_frame_device_buffer": ([P, ctypes.POINTER(P)], i32),
        "aur_frame_rays_traced": ([P], u64),
        "aur_frame_ray_counts": ([P, ctypes.POINTER(u64), ctypes.POINTER(u64)], i32),
        "aur_frame_copy": ([P, P], i32),
        "aur_frame_set_region": ([P, u32, u32, u32, u32], i32),
        "aur_frame_get_region": ([P, ctypes.POINTER(u32), ctypes.POINTER(u32), ctypes.POINTER(u32),
                                  ctypes.POINTER(u32)], i32),
        "aur_frame_filter": ([P], i32),
        "aur_frame_write": ([P, P], i32),
        "aur_geometry_upload_objects": ([P, u32, P, P, P, P, u32, P], i32),
        "aur_scene_add": ([P, u32, P, P, P, u32, P, u32, P, ctypes.POINTER(u32)], i32),
        "aur_scene_clear": ([P], i32),
        "aur_scene_count": ([P], u32),
        "aur_batch_render": ([P, u32, P, P, P], i32),
        "aur_frame_adjoint": ([P, P, ctypes.POINTER(RenderOpts), P, P, P], i32),
        "aur_scene_bounds": ([P, P, P], i32),
        "aur_create_lights_cdf": ([P, P, ctypes.POINTER(u32)], i32),
        "aur_create_emitters": ([P, P, u64, ctypes.POINTER(u32)], i32),
        "aur_create_emitters_noalloc": ([P, P, u64, ctypes.POINTER(u32)], i32),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def set_alloc_hook(alloc=None, free=None) -> None:
    """Routes every later device allocation of the library through
    alloc(nbytes) -> int device pointer (0 when out of memory) and
    free(ptr); with no arguments restores hipMalloc / hipFree.  Buffers are
    freed by the allocator that made them."""
    lib = load()
    if alloc is None:
        st = lib.aur_set_alloc_hook(None)
        if st != AUR_OK:
            raise AuroraError(st, "aur_set_alloc_hook")
        return
    h = AllocHook(ALLOC_FN(lambda n, user: alloc(int(n)) or None), FREE_FN(lambda p, user: free(int(p))), None)
    st = lib.aur_set_alloc_hook(ctypes.byref(h))
    if st != AUR_OK:
        raise AuroraError(st, "aur_set_alloc_hook")
    _hooks.append(h)


def torch_alloc_hook():
    """(alloc, free) over PyTorch's caching allocator on the current device,
    for set_alloc_hook: the library's buffers then come from torch's pool."""
    import torch

    def alloc(n: int) -> int:
        try:
            return int(torch.cuda.caching_allocator_alloc(n, torch.cuda.current_device()))
        except RuntimeError:
            return 0

    def free(p: int) -> None:
        torch.cuda.caching_allocator_delete(p)

    return alloc, free


def _ptr(a: np.ndarray | None):
    if a is None:
        return None
    return ctypes.c_void_p(a.ctypes.data)


class Context:
    """One aur_context: a device, a stream and the scene/frame it owns."""

    def __init__(self, device: int = 0):
        self._lib = load()
        h = ctypes.c_void_p()
        st = self._lib.aur_create(device, ctypes.byref(h))
        if st != AUR_OK:
            raise AuroraError(st, f"aur_create(device={device}) failed")
        self._h = h
        self.count = 0
        self.num_lights = 0

    # -- plumbing --
    def _check(self, st: int, what: str) -> None:
        if st != AUR_OK:
            msg = self._lib.aur_last_error(self._h)
            raise AuroraError(st, f"{what}: {msg.decode() if msg else ''}")

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.aur_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self):
        return self._h

    @property
    def stream(self) -> int:
        return int(self._lib.aur_stream(self._h) or 0)

    def synchronize(self) -> None:
        self._check(self._lib.aur_synchronize(self._h), "synchronize")

    def cancel(self) -> None:
        """Stops the render running on this context (another thread) at its
        next pass boundary, or the next render at its first; that render
        raises AuroraError with status CANCELLED."""
        self._check(self._lib.aur_cancel(self._h), "cancel")

    # -- scene --
    def upload_geometry(self, vertices, normals=None, texcoords=None, shader_ids=None) -> None:
        v = np.ascontiguousarray(vertices, dtype=np.float32)
        n = v.shape[0]
        nrm = None if normals is None else np.ascontiguousarray(normals, dtype=np.float32)
        uv = None if texcoords is None else np.ascontiguousarray(texcoords, dtype=np.float32)
        ids = None if shader_ids is None else np.ascontiguousarray(shader_ids, dtype=np.uint16)
        self._check(self._lib.aur_geometry_upload(self._h, n, _ptr(v), _ptr(nrm), _ptr(uv), _ptr(ids)),
                    "geometry upload")
        self.count = n

    def upload_objects(self, vertices, normals, transforms, texcoords=None, shader_ids=None) -> None:
        """Object-space triangles + per-object transforms (TRANSFORM_DTYPE);
        an odd count is padded to even on upload."""
        v = np.ascontiguousarray(vertices, np.float32)
        n = None if normals is None else np.ascontiguousarray(normals, np.float32)
        uv = None if texcoords is None else np.ascontiguousarray(texcoords, np.float32)
        ids = None if shader_ids is None else np.ascontiguousarray(shader_ids, np.uint16)
        xf = np.ascontiguousarray(transforms, dtype=TRANSFORM_DTYPE)
        self._check(self._lib.aur_geometry_upload_objects(self._h, v.shape[0], _ptr(v), _ptr(n), _ptr(uv),
                                                          _ptr(ids), xf.shape[0],
                                                          _ptr(xf) if xf.shape[0] else None),
                    "upload_objects")
        self.count = v.shape[0] + (v.shape[0] & 1)

    def upload_textures(self, textures: np.ndarray, texels: np.ndarray) -> None:
        """textures: scenes.TEXTURE_DTYPE records; texels: uint8 RGBA bytes."""
        t = np.ascontiguousarray(textures, dtype=scenes.TEXTURE_DTYPE)
        b = np.ascontiguousarray(texels, dtype=np.uint8)
        self._check(self._lib.aur_textures_upload(self._h, t.shape[0], _ptr(t) if t.shape[0] else None,
                                                  _ptr(b) if b.size else None, b.size), "upload_textures")

    def upload_shaders(self, shaders: np.ndarray) -> None:
        s = np.ascontiguousarray(shaders, dtype=scenes.SHADER_DTYPE)
        self._check(self._lib.aur_shaders_upload(self._h, s.shape[0], _ptr(s)), "shaders upload")

    def upload_lights(self, lights: np.ndarray) -> None:
        li = np.ascontiguousarray(lights, dtype=scenes.LIGHT_DTYPE)
        self._check(self._lib.aur_lights_upload(self._h, li.shape[0], _ptr(li) if li.shape[0] else None),
                    "lights upload")
        self.num_lights = li.shape[0]

    def upload_scene(self, scene: scenes.Scene) -> None:
        self.upload_geometry(scene.vertices, scene.normals, scene.texcoords, scene.shader_ids)
        self.upload_shaders(scene.shaders)
        self.upload_lights(scene.lights)

    def rebuild_nmh(self) -> BuildStats:
        st = BuildStats()
        self._check(self._lib.aur_rebuild_nmh(self._h, ctypes.byref(st)), "rebuild_nmh")
        return st

    def generate_tb(self) -> None:
        self._check(self._lib.aur_generate_tb(self._h), "generate_tb")

    def read_permutation(self) -> np.ndarray:
        out = np.empty(self.count, dtype=np.uint32)
        self._check(self._lib.aur_read_permutation(self._h, _ptr(out)), "read_permutation")
        return out

    def read_geometry(self) -> dict:
        n = self.count
        out = {
            "vertices": np.empty((n, 9), np.float32),
            "normals": np.empty((n, 9), np.float32),
            "texcoords": np.empty((n, 6), np.float32),
            "shader_ids": np.empty(n, np.uint16),
            "tangents": np.empty((n, 9), np.float32),
            "bitangents": np.empty((n, 9), np.float32),
        }
        self._check(self._lib.aur_read_geometry(self._h, *(_ptr(out[k]) for k in (
            "vertices", "normals", "texcoords", "shader_ids", "tangents", "bitangents"))),
            "read_geometry")
        return out

    # -- light sampling tables --
    def scene_bounds(self) -> tuple[np.ndarray, np.ndarray]:
        lo, hi = np.zeros(3, np.float32), np.zeros(3, np.float32)
        self._check(self._lib.aur_scene_bounds(self._h, _ptr(lo), _ptr(hi)), "scene_bounds")
        return lo, hi

    def lights_cdf_ptr(self, cdf_ptr: int) -> int:
        """Writes the light-power CDF to the device buffer; returns the count."""
        n = ctypes.c_uint32()
        self._check(self._lib.aur_create_lights_cdf(self._h, ctypes.c_void_p(cdf_ptr), ctypes.byref(n)),
                    "create_lights_cdf")
        return int(n.value)

    def emitters_ptr(self, out_ptr: int, capacity: int) -> int:
        """Writes the emitter table (EMITTER_DTYPE records) to the device
        buffer; returns the emitter count."""
        n = ctypes.c_uint32()
        self._check(self._lib.aur_create_emitters(self._h, ctypes.c_void_p(out_ptr), capacity, ctypes.byref(n)),
                    "create_emitters")
        return int(n.value)

    def emitters_noalloc_ptr(self, out_ptr: int, capacity: int) -> int:
        """aur_create_emitters_noalloc: the same table with the library's
        fixed device scratch (no device allocation); returns the count."""
        n = ctypes.c_uint32()
        self._check(self._lib.aur_create_emitters_noalloc(self._h, ctypes.c_void_p(out_ptr), capacity,
                                                          ctypes.byref(n)), "create_emitters_noalloc")
        return int(n.value)

    def lights_cdf(self) -> np.ndarray:
        import torch

        buf = torch.zeros(max(1, self.num_lights), dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        n = self.lights_cdf_ptr(buf.data_ptr())
        return buf[:n].cpu().numpy()

    def emitters(self, capacity: int | None = None) -> np.ndarray:
        import torch

        cap = self.count if capacity is None else capacity
        buf = torch.zeros(max(1, cap) * EMITTER_DTYPE.itemsize, dtype=torch.uint8, device="cuda")
        torch.cuda.synchronize()
        n = self.emitters_ptr(buf.data_ptr(), cap)
        return buf[: n * EMITTER_DTYPE.itemsize].cpu().numpy().view(EMITTER_DTYPE)

    # -- traversal (device pointers) --
    def trace_closest_ptr(self, n: int, rays_ptr: int, hits_ptr: int) -> None:
        self._check(self._lib.aur_trace_closest(self._h, n, ctypes.c_void_p(rays_ptr),
                                                ctypes.c_void_p(hits_ptr)), "trace_closest")

    def set_trace_params(self, persistent: int = 1, block: int = 256, refill: int = 24,
                         waves_per_cu: int = 0, stack_lds_rows: int = 0) -> None:
        p = TraceParams(persistent, block, refill, waves_per_cu, stack_lds_rows)
        self._check(self._lib.aur_set_trace_params(self._h, ctypes.byref(p)), "set_trace_params")

    def get_trace_params(self) -> dict:
        p = TraceParams()
        self._check(self._lib.aur_get_trace_params(self._h, ctypes.byref(p)), "get_trace_params")
        return {k: getattr(p, k) for k, _ in TraceParams._fields_}

    def set_render_params(self, max_paths: int = 0, graph: int = 0, refill: int = 0,
                          stack_lds_rows: int = 0) -> None:
        p = RenderParams(max_paths, graph, refill, stack_lds_rows)
        self._check(self._lib.aur_set_render_params(self._h, ctypes.byref(p)), "set_render_params")

    def get_render_params(self) -> dict:
        p = RenderParams()
        self._check(self._lib.aur_get_render_params(self._h, ctypes.byref(p)), "get_render_params")
        return {"max_paths": p.max_paths, "graph": p.graph, "refill": p.refill,
                "stack_lds_rows": p.stack_lds_rows}

    def trace_closest_stats_ptr(self, n: int, rays_ptr: int, hits_ptr: int) -> tuple[int, int]:
        p, q = ctypes.c_uint64(), ctypes.c_uint64()
        self._check(self._lib.aur_trace_closest_stats(self._h, n, ctypes.c_void_p(rays_ptr),
                                                      ctypes.c_void_p(hits_ptr), ctypes.byref(p),
                                                      ctypes.byref(q)), "trace_closest_stats")
        return int(p.value), int(q.value)

    def check_division(self, samples: int, seed: int = 1) -> tuple[int, int, int]:
        """(floats checked, reciprocal mismatches, quotient mismatches) of the
        traversal's division shortcuts against true division, on the device."""
        c, r, q = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        self._check(self._lib.aur_check_division(self._h, samples, seed, ctypes.byref(c),
                                                 ctypes.byref(r), ctypes.byref(q)),
                    "check_division")
        return int(c.value), int(r.value), int(q.value)

    def trace_any_ptr(self, n: int, rays_ptr: int, ids_ptr: int) -> None:
        self._check(self._lib.aur_trace_any(self._h, n, ctypes.c_void_p(rays_ptr),
                                            ctypes.c_void_p(ids_ptr)), "trace_any")

    def trace(self, rays: np.ndarray, any_hit: bool = False):
        """Host convenience: copies rays (n, 8) float32 to the device with torch,
        traces, and returns (ids int32 (n,), t/u/v float32 (n, 3))."""
        import torch

        r = torch.from_numpy(np.ascontiguousarray(rays, dtype=np.float32)).to("cuda")
        n = r.shape[0]
        if any_hit:
            ids = torch.empty(n, dtype=torch.int32, device="cuda")
            self.trace_any_ptr(n, r.data_ptr(), ids.data_ptr())
            self.synchronize()
            return ids.cpu().numpy(), None
        hits = torch.empty((n, 4), dtype=torch.float32, device="cuda")
        self.trace_closest_ptr(n, r.data_ptr(), hits.data_ptr())
        self.synchronize()
        h = hits.cpu().numpy()
        return h[:, 0].view(np.int32).copy(), h[:, 1:4].copy()

    # -- frame --
    def frame_create(self, width: int, height: int, spp: int) -> None:
        self._check(self._lib.aur_frame_create(self._h, width, height, spp), "frame_create")
        self.width, self.height = width, height

    def frame_destroy(self) -> None:
        self._check(self._lib.aur_frame_destroy(self._h), "frame_destroy")

    def frame_clear(self) -> None:
        self._check(self._lib.aur_frame_clear(self._h), "frame_clear")

    @staticmethod
    def opts(mode: int, spp: int, max_depth: int = 4, rr_start: int = 3, seed: int = 1) -> RenderOpts:
        return RenderOpts(ABI_VERSION, mode, spp, max_depth, rr_start, seed)

    def frame_render(self, camera: np.ndarray, opts: RenderOpts) -> None:
        cam = np.ascontiguousarray(camera, dtype=scenes.CAMERA_DTYPE)
        self._check(self._lib.aur_frame_render(self._h, _ptr(cam), ctypes.byref(opts)), "frame_render")

    def frame_render_samples(self, camera: np.ndarray, opts: RenderOpts, s0: int, s1: int) -> None:
        cam = np.ascontiguousarray(camera, dtype=scenes.CAMERA_DTYPE)
        self._check(self._lib.aur_frame_render_samples(self._h, _ptr(cam), ctypes.byref(opts), s0, s1),
                    "frame_render_samples")

    def frame_read(self) -> np.ndarray:
        out = np.empty((self.height * self.width, 4), dtype=np.float32)
        self._check(self._lib.aur_frame_read(self._h, _ptr(out)), "frame_read")
        return out

    def frame_set_region(self, left: int, right: int, bottom: int, top: int) -> None:
        self._check(self._lib.aur_frame_set_region(self._h, left, right, bottom, top), "frame_set_region")

    def frame_get_region(self) -> tuple[int, int, int, int]:
        v = [ctypes.c_uint32() for _ in range(4)]
        self._check(self._lib.aur_frame_get_region(self._h, *(ctypes.byref(x) for x in v)), "frame_get_region")
        return tuple(int(x.value) for x in v)

    def frame_filter(self) -> None:
        self._check(self._lib.aur_frame_filter(self._h), "frame_filter")

    def frame_write(self, rgba: np.ndarray) -> None:
        a = np.ascontiguousarray(rgba, dtype=np.float32)
        if a.size != self.width * self.height * 4:
            raise ValueError("frame_write: need width*height*4 floats")
        self._check(self._lib.aur_frame_write(self._h, _ptr(a)), "frame_write")

    def frame_device_buffer(self) -> int:
        p = ctypes.c_void_p()
        self._check(self._lib.aur_frame_device_buffer(self._h, ctypes.byref(p)), "frame_device_buffer")
        return int(p.value)

    def rays_traced(self) -> int:
        return int(self._lib.aur_frame_rays_traced(self._h))

    def ray_counts(self) -> tuple[int, int]:
        """(camera + extension rays, shadow rays) of the last render."""
        e, sh = ctypes.c_uint64(), ctypes.c_uint64()
        self._check(self._lib.aur_frame_ray_counts(self._h, ctypes.byref(e), ctypes.byref(sh)), "ray_counts")
        return int(e.value), int(sh.value)

    # ---- batches ----
    def scene_add(self, scene: scenes.Scene) -> int:
        """Builds and keeps a scene for batches; returns its id."""
        v = np.ascontiguousarray(scene.vertices, np.float32)
        n = np.ascontiguousarray(scene.normals, np.float32)
        ids = np.ascontiguousarray(scene.shader_ids, np.uint16)
        sh = np.ascontiguousarray(scene.shaders)
        li = np.ascontiguousarray(scene.lights)
        out = ctypes.c_uint32()
        self._check(self._lib.aur_scene_add(self._h, v.shape[0], _ptr(v), _ptr(n), _ptr(ids), sh.shape[0],
                                            _ptr(sh), li.shape[0], _ptr(li) if li.shape[0] else None,
                                            ctypes.byref(out)), "scene_add")
        self.count = 0
        self.num_lights = 0
        return int(out.value)

    def scene_clear(self) -> None:
        self._check(self._lib.aur_scene_clear(self._h), "scene_clear")

    def scene_count(self) -> int:
        return int(self._lib.aur_scene_count(self._h))

    @staticmethod
    def make_jobs(specs) -> np.ndarray:
        """specs: iterable of dicts with scene, width, height, spp, max_depth,
        rr_start, seed, camera; out_offset is assigned densely in order."""
        specs = list(specs)
        jobs = np.zeros(len(specs), dtype=JOB_DTYPE)
        off = 0
        for i, sp in enumerate(specs):
            for k in ("scene", "width", "height", "spp", "max_depth", "rr_start", "seed"):
                jobs[i][k] = sp[k]
            jobs[i]["abi_version"] = ABI_VERSION
            jobs[i]["camera"] = sp["camera"]
            jobs[i]["out_offset"] = off
            off += int(sp["width"]) * int(sp["height"])
        return jobs

    def batch_render_ptr(self, jobs: np.ndarray, rgba_ptr: int, with_times: bool = False):
        """Renders the jobs into the device buffer at rgba_ptr; returns the
        per-job completion times (ms) when with_times."""
        j = np.ascontiguousarray(jobs, dtype=JOB_DTYPE)
        times = np.zeros(j.shape[0], np.float32) if with_times else None
        self._check(self._lib.aur_batch_render(self._h, j.shape[0], _ptr(j), ctypes.c_void_p(rgba_ptr),
                                               _ptr(times) if times is not None else None), "batch_render")
        return times

    def frame_copy_ptr(self, rgba_ptr: int) -> None:
        """Copies the float4 accumulation buffer to device memory at rgba_ptr."""
        self._check(self._lib.aur_frame_copy(self._h, ctypes.c_void_p(rgba_ptr)), "frame_copy")

    def frame_adjoint_ptr(self, camera: np.ndarray, opts: RenderOpts, d_image_ptr: int,
                          grad_albedo_ptr: int, grad_radiance_ptr: int) -> None:
        """Device pointers: dLoss/dI float4 per pixel in, gradients out."""
        cam = np.ascontiguousarray(camera, dtype=scenes.CAMERA_DTYPE)
        self._check(self._lib.aur_frame_adjoint(self._h, _ptr(cam), ctypes.byref(opts),
                                                ctypes.c_void_p(d_image_ptr),
                                                ctypes.c_void_p(grad_albedo_ptr),
                                                ctypes.c_void_p(grad_radiance_ptr or None)),
                    "frame_adjoint")
