"""ctypes bindings to the CPU oracle (oracle/liboracle.so).

The oracle is the checker: tests, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use it; the product path never does.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "liboracle.so")

_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def use_native() -> str:
    """Builds the oracle with -march=native for this machine
    (oracle/_native/liboracle.so) and binds it in place of the portable
    build; the CPU baseline times this one.  Returns the library path."""
    global _lib
    subprocess.run(["make", "-s", "-B", "-C", ORACLE_DIR, "native"], check=True)
    _lib = None
    return lib(os.path.join(ORACLE_DIR, "_native", "liboracle.so"))._name


def cpu_model() -> str:
    """The host CPU's model name, with its family/model numbers and the
    number of logical CPUs from /proc/cpuinfo (a virtual machine may report
    only a short name)."""
    info: dict[str, str] = {}
    ncpu = 0
    try:
        for line in open("/proc/cpuinfo"):
            k, _, v = line.partition(":")
            k = k.strip()
            if k == "processor":
                ncpu += 1
            elif k in ("model name", "cpu family", "model") and k not in info:
                info[k] = v.strip()
    except OSError:
        return "unknown"
    name = info.get("model name", "unknown")
    if "cpu family" in info and "model" in info:
        name += f" (family {info['cpu family']}, model {info['model']})"
    return f"{name}, {ncpu} logical CPUs" if ncpu else name


def lib(path: str | None = None):
    global _lib
    if _lib is None:
        if path is None and not os.path.exists(ORACLE_SO):
            build()
        L = ctypes.CDLL(path or ORACLE_SO)
        P = ctypes.c_void_p
        u32, i32, u64 = ctypes.c_uint32, ctypes.c_int, ctypes.c_uint64
        L.orc_nmh_build.argtypes = [u32, P, P]
        L.orc_nmh_build.restype = i32
        L.orc_nmh_build_mt.argtypes = [u32, P, P, i32]
        L.orc_nmh_build_mt.restype = i32
        L.orc_nmh_levels.argtypes = [u32]
        L.orc_nmh_levels.restype = u32
        L.orc_calc_partition.argtypes = [u32, P, P]
        L.orc_trace.argtypes = [u32, P, u64, P, i32, P, P, i32]
        L.orc_trace.restype = i32
        L.orc_ray_triangle.argtypes = [P, P, P]
        L.orc_ray_triangle.restype = i32
        L.orc_generate_tb.argtypes = [u32, P, P, P]
        L.orc_camera_ray.argtypes = [P, u32, u32, u32, u32, u32, u32, P]
        L.orc_render.argtypes = [u32, P, P, P, P, P, u32, P, u32, P, P, u32, u32, P, P, u32, P, i32]
        L.orc_render.restype = i32
        L.orc_adjoint.argtypes = [u32, P, P, P, P, P, u32, P, u32, P, P, u32, u32, P, P, P, P, i32]
        L.orc_adjoint.restype = i32
        L.orc_transform_objects.argtypes = [u32, P, P, u32, P, P, P]
        L.orc_transform_objects.restype = i32
        L.orc_filter_pixels.argtypes = [u32, u32, P, P, P]
        L.orc_render_tex.argtypes = [u32, P, P, P, P, P, u32, P, u32, P, P, u32, u32, P, P, u32, P, i32,
                                     P, u32, P, P]
        L.orc_render_tex.restype = i32
        L.orc_tex_sample.argtypes = [P, P, ctypes.c_float, ctypes.c_float, P]
        L.orc_concentric_disk.argtypes = [ctypes.c_float, ctypes.c_float, P]
        L.orc_refract.argtypes = [P, P, ctypes.c_float, ctypes.c_float, P]
        L.orc_onb.argtypes = [P, P, P]
        L.orc_reflect.argtypes = [P, P, P]
        L.orc_sample_cosine.argtypes = [P, ctypes.c_float, ctypes.c_float, P]
        L.orc_refract.restype = ctypes.c_int
        L.orc_philox.argtypes = [P, u32, u32, P]
        L.orc_sobol_owen.argtypes = [u32, u32, u32, u32, P]
        L.orc_sobol_dim1.argtypes = [u32]
        L.orc_sobol_dim1.restype = u32
        L.orc_scene_bounds.argtypes = [u32, P, P, P]
        L.orc_light_powers.argtypes = [u32, P, P, P, P]
        L.orc_lights_cdf.argtypes = [u32, P, P, P, P]
        L.orc_lights_cdf.restype = u32
        L.orc_emitters.argtypes = [u32, P, P, u32, P, P]
        L.orc_emitters.restype = u32
        L.orc_nested_uniform_scramble.argtypes = [u32, u32]
        L.orc_nested_uniform_scramble.restype = u32
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def calc_partition(size: int) -> tuple[int, int]:
    L = np.zeros(1, np.uint32)
    R = np.zeros(1, np.uint32)
    lib().orc_calc_partition(size, _p(L), _p(R))
    return int(L[0]), int(R[0])


def nmh_levels(count: int) -> int:
    return int(lib().orc_nmh_levels(count))


def nmh_build(vertices: np.ndarray, threads: int = 1) -> np.ndarray:
    """Literal restatement of the reference build; threads > 1 parallelises
    its stable sorts (identical result)."""
    v = np.ascontiguousarray(vertices, dtype=np.float32)
    n = v.shape[0]
    out = np.empty(n, dtype=np.uint32)
    rc = lib().orc_nmh_build_mt(n, _p(v), _p(out), threads)
    if rc != 0:
        raise RuntimeError(f"orc_nmh_build failed: {rc}")
    return out


def trace(tree_vertices: np.ndarray, rays: np.ndarray, any_hit: bool = False, threads: int = 0,
          counters: bool = False):
    v = np.ascontiguousarray(tree_vertices, dtype=np.float32)
    r = np.ascontiguousarray(rays, dtype=np.float32)
    n = r.shape[0]
    hits = np.zeros((n, 4), dtype=np.uint32)
    cnt = np.zeros((n, 2), dtype=np.uint32) if counters else None
    th = threads or os.cpu_count() or 1
    rc = lib().orc_trace(v.shape[0], _p(v), n, _p(r), int(any_hit), _p(hits),
                         _p(cnt) if counters else None, th)
    if rc != 0:
        raise RuntimeError(f"orc_trace failed: {rc}")
    ids = hits[:, 0].view(np.int32).copy()
    tuv = hits[:, 1:4].view(np.float32).copy()
    return (ids, tuv, cnt) if counters else (ids, tuv)


def ray_triangle(ray8: np.ndarray, tri9: np.ndarray):
    r = np.ascontiguousarray(ray8, dtype=np.float32)
    t = np.ascontiguousarray(tri9, dtype=np.float32)
    out = np.zeros(3, dtype=np.float32)
    ok = lib().orc_ray_triangle(_p(r), _p(t), _p(out))
    return bool(ok), out


def generate_tb(normals: np.ndarray):
    n = np.ascontiguousarray(normals, dtype=np.float32)
    T = np.empty_like(n)
    B = np.empty_like(n)
    lib().orc_generate_tb(n.shape[0], _p(n), _p(T), _p(B))
    return T, B


def camera_ray(camera: np.ndarray, w: int, h: int, x: int, y: int, sample: int, samples: int):
    cam = np.ascontiguousarray(camera)
    out = np.zeros(3, dtype=np.float32)
    lib().orc_camera_ray(cam.ctypes.data_as(ctypes.c_void_p), w, h, x, y, sample, samples, _p(out))
    return out


def render(tree, scene, mode: int, spp: int, max_depth: int, rr_start: int, seed: int,
           pixels: np.ndarray | None = None, threads: int = 0, textures=None) -> np.ndarray:
    """tree: dict with tree-order arrays (vertices, normals, tangents,
    bitangents, shader_ids, texcoords).  textures: (TEXTURE_DTYPE records,
    uint8 texels) or None.  Returns float32 (npix, 4) accumulations."""
    opts = np.array([mode, spp, max_depth, rr_start, seed], dtype=np.uint32)
    w, h = scene.width, scene.height
    px = None if pixels is None else np.ascontiguousarray(pixels, dtype=np.uint32)
    npix = w * h if px is None else px.shape[0]
    out = np.zeros((npix, 4), dtype=np.float32)
    v = np.ascontiguousarray(tree["vertices"], np.float32)
    nrm = np.ascontiguousarray(tree["normals"], np.float32)
    T = np.ascontiguousarray(tree["tangents"], np.float32)
    B = np.ascontiguousarray(tree["bitangents"], np.float32)
    ids = np.ascontiguousarray(tree["shader_ids"], np.uint16)
    sh = np.ascontiguousarray(scene.shaders)
    li = np.ascontiguousarray(scene.lights)
    cam = np.ascontiguousarray(scene.camera)
    th = threads or os.cpu_count() or 1
    uv = np.ascontiguousarray(tree["texcoords"], np.float32)
    tdesc, texels = (None, None) if textures is None else (np.ascontiguousarray(textures[0]),
                                                            np.ascontiguousarray(textures[1], np.uint8))
    ntex = 0 if tdesc is None else tdesc.shape[0]
    rc = lib().orc_render_tex(v.shape[0], _p(v), _p(nrm), _p(T), _p(B), _p(ids), sh.shape[0],
                              sh.ctypes.data_as(ctypes.c_void_p), li.shape[0],
                              li.ctypes.data_as(ctypes.c_void_p) if li.shape[0] else None,
                              cam.ctypes.data_as(ctypes.c_void_p), w, h, _p(opts),
                              _p(px) if px is not None else None, npix, _p(out), th, _p(uv), ntex,
                              tdesc.ctypes.data_as(ctypes.c_void_p) if ntex else None,
                              _p(texels) if ntex else None)
    if rc != 0:
        raise RuntimeError(f"orc_render failed: {rc}")
    return out


def adjoint(tree, scene, spp: int, max_depth: int, rr_start: int, seed: int, dimg: np.ndarray,
            threads: int = 0):
    """Gradients (albedo (n_shaders, 3), radiance (n_lights, 3)), float64, of
    sum <dimg, I> for the PBR image; dimg float32 (npix, 4)."""
    opts = np.array([1, spp, max_depth, rr_start, seed], dtype=np.uint32)
    w, h = scene.width, scene.height
    d = np.ascontiguousarray(dimg, np.float32).reshape(w * h, 4)
    v = np.ascontiguousarray(tree["vertices"], np.float32)
    nrm = np.ascontiguousarray(tree["normals"], np.float32)
    T = np.ascontiguousarray(tree["tangents"], np.float32)
    B = np.ascontiguousarray(tree["bitangents"], np.float32)
    ids = np.ascontiguousarray(tree["shader_ids"], np.uint16)
    sh = np.ascontiguousarray(scene.shaders)
    li = np.ascontiguousarray(scene.lights)
    cam = np.ascontiguousarray(scene.camera)
    ga = np.zeros((sh.shape[0], 3), np.float64)
    gl = np.zeros((max(li.shape[0], 1), 3), np.float64)
    th = threads or os.cpu_count() or 1
    rc = lib().orc_adjoint(v.shape[0], _p(v), _p(nrm), _p(T), _p(B), _p(ids), sh.shape[0],
                           sh.ctypes.data_as(ctypes.c_void_p), li.shape[0],
                           li.ctypes.data_as(ctypes.c_void_p) if li.shape[0] else None,
                           cam.ctypes.data_as(ctypes.c_void_p), w, h, _p(opts), _p(d), _p(ga),
                           _p(gl), th)
    if rc != 0:
        raise RuntimeError(f"orc_adjoint failed: {rc}")
    return ga, gl[: li.shape[0]]


def transform_objects(vertices: np.ndarray, normals: np.ndarray, transforms: np.ndarray):
    """Padded, transformed (vertices, normals), float32 (N, 9) each."""
    v = np.ascontiguousarray(vertices, np.float32)
    n = np.ascontiguousarray(normals, np.float32)
    xf = np.ascontiguousarray(transforms)
    count = v.shape[0]
    N = count + (count & 1)
    ov = np.zeros((N, 9), np.float32)
    on = np.zeros((N, 9), np.float32)
    rc = lib().orc_transform_objects(count, _p(v), _p(n), xf.shape[0],
                                     xf.ctypes.data_as(ctypes.c_void_p) if xf.shape[0] else None,
                                     _p(ov), _p(on))
    if rc != 0:
        raise RuntimeError(f"orc_transform_objects failed: {rc}")
    return ov, on


def filter_pixels(img: np.ndarray, w: int, h: int, region) -> np.ndarray:
    """img float32 (w*h, 4); region (left, right, bottom, top) inclusive."""
    a = np.ascontiguousarray(img, np.float32).reshape(w * h, 4)
    out = np.empty_like(a)
    r = np.ascontiguousarray(region, np.uint32)
    lib().orc_filter_pixels(w, h, _p(r), _p(a), _p(out))
    return out


def tree_arrays(scene, perm: np.ndarray) -> dict:
    """Gather a scene's triangle arrays into tree order (applyIndices,
    src/kernels/rebuilder.cu:217-245) and add the tangent frames."""
    nrm = scene.normals[perm]
    T, B = generate_tb(nrm)
    return {
        "vertices": scene.vertices[perm],
        "normals": nrm,
        "texcoords": scene.texcoords[perm],
        "shader_ids": scene.shader_ids[perm],
        "tangents": T,
        "bitangents": B,
    }


def tex_sample(desc, texels: np.ndarray, u: float, v: float) -> np.ndarray:
    d = np.ascontiguousarray(desc)
    b = np.ascontiguousarray(texels, np.uint8)
    out = np.zeros(3, np.float32)
    lib().orc_tex_sample(d.ctypes.data_as(ctypes.c_void_p), _p(b), u, v, _p(out))
    return out


def concentric_disk(u1: float, u2: float):
    out = np.zeros(2, dtype=np.float32)
    lib().orc_concentric_disk(u1, u2, _p(out))
    return out


def refract(i, n, eta1: float, eta2: float):
    """The reference's refract: float32[3] direction, or None on total
    internal reflection."""
    iv = np.ascontiguousarray(i, dtype=np.float32)
    nv = np.ascontiguousarray(n, dtype=np.float32)
    out = np.zeros(3, dtype=np.float32)
    return out if lib().orc_refract(_p(iv), _p(nv), eta1, eta2, _p(out)) else None


def reflect(i, n) -> np.ndarray:
    iv = np.ascontiguousarray(i, dtype=np.float32)
    nv = np.ascontiguousarray(n, dtype=np.float32)
    out = np.zeros(3, np.float32)
    lib().orc_reflect(_p(iv), _p(nv), _p(out))
    return out


def onb(n):
    nv = np.ascontiguousarray(n, dtype=np.float32)
    t1, t2 = np.zeros(3, np.float32), np.zeros(3, np.float32)
    lib().orc_onb(_p(nv), _p(t1), _p(t2))
    return t1, t2


def sample_cosine(n, u1: float, u2: float) -> np.ndarray:
    nv = np.ascontiguousarray(n, dtype=np.float32)
    out = np.zeros(3, np.float32)
    lib().orc_sample_cosine(_p(nv), u1, u2, _p(out))
    return out


def philox(ctr, k0: int, k1: int) -> np.ndarray:
    c = np.asarray(ctr, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    lib().orc_philox(_p(c), k0, k1, _p(out))
    return out


def sobol_owen(seed: int, pixel: int, sample: int, dims: int) -> np.ndarray:
    """PBR sample point (x, y, z) as 32-bit fixed-point fractions."""
    out = np.zeros(3, dtype=np.uint32)
    lib().orc_sobol_owen(seed, pixel, sample, dims, _p(out))
    return out


def sobol_dim1(i: int) -> int:
    return int(lib().orc_sobol_dim1(i))


def nested_uniform_scramble(x: int, seed: int) -> int:
    return int(lib().orc_nested_uniform_scramble(x, seed))


EMITTER_DTYPE = np.dtype([("area", "<f4"), ("pdf", "<f4"), ("cdf", "<f4"), ("power", "<f4", 3),
                          ("triangle", "<u4")])


def scene_bounds(vertices: np.ndarray):
    v = np.ascontiguousarray(vertices, np.float32)
    lo, hi = np.zeros(3, np.float32), np.zeros(3, np.float32)
    lib().orc_scene_bounds(v.shape[0], _p(v), _p(lo), _p(hi))
    return lo, hi


def light_powers(lights: np.ndarray, bmin, bmax) -> np.ndarray:
    li = np.ascontiguousarray(lights)
    lo, hi = np.asarray(bmin, np.float32), np.asarray(bmax, np.float32)
    out = np.zeros(li.shape[0], np.float32)
    lib().orc_light_powers(li.shape[0], li.ctypes.data_as(ctypes.c_void_p), _p(lo), _p(hi), _p(out))
    return out


def lights_cdf(lights: np.ndarray, bmin, bmax) -> np.ndarray:
    li = np.ascontiguousarray(lights)
    lo, hi = np.asarray(bmin, np.float32), np.asarray(bmax, np.float32)
    out = np.zeros(li.shape[0], np.float32)
    lib().orc_lights_cdf(li.shape[0], li.ctypes.data_as(ctypes.c_void_p), _p(lo), _p(hi), _p(out))
    return out


def emitters(tree_vertices: np.ndarray, tree_shader_ids: np.ndarray, shaders: np.ndarray) -> np.ndarray:
    v = np.ascontiguousarray(tree_vertices, np.float32)
    ids = np.ascontiguousarray(tree_shader_ids, np.uint16)
    sh = np.ascontiguousarray(shaders)
    out = np.zeros(v.shape[0], EMITTER_DTYPE)
    n = lib().orc_emitters(v.shape[0], _p(v), _p(ids), sh.shape[0], sh.ctypes.data_as(ctypes.c_void_p),
                           out.ctypes.data_as(ctypes.c_void_p))
    return out[:n]
