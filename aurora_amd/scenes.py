"""Seeded synthetic scenes and ray sets for the render core.

These stand in for the reference's Maya scene export (src/core/scene.cpp:85-396):
they emit the same per-triangle arrays the reference uploads
(src/data/geometry.h:13-55):

* vertices: "points" layout per triangle, x1 x2 x3 y1 y2 y3 z1 z2 z3
  (src/core/scene.cpp:186-192);
* normals: "values" layout v1xyz v2xyz v3xyz;
* texcoords: 6 floats per triangle;
* shader ids: uint16, 1-based, 0 = default (src/util/array.h:12-18);

plus shader and light records with the reference's field meaning
(src/util/shader.h:23-37, src/util/light.h:24-35, src/core/scene.cpp:242-396)
and a camera built like src/util/camera.h:18-27.

Triangles are one-sided (src/kernels/lib/ray.cuh:33-34 culls det < 1e-7), so
every generator winds faces toward where they must be seen from: box walls
face inward, spheres outward, the ground plane up.

All randomness comes from numpy's PCG64 generator seeded per config, with
geometry computed in float64 and rounded once to float32.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

SHADER_DTYPE = np.dtype(
    [
        ("type", "<i4"),
        ("diffuse_color", "<f4", 3),
        ("emission_color", "<f4", 3),
        ("specular_color", "<f4", 3),
        ("reflection_color", "<f4", 3),
        ("diffuse", "<f4"),
        ("emission", "<f4"),
        ("exponent", "<f4"),
        ("reflectivity", "<f4"),
        ("translucence", "<f4"),
        ("refractive_index", "<f4"),
        ("texture", "<u4"),
    ]
)
assert SHADER_DTYPE.itemsize == 80

LIGHT_AMBIENT, LIGHT_POINT, LIGHT_DIRECTIONAL, LIGHT_AREA = 0, 1, 2, 3
LIGHT_DTYPE = np.dtype(
    [
        ("type", "<i4"),
        ("intensity", "<f4"),
        ("ambient", "<f4"),
        ("area", "<f4"),
        ("samples", "<i2"),
        ("pad", "<i2"),
        ("position", "<f4", 3),
        ("color", "<f4", 3),
        ("direction", "<f4", 3),
        ("e1", "<f4", 3),
        ("e2", "<f4", 3),
    ]
)
assert LIGHT_DTYPE.itemsize == 80

CAMERA_DTYPE = np.dtype(
    [
        ("position", "<f4", 3),
        ("forward", "<f4", 3),
        ("right", "<f4", 3),
        ("up", "<f4", 3),
        ("tanfov", "<f4", 2),
        ("aspect", "<f4"),
    ]
)
assert CAMERA_DTYPE.itemsize == 60

SHADER_LAMBERT, SHADER_PHONG = 0, 1


@dataclasses.dataclass
class Scene:
    """One scene in the reference's upload layout (tree order not yet applied)."""

    vertices: np.ndarray  # (N, 9) float32, points layout
    normals: np.ndarray  # (N, 9) float32, values layout
    texcoords: np.ndarray  # (N, 6) float32
    shader_ids: np.ndarray  # (N,) uint16, 1-based
    shaders: np.ndarray  # SHADER_DTYPE records
    lights: np.ndarray  # LIGHT_DTYPE records
    camera: np.ndarray  # CAMERA_DTYPE record (shape ())
    width: int = 256
    height: int = 256
    spp: int = 16
    max_depth: int = 4
    rr_start: int = 3
    name: str = ""

    @property
    def count(self) -> int:
        return int(self.vertices.shape[0])


# --------------------------------------------------------------------------
# geometry helpers


def _unit_rows(v: np.ndarray) -> np.ndarray:
    return v / np.sqrt(np.sum(v * v, axis=-1, keepdims=True))


def icosphere(subdivisions: int) -> tuple[np.ndarray, np.ndarray]:
    """Unit icosphere by midpoint subdivision; faces wound CCW seen from outside.

    Returns (unit vertices float64 (V,3), faces int64 (F,3)); F = 20 * 4**s.
    """
    t = (1.0 + math.sqrt(5.0)) / 2.0
    verts = [
        (-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0),
        (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
        (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1),
    ]
    faces = [
        (0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11),
        (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6), (7, 1, 8),
        (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9),
        (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1),
    ]
    V = _unit_rows(np.asarray(verts, dtype=np.float64))
    F = np.asarray(faces, dtype=np.int64)
    for _ in range(subdivisions):
        edges = np.concatenate([F[:, [0, 1]], F[:, [1, 2]], F[:, [2, 0]]], axis=0)
        key = np.sort(edges, axis=1)
        uniq, inv = np.unique(key, axis=0, return_inverse=True)
        mids = _unit_rows(0.5 * (V[uniq[:, 0]] + V[uniq[:, 1]]))
        base = V.shape[0]
        V = np.concatenate([V, mids], axis=0)
        nf = F.shape[0]
        m01 = base + inv[:nf]
        m12 = base + inv[nf : 2 * nf]
        m20 = base + inv[2 * nf :]
        F = np.concatenate(
            [
                np.stack([F[:, 0], m01, m20], axis=1),
                np.stack([F[:, 1], m12, m01], axis=1),
                np.stack([F[:, 2], m20, m12], axis=1),
                np.stack([m01, m12, m20], axis=1),
            ],
            axis=0,
        )
    # enforce outward winding: (v2-v1)x(v3-v1) points away from the centre
    a, b, c = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    n = np.cross(b - a, c - a)
    flip = np.sum(n * (a + b + c), axis=1) < 0
    F[flip] = F[flip][:, [0, 2, 1]]
    return V, F


def _random_rotation(rng: np.random.Generator) -> np.ndarray:
    """Uniform rotation from a unit quaternion drawn by Marsaglia's method."""
    while True:
        x1, x2 = rng.uniform(-1.0, 1.0, 2)
        s1 = x1 * x1 + x2 * x2
        if s1 < 1.0:
            break
    while True:
        x3, x4 = rng.uniform(-1.0, 1.0, 2)
        s2 = x3 * x3 + x4 * x4
        if 0.0 < s2 < 1.0:
            break
    k = math.sqrt((1.0 - s1) / s2)
    w, x, y, z = x1, x2, x3 * k, x4 * k
    return np.array(
        [
            [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
            [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
            [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
        ]
    )


def _pack_triangles(P: np.ndarray, Nrm: np.ndarray):
    """P, Nrm: (T, 3, 3) float64 [triangle, vertex, xyz] -> points/values layouts."""
    verts = np.ascontiguousarray(np.transpose(P, (0, 2, 1)).reshape(-1, 9)).astype(np.float32)
    norms = np.ascontiguousarray(Nrm.reshape(-1, 9)).astype(np.float32)
    return verts, norms


def _quad(corner, e1, e2):
    """Two triangles covering corner + [0,1]e1 + [0,1]e2, normal e1 x e2."""
    c = np.asarray(corner, dtype=np.float64)
    e1 = np.asarray(e1, dtype=np.float64)
    e2 = np.asarray(e2, dtype=np.float64)
    p00, p10, p01, p11 = c, c + e1, c + e2, c + e1 + e2
    tris = np.array([[p00, p10, p11], [p00, p11, p01]])
    n = np.cross(e1, e2)
    n = n / np.linalg.norm(n)
    return tris, np.broadcast_to(n, (2, 3, 3)).copy()


def _box_quads(lo: float, hi: float):
    """12 inward-facing axis-aligned quads (each box face split in two)."""
    s = hi - lo
    h = s / 2.0
    quads = []
    # (corner, e1, e2) with e1 x e2 pointing into the box; each face is split
    # along e1 into two halves.
    faces = [
        ((lo, lo, lo), (0, 0, s), (s, 0, 0)),  # floor y=lo, normal +y
        ((lo, hi, lo), (s, 0, 0), (0, 0, s)),  # ceiling y=hi, normal -y
        ((lo, lo, hi), (0, s, 0), (s, 0, 0)),  # back z=hi, normal -z
        ((lo, lo, lo), (s, 0, 0), (0, s, 0)),  # front z=lo, normal +z
        ((lo, lo, lo), (0, s, 0), (0, 0, s)),  # left x=lo, normal +x
        ((hi, lo, lo), (0, 0, s), (0, s, 0)),  # right x=hi, normal -x
    ]
    for corner, e1, e2 in faces:
        e1 = np.asarray(e1, dtype=np.float64)
        e2 = np.asarray(e2, dtype=np.float64)
        c = np.asarray(corner, dtype=np.float64)
        half = e1 * (h / s)
        quads.append(_quad(c, half, e2))
        quads.append(_quad(c + half, half, e2))
    return quads


def _sphere_tris(V: np.ndarray, F: np.ndarray, centre, radius, R=None):
    U = V if R is None else V @ R.T
    P = np.asarray(centre, dtype=np.float64) + radius * U
    return P[F], U[F]


# The reference's Transform (src/util/transform.h:14-27): first triangle and
# triangle count of an object, row-vector world matrix (translation in row 3)
# and its inverse transpose for normals.
TRANSFORM_DTYPE = np.dtype([("offset", np.uint32), ("size", np.uint32), ("world", np.float32, (4, 4)),
                            ("normal", np.float32, (4, 4))])
assert TRANSFORM_DTYPE.itemsize == 136


def make_transform(offset: int, size: int, world: np.ndarray) -> np.void:
    """world: 4x4 row-vector matrix (p' = [p 1] @ world)."""
    t = np.zeros((), dtype=TRANSFORM_DTYPE)
    t["offset"] = offset
    t["size"] = size
    t["world"] = world
    t["normal"] = np.linalg.inv(np.asarray(world, np.float64)).T
    return t


def instanced_spheres(seed: int = 9, n_objects: int = 12, subdivisions: int = 2, box: float = 20.0,
                      drop_last: bool = True):
    """Object-space unit icospheres with per-object world matrices (scale,
    rotation, translation), as Scene::updateMeshes stages them
    (src/core/scene.cpp:85-240).  drop_last removes the last triangle so the
    count is odd.  Returns (vertices, normals, shader_ids, transforms)."""
    rng = np.random.default_rng(seed)
    V, F = icosphere(subdivisions)
    P, Nm = _sphere_tris(V, F, (0.0, 0.0, 0.0), 1.0)
    v1, n1 = _pack_triangles(P, Nm)
    verts, norms, ids, xfs = [], [], [], []
    off = 0
    for k in range(n_objects):
        S = np.diag(rng.uniform(0.5, 3.0, 3))  # non-uniform scale: normals need the inverse transpose
        R = _random_rotation(rng)
        M = np.eye(4)
        M[:3, :3] = S @ R.T  # row vectors: p' = p S R^T
        M[3, :3] = rng.uniform(0.0, box, 3)
        verts.append(v1)
        norms.append(n1)
        ids.append(np.full(v1.shape[0], 1 + k % 4, np.uint16))
        xfs.append(make_transform(off, v1.shape[0], M))
        off += v1.shape[0]
    v = np.concatenate(verts)
    n = np.concatenate(norms)
    i = np.concatenate(ids)
    if drop_last:
        v, n, i = v[:-1], n[:-1], i[:-1]
        xfs[-1]["size"] -= 1
    return v, n, i, np.array(xfs, dtype=TRANSFORM_DTYPE)


def _lambert(albedo, emission=(0.0, 0.0, 0.0)) -> np.void:
    s = np.zeros((), dtype=SHADER_DTYPE)
    s["type"] = SHADER_LAMBERT
    s["diffuse_color"] = albedo
    s["emission_color"] = emission
    s["diffuse"] = 1.0
    s["refractive_index"] = 1.0
    return s


def _phong(albedo, specular, exponent, reflectivity) -> np.void:
    s = _lambert(albedo)
    s["type"] = SHADER_PHONG
    s["specular_color"] = specular
    s["exponent"] = exponent
    s["reflectivity"] = reflectivity
    return s


def area_light(corner, e1, e2, direction, radiance, samples=1) -> np.void:
    """Area light record as built by src/core/scene.cpp:376-389."""
    L = np.zeros((), dtype=LIGHT_DTYPE)
    L["type"] = LIGHT_AREA
    L["intensity"] = 1.0
    L["color"] = radiance
    L["samples"] = samples
    L["position"] = corner
    L["e1"] = e1
    L["e2"] = e2
    L["direction"] = direction
    L["area"] = np.float32(np.linalg.norm(e1) * np.linalg.norm(e2))
    return L


def _light(kind, color, intensity) -> np.void:
    # fields every light type sets (src/core/scene.cpp:340-345)
    L = np.zeros((), dtype=LIGHT_DTYPE)
    L["type"] = kind
    L["color"] = color
    L["intensity"] = intensity
    L["samples"] = 1
    return L


def point_light(position, color, intensity=1.0) -> np.void:
    """Point light record (src/core/scene.cpp:364-368): radiant intensity
    color * intensity, falling off as 1/d^2."""
    L = _light(LIGHT_POINT, color, intensity)
    L["position"] = position
    return L


def directional_light(direction, color, intensity=1.0) -> np.void:
    """Directional light record (src/core/scene.cpp:370-374): `direction` is
    the direction the light travels (unit length)."""
    L = _light(LIGHT_DIRECTIONAL, color, intensity)
    d = np.asarray(direction, dtype=np.float64)
    L["direction"] = d / np.linalg.norm(d)
    return L


def ambient_light(position, color, ambient, intensity=1.0) -> np.void:
    """Ambient light record (src/core/scene.cpp:356-362): the fraction
    `ambient` (the ambient shade) lights every point unshadowed, the rest
    comes from `position` with a shadow ray."""
    L = _light(LIGHT_AMBIENT, color, intensity)
    L["position"] = position
    L["ambient"] = ambient
    return L


def make_camera(position, direction, tanfov) -> np.ndarray:
    """Camera of src/util/camera.h:18-27 in float32 (tanfov given directly)."""
    f32 = np.float32
    pos = np.asarray(position, dtype=f32)
    d = np.asarray(direction, dtype=f32)
    fwd = d * (f32(1.0) / np.sqrt(f32(d[0] * d[0] + d[1] * d[1] + d[2] * d[2])))
    up0 = np.array([0.0, 1.0, 0.0], dtype=f32)

    def cross(a, b):
        return np.array(
            [a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]],
            dtype=f32,
        )

    right = cross(fwd, up0)
    up = cross(right, fwd)
    cam = np.zeros((), dtype=CAMERA_DTYPE)
    cam["position"] = pos
    cam["forward"] = fwd
    cam["right"] = right
    cam["up"] = up
    cam["tanfov"] = np.asarray(tanfov, dtype=f32)
    cam["aspect"] = f32(tanfov[0] / tanfov[1])
    return cam


def _finish(parts, ids, shaders, lights, camera, **kw) -> Scene:
    P = np.concatenate([p for p, _ in parts], axis=0)
    Nm = np.concatenate([n for _, n in parts], axis=0)
    verts, norms = _pack_triangles(P, Nm)
    ids = np.concatenate(ids).astype(np.uint16)
    if verts.shape[0] % 2:
        # pad to an even count by duplicating the last triangle (the
        # reference's padToEven, src/data/geometry.cpp:164-176, with its
        # float-instead-of-triangle indexing fixed)
        verts = np.concatenate([verts, verts[-1:]], axis=0)
        norms = np.concatenate([norms, norms[-1:]], axis=0)
        ids = np.concatenate([ids, ids[-1:]])
    uv = np.zeros((verts.shape[0], 6), dtype=np.float32)
    return Scene(
        vertices=canonical(verts),
        normals=canonical(norms),
        texcoords=uv,
        shader_ids=ids,
        shaders=np.array(shaders, dtype=SHADER_DTYPE),
        lights=np.array(lights, dtype=LIGHT_DTYPE),
        camera=camera,
        **kw,
    )


# RGBA8 texture descriptor (include/aurora_amd.h aur_texture)
TEXTURE_DTYPE = np.dtype([("width", np.uint32), ("height", np.uint32), ("offset", np.uint64)])
assert TEXTURE_DTYPE.itemsize == 16


def textured(scene: Scene, seed: int = 4):
    """Adds texture maps to a scene (as Scene::updateTextures feeds them,
    src/core/scene.cpp:298-326): per-vertex uvs from a planar projection of
    the positions (values outside [0,1] exercise wrap addressing), a
    checkerboard 8x8, a 5x3 noise texture and, for one shader, a texture
    index past the uploaded count (sampled as white).
    Returns (scene, textures (TEXTURE_DTYPE), texels uint8)."""
    rng = np.random.default_rng(seed)
    s = dataclasses.replace(scene)
    P = np.transpose(s.vertices.reshape(-1, 3, 3), (0, 2, 1))  # (T, vertex, xyz)
    uv = np.stack([0.37 * P[..., 0] + 0.11 * P[..., 1], 0.29 * P[..., 2] - 0.23 * P[..., 1]], axis=2)
    s.texcoords = uv.reshape(-1, 6).astype(np.float32)
    check = np.zeros((8, 8, 4), np.uint8)
    yy, xx = np.mgrid[0:8, 0:8]
    on = (xx + yy) % 2 == 0
    check[on] = (230, 200, 40, 255)
    check[~on] = (30, 60, 210, 255)
    noise = rng.integers(0, 256, (3, 5, 4)).astype(np.uint8)
    texels = np.concatenate([check.reshape(-1), noise.reshape(-1)])
    tex = np.array([(8, 8, 0), (5, 3, check.size)], dtype=TEXTURE_DTYPE)
    sh = s.shaders.copy()
    for i in range(sh.shape[0]):
        sh[i]["texture"] = (1, 2, 0, 7)[i % 4]  # 7: no such texture -> white
    s.shaders = sh
    return s, tex, texels


def canonical(a: np.ndarray) -> np.ndarray:
    """-0.0 -> +0.0 (float key order then equals the radix order)."""
    return (a + np.float32(0.0)).astype(np.float32)


# --------------------------------------------------------------------------
# configs


def cornell(seed: int = 1, n_spheres: int = 2, subdivisions: int = 3, width: int = 256,
            height: int = 256, spp: int = 16, max_depth: int = 4, n_lights: int = 1,
            split_materials: bool = False) -> Scene:
    """Cornell-style [0,5]^3 box + icospheres (2 at subdivision 3 by default;
    the batch and adjoint scenes use 2-40 at subdivision 4)."""
    rng = np.random.default_rng(seed)
    parts, ids, shaders = [], [], []
    for q in _box_quads(0.0, 5.0):
        tris, nrm = q
        if split_materials:
            for k in range(2):
                shaders.append(_lambert(rng.uniform(0.2, 0.8, 3)))
                parts.append((tris[k : k + 1], nrm[k : k + 1]))
                ids.append(np.full(1, len(shaders)))
        else:
            shaders.append(_lambert(rng.uniform(0.2, 0.8, 3)))
            parts.append((tris, nrm))
            ids.append(np.full(2, len(shaders)))
    V, F = icosphere(subdivisions)
    radius = 0.8 if n_spheres <= 6 else 0.8 * (6.0 / n_spheres) ** (1.0 / 3.0)
    lo, hi = radius, 5.0 - radius
    centres = []
    for _ in range(n_spheres):
        for _attempt in range(200):
            c = rng.uniform(lo, hi, 3)
            if all(np.linalg.norm(c - o) > 2.0 * radius for o in centres):
                break
        centres.append(c)
        R = _random_rotation(rng)
        P, Nm = _sphere_tris(V, F, c, radius, R)
        if split_materials:
            up = (P.mean(axis=1) - c)[:, 1] >= 0.0
            for half in (up, ~up):
                shaders.append(_lambert(rng.uniform(0.2, 0.8, 3)))
                parts.append((P[half], Nm[half]))
                ids.append(np.full(int(half.sum()), len(shaders)))
        else:
            shaders.append(_lambert(rng.uniform(0.2, 0.8, 3)))
            parts.append((P, Nm))
            ids.append(np.full(P.shape[0], len(shaders)))
    lights = []
    y = 5.0 - 1e-3
    if n_lights == 1:
        lights.append(area_light((2.0, y, 2.0), (1.0, 0.0, 0.0), (0.0, 0.0, 1.0), (0.0, -1.0, 0.0),
                                 (15.0, 15.0, 15.0)))
    else:
        for k in range(n_lights):
            cx = 1.25 + 2.5 * (k % 2)
            cz = 1.25 + 2.5 * ((k // 2) % 2)
            lights.append(area_light((cx - 0.25, y, cz - 0.25), (0.5, 0.0, 0.0), (0.0, 0.0, 0.5),
                                     (0.0, -1.0, 0.0), (15.0, 15.0, 15.0)))
    cam = make_camera((2.5, 2.5, -6.0), (0.0, 0.0, 1.0), (0.9 * width / height, 0.9))
    return _finish(parts, ids, shaders, lights, cam, width=width, height=height, spp=spp,
                   max_depth=max_depth, rr_start=3, name=f"cornell-s{seed}-k{n_spheres}")


def spheres(seed: int = 2, n_spheres: int = 200, subdivisions: int = 4, box: float = 100.0,
            width: int = 1920, height: int = 1080, spp: int = 64, max_depth: int = 8) -> Scene:
    """200 icospheres in a 100 m box, ground plane, 4 quad lights, 32 materials."""
    rng = np.random.default_rng(seed)
    shaders = []
    for m in range(32):
        albedo = rng.uniform(0.2, 0.8, 3)
        if m % 2 == 0:
            shaders.append(_lambert(albedo))
        else:
            shaders.append(_phong(albedo, rng.uniform(0.2, 0.8, 3), 10.0 ** (3.0 * rng.uniform()),
                                  0.2))
    V, F = icosphere(subdivisions)
    parts, ids = [], []
    for _ in range(n_spheres):
        c = rng.uniform(0.0, box, 3)
        radius = 0.5 * 10.0 ** rng.uniform()
        R = _random_rotation(rng)
        P, Nm = _sphere_tris(V, F, c, radius, R)
        parts.append((P, Nm))
        ids.append(np.full(P.shape[0], 1 + int(rng.integers(0, 32))))
    g = _quad((-0.5 * box, 0.0, 1.5 * box), (2.0 * box, 0.0, 0.0), (0.0, 0.0, -2.0 * box))
    parts.append(g)
    ids.append(np.full(2, 1))
    lights = []
    for k in range(4):
        cx = 0.25 * box + 0.5 * box * (k % 2)
        cz = 0.25 * box + 0.5 * box * (k // 2)
        s = 0.1 * box
        rad = float(rng.uniform(5.0, 20.0))
        lights.append(area_light((cx - 0.5 * s, 1.2 * box, cz - 0.5 * s), (s, 0.0, 0.0),
                                 (0.0, 0.0, s), (0.0, -1.0, 0.0), (rad, rad, rad)))
    cam = make_camera((0.5 * box, 0.6 * box, -0.9 * box), (0.0, -0.3 * box, 1.4 * box),
                      (0.9 * width / height, 0.9))
    return _finish(parts, ids, shaders, lights, cam, width=width, height=height, spp=spp,
                   max_depth=max_depth, rr_start=3, name=f"spheres-s{seed}-k{n_spheres}")


def c4_jobs(n_jobs: int = 256, seed: int = 44, sides=(64, 128, 256, 384, 512), spps=(4, 16, 64)) -> list:
    """Preview jobs of the batched workload: Cornell scenes with 2..40
    subdivision-4 spheres (10,264-204,824 triangles), square sides drawn from
    `sides`, spp from `spps`, depth 6, Russian roulette from bounce 3."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n_jobs):
        side = int(rng.choice(list(sides)))
        spp = int(rng.choice(list(spps)))
        n_sph = int(rng.integers(2, 41))
        out.append(cornell(seed=1000 + k, n_spheres=n_sph, subdivisions=4, width=side, height=side,
                           spp=spp, max_depth=6))
    return out


def c5_scene(width: int = 512, height: int = 512) -> Scene:
    """The differentiable-rendering scene: Cornell box with 20 subdivision-4
    spheres (102,424 triangles), every wall triangle and sphere half its own
    Lambert shader (64), 4 ceiling lights, depth 6."""
    return cornell(n_spheres=20, subdivisions=4, width=width, height=height, spp=64, max_depth=6,
                   n_lights=4, split_materials=True)


def soup(seed: int = 3, n: int = 10_000_000, box: float = 1000.0, chunk: int = 1 << 20) -> Scene:
    """Random equilateral-triangle soup (centroids uniform in the box,
    edge ~ LogNormal(0, 1) m, uniformly random orientation)."""
    rng = np.random.default_rng(seed)
    verts = np.empty((n, 9), dtype=np.float32)
    norms = np.empty((n, 9), dtype=np.float32)
    base = np.array([[1.0, 0.0, 0.0], [-0.5, math.sqrt(3) / 2, 0.0], [-0.5, -math.sqrt(3) / 2, 0.0]])
    base = base / math.sqrt(3.0)  # unit-edge equilateral triangle around the origin
    for s in range(0, n, chunk):
        m = min(chunk, n - s)
        c = rng.uniform(0.0, box, (m, 3))
        edge = rng.lognormal(0.0, 1.0, m)
        q = rng.normal(size=(m, 4))
        q /= np.sqrt(np.sum(q * q, axis=1, keepdims=True))
        w, x, y, z = q.T
        R = np.stack(
            [
                np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)], -1),
                np.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], -1),
                np.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], -1),
            ],
            axis=1,
        )  # (m, 3, 3)
        P = c[:, None, :] + edge[:, None, None] * np.einsum("mij,kj->mki", R, base)
        nrm = np.cross(P[:, 1] - P[:, 0], P[:, 2] - P[:, 0])
        nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
        v, nn = _pack_triangles(P, np.broadcast_to(nrm[:, None, :], (m, 3, 3)))
        verts[s : s + m] = v
        norms[s : s + m] = nn
    sh = [_lambert((0.5, 0.5, 0.5))]
    cam = make_camera((0.5 * box, 0.5 * box, -0.6 * box), (0.0, 0.0, 1.0), (1.0, 1.0))
    return Scene(vertices=canonical(verts), normals=canonical(norms),
                 texcoords=np.zeros((n, 6), dtype=np.float32),
                 shader_ids=np.ones(n, dtype=np.uint16), shaders=np.array(sh, dtype=SHADER_DTYPE),
                 lights=np.zeros(0, dtype=LIGHT_DTYPE), camera=cam, width=4096, height=4096,
                 spp=1, max_depth=1, name=f"soup-s{seed}-n{n}")


def random_soup(seed: int, n: int, box: float = 10.0, grid: float | None = None) -> Scene:
    """Small random soups for parity tests.  With `grid`, every coordinate is
    snapped to a multiple of `grid`, which makes equal sort keys common."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(0.0, box, (n, 1, 3))
    P = c + rng.uniform(-0.05 * box, 0.05 * box, (n, 3, 3))
    if grid is not None:
        P = np.round(P / grid) * grid
    nrm = np.cross(P[:, 1] - P[:, 0], P[:, 2] - P[:, 0])
    ln = np.linalg.norm(nrm, axis=1, keepdims=True)
    nrm = np.where(ln > 0, nrm / np.where(ln > 0, ln, 1.0), np.array([0.0, 1.0, 0.0]))
    v, nn = _pack_triangles(P, np.broadcast_to(nrm[:, None, :], (n, 3, 3)))
    cam = make_camera((0.5 * box, 0.5 * box, -box), (0.0, 0.0, 1.0), (1.0, 1.0))
    return Scene(vertices=canonical(v), normals=canonical(nn), texcoords=np.zeros((n, 6), np.float32),
                 shader_ids=np.ones(n, dtype=np.uint16),
                 shaders=np.array([_lambert((0.5, 0.5, 0.5))], dtype=SHADER_DTYPE),
                 lights=np.zeros(0, dtype=LIGHT_DTYPE), camera=cam, name=f"soup-s{seed}-n{n}")


# --------------------------------------------------------------------------
# analytic scenes (closed-form expectations exist for these)


def plane_under_light(albedo: float = 0.5, radiance: float = 10.0, light_height: float = 2.0,
                      light_size: float = 1.0, width: int = 32, height: int = 32) -> Scene:
    """A 10 m x 10 m Lambert floor (y = 0, facing up) under one downward
    facing square area light: nothing occludes, so a floor point's direct
    light has closed forms (polygon irradiance; compat estimator quadrature)."""
    parts = [_quad((-5.0, 0.0, -5.0), (0.0, 0.0, 10.0), (10.0, 0.0, 0.0))]
    ids = [np.full(2, 1)]
    h = 0.5 * light_size
    light = area_light((-h, light_height, -h), (light_size, 0.0, 0.0), (0.0, 0.0, light_size),
                       (0.0, -1.0, 0.0), (radiance, radiance, radiance))
    cam = make_camera((0.0, 3.0, -4.0), (0.0, -3.0, 4.0), (0.6, 0.6 * height / width))
    return _finish(parts, ids, [_lambert((albedo, albedo, albedo))], [light], cam, width=width,
                   height=height, spp=16, max_depth=1, name="plane-under-light")


def furnace(albedo: float = 0.5, emission: float = 1.0, subdivisions: int = 2, width: int = 32,
            height: int = 32, max_depth: int = 5) -> Scene:
    """Closed, inward-facing emissive Lambert sphere with the camera at its
    centre and no lights: every path hits at every bounce, so the PBR path
    loop's expected radiance is emission * sum_{k<depth} albedo^k."""
    V, F = icosphere(subdivisions)
    F = F[:, [0, 2, 1]]  # inward winding
    P, _ = _sphere_tris(V, F, (0.0, 0.0, 0.0), 1.0)
    # flat (face) normals: a sampled direction never dips below its own
    # triangle, so no path can leave the closed sphere
    n = np.cross(P[:, 1] - P[:, 0], P[:, 2] - P[:, 0])
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    Nm = np.broadcast_to(n[:, None, :], P.shape).copy()
    cam = make_camera((0.0, 0.0, 0.0), (0.0, 0.0, 1.0), (1.0, 1.0 * height / width))
    sh = _lambert((albedo, albedo, albedo), (emission, emission, emission))
    return _finish([(P, Nm)], [np.full(P.shape[0], 1)], [sh], [], cam, width=width, height=height,
                   spp=4, max_depth=max_depth, name="furnace")


def surface(albedo, emission=(0.0, 0.0, 0.0), reflectivity: float = 0.0, translucence: float = 0.0,
            refractive_index: float = 1.0, diffuse: float = 1.0, texture: int = 0) -> np.void:
    """Shader record for the analytic scenes: Lambert albedo `albedo`
    (diffuse_color, scaled by `diffuse`), emission colour, and compat mode's
    mirror weight (reflectivity) and refraction weight (translucence)."""
    s = _lambert(albedo, emission)
    s["diffuse"] = diffuse
    s["reflectivity"] = reflectivity
    s["translucence"] = translucence
    s["refractive_index"] = refractive_index
    s["texture"] = texture
    return s


def horizontal(y: float, up: bool, half: float = 128.0, flip_normals: bool = False):
    """Square of side 2*half in the plane at height y, centred on the y axis,
    facing +y (up) or -y.  Corners are integers and the edges powers of two,
    so hit heights on it are exact up to the ray's own rounding.  With
    flip_normals the vertex (shading) normals point away from the face."""
    s = 2.0 * half
    if up:
        tris, nrm = _quad((-half, y, -half), (0.0, 0.0, s), (s, 0.0, 0.0))
    else:
        tris, nrm = _quad((-half, y, -half), (s, 0.0, 0.0), (0.0, 0.0, s))
    return tris, (-nrm if flip_normals else nrm)


def planes(parts, shaders, lights, camera, width: int = 24, height: int = 16, spp: int = 16,
           max_depth: int = 1, name: str = "planes") -> Scene:
    """Scene of the given (triangles, normals) parts; part k uses shader k."""
    ids = [np.full(p[0].shape[0], k + 1) for k, p in enumerate(parts)]
    return _finish(list(parts), ids, list(shaders), list(lights), camera, width=width, height=height,
                   spp=spp, max_depth=max_depth, name=name)


# --------------------------------------------------------------------------
# ray sets: records of 8 float32 (ox oy oz tmax dx dy dz 0)


def incoherent_rays(n: int, lo, hi, seed: int, tmax: float = 1_000_000.0) -> np.ndarray:
    """Origins uniform in the box [lo, hi]^3, directions uniform on the sphere
    (z = 1 - 2 u1, phi = 2 pi u2)."""
    rng = np.random.default_rng(seed)
    rays = np.empty((n, 8), dtype=np.float32)
    chunk = 1 << 22
    lo = np.broadcast_to(np.asarray(lo, dtype=np.float64), (3,))
    hi = np.broadcast_to(np.asarray(hi, dtype=np.float64), (3,))
    for s in range(0, n, chunk):
        m = min(chunk, n - s)
        o = lo + (hi - lo) * rng.random((m, 3))
        u = rng.random((m, 2))
        z = 1.0 - 2.0 * u[:, 0]
        phi = 2.0 * math.pi * u[:, 1]
        r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
        rays[s : s + m, 0:3] = o
        rays[s : s + m, 3] = tmax
        rays[s : s + m, 4] = r * np.cos(phi)
        rays[s : s + m, 5] = r * np.sin(phi)
        rays[s : s + m, 6] = z
        rays[s : s + m, 7] = 0.0
    return rays


def coherent_rays(camera: np.ndarray, width: int, height: int, tmax: float = 1_000_000.0) -> np.ndarray:
    """Pinhole primary rays through pixel centres (one per pixel)."""
    f32 = np.float32
    xs = (np.arange(width, dtype=f32) + f32(0.5)) / f32(width) - f32(0.5)
    ys = (np.arange(height, dtype=f32) + f32(0.5)) / f32(height) - f32(0.5)
    px, py = np.meshgrid(xs, ys)
    r = camera["right"].astype(f32)
    u = camera["up"].astype(f32)
    fw = camera["forward"].astype(f32)
    tx, ty = f32(camera["tanfov"][0]), f32(camera["tanfov"][1])
    d = (px[..., None] * r) * tx + (py[..., None] * u) * ty + fw
    d = d * (f32(1.0) / np.sqrt(np.sum(d * d, axis=-1, keepdims=True)))
    rays = np.zeros((width * height, 8), dtype=f32)
    rays[:, 0:3] = camera["position"]
    rays[:, 3] = tmax
    rays[:, 4:7] = d.reshape(-1, 3)
    return rays


def scene_bounds(scene: Scene):
    v = scene.vertices.reshape(-1, 3, 3)  # [tri, axis, vertex]
    return v.min(axis=(0, 2)), v.max(axis=(0, 2))
