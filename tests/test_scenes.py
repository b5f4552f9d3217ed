"""Scene generators: sizes of the named configs, winding, records, determinism."""
from __future__ import annotations

import numpy as np

from aurora_amd import scenes


def _face_normals(s):
    P = s.vertices.reshape(-1, 3, 3).transpose(0, 2, 1).astype(np.float64)  # [tri, vertex, xyz]
    return np.cross(P[:, 1] - P[:, 0], P[:, 2] - P[:, 0]), P


def test_cornell_counts_and_records():
    s = scenes.cornell()
    assert s.count == 24 + 2 * 1280 == 2584
    assert s.shaders.shape == (14,) and s.lights.shape == (1,)
    L = s.lights[0]
    assert L["type"] == scenes.LIGHT_AREA and L["area"] == 1.0
    assert np.allclose(L["color"] * L["intensity"], 15.0)
    assert np.all((s.shaders["diffuse_color"] >= 0.2) & (s.shaders["diffuse_color"] <= 0.8))
    assert s.shader_ids.min() >= 1 and s.shader_ids.max() <= 14
    lo, hi = scenes.scene_bounds(s)
    assert np.allclose(lo, 0.0) and np.allclose(hi, 5.0)


def test_cornell_winding_faces_inward_and_spheres_outward():
    s = scenes.cornell()
    n, P = _face_normals(s)
    c = P.mean(axis=1)
    box = slice(0, 24)
    assert np.all(np.sum(n[box] * (np.array([2.5, 2.5, 2.5]) - c[box]), axis=1) > 0)
    # stored vertex normals agree with the winding
    vn = s.normals.reshape(-1, 3, 3).mean(axis=1)
    assert np.all(np.sum(vn * n, axis=1) > 0)


def test_spheres_config_size():
    s = scenes.spheres()
    assert s.count == 200 * 5120 + 2 == 1_024_002
    assert s.shaders.shape == (32,) and s.lights.shape == (4,)
    rad = s.lights["color"][:, 0] * s.lights["intensity"]
    assert np.all((rad >= 5) & (rad <= 20))
    n, _ = _face_normals(s)
    vn = s.normals.reshape(-1, 3, 3).mean(axis=1)
    assert np.all(np.sum(vn * n, axis=1) > 0)
    assert (s.width, s.height, s.spp, s.max_depth) == (1920, 1080, 64, 8)


def test_generators_are_deterministic():
    a, b = scenes.spheres(n_spheres=20), scenes.spheres(n_spheres=20)
    assert np.array_equal(a.vertices, b.vertices) and np.array_equal(a.shader_ids, b.shader_ids)
    r1 = scenes.incoherent_rays(1000, 0, 100, seed=3)
    r2 = scenes.incoherent_rays(1000, 0, 100, seed=3)
    assert np.array_equal(r1, r2)


def test_no_negative_zero():
    s = scenes.spheres(n_spheres=10)
    assert not np.any(np.signbit(s.vertices) & (s.vertices == 0))


def test_incoherent_rays_distribution():
    r = scenes.incoherent_rays(200_000, 0.0, 100.0, seed=1)
    d = r[:, 4:7].astype(np.float64)
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-6)
    assert np.all(np.abs(d.mean(axis=0)) < 0.01)  # isotropic
    assert np.all((r[:, 0:3] >= 0) & (r[:, 0:3] <= 100))
    assert np.all(r[:, 3] == 1e6)


def test_soup_small():
    s = scenes.soup(n=10_000)
    assert s.count == 10_000
    P = s.vertices.reshape(-1, 3, 3).transpose(0, 2, 1).astype(np.float64)
    e = np.linalg.norm(P[:, 1] - P[:, 0], axis=1)
    # equilateral, log-normal edge lengths (median ~ 1 m)
    # equal edges up to float32 rounding of coordinates near 1000 m (ulp 6e-5)
    assert np.allclose(e, np.linalg.norm(P[:, 2] - P[:, 1], axis=1), rtol=1e-4, atol=3e-4)
    assert 0.8 < np.median(e) < 1.25


def test_odd_counts_are_padded():
    s = scenes.random_soup(1, 3)
    assert s.count == 3  # random_soup does not pad; generators built by _finish do
    c = scenes.cornell(n_spheres=1, subdivisions=0)
    assert c.count % 2 == 0


def test_delta_light_records():
    p = scenes.point_light((1.0, 2.0, 3.0), (0.5, 0.25, 1.0), 8.0)
    assert p["type"] == scenes.LIGHT_POINT and p["intensity"] == 8.0 and p["samples"] == 1
    assert np.array_equal(p["position"], np.float32([1.0, 2.0, 3.0]))
    assert p["area"] == 0.0 and p["ambient"] == 0.0  # src/core/scene.cpp:344-345
    d = scenes.directional_light((0.0, -3.0, 4.0), (1.0, 1.0, 1.0))
    assert d["type"] == scenes.LIGHT_DIRECTIONAL
    assert np.allclose(d["direction"], [0.0, -0.6, 0.8]) and abs(np.linalg.norm(d["direction"]) - 1) < 1e-6
    a = scenes.ambient_light((0.0, 5.0, 0.0), (1.0, 0.9, 0.8), 0.45, 2.0)
    assert a["type"] == scenes.LIGHT_AMBIENT and a["ambient"] == np.float32(0.45) and a["intensity"] == 2.0
    recs = np.array([p, d, a], dtype=scenes.LIGHT_DTYPE)
    assert recs.itemsize == 80 and recs["pad"].max() == 0
