"""Pins the oracle's texture fetch (texfetch with the reference's linear /
wrap / normalised-coordinate configuration, src/kernels/lib/texunit.cuh:12-19,
54-94, as an own float bilinear) by closed forms: texel centres return the
texel, midpoints the mean of the two neighbours, coordinates repeat with
period 1, and a ramp texture interpolates linearly.  The texture unit's 1/256
weight quantisation is not reproduced (parity unpinned)."""
from __future__ import annotations

import numpy as np

from aurora_amd import scenes


def _tex(w, h, data):
    d = np.array([(w, h, 0)], dtype=scenes.TEXTURE_DTYPE)[0]
    return d, np.ascontiguousarray(data, np.uint8).reshape(-1)


def test_texel_centres_and_midpoints(oracle):
    rng = np.random.default_rng(0)
    w, h = 5, 3
    img = rng.integers(0, 256, (h, w, 4)).astype(np.uint8)
    d, b = _tex(w, h, img)
    k = np.float32(1.0) / np.float32(255.0)
    for j in range(h):
        for i in range(w):
            got = oracle.tex_sample(d, b, (i + 0.5) / w, (j + 0.5) / h)
            assert np.array_equal(got, img[j, i, :3].astype(np.float32) * k)
    # midpoint of two horizontal neighbours (interior)
    got = oracle.tex_sample(d, b, 2.0 / w, 1.5 / h)
    mean = (img[1, 1, :3].astype(np.float64) + img[1, 2, :3]) / 2 / 255
    assert np.allclose(got, mean, atol=1e-6)
    # wrap: the left edge blends the last and first columns
    got = oracle.tex_sample(d, b, 0.0, 0.5)
    mean = (img[1, 0, :3].astype(np.float64) + img[1, w - 1, :3]) / 2 / 255
    assert np.allclose(got, mean, atol=1e-6)


def test_wrap_period_and_ramp(oracle):
    w, h = 8, 4
    img = np.zeros((h, w, 4), np.uint8)
    img[..., 0] = (np.arange(w) * 30)[None, :]  # ramp in x
    img[..., 1] = (np.arange(h) * 60)[:, None]  # ramp in y
    d, b = _tex(w, h, img)
    for u, v in [(0.3125, 0.40625), (0.609375, 0.875), (0.453125, 0.5625)]:  # dyadic: u + k is exact
        a = oracle.tex_sample(d, b, u, v)
        for du, dv in [(1.0, 0.0), (-3.0, 2.0), (0.0, -1.0)]:
            assert np.array_equal(oracle.tex_sample(d, b, u + du, v + dv), a)
    # inside the texel-centre span the ramp is reproduced linearly
    for u in np.linspace(0.5 / w, (w - 0.5) / w, 9):
        got = oracle.tex_sample(d, b, u, 0.5)
        assert abs(got[0] - (u * w - 0.5) * 30 / 255) < 1e-5
