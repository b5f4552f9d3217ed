"""Pins the oracle's pixel filter (cudaFilterPixelsKernel restated,
src/kernels/utility.cu:231-258) by closed forms: the 3x3 tent's impulse
response, exact invariance of constant images (the weights are powers of two
summing to 1), clamping at region edges (a one-pixel region is the identity,
a one-row region the 1-D tent 1/4 1/2 1/4), the fixed clamp of the row above
(the reference clamps it to region.bottom), and untouched pixels outside
the region."""
from __future__ import annotations

import numpy as np


def test_impulse_response(oracle):
    w, h = 9, 7
    img = np.zeros((w * h, 4), np.float32)
    img[3 * w + 4] = (16.0, 32.0, 48.0, 64.0)
    out = oracle.filter_pixels(img, w, h, (0, w - 1, 0, h - 1))
    k = np.array([[1, 2, 1], [2, 4, 2], [1, 2, 1]], np.float32) / 16.0
    exp = np.zeros((h, w), np.float32)
    exp[2:5, 3:6] = k
    assert np.array_equal(out[:, 0].reshape(h, w), exp * 16.0)
    assert np.array_equal(out[:, 3].reshape(h, w), exp * 64.0)


def test_constant_and_edges(oracle):
    w, h = 8, 6
    img = np.full((w * h, 4), 3.0, np.float32)
    out = oracle.filter_pixels(img, w, h, (0, w - 1, 0, h - 1))
    assert np.array_equal(out, img)
    rng = np.random.default_rng(1)
    img = rng.integers(0, 64, (w * h, 4)).astype(np.float32)
    # one-pixel region: identity there, everything else copied
    out = oracle.filter_pixels(img, w, h, (5, 5, 2, 2))
    assert np.array_equal(out, img)
    # one-row region: taps above and below clamp into the row -> 1/4 1/2 1/4
    out = oracle.filter_pixels(img, w, h, (0, w - 1, 3, 3)).reshape(h, w, 4)
    row = img.reshape(h, w, 4)[3]
    left = np.concatenate([row[:1], row[:-1]])
    right = np.concatenate([row[1:], row[-1:]])
    assert np.array_equal(out[3], 0.25 * left + 0.5 * row + 0.25 * right)
    assert np.array_equal(np.delete(out, 3, axis=0), np.delete(img.reshape(h, w, 4), 3, axis=0))


def test_row_above_clamps_to_region_top(oracle):
    """At the region's top row the row above is the top row itself, not the
    region's bottom row as in the reference."""
    w, h = 5, 5
    img = np.zeros((w * h, 4), np.float32)
    img[0 * w:1 * w] = 64.0  # bottom row bright
    out = oracle.filter_pixels(img, w, h, (0, w - 1, 0, 4)).reshape(h, w, 4)
    assert np.all(out[4] == 0.0)  # top row unaffected by the bottom row
    assert np.all(out[1, :, 0] == 16.0)  # one row above the bright row: 1/4 of it
