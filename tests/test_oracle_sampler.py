"""PBR sample points (Owen-scrambled Sobol) pinned by what the mathematics
fixes, not by restating the code:

* Sobol dimension 1 against its published direction numbers (Joe & Kuo's
  dimension 2, polynomial x + 1: m_k = 1, 3, 5, 15, 17, 51, ... = OEIS
  A001317, Pascal's triangle mod 2 read as binary) and GF(2) linearity;
* the nested uniform scramble is an Owen scramble: two inputs whose first
  difference is binary digit t map to outputs whose first difference is
  digit t (so it is injective and keeps every elementary interval whole);
* every power-of-two prefix of a pixel's points is a (0, m, 2)-net, and the
  z coordinate puts one point in each interval of width 2^-m;
* different seeds give independent, uniform randomisations.

The GPU replay tests then compare the device sampler (a superset-sum form of
dimension 1) with this direct form bit for bit.
"""
from __future__ import annotations

import numpy as np
import pytest

A001317 = [1, 3, 5, 15, 17, 51, 85, 255, 257, 771, 1285, 3855, 4369, 13107, 21845, 65535]


def test_sobol_dim1_direction_numbers(oracle):
    for k, m in enumerate(A001317, start=1):
        assert oracle.sobol_dim1(1 << (k - 1)) == m << (32 - k), k
    # the direct-order start of the sequence: 0, 1/2, 3/4, 1/4, 5/8, 1/8, 3/8, 7/8
    first = [oracle.sobol_dim1(i) / 2.0 ** 32 for i in range(8)]
    assert first == [0.0, 0.5, 0.75, 0.25, 0.625, 0.125, 0.375, 0.875]
    # unit-diagonal triangular generator: index bit k (from 0) sets binary
    # digit k + 1 (from the most significant) and none after it, so the first
    # 2^m points are the multiples of 2^-m
    for k in range(32):
        v = oracle.sobol_dim1(1 << k)
        assert (v >> (31 - k)) & 1 == 1 and v & ((1 << (31 - k)) - 1) == 0, k
    assert sorted(oracle.sobol_dim1(i) >> 22 for i in range(1024)) == list(range(1024))


def test_sobol_dim1_is_linear_over_gf2(oracle):
    rng = np.random.default_rng(1)
    for a, b in rng.integers(0, 2 ** 32, size=(200, 2), dtype=np.uint64):
        a, b = int(a), int(b)
        assert oracle.sobol_dim1(a ^ b) == oracle.sobol_dim1(a) ^ oracle.sobol_dim1(b)


def _first_diff(a: int, b: int) -> int:
    """Index (0 = most significant) of the first differing binary digit."""
    return 32 - (a ^ b).bit_length()


def test_scramble_is_owen_scramble(oracle):
    rng = np.random.default_rng(2)
    for seed in rng.integers(0, 2 ** 32, size=8, dtype=np.uint64):
        seed = int(seed)
        for _ in range(200):
            x = int(rng.integers(0, 2 ** 32, dtype=np.uint64))
            t = int(rng.integers(0, 32))
            y = x ^ (1 << (31 - t)) ^ int(rng.integers(0, 2 ** (31 - t), dtype=np.uint64))  # differ first at t
            sx, sy = oracle.nested_uniform_scramble(x, seed), oracle.nested_uniform_scramble(y, seed)
            assert _first_diff(sx, sy) == t, (seed, x, y, t)
        # a bijection: 4096 consecutive inputs, 4096 distinct outputs
        outs = {oracle.nested_uniform_scramble(i, seed) for i in range(4096)}
        assert len(outs) == 4096
    # the scramble does depend on the seed
    assert oracle.nested_uniform_scramble(12345, 1) != oracle.nested_uniform_scramble(12345, 2)


def _points(oracle, seed, pixel, dims, n):
    return np.array([oracle.sobol_owen(seed, pixel, s, dims) for s in range(n)], dtype=np.uint64)


@pytest.mark.parametrize("seed,pixel,dims", [(7, 0, 0), (8, 12345, 1), (1, 99, (3 << 16) | 17),
                                             (0xDEADBEEF, 2073599, (7 << 16) | 1)] +
                         [(s, 31 * s + 5, (s % 9 << 16) | (16 + s % 5)) for s in range(100, 124)])
def test_prefixes_are_nets(oracle, seed, pixel, dims):
    pts = _points(oracle, seed, pixel, dims, 1 << 10)
    for m in range(1, 11):
        p = pts[: 1 << m]
        for a in range(m + 1):
            b = m - a
            cx = (p[:, 0] >> np.uint64(32 - a)) if a else np.zeros(len(p), np.uint64)
            cy = (p[:, 1] >> np.uint64(32 - b)) if b else np.zeros(len(p), np.uint64)
            cells = cx * np.uint64(1 << b) + cy
            assert len(np.unique(cells)) == len(p), (m, a)  # one point per 2^-a x 2^-b cell
        z = p[:, 2] >> np.uint64(32 - m)
        assert len(np.unique(z)) == len(p), m


def test_seeds_are_independent_randomisations(oracle):
    n = 4096
    xs = np.array([oracle.sobol_owen(seed, 5, 3, 1)[:3] for seed in range(n)], dtype=np.float64) / 2.0 ** 32
    # one fixed point across seeds: uniform on [0,1)^3, uncorrelated coordinates
    assert np.all(np.abs(xs.mean(axis=0) - 0.5) < 4 * np.sqrt(1 / 12 / n))
    c = np.corrcoef(xs.T)
    assert np.all(np.abs(c[np.triu_indices(3, 1)]) < 4 / np.sqrt(n))
    # different dimension pairs of one pixel are not the same points
    a = _points(oracle, 7, 5, (1 << 16) | 1, 64)
    b = _points(oracle, 7, 5, (2 << 16) | 1, 64)
    assert not np.array_equal(a[:, :2], b[:, :2])
