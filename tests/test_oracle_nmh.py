"""Pins the CPU oracle's NMH build (src/kernels/rebuilder.cu:311-450).

The reference has no tests or golden vectors, and it cannot be built here
(CUDA 4.2, Thrust, cuRAND, Maya SDK), so the oracle is pinned by:
  * calcPartition's published table (rebuilder.cu:40-50, worked by hand);
  * the closed form of a left-complete binary tree's subtree sizes;
  * agreement with an independent top-down formulation (tests/nmh_model.py)
    on tie-heavy inputs, which fixes the stable sort, the first-maximum swap
    and the re-sorting of finished pairs by later levels;
  * structural invariants of the finished tree on the oracle's output.
A plausible oracle mistake (unstable sort, last-maximum swap, skipping the
finished-pair reordering, wrong child offsets) fails at least one of these.
"""
from __future__ import annotations

import numpy as np
import pytest

from aurora_amd import scenes
from tests import nmh_model

# (triangles in split) -> (left, right), from the left-complete tree shapes
PARTITION_TABLE = {4: (2, 0), 6: (2, 2), 8: (4, 2), 10: (6, 2), 12: (6, 4), 14: (6, 6), 16: (8, 6),
                   18: (10, 6), 30: (14, 14), 32: (16, 14), 48: (30, 16), 64: (32, 30)}


def test_calc_partition_table(oracle):
    for size, lr in PARTITION_TABLE.items():
        assert oracle.calc_partition(size) == lr, size


def test_calc_partition_equals_subtree_sizes(oracle):
    # the left and right subtrees of the root of an n-node left-complete tree
    for n in range(2, 3000):
        L, R = oracle.calc_partition(2 * n)
        assert L == 2 * nmh_model.subtree_nodes(1, n)
        assert R == 2 * nmh_model.subtree_nodes(2, n)
        assert L + R == 2 * (n - 1)


def test_levels(oracle):
    for n, levels in [(2, 1), (4, 2), (6, 2), (8, 3), (2584, 11), (1_024_002, 19), (10_000_000, 23)]:
        assert oracle.nmh_levels(n) == levels


@pytest.mark.parametrize("n", [2, 4, 6, 10, 16, 18, 34, 62, 64, 100, 128, 130, 250, 512, 1022])
@pytest.mark.parametrize("grid", [None, 1.0, 0.25])
def test_oracle_matches_independent_model(oracle, n, grid):
    for seed in range(3):
        s = scenes.random_soup(seed * 1000 + n, n, box=10.0, grid=grid)
        assert np.array_equal(oracle.nmh_build(s.vertices), nmh_model.build(s.vertices))


def test_oracle_matches_model_cornell(oracle):
    s = scenes.cornell()
    assert np.array_equal(oracle.nmh_build(s.vertices), nmh_model.build(s.vertices))


def _bounds(vertices):
    v = vertices.reshape(-1, 3, 3).astype(np.float64)  # [tri, axis, vertex]
    return v.min(axis=2), v.max(axis=2)


def _check_tree(vertices, perm):
    N = vertices.shape[0]
    n = N // 2
    assert np.array_equal(np.sort(perm), np.arange(N))  # bijection
    bmin, bmax = _bounds(vertices[perm])  # tree order
    # subtree extents bottom-up
    lo = np.minimum(bmin[0::2], bmin[1::2])  # per node, 3 axes
    hi = np.maximum(bmax[0::2], bmax[1::2])
    sub_lo, sub_hi = lo.copy(), hi.copy()
    sub_keys_max = [None] * n
    for k in range(n - 1, -1, -1):
        for c in (2 * k + 1, 2 * k + 2):
            if c < n:
                sub_lo[k] = np.minimum(sub_lo[k], sub_lo[c])
                sub_hi[k] = np.maximum(sub_hi[k], sub_hi[c])
    depth = np.floor(np.log2(np.arange(n) + 1)).astype(int)
    for k in range(n):
        a = depth[k] % 3
        # (i) node k's slab on its axis equals its subtree extent
        assert lo[k, a] == sub_lo[k, a] and hi[k, a] == sub_hi[k, a], k
    # (ii) left subtree's min keys <= right subtree's min keys on the children's axis
    tri_min = bmin  # [tri, axis]
    members = [[2 * k, 2 * k + 1] for k in range(n)]
    for k in range(n - 1, -1, -1):
        for c in (2 * k + 1, 2 * k + 2):
            if c < n:
                members[k] = members[k] + members[c]
    for k in range(n):
        l, r = 2 * k + 1, 2 * k + 2
        if r < n:
            b = (depth[k] + 1) % 3
            assert tri_min[members[l], b].max() <= tri_min[members[r], b].min(), k
    del sub_keys_max


@pytest.mark.parametrize("n,grid", [(2, None), (6, None), (64, None), (300, None), (1000, 0.5),
                                    (2048, 0.25), (4000, None)])
def test_tree_invariants(oracle, n, grid):
    s = scenes.random_soup(17 + n, n, box=10.0, grid=grid)
    _check_tree(s.vertices, oracle.nmh_build(s.vertices))


def test_tree_invariants_cornell(oracle):
    s = scenes.cornell()
    _check_tree(s.vertices, oracle.nmh_build(s.vertices))


def test_pair_order_without_ties(oracle):
    """With all keys distinct, a finished pair ends up ordered by the min key
    of the last level's axis (every later level re-sorts it; the last wins)."""
    s = scenes.random_soup(5, 512, box=10.0)
    perm = oracle.nmh_build(s.vertices)
    L = oracle.nmh_levels(512)
    bmin, _ = _bounds(s.vertices[perm])
    a = (L - 1) % 3
    assert np.all(bmin[0::2, a] <= bmin[1::2, a])


def test_rejects_bad_counts(oracle):
    with pytest.raises(RuntimeError):
        oracle.nmh_build(np.zeros((3, 9), np.float32))
    with pytest.raises(RuntimeError):
        oracle.nmh_build(np.zeros((0, 9), np.float32))
