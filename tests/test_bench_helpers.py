"""bench.py's measurement helpers on the CPU, pinned by what fixes them:

* the two-seed error estimate equals the true rel-RMSE in expectation (two
  independent noisy renders of a known image, many pixels);
* the build's algorithmic bytes follow the per-level pending counts
  P_l = N - 2(2^l - 1): 0.63 GB at 1,024,002 and 6.81 GB at 10M triangles,
  the figures the build's roofline is quoted on, and a hand count for 8;
* the traffic figure is read from the committed rocprof summary.
"""
from __future__ import annotations

import numpy as np

import bench


def test_two_seed_estimate_matches_true_error():
    rng = np.random.default_rng(5)
    n = 200_000
    mean = rng.uniform(0.2, 2.0, (n, 3))
    sigma = 0.05 * mean.mean()
    a = np.concatenate([mean + rng.normal(0, sigma, (n, 3)), np.ones((n, 1))], 1).astype(np.float32)
    b = np.concatenate([mean + rng.normal(0, sigma, (n, 3)), np.ones((n, 1))], 1).astype(np.float32)
    true = sigma * np.sqrt(3 * n) / np.linalg.norm(mean)  # ||I - mu|| / ||mu||
    est = bench.two_seed_rel_rmse(a, b)
    assert abs(est / true - 1) < 0.01
    assert bench.two_seed_rel_rmse(a, a) == 0.0
    # alpha (coverage) is not part of the error
    b2 = b.copy()
    b2[:, 3] = 7.0
    assert bench.two_seed_rel_rmse(a, b2) == est


def test_build_bytes():
    # 8 triangles: 4 nodes, 3 levels; pending 8, 6, 2
    assert bench.nmh_build_bytes(8) == 60 * 8 + 12 * (8 + 6 + 2) + 8 * (8 + 6) + 196 * 8
    assert round(bench.nmh_build_bytes(1_024_002) / 1e9, 2) == 0.63
    assert round(bench.nmh_build_bytes(10_000_000) / 1e9, 2) == 6.81


def test_traffic_from_committed_summary():
    t = bench.parse_traffic("k_trace_pt", "k_trace_incoherent_1M.md")
    assert isinstance(t, int) and 1e9 < t < 1e11
    assert bench.parse_traffic("no_such_kernel", "k_trace_incoherent_1M.md") is None
    assert bench.parse_traffic("k_trace_pt", "no_such_file.md") is None
