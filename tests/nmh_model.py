"""Independent top-down formulation of the NMH build, used only by tests.

The reference (src/kernels/rebuilder.cu:311-450) sorts the whole array by
(node label, min key) at every level, which also reorders the two triangles
of every finished node by every later level's key.  This module states the
same result differently:

* each node's subtree size comes from the closed form of a left-complete
  binary tree (cnt below) instead of calcPartition's split bookkeeping;
* a node's triangles are sorted (stably) by the current axis' min key, the
  first is kept, the first maximum of the max key among the rest is swapped
  into second place, and the remainder is handed to the two children;
* the final order inside a finished pair is computed once from the keys of
  the levels after it (newest key most significant), not by re-sorting.

Agreement between this model and the literal oracle on tie-heavy inputs pins
the oracle's per-level sort, its stability and its pair reordering.
"""
from __future__ import annotations

import numpy as np


def subtree_nodes(h: int, n: int) -> int:
    """Nodes in the subtree of heap node h in a left-complete tree of n nodes."""
    total, first, width = 0, h, 1
    while first < n:
        total += min(width, n - first)
        first = 2 * first + 1
        width *= 2
    return total


def build(vertices: np.ndarray) -> np.ndarray:
    N = vertices.shape[0]
    n = N // 2
    L = n.bit_length()  # log2i(n) + 1
    v = vertices.reshape(N, 3, 3).astype(np.float64)
    bmin = v.min(axis=2).T  # [axis, tri]
    bmax = v.max(axis=2).T
    P = [None] * N
    groups = [(list(range(N)), [0])]
    for lev in range(L):
        a = lev % 3
        nxt = []
        for E, children in groups:
            Es = sorted(E, key=lambda t: bmin[a][t])
            off = 0
            for h in children:
                s = 2 * subtree_nodes(h, n)
                C = Es[off : off + s]
                off += s
                if lev == L - 1:
                    P[2 * h], P[2 * h + 1] = C[0], C[1]
                    continue
                m = 1
                for i in range(2, s):
                    if bmax[a][C[i]] > bmax[a][C[m]]:
                        m = i
                C[1], C[m] = C[m], C[1]
                x, y = C[0], C[1]
                y_first = False
                for later in range(L - 1, max(lev + 1, L - 3) - 1, -1):
                    b = later % 3
                    if bmin[b][y] < bmin[b][x]:
                        y_first = True
                        break
                    if bmin[b][x] < bmin[b][y]:
                        break
                P[2 * h], P[2 * h + 1] = (y, x) if y_first else (x, y)
                rest = C[2:]
                if rest:
                    kids = [c for c in (2 * h + 1, 2 * h + 2) if c < n]
                    nxt.append((rest, kids))
        groups = nxt
    return np.asarray(P, dtype=np.uint32)
