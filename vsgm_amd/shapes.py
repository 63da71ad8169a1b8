"""All connected non-isomorphic simple graphs on 3..6 vertices (2+6+21+112 = 141).

Canonical form = the smallest edge bitmask over all vertex permutations,
computed by brute force with numpy. The random pattern mix of the batched
workload draws from these, and the learned planner's training records cover
them; the plan-fixture generator and the tests enumerate them too.
"""
from __future__ import annotations

import itertools
from functools import lru_cache

import numpy as np


def _edges(n):
    return [(i, j) for i in range(n) for j in range(i + 1, n)]


def _connected(n, mask, edges):
    adj = [[] for _ in range(n)]
    for k, (i, j) in enumerate(edges):
        if mask >> k & 1:
            adj[i].append(j)
            adj[j].append(i)
    seen = {0}
    stack = [0]
    while stack:
        v = stack.pop()
        for w in adj[v]:
            if w not in seen:
                seen.add(w)
                stack.append(w)
    return len(seen) == n


@lru_cache(maxsize=None)
def connected_shapes(n: int):
    """Canonical edge masks of the connected graphs on n vertices, ascending."""
    edges = _edges(n)
    m = len(edges)
    index = {e: k for k, e in enumerate(edges)}
    masks = np.arange(1 << m, dtype=np.int64)
    bits = (masks[:, None] >> np.arange(m)) & 1
    canon = np.full(1 << m, np.iinfo(np.int64).max, dtype=np.int64)
    for perm in itertools.permutations(range(n)):
        dest = np.array([index[tuple(sorted((perm[i], perm[j])))] for i, j in edges], dtype=np.int64)
        img = (bits << dest).sum(axis=1)
        np.minimum(canon, img, out=canon)
    out = sorted({int(c) for c in np.unique(canon) if _connected(n, int(c), edges)})
    return tuple(out)


def shape_matrix(n: int, mask: int) -> np.ndarray:
    a = np.zeros((n, n), np.uint8)
    for k, (i, j) in enumerate(_edges(n)):
        if mask >> k & 1:
            a[i, j] = a[j, i] = 1
    return a


def all_shapes(sizes=(3, 4, 5, 6)):
    return [(n, mask, shape_matrix(n, mask)) for n in sizes for mask in connected_shapes(n)]


if __name__ == "__main__":
    print({n: len(connected_shapes(n)) for n in (3, 4, 5, 6)})
