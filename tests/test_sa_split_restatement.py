"""The smoothed-aggregation restatement's cell split, pinned on the CPU.

`split_aggregates` (tests/test_gpu_smoothed_aggregation.py) is the checker the
GPU hierarchy's level sizes and PCG paths are compared with on cracked
lattices, so it is pinned here by hand-worked graphs and by a brute-force
breadth-first search per cell: every cell becomes the connected components of
the edges inside it, cells keep their order, a cell's components follow their
smallest node, and a split that would leave more than 80% of the nodes as
aggregates is not made.  No reference counterpart (LPM-CPP has no multigrid).
"""
from collections import deque

import numpy as np

from test_gpu_smoothed_aggregation import _aggregate, coarse_edges, split_aggregates


def _e(pairs):
    return np.array(pairs, dtype=np.int64).reshape(-1, 2)


def test_hand_worked_split_and_coarse_edges():
    # cells {0,1,2,3} and {4,5,6,7}; inside the first only 0-1 and 2-3 are joined
    agg = np.array([0, 0, 0, 0, 1, 1, 1, 1])
    e = _e([(0, 1), (2, 3), (1, 4), (4, 5), (5, 6), (6, 7), (3, 4)])
    agg2, nc = split_aggregates(agg, 2, e)
    assert nc == 3
    np.testing.assert_array_equal(agg2, [0, 0, 1, 1, 2, 2, 2, 2])
    np.testing.assert_array_equal(coarse_edges(agg2, e), [[0, 2], [1, 2]])
    # a cell's components follow their smallest node, not the edge order
    agg2, nc = split_aggregates(np.zeros(4, int), 1, _e([(1, 3), (0, 2)]))
    np.testing.assert_array_equal(agg2, [0, 1, 0, 1])


def test_split_that_leaves_too_many_aggregates_is_not_made():
    agg = np.array([0, 0, 0, 0, 1, 1, 1, 1])
    agg2, nc = split_aggregates(agg, 2, _e([(0, 1)]))  # 7 components > 0.8 * 8
    assert nc == 2
    np.testing.assert_array_equal(agg2, agg)
    agg2, nc = split_aggregates(agg, 2, _e([(0, 1), (2, 3), (4, 5), (6, 7)]))  # 4 <= 6.4
    assert nc == 4


def _brute(agg, e):
    n = len(agg)
    adj = [[] for _ in range(n)]
    for u, v in e:
        if agg[u] == agg[v]:
            adj[u].append(v)
            adj[v].append(u)
    comp = [-1] * n
    order = []  # (cell, smallest node) of each component
    for s in range(n):
        if comp[s] >= 0:
            continue
        comp[s] = len(order)
        q = deque([s])
        while q:
            x = q.popleft()
            for y in adj[x]:
                if comp[y] < 0:
                    comp[y] = comp[s]
                    q.append(y)
        order.append((agg[s], s))  # s is the smallest node: nodes are visited in order
    rank = {c: r for r, c in enumerate(sorted(range(len(order)), key=lambda c: order[c]))}
    return np.array([rank[comp[i]] for i in range(n)]), len(order)


def test_random_geometric_graphs_match_breadth_first_search():
    rng = np.random.default_rng(7)
    for trial in range(20):
        n = int(rng.integers(50, 400))
        pos = np.zeros((n, 3))
        pos[:, :2] = rng.random((n, 2)) * 10
        d = np.linalg.norm(pos[:, None, :] - pos[None, :, :], axis=2)
        r = 0.6 + 0.4 * rng.random()
        iu, ju = np.nonzero(np.triu(d < r, 1))
        keep = rng.random(len(iu)) < 0.8  # some bonds broken
        e = np.stack([iu[keep], ju[keep]], 1)
        agg, nc = _aggregate(pos, 2.0 + rng.random())
        want, nw = _brute(agg, e)
        got, ng = split_aggregates(agg, nc, e)
        if nw > 0.8 * n:
            assert ng == nc
            np.testing.assert_array_equal(got, agg)
        else:
            assert ng == nw
            np.testing.assert_array_equal(got, want)
            ce = coarse_edges(got, e)
            a, b = got[e[:, 0]], got[e[:, 1]]
            pairs = {(min(x, y), max(x, y)) for x, y in zip(a, b) if x != y}
            assert sorted(pairs) == [tuple(p) for p in ce]
