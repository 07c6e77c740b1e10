"""Oracle graph build pinned on the CPU: KNN / weights / edges / FPS.

Each oracle function is checked against an independent pure-Python (numpy
scalar) restatement of the reference rule on small inputs, against hand-built
tie cases and closed forms, and on the stripes graph fixture
(datasets/stripes/.../graph.obj, 200 nodes):

  findNearestNodes   Deform.hpp:153-185  selection sort, j scanned from the end,
                     strict '<', slots swapped after every pick
  computeWeights     Deform.hpp:187-208  w = (1 - d/d_K)^2 / sum, K=1 -> 1
  setupEdges         Deform.hpp:84-95    (K+1)-NN of each node, slot 0 dropped
  FPS                helper.cpp:139-195  first = argmax(x+y+z) (first max),
                     then first argmax of the running min distance
                     (pts_distance in double, stored as float)
"""
import json
import math
import os

import numpy as np
import pytest

import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
f32 = np.float32


def stripes():
    with open(os.path.join(GOLDEN, "stripes_graph.json")) as f:
        d = json.load(f)
    return np.array(d["nodes"], np.float32), d


def knn_brute(nodes, q, k):
    """findNearestNodes restated: float differences and sum of squares, float
    sqrt stored as double, selection sort with swaps."""
    M = nodes.shape[0]
    d = []
    for j in range(M):
        t = (f32(q[0]) - nodes[j, 0], f32(q[1]) - nodes[j, 1], f32(q[2]) - nodes[j, 2])
        s = f32(f32(t[0] * t[0] + t[1] * t[1]) + t[2] * t[2])
        d.append(float(np.sqrt(s)))
    index = list(range(M))
    out = []
    for i in range(k):
        m = i
        for j in range(M - 1, i, -1):
            if d[j] < d[m]:
                m = j
        out.append(index[m])
        d[i], d[m] = d[m], d[i]
        index[i], index[m] = index[m], index[i]
    return out


def weights_brute(nodes, q, idx, k):
    def dist(j):
        t = (f32(q[0]) - nodes[j, 0], f32(q[1]) - nodes[j, 1], f32(q[2]) - nodes[j, 2])
        return float(np.sqrt(f32(f32(t[0] * t[0] + t[1] * t[1]) + t[2] * t[2])))
    dmax = dist(idx[k])
    w = [(1.0 - dist(idx[j]) / dmax) ** 2 for j in range(k)]
    if k == 1:
        return [1.0]
    s = sum(w)
    return [v / s for v in w]


@pytest.mark.parametrize("k", [1, 4, 8, 12])
def test_knn_matches_selection_sort(k):
    rng = np.random.default_rng(k)
    nodes = rng.uniform(-1, 1, size=(60, 3)).astype(np.float32)
    qs = rng.uniform(-1.2, 1.2, size=(40, 3)).astype(np.float32)
    idx, w = O.knn_weights(nodes, qs, k)
    for i in range(qs.shape[0]):
        ref = knn_brute(nodes, qs[i], k + 1)
        assert list(idx[i]) == ref
        np.testing.assert_allclose(w[i], weights_brute(nodes, qs[i], ref, k), rtol=0, atol=1e-15)
        assert abs(w[i].sum() - 1.0) < 1e-12


def test_knn_ties_duplicates_and_lattice():
    # duplicate nodes and a symmetric lattice: equal distances everywhere
    lat = np.array([[x, y, z] for x in (-1, 0, 1) for y in (-1, 0, 1) for z in (-1, 0, 1)], np.float32)
    nodes = np.concatenate([lat, lat[:5], lat[10:12]])
    qs = np.array([[0, 0, 0], [0.5, 0.5, 0.5], [1, 1, 1], [0.25, 0, 0], [-1, -1, -1]], np.float32)
    for k in (1, 3, 8, 12):
        idx, _ = O.knn_weights(nodes, qs, k)
        for i in range(qs.shape[0]):
            assert list(idx[i]) == knn_brute(nodes, qs[i], k + 1)


def test_knn_self_first_and_k1_weight():
    rng = np.random.default_rng(5)
    nodes = rng.normal(size=(30, 3)).astype(np.float32)
    idx, w = O.knn_weights(nodes, nodes, 1)
    assert np.all(idx[:, 0] == np.arange(30))   # distance 0 to itself
    assert np.all(w[:, 0] == 1.0)


def test_graph_edges_drop_slot0():
    nodes, _ = stripes()
    for k in (4, 8):
        nbr = O.graph_edges(nodes, k)
        for i in range(0, nodes.shape[0], 7):
            assert list(nbr[i]) == knn_brute(nodes, nodes[i], k + 1)[1:]


def fps_brute(pos, m):
    N = pos.shape[0]
    m = min(m, N)
    sums = [float(f32(f32(f32(0) + pos[i, 0]) + pos[i, 1]) + pos[i, 2]) for i in range(N)]
    first = max(range(N), key=lambda i: (sums[i], -i))
    out = [first]
    dist = [float(np.finfo(np.float32).max)] * N
    for i in range(m):
        if i == first:
            continue
        last = pos[out[-1]]
        for j in range(N):
            dd = f32(math.sqrt(sum(float(f32(last[c] - pos[j, c])) ** 2 for c in range(3))))
            if dd < dist[j]:
                dist[j] = float(dd)
        best = max(range(N), key=lambda j: (dist[j], -j))
        out.append(best)
        if len(out) == m:
            break
    return out


@pytest.mark.parametrize("m", [1, 2, 17, 64])
def test_fps_matches_restatement(m):
    rng = np.random.default_rng(m)
    pos = rng.uniform(-1, 1, size=(300, 3)).astype(np.float32)
    assert list(O.fps(pos, m)) == fps_brute(pos, m)


def test_fps_first_node_and_stripes_fixture():
    nodes, _ = stripes()
    sel = O.fps(nodes, 50)
    s = nodes.astype(np.float64).sum(1)
    assert sel[0] == int(np.argmax(s))            # first max of x+y+z wins
    assert len(set(sel.tolist())) == 50          # no repeats while distances are > 0
    assert list(sel) == fps_brute(nodes, 50)
    # m > N: every point exactly once
    allp = O.fps(nodes, 500)
    assert sorted(allp.tolist()) == list(range(nodes.shape[0]))


def test_fps_single_node_corner():
    """For m == 1 the reference loop still appends the farthest point when the
    first node is not index 0 (helper.cpp:157-174)."""
    pos = np.array([[0, 0, 0], [1, 1, 1], [-1, 0, 0]], np.float32)
    assert list(O.fps(pos, 1)) == fps_brute(pos, 1) == [1, 2]
    pos2 = np.array([[1, 1, 1], [0, 0, 0]], np.float32)
    assert list(O.fps(pos2, 1)) == fps_brute(pos2, 1) == [0]
