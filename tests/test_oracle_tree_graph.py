"""The oracle's tree, kNN, depth and order restatements against the reference's own
known-answer tests (tests/test_tree.py, tests/test_graph.py) and brute force."""
import numpy as np
import pytest

import oracle as O
from helpers import ref_points


def test_build_tree_simple():
    # tests/test_tree.py:19-24
    pts, sd, idx = O.build_tree(np.arange(1, 11, dtype=np.float64)[:, None])
    assert np.array_equal(pts[:, 0], [6, 3, 9, 2, 8, 5, 10, 1, 7, 4])
    assert np.array_equal(sd, np.zeros(10))
    assert np.array_equal(idx, [5, 2, 8, 1, 7, 4, 9, 0, 6, 3])


@pytest.mark.parametrize("n,d,seed,kind", [(1, 3, 0, "c"), (2, 2, 1, "c"), (17, 3, 3, "c"), (1000, 6, 4, "c"),
                                           (65537, 3, 5, "c"), (40000, 2, 6, "c"), (20000, 3, 7, "lattice"),
                                           (5000, 2, 8, "ref")])
def test_build_tree_runs_equals_literal(n, d, seed, kind):
    # the default build sorts each node's contiguous run independently; the literal build
    # sorts the whole array by (node, value, position) per level as tree.py:153-184 does
    from helpers import clustered_points

    if kind == "c":
        p = clustered_points(n, d=d, seed=seed)
    elif kind == "ref":
        p = ref_points((n, d))
    else:
        rng = np.random.default_rng(seed)
        p = (rng.integers(0, 5, (n, d)) * 0.5).astype(np.float64)
        p[::7, 1] = -0.0
    a, b = O.build_tree(p), O.build_tree(p, literal=True)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_rng_canary():
    # tests/test_tree.py:29 pins the JAX RNG restatement used below
    assert np.isclose(ref_points((100000, 3))[0, 0], -0.61777326, rtol=1e-8)


def test_build_tree_random():
    # tests/test_tree.py:27-32
    pts, sd, idx = O.build_tree(ref_points((100000, 3)))
    assert np.array_equal(sd[1000:1005], [2, 0, 0, 1, 1])
    assert idx[-1] == 59735


@pytest.mark.parametrize("d", [2, 6])
def test_build_tree_shapes(d):
    # tests/test_tree.py:35-41
    pts, sd, idx = O.build_tree(ref_points((1000, d)))
    assert pts.shape == (1000, d) and sd.shape == (1000,) and idx.shape == (1000,)
    assert np.array_equal(np.sort(idx), np.arange(1000))


def test_tree_invariant_split_planes():
    # every node separates its subtrees along its split dimension
    rng = np.random.default_rng(3)
    p = rng.standard_normal((777, 3))
    pts, sd, idx = O.build_tree(p)
    n = len(pts)

    def subtree(c):
        out, stack = [], [c]
        while stack:
            x = stack.pop()
            if x >= n:
                continue
            out.append(x)
            lvl = int(np.floor(np.log2(x + 1)))
            stack += [x + (1 << lvl), x + (2 << lvl)]
        return out

    for c in range(0, 200):
        lvl = int(np.floor(np.log2(c + 1)))
        dim = sd[c]
        for x in subtree(c + (1 << lvl)):
            assert pts[x, dim] <= pts[c, dim]
        for x in subtree(c + (2 << lvl)):
            assert pts[x, dim] >= pts[c, dim]


def _brute_preceding(points, n0, k):
    dist = np.linalg.norm(points[:, None, :] - points[None, :, :], axis=-1)
    i, j = np.indices(dist.shape)
    dist[i <= j] = np.inf
    return np.argsort(dist, axis=-1, kind="stable")[n0:, :k]


def test_query_preceding_bruteforce():
    # tests/test_tree.py:44-59
    pts, sd, idx = O.build_tree(ref_points((1000, 3)))
    nb = O.query_preceding_neighbors(pts, sd, n0=30, k=30)
    assert np.array_equal(nb, _brute_preceding(pts, 30, 30))


def test_query_preceding_ties():
    # Integer lattice: many exact distance ties.  The reference keeps, among candidates
    # tied at the k-th distance, the one its traversal offers first (tree.py:84-95), so
    # the check is: the k distances equal the k smallest preceding distances, and the
    # list is ordered by (distance, index) (tree.py:102-104).
    rng = np.random.default_rng(5)
    p = rng.integers(0, 6, (600, 2)).astype(np.float64)
    pts, sd, idx = O.build_tree(p)
    n0 = k = 12
    nb = O.query_preceding_neighbors(pts, sd, n0=n0, k=k)
    for r in range(len(nb)):
        q = n0 + r
        dall = np.sort(np.linalg.norm(pts[:q] - pts[q], axis=1))[:k]
        dsel = np.linalg.norm(pts[nb[r]] - pts[q], axis=1)
        assert np.array_equal(dsel, dall)
        assert np.all(nb[r] < q) and len(set(nb[r].tolist())) == k
        order = np.lexsort((nb[r], dsel))
        assert np.array_equal(order, np.arange(k))


def test_query_n0_less_than_k():
    pts, sd, idx = O.build_tree(ref_points((100, 2)))
    with pytest.raises(ValueError):
        O.query_preceding_neighbors(pts, sd, n0=5, k=10)


def test_compute_depths():
    # tests/test_graph.py:40-52
    nb = np.array([[0, 1, 2], [0, 1, 2], [1, 3, 4], [2, 4, 5]])
    assert np.array_equal(O.compute_depths(nb, n0=3), [0, 0, 0, 1, 1, 2, 3])
    assert np.array_equal(O.compute_depths_jacobi(nb, n0=3), [0, 0, 0, 1, 1, 2, 3])


def test_depths_forward_equals_jacobi():
    pts, sd, idx = O.build_tree(ref_points((3000, 2)))
    nb = O.query_preceding_neighbors(pts, sd, n0=50, k=8)
    assert np.array_equal(O.compute_depths(nb, n0=50), O.compute_depths_jacobi(nb, n0=50))


def test_order_by_depth():
    # tests/test_graph.py:55-82
    points = np.array([0.0, 1.0, 2.0, 3.0, 4.0, 5.0])[:, None]
    indices = np.arange(6)
    neighbors = np.array([[0, 1], [1, 2], [2, 3], [0, 1]])
    depths = np.array([0, 0, 1, 2, 3, 1])
    p, i, n, d = O.order_by_depth(points, indices, neighbors, depths)
    assert np.array_equal(p[:, 0], [0, 1, 2, 5, 3, 4])
    assert np.array_equal(i, [0, 1, 2, 5, 3, 4])
    assert np.array_equal(n, [[0, 1], [0, 1], [1, 2], [2, 4]])
    assert np.array_equal(d, [0, 0, 1, 1, 2, 3])


def test_graph_random():
    # tests/test_graph.py:14-27
    g = O.build_graph(ref_points((100000, 3)), n0=1000, k=10)
    assert len(g.offsets) == 61
    assert g.offsets[1] == 1115
    assert g.neighbors[0, 0] == 488
    assert np.isclose(g.points[0, 0], -1.38417124, rtol=1e-8)


def test_graph_shapes():
    # tests/test_graph.py:30-37
    g = O.build_graph(ref_points((5000, 6)), n0=500, k=15)
    assert g.points.shape == (5000, 6) and g.neighbors.shape == (4500, 15) and g.indices.shape == (5000,)
    assert g.offsets[0] == 500 and g.offsets[-1] == 5000
