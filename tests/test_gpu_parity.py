"""Parity of the gfx950 kernels (through the C ABI) with the oracle and with the reference's
own known answers.  Integer outputs (tree permutation, split dims, neighbours, depths,
order, offsets) must match bit for bit; floating-point outputs within the tolerance
stated in each test."""
import numpy as np
import pytest
import torch

import graphgp_amd as gp
import oracle as O
from helpers import clustered_points, ref_points

pytestmark = pytest.mark.gpu


def cpu(t):
    return t.detach().cpu().numpy()


# ---------------------------------------------------------------- tree
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_build_tree_simple(dtype):
    # tests/test_tree.py:19-24
    pts, sd, idx = gp.build_tree(np.arange(1, 11, dtype=dtype)[:, None])
    assert np.array_equal(cpu(pts)[:, 0], [6, 3, 9, 2, 8, 5, 10, 1, 7, 4])
    assert np.array_equal(cpu(sd), np.zeros(10))
    assert np.array_equal(cpu(idx), [5, 2, 8, 1, 7, 4, 9, 0, 6, 3])


def test_build_tree_random_reference_kat():
    # tests/test_tree.py:27-32, whole arrays against the oracle
    p = ref_points((100000, 3))
    pts, sd, idx = gp.build_tree(p)
    assert np.array_equal(cpu(sd)[1000:1005], [2, 0, 0, 1, 1])
    assert int(cpu(idx)[-1]) == 59735
    op, osd, oidx = O.build_tree(p)
    assert np.array_equal(cpu(idx), oidx)
    assert np.array_equal(cpu(sd), osd)
    assert np.array_equal(cpu(pts), op)


@pytest.mark.parametrize("n,d,seed", [(1, 3, 0), (2, 2, 1), (3, 1, 2), (17, 3, 3), (1000, 6, 4), (65537, 3, 5),
                                      (40000, 2, 6)])
def test_build_tree_vs_oracle_f32_clustered(n, d, seed):
    p = clustered_points(n, d=d, seed=seed)
    pts, sd, idx = gp.build_tree(p)
    op, osd, oidx = O.build_tree(p)
    assert np.array_equal(cpu(idx), oidx)
    assert np.array_equal(cpu(sd), osd)
    assert np.array_equal(cpu(pts), op)


@pytest.mark.parametrize("n", [3000, 30000])
def test_build_tree_f64_lattice(n):
    # float64 keys with many ties, all levels in one workgroup (3000) and the hybrid (30000)
    rng = np.random.default_rng(n)
    p = rng.integers(0, 9, (n, 3)).astype(np.float64) * 0.25
    pts, sd, idx = gp.build_tree(p)
    op, osd, oidx = O.build_tree(p)
    assert np.array_equal(cpu(idx), oidx) and np.array_equal(cpu(sd), osd)


def test_build_tree_duplicates_and_ties():
    rng = np.random.default_rng(11)
    p = rng.integers(0, 7, (20000, 3)).astype(np.float32)
    p[::5] = -0.0  # signed zeros sort equal to +0.0
    pts, sd, idx = gp.build_tree(p)
    op, osd, oidx = O.build_tree(p)
    assert np.array_equal(cpu(idx), oidx) and np.array_equal(cpu(sd), osd)


# ---------------------------------------------------------------- kNN
@pytest.mark.parametrize("k,n0,n,d,dtype", [(30, 30, 1000, 3, np.float64), (10, 100, 20000, 2, np.float32),
                                            (30, 1000, 60000, 3, np.float32), (8, 8, 5000, 1, np.float64),
                                            (16, 50, 8000, 6, np.float32), (70, 70, 1500, 3, np.float32)])
def test_query_preceding_vs_oracle(k, n0, n, d, dtype):
    p = clustered_points(n, d=d, seed=k + n0, dtype=dtype)
    op, osd, _ = O.build_tree(p)
    nb = gp.query_preceding_neighbors(op, osd, n0=n0, k=k)
    onb = O.query_preceding_neighbors(op, osd, n0=n0, k=k)
    assert np.array_equal(cpu(nb), onb)


def test_query_preceding_bruteforce_reference():
    # tests/test_tree.py:44-59
    pts, sd, _ = gp.build_tree(ref_points((1000, 3)))
    nb = cpu(gp.query_preceding_neighbors(pts, sd, n0=30, k=30))
    P = cpu(pts)
    dist = np.linalg.norm(P[:, None] - P[None], axis=-1)
    i, j = np.indices(dist.shape)
    dist[i <= j] = np.inf
    assert np.array_equal(nb, np.argsort(dist, axis=-1, kind="stable")[30:, :30])


def test_query_preceding_lattice_ties():
    rng = np.random.default_rng(5)
    p = rng.integers(0, 6, (3000, 2)).astype(np.float32)
    op, osd, _ = O.build_tree(p)
    nb = gp.query_preceding_neighbors(op, osd, n0=12, k=12)
    assert np.array_equal(cpu(nb), O.query_preceding_neighbors(op, osd, n0=12, k=12))


def test_query_neighbors_explicit_indices():
    p = clustered_points(5000, d=3, seed=9)
    op, osd, _ = O.build_tree(p)
    rng = np.random.default_rng(0)
    q = rng.integers(0, 5000, 700)
    m = np.maximum(q, rng.integers(20, 5000, 700))
    nb = gp.query_neighbors(op, osd, q, m, k=12)
    assert np.array_equal(cpu(nb), O.query_neighbors(op, osd, q, m, k=12))


def test_query_preceding_under_cu_budget():
    # a CU budget turns the kNN launch into a ticketed persistent grid; neighbours unchanged
    p = clustered_points(60000, d=3, seed=77, dtype=np.float32)
    op, osd, _ = O.build_tree(p)
    lib = gp._lib.lib()
    prev = lib.ggp_set_cu_budget(2)
    try:
        nb = gp.query_preceding_neighbors(op, osd, n0=32, k=32)
        nb8 = gp.query_preceding_neighbors(op, osd, n0=8, k=8)
    finally:
        lib.ggp_set_cu_budget(prev)
    assert np.array_equal(cpu(nb), O.query_preceding_neighbors(op, osd, n0=32, k=32))
    assert np.array_equal(cpu(nb8), O.query_preceding_neighbors(op, osd, n0=8, k=8))


def test_query_n0_less_than_k_raises():
    pts, sd, _ = gp.build_tree(ref_points((100, 2)))
    with pytest.raises(ValueError, match="n0 must be at least k"):
        gp.query_preceding_neighbors(pts, sd, n0=5, k=10)


# ---------------------------------------------------------------- graph
@pytest.mark.parametrize("method", ["flow", "rounds", "sweeps"])
def test_compute_depths_kat(method):
    # tests/test_graph.py:40-52
    nb = np.array([[0, 1, 2], [0, 1, 2], [1, 3, 4], [2, 4, 5]])
    assert np.array_equal(cpu(gp.compute_depths(nb, n0=3, method=method)), [0, 0, 0, 1, 1, 2, 3])


@pytest.mark.parametrize("method", ["flow", "rounds", "sweeps"])
def test_compute_depths_random_vs_oracle(method):
    # random graphs in topological order with repeated neighbours, and a 3000-long chain
    # (3000 rounds: more than one batch of rounds between host checks)
    rng = np.random.default_rng(5)
    for n0, n_ref, k in [(5, 2000, 4), (50, 30000, 8), (1, 3000, 1)]:
        hi = n0 + np.arange(n_ref)[:, None]
        if k == 1:
            nb = (hi - 1).astype(np.int32)
        else:
            nb = np.minimum(rng.integers(0, hi, (n_ref, k)), hi - 1 - rng.integers(0, 3, (n_ref, k)))
            nb = np.maximum(nb, 0).astype(np.int32)
        d = cpu(gp.compute_depths(nb, n0=n0, method=method))
        assert np.array_equal(d, O.compute_depths(nb, n0=n0)), (n0, n_ref, k)


@pytest.mark.parametrize("method", ["flow", "rounds", "sweeps"])
def test_compute_depths_hub_and_chain(method):
    # one refined point is a neighbour of every later point (thousands of children), the
    # second column chains the points, the third is random
    rng = np.random.default_rng(11)
    n0, n_ref = 4, 5000
    i = n0 + np.arange(n_ref)
    nb = np.stack([np.where(i > n0, n0, 0), np.maximum(i - 1, 0), rng.integers(0, i)], axis=1).astype(np.int32)
    d = cpu(gp.compute_depths(nb, n0=n0, method=method))
    assert np.array_equal(d, O.compute_depths(nb, n0=n0))
    assert d.max() == n_ref


def test_order_by_depth_kat():
    # tests/test_graph.py:55-82
    points = np.array([0.0, 1.0, 2.0, 3.0, 4.0, 5.0])[:, None]
    p, i, n, d = gp.order_by_depth(points, np.arange(6), np.array([[0, 1], [1, 2], [2, 3], [0, 1]]),
                                   np.array([0, 0, 1, 2, 3, 1]))
    assert np.array_equal(cpu(p)[:, 0], [0, 1, 2, 5, 3, 4])
    assert np.array_equal(cpu(i), [0, 1, 2, 5, 3, 4])
    assert np.array_equal(cpu(n), [[0, 1], [0, 1], [1, 2], [2, 4]])
    assert np.array_equal(cpu(d), [0, 0, 1, 1, 2, 3])


@pytest.mark.parametrize("method", ["flow", "rounds", "sweeps"])
def test_compute_depths_rejects_malformed(method):
    with pytest.raises(ValueError):
        gp.compute_depths(np.array([[0, 4], [1, 2]]), n0=3, method=method)
    with pytest.raises(ValueError):
        gp.compute_depths(np.array([[0, 1], [-1, 2]]), n0=3, method=method)


def test_build_graph_reference_kat():
    # tests/test_graph.py:14-27, plus whole arrays against the oracle
    p = ref_points((100000, 3))
    g = gp.build_graph(p, n0=1000, k=10)
    assert len(g.offsets) == 61
    assert g.offsets[1] == 1115
    assert int(g.neighbors[0, 0]) == 488
    assert np.isclose(float(g.points[0, 0]), -1.38417124, rtol=1e-8)
    og = O.build_graph(p, n0=1000, k=10)
    assert g.offsets == og.offsets
    assert np.array_equal(cpu(g.neighbors), og.neighbors)
    assert np.array_equal(cpu(g.indices), og.indices)
    assert np.array_equal(cpu(g.points), og.points)
    gp.check_graph(g)


@pytest.mark.parametrize("n,d,n0,k", [(5000, 6, 500, 15), (50000, 3, 1000, 30), (20000, 2, 100, 10)])
def test_build_graph_vs_oracle_f32(n, d, n0, k):
    p = clustered_points(n, d=d, seed=n)
    g = gp.build_graph(p, n0=n0, k=k)
    og = O.build_graph(p, n0=n0, k=k)
    assert g.offsets == og.offsets
    assert g.offsets[0] == n0 and g.offsets[-1] == n
    assert np.array_equal(cpu(g.neighbors), og.neighbors)
    assert np.array_equal(cpu(g.indices), og.indices)
    assert np.array_equal(cpu(g.points), og.points)


def test_check_graph_detects_errors():
    p = clustered_points(3000, d=2, seed=1)
    g = gp.build_graph(p, n0=50, k=6)
    gp.check_graph(g)
    bad = g.neighbors.clone()
    bad[-1, 0] = g.points.shape[0] - 1
    with pytest.raises(AssertionError, match="topological"):
        gp.check_graph(gp.Graph(g.points, bad, g.offsets, g.indices))
    with pytest.raises(AssertionError, match="first offset"):
        gp.check_graph(gp.Graph(g.points, g.neighbors, (49,) + g.offsets[1:], g.indices))
    # graph.py:58-60: non-decreasing offsets, last offset within the points
    off = list(g.offsets)
    off[2], off[3] = off[3], off[2]
    with pytest.raises(AssertionError, match="non-decreasing"):
        gp.check_graph(gp.Graph(g.points, g.neighbors, tuple(off), g.indices))
    with pytest.raises(AssertionError, match="Last offset"):
        gp.check_graph(gp.Graph(g.points, g.neighbors, g.offsets[:-1] + (3001,), g.indices))
    # graph.py:68-70: a neighbour inside its own batch (but preceding, so the order holds)
    b = 3
    i = g.offsets[b] + 5
    assert i < g.offsets[b + 1]
    bad = g.neighbors.clone()
    bad[i - g.offsets[0], 0] = g.offsets[b]
    with pytest.raises(AssertionError, match="same batch"):
        gp.check_graph(gp.Graph(g.points, bad, g.offsets, g.indices))


def test_check_graph_device_status_bits():
    # the device rules through the C ABI, including the offsets bit that the Python layer
    # asserts on the host before it launches (bit 0 malformed offsets, 1 order, 2 same batch)
    import ctypes

    from graphgp_amd import _lib

    p = clustered_points(3000, d=2, seed=1)
    g = gp.build_graph(p, n0=50, k=6)
    n_ref, k = g.neighbors.shape

    def status(nb, offsets):
        off = torch.tensor(offsets, dtype=torch.int64, device="cuda")
        st = torch.zeros(1, dtype=torch.int32, device="cuda")
        _lib.check(_lib.lib().ggp_check_graph(_lib.ptr(nb), n_ref, k, 3000, _lib.ptr(off), len(offsets),
                                              _lib.ptr(st), _lib.stream()), "check_graph")
        return int(st.cpu()[0])

    assert status(g.neighbors, g.offsets) == 0
    off = list(g.offsets)
    off[2], off[3] = off[3], off[2]
    assert status(g.neighbors, off) & 1
    assert status(g.neighbors, list(g.offsets[:-1]) + [3001]) & 1
    assert status(g.neighbors, [49] + list(g.offsets[1:])) & 1
    bad = g.neighbors.clone()
    bad[-1, 0] = 2999
    assert status(bad, g.offsets) & 2
    bad = g.neighbors.clone()
    bad[g.offsets[3] + 5 - 50, 0] = g.offsets[3]
    assert status(bad, g.offsets) == 4


# ---------------------------------------------------------------- refine (FP64)
@pytest.fixture(scope="module")
def ref_setup():
    # tests/test_refine.py:15-26
    points = ref_points((1000, 3))
    og = O.build_graph(points, n0=100, k=10)
    g = gp.build_graph(points, n0=100, k=10)
    cov = gp.extras.matern_kernel(p=0, variance=1.0, cutoff=1.0, r_min=1e-4, r_max=10, n_bins=1000, jitter=1e-5)
    ocov = O.matern_kernel(p=0, variance=1.0, cutoff=1.0, r_min=1e-4, r_max=10, n_bins=1000, jitter=1e-5)
    return g, og, cov, ocov


def test_logdet_reference_kat(ref_setup):
    # tests/test_refine.py:29-43 (rtol 1e-8)
    g, og, cov, _ = ref_setup
    assert np.isclose(float(g.points[0, 0]), -1.95624711, rtol=1e-8)
    assert np.isclose(float(gp.generate_logdet(g, cov, precision="fp64")), -600.36165088, rtol=1e-8)
    assert np.isclose(float(gp.generate_dense_logdet(g.points, cov)), -610.90538067, rtol=1e-8)


def test_generate_fp64_vs_oracle(ref_setup):
    g, og, cov, ocov = ref_setup
    xi = ref_points((1000,))
    v = cpu(gp.generate(g, cov, xi))
    ov = O.generate(og, ocov, xi)
    assert v.dtype == np.float64
    assert np.allclose(v, ov, rtol=1e-10, atol=1e-12)


def test_inverse_round_trip_fp64(ref_setup):
    # tests/test_refine.py:46-54
    g, og, cov, ocov = ref_setup
    xi = ref_points((1000,))
    v = gp.generate(g, cov, xi)
    xb = gp.generate_inv(g, cov, v)
    assert np.allclose(cpu(xb), O.generate_inv(og, ocov, cpu(v)), rtol=1e-9, atol=1e-10)
    assert np.allclose(cpu(gp.generate(g, cov, xb)), cpu(v), rtol=1e-12, atol=1e-13)


def test_raw_refine_api_fp64(ref_setup):
    g, og, cov, ocov = ref_setup
    n0 = g.offsets[0]
    xi = ref_points((1000,))[og.indices]
    init = O.generate_dense(og.points[:n0], ocov, xi[:n0])
    v = cpu(gp.refine(g.points, g.neighbors, g.offsets, cov, init, xi[n0:]))
    ov = O.refine(og.points, og.neighbors, og.offsets, ocov, init, xi[n0:])
    assert np.allclose(v, ov, rtol=1e-10, atol=1e-12)
    i0, xr = gp.refine_inv(g.points, g.neighbors, g.offsets, cov, v)
    assert np.allclose(cpu(i0), init, rtol=1e-14) and np.allclose(cpu(xr), xi[n0:], rtol=1e-8, atol=1e-9)
    ld = float(gp.refine_logdet(g.points, g.neighbors, g.offsets, cov))
    assert np.isclose(ld, O.refine_logdet(og.points, og.neighbors, og.offsets, ocov), rtol=1e-12)
    assert np.allclose(cpu(gp.generate_dense(g.points[:n0], cov, xi[:n0])), init, rtol=1e-10, atol=1e-12)
    assert np.allclose(cpu(gp.generate_dense_inv(g.points[:n0], cov, init)), xi[:n0], rtol=1e-8, atol=1e-9)
    K = cpu(gp.compute_cov_matrix(cov, g.points[:50], g.points[:70]))
    assert np.allclose(K, O.compute_cov_matrix(ocov, og.points[:50], og.points[:70]), rtol=1e-13)


def test_approaches_dense_generic_k():
    # tests/test_refine.py:66-80 (k = n0 = 200 runs the generic-k kernel)
    points = ref_points((1000, 3))
    g = gp.build_graph(points, n0=200, k=200)
    g = gp.Graph(g.points, g.neighbors, g.offsets)
    cov = gp.extras.matern_kernel(p=0, variance=1.0, cutoff=1.0, r_min=1e-4, r_max=10, n_bins=1000, jitter=1e-5)
    xi = ref_points((1000,))
    true_v = cpu(gp.generate_dense(g.points, cov, xi))
    v = cpu(gp.generate(g, cov, xi))
    assert np.allclose(true_v, v, atol=0.02)
    ocov = O.matern_kernel(p=0, variance=1.0, cutoff=1.0, r_min=1e-4, r_max=10, n_bins=1000, jitter=1e-5)
    og = O.OGraph(cpu(g.points), cpu(g.neighbors), g.offsets, None)
    assert np.allclose(v, O.generate(og, ocov, xi), rtol=1e-9, atol=1e-10)


# ---------------------------------------------------------------- every register configuration
# One case per register configuration of the refinement kernel: k + 1 = 4, 8, ..., 40 runs
# one, two, four and eight lanes per point in FP32 and in FP64, each with its own lane
# broadcasts and group sums.
@pytest.mark.parametrize("k", [3, 7, 11, 15, 19, 23, 27, 31, 39])
def test_generate_each_register_configuration_vs_oracle(k):
    rng = np.random.default_rng(k)
    n = 3000
    p = rng.uniform(0, 1, (n, 3)).astype(np.float32)
    g = gp.build_graph(p, n0=max(k, 64), k=k)
    og = O.OGraph(cpu(g.points), cpu(g.neighbors), g.offsets, cpu(g.indices))
    kw = dict(p=1, variance=1.0, cutoff=0.2, r_min=1e-5, r_max=2.0, n_bins=2001, jitter=1e-5)
    cov, ocov = gp.extras.matern_kernel(**kw), O.matern_kernel(**kw)
    xi = rng.standard_normal(n)
    ov = O.generate(og, ocov, xi)
    scale = float(np.abs(ov).max())
    v64 = cpu(gp.generate(g, cov, xi, precision="fp64"))
    assert np.max(np.abs(v64 - ov)) <= 1e-9 * scale, (k, float(np.max(np.abs(v64 - ov))))
    for prec, tol in (("fp32", 2e-2), ("fp32_increment", 2e-3)):
        v = cpu(gp.generate(g, cov, xi.astype(np.float32), precision=prec))
        assert np.max(np.abs(v - ov)) <= tol * scale, (k, prec, float(np.max(np.abs(v - ov))))
    xb = cpu(gp.generate_inv(g, cov, ov, precision="fp64"))
    assert np.max(np.abs(xb - xi)) <= 1e-6, (k, float(np.max(np.abs(xb - xi))))


# ---------------------------------------------------------------- refine (FP32)
@pytest.fixture(scope="module")
def c1_setup():
    # the 2-D plain-method workload at 1/5 size: N(0,1) points, RBF var 1 scale 0.3, jitter 1e-4
    rng = np.random.default_rng(99)
    p = rng.standard_normal((20000, 2)).astype(np.float32)
    g = gp.build_graph(p, n0=100, k=10)
    og = O.OGraph(cpu(g.points), cpu(g.neighbors), g.offsets, cpu(g.indices))
    cov = gp.extras.rbf_kernel(variance=1.0, scale=0.3, r_min=1e-4, r_max=10, n_bins=1001, jitter=1e-4)
    ocov = O.rbf_kernel(variance=1.0, scale=0.3, r_min=1e-4, r_max=10, n_bins=1001, jitter=1e-4)
    xi = rng.standard_normal(20000).astype(np.float32)
    return g, og, cov, ocov, xi


@pytest.mark.parametrize("precision,tol", [("fp32", 2e-3), ("fp32_increment", 2e-4), ("fp64", 1e-10)])
def test_generate_precisions_vs_fp64_oracle(c1_setup, precision, tol):
    g, og, cov, ocov, xi = c1_setup
    v = cpu(gp.generate(g, cov, xi, precision=precision))
    ov = O.generate(og, ocov, xi)
    err = np.abs(v - ov)
    assert np.max(err) <= tol, (precision, float(np.max(err)), float(np.sqrt(np.mean(err ** 2))))


# log-det tolerance: plain FP32 loses up to ~1e-3 relative in sigma where sigma^2 = c - b'A^-1 b
# cancels (near-duplicate neighbours); the increment formulation keeps it to ~1e-6 relative.
@pytest.mark.parametrize("precision,tol,ld_rtol", [("fp32", 2e-3, 1e-4), ("fp32_increment", 2e-4, 1e-5)])
def test_inverse_and_logdet_fp32(c1_setup, precision, tol, ld_rtol):
    g, og, cov, ocov, xi = c1_setup
    ov = O.generate(og, ocov, xi)
    xb = cpu(gp.generate_inv(g, cov, ov.astype(np.float32), precision=precision))
    oxb = O.generate_inv(og, ocov, ov.astype(np.float32).astype(np.float64))
    assert np.max(np.abs(xb - oxb)) <= 50 * tol
    ld = float(gp.generate_logdet(g, cov, precision=precision))
    assert np.isclose(ld, O.generate_logdet(og, ocov), rtol=ld_rtol)


def test_nan_rule_non_pd():
    # duplicate points with zero jitter make the local covariance singular: all NaN
    p = np.repeat(clustered_points(500, d=2, seed=2), 2, axis=0)
    g = gp.build_graph(p, n0=10, k=5)
    cov = gp.extras.rbf_kernel(variance=1.0, scale=0.3, r_min=1e-4, r_max=10, n_bins=100, jitter=0.0)
    v = cpu(gp.generate(g, cov, np.ones(1000, np.float32)))
    assert np.all(np.isnan(v))
    x = cpu(gp.generate_inv(g, cov, np.ones(1000, np.float32)))
    assert np.all(np.isnan(x))


def test_indefinite_seed_block_reference_rules():
    # the 1000 seeds of a 4,194,304-point clustered graph (tests/golden/make_indefinite_seeds.py);
    # with jitter 1e-5 their interpolated covariance is indefinite.  As in the reference,
    # the failed Cholesky makes every value of generate_dense / generate_dense_inv NaN
    # (refine.py:74, :193), while the log-det is slogdet's finite log|det K| / 2 (refine.py:248)
    import os

    p = np.load(os.path.join(os.path.dirname(__file__), "golden", "indefinite_seeds_4m.npy"))
    kw = dict(p=1, variance=1.0, cutoff=0.05, r_min=1e-6, r_max=2.0, n_bins=2001, jitter=1e-5)
    cov, ocov = gp.extras.matern_kernel(**kw), O.matern_kernel(**kw)
    xi = np.ones(len(p))
    assert np.all(np.isnan(O.generate_dense(p, ocov, xi)))
    assert np.all(np.isnan(cpu(gp.generate_dense(p, cov, xi))))
    assert np.all(np.isnan(cpu(gp.generate_dense_inv(p, cov, xi))))
    ld = float(gp.generate_dense_logdet(p, cov))
    old = O.generate_dense_logdet(p, ocov)
    assert np.isfinite(ld) and np.isfinite(old)
    assert np.isclose(ld, old, rtol=1e-9), (ld, old)


def test_logdet_with_indefinite_seed_block_is_finite():
    # generate_logdet = seed block's slogdet / 2 + sum of log std: finite although the seed
    # block is indefinite; generate and generate_inv are all NaN.  The graph is assembled
    # here so that its first 1000 points are exactly the indefinite seeds: 4000 refined
    # points near them, each with its 10 nearest preceding points (brute force), in the
    # reference's depth order (the oracle's depths, order and offsets).
    import os

    seeds = np.load(os.path.join(os.path.dirname(__file__), "golden", "indefinite_seeds_4m.npy")).astype(np.float32)
    rng = np.random.default_rng(5)
    extra = (seeds[rng.integers(0, len(seeds), 4000)] + rng.normal(0, 0.01, (4000, 3))).astype(np.float32)
    pts = np.concatenate([seeds, extra])
    n0, k, n = len(seeds), 10, len(seeds) + len(extra)
    nb = np.empty((n - n0, k), np.int32)
    for r in range(n - n0):
        d2 = ((pts[: n0 + r] - pts[n0 + r]) ** 2).sum(axis=1)
        nb[r] = np.argsort(d2, kind="stable")[:k]
    dep = O.compute_depths(nb, n0=n0)
    po, io, no, do = O.order_by_depth(pts, np.arange(n), nb, dep)
    assert np.array_equal(po[:n0], seeds)  # the seeds keep their places (depth 0)
    g = gp.Graph(torch.from_numpy(po).cuda(), torch.from_numpy(no).cuda(), O.offsets_from_depths(do),
                 torch.from_numpy(io.astype(np.int32)).cuda())
    gp.check_graph(g)
    kw = dict(p=1, variance=1.0, cutoff=0.05, r_min=1e-6, r_max=2.0, n_bins=2001, jitter=1e-5)
    cov, ocov = gp.extras.matern_kernel(**kw), O.matern_kernel(**kw)
    og = O.OGraph(po, no, g.offsets, io)
    sd = O.generate_dense_logdet(seeds, ocov)
    assert np.isfinite(sd) and np.all(np.isnan(O.generate_dense(seeds, ocov, np.ones(n0))))
    ld = float(gp.generate_logdet(g, cov, precision="fp64"))
    assert np.isfinite(ld) and np.isclose(ld, O.generate_logdet(og, ocov), rtol=1e-9)
    assert np.all(np.isnan(cpu(gp.generate(g, cov, np.ones(n), precision="fp64"))))
    assert np.all(np.isnan(cpu(gp.generate_inv(g, cov, np.ones(n), precision="fp64"))))


def test_argument_errors():
    p = clustered_points(500, d=2, seed=3)
    g = gp.build_graph(p, n0=20, k=5)
    cov = gp.extras.rbf_kernel(variance=1.0, scale=0.3, r_min=1e-4, r_max=10, n_bins=100)
    with pytest.raises(ValueError, match="Length of xi"):
        gp.generate(g, cov, np.ones(499, np.float32))
    with pytest.raises(ValueError, match="Invalid covariance"):
        gp.generate(g, [cov[0], cov[1]], np.ones(500, np.float32))
    with pytest.raises(ValueError, match="n0 must be at least k"):
        gp.build_graph(p, n0=4, k=5)


# ---------------------------------------------------------------- seed block (FP64)
@pytest.mark.parametrize("n", [1, 2, 63, 64, 65, 130, 1000])
def test_dense_seed_block_vs_oracle(n):
    rng = np.random.default_rng(n)
    p = rng.uniform(0, 1, (n, 3)).astype(np.float32)
    cov = gp.extras.matern_kernel(p=1, variance=1.0, cutoff=0.05, r_min=1e-6, r_max=2.0, n_bins=2001, jitter=1e-5)
    ocov = O.matern_kernel(p=1, variance=1.0, cutoff=0.05, r_min=1e-6, r_max=2.0, n_bins=2001, jitter=1e-5)
    xi = rng.standard_normal(n)
    v = cpu(gp.generate_dense(p, cov, xi))
    ov = O.generate_dense(p, ocov, xi)
    assert np.allclose(v, ov, rtol=1e-9, atol=1e-11)
    assert np.allclose(cpu(gp.generate_dense_inv(p, cov, ov)), xi, rtol=1e-7, atol=1e-8)
    assert np.isclose(float(gp.generate_dense_logdet(p, cov)), O.generate_dense_logdet(p, ocov), rtol=1e-11)


def test_scheduled_generate_equals_unscheduled(monkeypatch):
    # the dataflow schedule changes the order of evaluation and storage, and lists each
    # point's neighbours (after the nearest) in slot order: the local systems are the same
    # up to a permutation, so the values agree to rounding; the schedule is a permutation
    # that keeps the batch order
    import sys

    R = sys.modules["graphgp_amd.graph"]
    p = clustered_points(200_000, d=3, seed=31)
    monkeypatch.setattr(R, "SCHEDULE_MIN_REFINED", 1 << 40)  # built and first run unscheduled
    g = gp.build_graph(p, n0=1000, k=30)
    cov = gp.extras.matern_kernel(p=1, variance=1.0, cutoff=0.05, r_min=1e-6, r_max=2, n_bins=2001, jitter=1e-5)
    xi = torch.randn(200_000, device="cuda", generator=torch.Generator(device="cuda").manual_seed(3))
    a = gp.generate(g, cov, xi, precision="fp32_increment")
    assert getattr(g, "_ggp_schedule", None) is None
    monkeypatch.setattr(R, "SCHEDULE_MIN_REFINED", 0)
    b = gp.generate(g, cov, xi, precision="fp32_increment")
    c = gp.generate(g, cov, xi, precision="fp32_increment")  # cached schedule
    d64 = float((a - b).abs().max())
    print(f"fp32_increment scheduled vs unscheduled: max |diff| {d64:.3e}", flush=True)
    assert d64 <= 1e-4  # the mode's accuracy bound against FP64
    assert torch.equal(b.view(torch.int32), c.view(torch.int32))
    order, nbv, pts_s = g._ggp_schedule[1]
    o = cpu(order).astype(np.int64)
    n0 = g.offsets[0]
    assert np.array_equal(np.sort(o), np.arange(len(o)))
    batch = np.searchsorted(np.asarray(g.offsets), n0 + o, side="right") - 1
    assert np.all(np.diff(batch) >= 0)
    slot = np.empty_like(o)
    slot[o] = np.arange(len(o))
    nb = cpu(g.neighbors).astype(np.int64)[o]
    expect = np.where(nb < n0, nb, n0 + slot[np.maximum(nb - n0, 0)])
    expect[:, 1:] = np.sort(expect[:, 1:], axis=1)
    assert np.array_equal(cpu(nbv), expect)
    P = cpu(g.points)
    assert np.array_equal(cpu(pts_s), np.concatenate([P[:n0], P[n0 + o]]))
    # the fused kernels in slot space (plain FP32 and FP64 generate)
    monkeypatch.setattr(R, "SCHEDULE_MIN_REFINED", 1 << 40)
    ref = {pr: gp.generate(g, cov, xi, precision=pr) for pr in ("fp32", "fp64")}
    monkeypatch.setattr(R, "SCHEDULE_MIN_REFINED", 0)
    for pr, tol in (("fp32", 2e-2), ("fp64", 1e-9)):
        v = gp.generate(g, cov, xi, precision=pr)
        dv = float((v - ref[pr]).abs().max())
        print(f"{pr} scheduled vs unscheduled: max |diff| {dv:.3e}", flush=True)
        assert dv <= tol, (pr, dv)
    # points changed in place: the cached schedule (slot-order copies) must not be reused
    old = g._ggp_schedule[1]
    with torch.no_grad():
        g.points[n0 + 7] += 1e-3
        g.points[-1] -= 1e-3
    a = gp.generate(g, cov, xi, precision="fp32_increment")
    assert g._ggp_schedule[1] is not old
    P = cpu(g.points)
    o2 = cpu(g._ggp_schedule[1][0]).astype(np.int64)
    assert np.array_equal(cpu(g._ggp_schedule[1][2]), np.concatenate([P[:n0], P[n0 + o2]]))
    monkeypatch.setattr(R, "SCHEDULE_MIN_REFINED", 1 << 40)
    b = gp.generate(g, cov, xi, precision="fp32_increment")
    assert float((a - b).abs().max()) <= 1e-4


@pytest.mark.parametrize("n", [60_000, 1_100_000])
def test_fp32_increment_list_overflow_falls_back_to_fp64(n, monkeypatch):
    # When more points need the FP64 recomputation than its list holds (a quarter of them),
    # generate recomputes the whole refinement in FP64 and returns those values, rounded.
    # A threshold of 1 flags every point.  1.1M points run with the dataflow schedule (slot
    # space), 60k points in graph order; the raw refine takes the same path.
    p = clustered_points(n, d=3, seed=71)
    g = gp.build_graph(p, n0=200, k=16)
    cov = gp.extras.matern_kernel(p=1, variance=1.0, cutoff=0.05, r_min=1e-6, r_max=2.0, n_bins=2001, jitter=1e-4)
    xi = torch.as_tensor(np.random.default_rng(72).standard_normal(n), dtype=torch.float32, device="cuda")
    ref = gp.generate(g, cov, xi, precision="fp64")
    out, cnt = gp.generate(g, cov, xi, precision="fp32_increment", return_counters=True)
    assert cnt["fp64_list_overflow"] == 0 and cnt["fp64_recomputed"] < (n - 200) // 4
    monkeypatch.setenv("GGP_FIX_TAU", "1")
    out, cnt = gp.generate(g, cov, xi, precision="fp32_increment", return_counters=True)
    assert cnt["fp64_list_overflow"] == 1 and cnt["nan"] == 0
    assert torch.equal(out, ref.float())
    # raw refine (graph order, seeds given): the refined rows are the FP64 refinement's
    n0 = g.offsets[0]
    gi = g.indices.long()
    init = ref[gi[:n0]].float()
    raw = gp.refine(g.points, g.neighbors, g.offsets, cov, init, xi[gi[n0:]], precision="fp32_increment")
    raw64 = gp.refine(g.points, g.neighbors, g.offsets, cov, init.double(), xi[gi[n0:]].double(), precision="fp64")
    assert torch.equal(raw[n0:], raw64[n0:].float())
