"""The oracle's refinement against the reference's known answers (tests/test_refine.py)
and against what the mathematics fixes (dense limit, Jacobian log-det, inverse, tables,
finite-difference gradients)."""
import math

import numpy as np
import pytest

import oracle as O
from helpers import ref_points


@pytest.fixture(scope="module")
def setup_graph():
    # tests/test_refine.py:15-26
    points = ref_points((1000, 3))
    graph = O.build_graph(points, n0=100, k=10)
    cov = O.matern_kernel(p=0, variance=1.0, cutoff=1.0, r_min=1e-4, r_max=10, n_bins=1000, jitter=1e-5)
    return graph, cov, points


def test_logdet_random(setup_graph):
    # tests/test_refine.py:29-43
    g, cov, _ = setup_graph
    assert np.isclose(g.points[0, 0], -1.95624711, rtol=1e-8)
    assert np.isclose(O.generate_logdet(g, cov), -600.36165088, rtol=1e-8)
    assert np.isclose(O.generate_dense_logdet(g.points, cov), -610.90538067, rtol=1e-8)


def test_inverse(setup_graph):
    # tests/test_refine.py:46-54
    g, cov, _ = setup_graph
    xi = ref_points((1000,))
    v = O.generate(g, cov, xi)
    xb = O.generate_inv(g, cov, v)
    assert np.allclose(O.generate(g, cov, xb), v, rtol=1e-12)
    assert np.allclose(xb, xi, rtol=1e-9, atol=1e-9)


def test_fast_jit_equals_loop(setup_graph):
    # tests/test_refine.py:57-63
    g, cov, _ = setup_graph
    xi = ref_points((1000,))[g.indices]
    n0 = g.offsets[0]
    init = O.generate_dense(g.points[:n0], cov, xi[:n0])
    a = O.refine(g.points, g.neighbors, g.offsets, cov, init, xi[n0:])
    b = O.refine(g.points, g.neighbors, g.offsets, cov, init, xi[n0:], loop=True)
    assert np.allclose(a, b, rtol=1e-12)


def test_full_conditioning_is_dense():
    # k = n0 = N - 1: the single refined point conditions on every other point, so
    # Vecchia equals the dense Cholesky exactly.
    rng = np.random.default_rng(1)
    p = rng.standard_normal((40, 2))
    g = O.build_graph(p, n0=39, k=39)
    cov = O.rbf_kernel(variance=1.0, scale=0.7, r_min=1e-4, r_max=10, n_bins=500, jitter=1e-4)
    xi = rng.standard_normal(40)
    dense = O.generate_dense(g.points, cov, xi[g.indices])
    v = O.generate(g, cov, xi)[g.indices]
    assert np.allclose(v, dense, rtol=1e-9, atol=1e-10)


def test_approaches_dense():
    # tests/test_refine.py:66-80 (k = n0 = 200, N = 1000)
    points = ref_points((1000, 3))
    g = O.build_graph(points, n0=200, k=200)
    cov = O.matern_kernel(p=0, variance=1.0, cutoff=1.0, r_min=1e-4, r_max=10, n_bins=1000, jitter=1e-5)
    xi = ref_points((1000,))
    gg = O.OGraph(g.points, g.neighbors, g.offsets, None)
    true_v = O.generate_dense(g.points, cov, xi)
    v = O.generate(gg, cov, xi)
    assert np.allclose(true_v, v, atol=0.02)


def _jacobian(g, cov, n):
    J = np.empty((n, n))
    base = O.generate(g, cov, np.zeros(n))
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1.0
        J[:, j] = O.generate(g, cov, e) - base
    return J


def test_logdet_is_log_jacobian():
    # generate is linear in xi, so its log-det is log|det J| of the dense Jacobian
    rng = np.random.default_rng(2)
    p = rng.standard_normal((60, 2))
    g = O.build_graph(p, n0=8, k=6)
    cov = O.rbf_kernel(variance=2.0, scale=0.5, r_min=1e-4, r_max=10, n_bins=300, jitter=1e-4)
    J = _jacobian(g, cov, 60)
    assert np.isclose(np.linalg.slogdet(J)[1], O.generate_logdet(g, cov), rtol=1e-10)


def test_jjt_approximates_k():
    # tests/test_refine.py:77-80 (J J^T ~ K), on a smaller case
    points = ref_points((300, 3))
    g = O.build_graph(points, n0=60, k=60)
    gg = O.OGraph(g.points, g.neighbors, g.offsets, None)
    cov = O.matern_kernel(p=0, variance=1.0, cutoff=1.0, r_min=1e-4, r_max=10, n_bins=1000, jitter=1e-5)
    J = _jacobian(gg, cov, 300)
    K = O.compute_cov_matrix(cov, g.points, g.points)
    assert np.allclose(J @ J.T, K, atol=0.02)


def test_dense_against_lapack():
    import scipy.linalg as sla

    rng = np.random.default_rng(4)
    p = rng.standard_normal((120, 3))
    cov = O.rbf_kernel(variance=1.0, scale=0.6, r_min=1e-4, r_max=10, n_bins=800, jitter=1e-3)
    K = O.compute_cov_matrix(cov, p, p)
    L = sla.cholesky(K, lower=True)
    xi = rng.standard_normal(120)
    assert np.allclose(O.generate_dense(p, cov, xi), L @ xi, rtol=1e-10, atol=1e-12)
    assert np.allclose(O.generate_dense_inv(p, cov, L @ xi), xi, rtol=1e-8, atol=1e-9)
    assert np.isclose(O.generate_dense_logdet(p, cov), np.linalg.slogdet(K)[1] / 2, rtol=1e-12)


def test_interp_matches_numpy():
    bins, vals = O.rbf_kernel(variance=1.0, scale=0.3, r_min=1e-4, r_max=10, n_bins=200)
    x = np.concatenate([[0.0, 1e-5, 10.0, 11.0], np.random.default_rng(0).uniform(0, 12, 500), bins[5:9]])
    assert np.allclose(O.interp(x, bins, vals), np.interp(x, bins, vals), rtol=1e-14, atol=1e-15)


def test_tables_closed_form():
    # Rasmussen & Williams eq. 4.16-4.17: Matern-1/2 = e^-x, Matern-3/2 = (1 + x) e^-x
    r, c = O.matern_kernel(p=1, variance=2.0, cutoff=0.5, r_min=1e-3, r_max=5, n_bins=100, jitter=0.1)
    x = math.sqrt(3) * r / 0.5
    ref = 2.0 * (1 + x) * np.exp(-x)
    ref[0] *= 1.1
    assert r[0] == 0 and np.isclose(r[1], 1e-3) and np.isclose(r[-1], 5.0) and len(r) == 100
    assert np.allclose(c, ref, rtol=1e-13)
    r, c = O.matern_kernel(p=0, variance=1.0, cutoff=1.0, r_min=1e-3, r_max=5, n_bins=50)
    assert np.allclose(c, np.exp(-r), rtol=1e-13)
    r, c = O.rbf_kernel(variance=3.0, scale=0.2, r_min=1e-3, r_max=5, n_bins=50, jitter=1e-2)
    ref = 3.0 * np.exp(-0.5 * (r / 0.2) ** 2)
    ref[0] *= 1.01
    assert np.allclose(c, ref, rtol=1e-13)


def test_refine_vjp_finite_differences():
    rng = np.random.default_rng(7)
    p = rng.uniform(0, 1, (300, 3))
    g = O.build_graph(p, n0=20, k=8)
    cov = O.rbf_kernel(variance=1.0, scale=0.2, r_min=1e-4, r_max=2, n_bins=120, jitter=1e-3)
    n0 = g.offsets[0]
    xi = rng.standard_normal(300 - n0)
    init = rng.standard_normal(n0)
    w = rng.standard_normal(300)

    def loss(xi_, init_, vals_):
        v = O.refine(g.points, g.neighbors, g.offsets, (cov[0], vals_), init_, xi_)
        return float(w @ v)

    v = O.refine(g.points, g.neighbors, g.offsets, cov, init, xi)
    gxi, ginit, gvals = O.refine_vjp(g.points, g.neighbors, cov, v, xi, w)
    h = 1e-6
    for j in rng.choice(len(xi), 10, replace=False):
        e = np.zeros_like(xi)
        e[j] = h
        fd = (loss(xi + e, init, cov[1]) - loss(xi - e, init, cov[1])) / (2 * h)
        assert np.isclose(gxi[j], fd, rtol=1e-6, atol=1e-8)
    for j in rng.choice(n0, 5, replace=False):
        e = np.zeros_like(init)
        e[j] = h
        fd = (loss(xi, init + e, cov[1]) - loss(xi, init - e, cov[1])) / (2 * h)
        assert np.isclose(ginit[j], fd, rtol=1e-6, atol=1e-8)
    nz = np.nonzero(np.abs(gvals) > 1e-3)[0]
    for j in rng.choice(nz, min(10, len(nz)), replace=False):
        e = np.zeros_like(cov[1])
        e[j] = 1e-7
        fd = (loss(xi, init, cov[1] + e) - loss(xi, init, cov[1] - e)) / 2e-7
        assert np.isclose(gvals[j], fd, rtol=1e-4, atol=1e-6)


def test_dense_vjp_finite_differences():
    rng = np.random.default_rng(8)
    p = rng.uniform(0, 1, (25, 2))
    cov = O.rbf_kernel(variance=1.0, scale=0.4, r_min=1e-4, r_max=2, n_bins=60, jitter=1e-2)
    xi = rng.standard_normal(25)
    w = rng.standard_normal(25)
    gxi, gvals = O.dense_vjp(p, cov, xi, w)
    loss = lambda x, vals: float(w @ O.generate_dense(p, (cov[0], vals), x))  # noqa: E731
    for j in range(0, 25, 5):
        e = np.zeros(25)
        e[j] = 1e-6
        assert np.isclose(gxi[j], (loss(xi + e, cov[1]) - loss(xi - e, cov[1])) / 2e-6, rtol=1e-6)
    nz = np.nonzero(np.abs(gvals) > 1e-3)[0]
    for j in nz[:: max(1, len(nz) // 8)]:
        e = np.zeros_like(cov[1])
        e[j] = 1e-7
        fd = (loss(xi, cov[1] + e) - loss(xi, cov[1] - e)) / 2e-7
        assert np.isclose(gvals[j], fd, rtol=1e-4, atol=1e-6)


def test_generate_vjp_composition_pinned():
    # generate is linear in xi, so <w, generate(u)> = <xi_bar(w), u> exactly (dot-product
    # test); the table cotangent is pinned by central differences of the whole generate.
    rng = np.random.default_rng(9)
    p = rng.uniform(0, 1, (400, 3))
    g = O.build_graph(p, n0=30, k=8)
    cov = O.rbf_kernel(variance=1.0, scale=0.2, r_min=1e-4, r_max=2, n_bins=150, jitter=1e-3)
    xi = rng.standard_normal(400)
    w = rng.standard_normal(400)
    gx, gv = O.generate_vjp(g, cov, xi, w)
    for _ in range(3):
        u = rng.standard_normal(400)
        lhs = float(w @ O.generate(g, cov, u))
        assert np.isclose(lhs, float(gx @ u), rtol=1e-10)
    loss = lambda vals: float(w @ O.generate(g, (cov[0], vals), xi))  # noqa: E731
    nz = np.nonzero(np.abs(gv) > 1e-3)[0]
    for j in nz[:: max(1, len(nz) // 8)]:
        e = np.zeros_like(cov[1])
        e[j] = 1e-7
        fd = (loss(cov[1] + e) - loss(cov[1] - e)) / 2e-7
        assert np.isclose(gv[j], fd, rtol=1e-4, atol=1e-6)


def test_indefinite_seed_fixture_is_indefinite():
    # pins the fixture of the GPU test of the NaN rule: the interpolated covariance of these
    # seeds has a negative eigenvalue, and the oracle's dense block turns some rows NaN
    import os

    p = np.load(os.path.join(os.path.dirname(__file__), "golden", "indefinite_seeds_4m.npy"))
    bins, vals = O.matern_kernel(p=1, variance=1.0, cutoff=0.05, r_min=1e-6, r_max=2.0, n_bins=2001, jitter=1e-5)
    P = p.astype(np.float64)
    K = np.interp(np.linalg.norm(P[:, None] - P[None], axis=-1), bins, vals)
    assert np.linalg.eigvalsh(K)[0] < 0
    assert np.isnan(O.generate_dense(p, (bins, vals), np.ones(len(p)))).any()


def test_seed_block_failure_rules_on_indefinite_block():
    # the 1000 seeds of a 4,194,304-point clustered graph (tests/golden/make_indefinite_seeds.py):
    # with jitter 1e-5 the interpolated covariance is indefinite.  jnp.linalg.cholesky then
    # returns an all-NaN factor (every value of generate_dense / generate_dense_inv is NaN),
    # while slogdet(K)[1] / 2 stays finite: numpy's LAPACK slogdet pins the oracle's LU.
    import os

    p = np.load(os.path.join(os.path.dirname(__file__), "golden", "indefinite_seeds_4m.npy"))
    cov = O.matern_kernel(p=1, variance=1.0, cutoff=0.05, r_min=1e-6, r_max=2.0, n_bins=2001, jitter=1e-5)
    K = O.compute_cov_matrix(cov, p, p)
    sign, logabs = np.linalg.slogdet(K)
    assert sign == -1.0
    assert np.isclose(O.generate_dense_logdet(p, cov), logabs / 2, rtol=1e-12)
    assert np.all(np.isnan(O.generate_dense(p, cov, np.ones(len(p)))))
    assert np.all(np.isnan(O.generate_dense_inv(p, cov, np.ones(len(p)))))


@pytest.mark.parametrize("n", [1, 2, 17, 300])
def test_dense_logdet_is_slogdet(n):
    # positive-definite blocks: the LU restatement agrees with LAPACK's slogdet and with the
    # Cholesky's log-diagonal; a singular block (duplicate points, no jitter) gives -inf
    rng = np.random.default_rng(n)
    p = rng.uniform(0, 1, (n, 3))
    cov = O.rbf_kernel(variance=1.0, scale=0.2, r_min=1e-5, r_max=2.0, n_bins=1000, jitter=1e-4)
    K = O.compute_cov_matrix(cov, p, p)
    assert np.isclose(O.generate_dense_logdet(p, cov), np.linalg.slogdet(K)[1] / 2, rtol=1e-11, atol=1e-13)
    assert np.isclose(O.generate_dense_logdet(p, cov), np.sum(np.log(np.diag(np.linalg.cholesky(K)))),
                      rtol=1e-11, atol=1e-13)
    if n >= 2:
        q = np.concatenate([p, p[:1]])
        cov0 = O.rbf_kernel(variance=1.0, scale=0.2, r_min=1e-5, r_max=2.0, n_bins=1000, jitter=0.0)
        assert O.generate_dense_logdet(q, cov0) == -np.inf


def _vjp_setup(seed=19, n=400, n0=30, k=8):
    rng = np.random.default_rng(seed)
    p = rng.uniform(0, 1, (n, 3))
    g = O.build_graph(p, n0=n0, k=k)
    cov = O.rbf_kernel(variance=1.0, scale=0.2, r_min=1e-4, r_max=2, n_bins=150, jitter=1e-3)
    return rng, g, cov


def test_generate_inv_vjp_pinned():
    # generate_inv is linear in the values: <t, inv(u)> = <values_bar(t), u> exactly; the
    # table cotangent is pinned by central differences of the whole inverse
    rng, g, cov = _vjp_setup()
    n = len(g.points)
    values = O.generate(g, cov, rng.standard_normal(n))
    t = rng.standard_normal(n)
    gv, gtab = O.generate_inv_vjp(g, cov, values, t)
    for _ in range(3):
        u = rng.standard_normal(n)
        assert np.isclose(float(t @ O.generate_inv(g, cov, u)), float(gv @ u), rtol=1e-10)
    loss = lambda vals: float(t @ O.generate_inv(g, (cov[0], vals), values))  # noqa: E731
    nz = np.nonzero(np.abs(gtab) > 1e-3 * np.abs(gtab).max())[0]
    assert len(nz) > 8
    for j in nz[:: max(1, len(nz) // 10)]:
        h = 1e-7 * max(abs(cov[1][j]), 1e-3)
        e = np.zeros_like(cov[1])
        e[j] = h
        fd = (loss(cov[1] + e) - loss(cov[1] - e)) / (2 * h)
        assert np.isclose(gtab[j], fd, rtol=1e-4, atol=1e-6 * np.abs(gtab).max())


def test_generate_logdet_vjp_pinned():
    # central differences of the whole log-det (seed block's slogdet plus the refined sum)
    _, g, cov = _vjp_setup(seed=23)
    gtab = O.generate_logdet_vjp(g, cov, 1.7)
    loss = lambda vals: 1.7 * O.generate_logdet(g, (cov[0], vals))  # noqa: E731
    nz = np.nonzero(np.abs(gtab) > 1e-3 * np.abs(gtab).max())[0]
    assert len(nz) > 8
    for j in nz[:: max(1, len(nz) // 10)]:
        h = 1e-7 * max(abs(cov[1][j]), 1e-3)
        e = np.zeros_like(cov[1])
        e[j] = h
        fd = (loss(cov[1] + e) - loss(cov[1] - e)) / (2 * h)
        assert np.isclose(gtab[j], fd, rtol=1e-4, atol=1e-6 * np.abs(gtab).max())
    # the seed block alone: d/dK (log|det K| / 2) = K^{-1} / 2 (slogdet's derivative)
    n0 = g.offsets[0]
    gd = O.dense_logdet_vjp(g.points[:n0], cov)
    dl = lambda vals: O.generate_dense_logdet(g.points[:n0], (cov[0], vals))  # noqa: E731
    j = int(np.argmax(np.abs(gd)))
    h = 1e-7 * abs(cov[1][j])
    e = np.zeros_like(cov[1])
    e[j] = h
    assert np.isclose(gd[j], (dl(cov[1] + e) - dl(cov[1] - e)) / (2 * h), rtol=1e-5)
