"""Reverse mode of generate on the GPU (adjoint kernel + FP64 seed-block adjoint) against
the oracle's hand-written reverse mode, the exact dot-product identity (generate is linear
in xi) and central finite differences of the GPU forward pass."""
import numpy as np
import pytest
import torch

import graphgp_amd as gp
import oracle as O
from helpers import clustered_points

pytestmark = pytest.mark.gpu


def cpu(t):
    return t.detach().cpu().numpy()


# Tables: a piecewise-linear table is not positive definite for every point set (coarse
# bins or tight clusters make local systems indefinite, and the NaN rule then fires), so
# the RBF tables use 1000 bins and the clustered case a Matern-3/2 table with jitter 1e-4.
def _setup(n, d, n0, k, seed, n_bins=1000, clustered=False):
    rng = np.random.default_rng(seed)
    p = clustered_points(n, d=d, seed=seed) if clustered else rng.uniform(0, 1, (n, d)).astype(np.float32)
    g = gp.build_graph(p, n0=n0, k=k)
    og = O.OGraph(cpu(g.points), cpu(g.neighbors), g.offsets, cpu(g.indices))
    if clustered:
        kw = dict(p=1, variance=1.0, cutoff=0.05, r_min=1e-6, r_max=2.0, n_bins=2001, jitter=1e-4)
        cov, ocov = gp.extras.matern_kernel(**kw), O.matern_kernel(**kw)
    else:
        kw = dict(variance=1.0, scale=0.1, r_min=1e-5, r_max=2, n_bins=n_bins, jitter=1e-5)
        cov, ocov = gp.extras.rbf_kernel(**kw), O.rbf_kernel(**kw)
    xi = rng.standard_normal(n)
    w = rng.standard_normal(n)
    return g, og, cov, ocov, xi, w


@pytest.mark.parametrize("n,d,n0,k,seed", [(3000, 3, 100, 8, 1), (5000, 3, 200, 20, 2), (4000, 2, 64, 12, 3),
                                           (2500, 3, 300, 30, 4), (2000, 1, 16, 4, 5)])
def test_vjp_fp64_vs_oracle(n, d, n0, k, seed):
    g, og, cov, ocov, xi, w = _setup(n, d, n0, k, seed)
    v = gp.generate(g, cov, xi, precision="fp64")
    gx, gv = gp.generate_vjp(g, cov, xi, v, w, precision="fp64")
    ogx, ogv = O.generate_vjp(og, ocov, xi, w)
    assert np.allclose(cpu(gx), ogx, rtol=1e-8, atol=1e-9 * np.abs(ogx).max())
    assert np.allclose(cpu(gv), ogv, rtol=1e-7, atol=1e-8 * np.abs(ogv).max())


def test_vjp_dot_product_identity_fp64():
    # <w, J u> = <J^T w, u> with J u = generate(u), exact up to rounding
    g, og, cov, ocov, xi, w = _setup(20000, 3, 500, 20, 7)
    v = gp.generate(g, cov, xi, precision="fp64")
    gx, _ = gp.generate_vjp(g, cov, xi, v, w, precision="fp64")
    rng = np.random.default_rng(8)
    for _ in range(3):
        u = rng.standard_normal(20000)
        lhs = float(np.dot(w, cpu(gp.generate(g, cov, u, precision="fp64"))))
        rhs = float(np.dot(cpu(gx), u))
        assert np.isclose(lhs, rhs, rtol=1e-10), (lhs, rhs)


def test_vjp_finite_differences_table_fp64():
    g, og, cov, ocov, xi, w = _setup(6000, 3, 200, 16, 9)
    bins, vals = cov
    vals = np.asarray(cpu(torch.as_tensor(vals)) if isinstance(vals, torch.Tensor) else vals, dtype=np.float64)
    bins = np.asarray(cpu(torch.as_tensor(bins)) if isinstance(bins, torch.Tensor) else bins, dtype=np.float64)
    v = gp.generate(g, (bins, vals), xi, precision="fp64")
    _, gv = gp.generate_vjp(g, (bins, vals), xi, v, w, precision="fp64")
    gv = cpu(gv)

    def loss(vv):
        return float(np.dot(w, cpu(gp.generate(g, (bins, vv), xi, precision="fp64"))))

    nz = np.nonzero(np.abs(gv) > 1e-2 * np.abs(gv).max())[0]
    for j in nz[:: max(1, len(nz) // 12)]:
        e = np.zeros_like(vals)
        e[j] = 1e-7
        fd = (loss(vals + e) - loss(vals - e)) / 2e-7
        assert np.isclose(gv[j], fd, rtol=1e-4, atol=1e-6 * np.abs(gv).max()), (j, gv[j], fd)


@pytest.mark.parametrize("precision", ["fp32", "fp32_increment"])
def test_vjp_fp32_vs_oracle(precision):
    g, og, cov, ocov, xi, w = _setup(20000, 3, 300, 20, 11, n_bins=1000)
    v = gp.generate(g, cov, xi.astype(np.float32), precision=precision)
    gx, gv = gp.generate_vjp(g, cov, xi.astype(np.float32), v, w.astype(np.float32), precision=precision)
    ogx, ogv = O.generate_vjp(og, ocov, xi.astype(np.float32), w.astype(np.float32))
    ex = np.abs(cpu(gx) - ogx).max() / np.abs(ogx).max()
    ev = np.linalg.norm(cpu(gv) - ogv) / np.linalg.norm(ogv)
    assert ex < 1e-3, ex
    assert ev < 1e-2, ev


def test_autograd_matches_vjp():
    g, og, cov, ocov, xi, w = _setup(4000, 3, 100, 10, 13)
    bins, vals = cov
    x = torch.tensor(xi, device="cuda", requires_grad=True)
    tv = torch.as_tensor(np.asarray(cpu(vals) if isinstance(vals, torch.Tensor) else vals), device="cuda",
                         dtype=torch.float64).requires_grad_(True)
    out = gp.generate(g, (bins, tv), x, precision="fp64")
    loss = (out * torch.tensor(w, device="cuda")).sum()
    loss.backward()
    gx, gv = gp.generate_vjp(g, (bins, tv.detach()), xi, out.detach(), w, precision="fp64")
    # neighbour cotangents are summed with atomics, so two backward passes agree to the
    # rounding of those sums: near-zero components get an absolute tolerance
    assert torch.allclose(x.grad, gx, rtol=1e-10, atol=1e-12 * float(gx.abs().max()))
    assert torch.allclose(tv.grad, gv, rtol=1e-10, atol=1e-12 * float(gv.abs().max()))
    # a second backward pass gives the same gradients (no state kept between calls); the
    # cotangents of shared neighbours are summed with atomics, so only the order of those
    # additions (last bits) may differ
    x2 = torch.tensor(xi, device="cuda", requires_grad=True)
    (gp.generate(g, (bins, tv.detach()), x2, precision="fp64") * torch.tensor(w, device="cuda")).sum().backward()
    assert torch.allclose(x2.grad, x.grad, rtol=1e-10, atol=1e-12 * float(x.grad.abs().max()))


def test_vjp_clustered_k30():
    g, og, cov, ocov, xi, w = _setup(30000, 3, 1000, 30, 17, clustered=True)
    v = gp.generate(g, cov, xi, precision="fp64")
    gx, gv = gp.generate_vjp(g, cov, xi, v, w, precision="fp64")
    ogx, ogv = O.generate_vjp(og, ocov, xi, w)
    assert np.allclose(cpu(gx), ogx, rtol=1e-6, atol=1e-7 * np.abs(ogx).max())
    assert np.linalg.norm(cpu(gv) - ogv) <= 1e-6 * np.linalg.norm(ogv)


@pytest.mark.parametrize("kind", ["rbf", "matern32"])
def test_grad_wrt_kernel_hyperparameters(kind):
    # The table producers run on the device with autograd, so the adjoint kernel's table
    # gradient reaches the hyperparameters.  Values scale as sqrt(variance) exactly (the
    # jitter multiplies the variance too): dL/dvariance = L / (2 variance).  The length
    # scale by central differences of the forward pass.
    rng = np.random.default_rng(21)
    n = 6000
    p = rng.uniform(0, 1, (n, 3)).astype(np.float32)
    g = gp.build_graph(p, n0=200, k=12)
    xi = torch.as_tensor(rng.standard_normal(n), device="cuda")
    w = torch.as_tensor(rng.standard_normal(n), device="cuda")
    var0, len0 = 1.3, 0.15

    def table(v, s):
        if kind == "rbf":
            return gp.extras.rbf_kernel(variance=v, scale=s, r_min=1e-5, r_max=2, n_bins=1000, jitter=1e-5)
        return gp.extras.matern_kernel(p=1, variance=v, cutoff=s, r_min=1e-5, r_max=2, n_bins=1000, jitter=1e-5)

    def loss(v, s):
        return (w * gp.generate(g, table(v, s), xi, precision="fp64")).sum()

    v = torch.tensor(var0, dtype=torch.float64, device="cuda", requires_grad=True)
    s = torch.tensor(len0, dtype=torch.float64, device="cuda", requires_grad=True)
    L = loss(v, s)
    gv, gs = torch.autograd.grad(L, [v, s])
    assert abs(gv.item() - L.item() / (2 * var0)) <= 1e-9 * abs(L.item() / (2 * var0))
    h = 1e-6 * len0
    with torch.no_grad():
        fd = (loss(var0, len0 + h) - loss(var0, len0 - h)).item() / (2 * h)
    assert abs(gs.item() - fd) <= 1e-5 * max(abs(fd), 1.0), (gs.item(), fd)


# ---------------------------------------------------------------- any k and d
@pytest.mark.parametrize("n,d,n0,k,seed", [(3000, 3, 100, 40, 31), (2500, 6, 64, 10, 32), (3000, 4, 100, 20, 33),
                                           (1500, 3, 60, 60, 34)])
def test_vjp_generic_configurations_vs_oracle(n, d, n0, k, seed):
    # k + 1 > 32 or d > 3: the one-lane-per-point adjoint (generate, inverse, log-det)
    g, og, cov, ocov, xi, w = _setup(n, d, n0, k, seed)
    v = gp.generate(g, cov, xi, precision="fp64")
    gx, gv = gp.generate_vjp(g, cov, xi, v, w, precision="fp64")
    ogx, ogv = O.generate_vjp(og, ocov, xi, w)
    assert np.allclose(cpu(gx), ogx, rtol=1e-8, atol=1e-9 * np.abs(ogx).max())
    assert np.allclose(cpu(gv), ogv, rtol=1e-7, atol=1e-8 * np.abs(ogv).max())
    gvv, gvt = gp.generate_inv_vjp(g, cov, cpu(v), w, precision="fp64")
    ogvv, ogvt = O.generate_inv_vjp(og, ocov, cpu(v), w)
    assert np.allclose(cpu(gvv), ogvv, rtol=1e-8, atol=1e-9 * np.abs(ogvv).max())
    assert np.allclose(cpu(gvt), ogvt, rtol=1e-7, atol=1e-8 * np.abs(ogvt).max())
    glt = gp.generate_logdet_vjp(g, cov, 0.7, precision="fp64")
    oglt = O.generate_logdet_vjp(og, ocov, 0.7)
    assert np.allclose(cpu(glt), oglt, rtol=1e-7, atol=1e-8 * np.abs(oglt).max())


# ---------------------------------------------------------------- inverse and log-det
@pytest.mark.parametrize("n,d,n0,k,seed", [(3000, 3, 100, 8, 41), (5000, 3, 200, 20, 42), (4000, 2, 64, 12, 43),
                                           (2500, 3, 300, 30, 44), (2000, 1, 16, 4, 45)])
def test_inv_and_logdet_vjp_fp64_vs_oracle(n, d, n0, k, seed):
    g, og, cov, ocov, xi, w = _setup(n, d, n0, k, seed)
    values = O.generate(og, ocov, xi)
    gvv, gvt = gp.generate_inv_vjp(g, cov, values, w, precision="fp64")
    ogvv, ogvt = O.generate_inv_vjp(og, ocov, values, w)
    assert np.allclose(cpu(gvv), ogvv, rtol=1e-8, atol=1e-9 * np.abs(ogvv).max())
    assert np.allclose(cpu(gvt), ogvt, rtol=1e-7, atol=1e-8 * np.abs(ogvt).max())
    glt = gp.generate_logdet_vjp(g, cov, precision="fp64")
    oglt = O.generate_logdet_vjp(og, ocov)
    assert np.allclose(cpu(glt), oglt, rtol=1e-7, atol=1e-8 * np.abs(oglt).max())


def test_inv_vjp_dot_product_identity_and_fd():
    # generate_inv is linear in the values: <t, J_inv u> = <J_inv^T t, u>; the table
    # cotangents of the inverse and the log-det by central differences of the GPU forward
    g, og, cov, ocov, xi, w = _setup(20000, 3, 500, 20, 47)
    values = cpu(gp.generate(g, cov, xi, precision="fp64"))
    gvv, gvt = gp.generate_inv_vjp(g, cov, values, w, precision="fp64")
    rng = np.random.default_rng(48)
    for _ in range(3):
        u = rng.standard_normal(20000)
        lhs = float(np.dot(w, cpu(gp.generate_inv(g, cov, u, precision="fp64"))))
        assert np.isclose(lhs, float(np.dot(cpu(gvv), u)), rtol=1e-10)
    bins, vals = cov
    glt = cpu(gp.generate_logdet_vjp(g, cov, precision="fp64"))
    gvt = cpu(gvt)
    inv_loss = lambda vv: float(np.dot(w, cpu(gp.generate_inv(g, (bins, vv), values, precision="fp64"))))  # noqa
    ld_loss = lambda vv: float(gp.generate_logdet(g, (bins, vv), precision="fp64"))  # noqa: E731
    for grad, loss in ((gvt, inv_loss), (glt, ld_loss)):
        nz = np.nonzero(np.abs(grad) > 1e-2 * np.abs(grad).max())[0]
        for j in nz[:: max(1, len(nz) // 10)]:
            h = 1e-8 * max(abs(float(vals[int(j)])), 1e-6)
            e = torch.zeros_like(vals)
            e[int(j)] = h
            fd = (loss(vals + e) - loss(vals - e)) / (2 * h)
            assert np.isclose(grad[j], fd, rtol=1e-4, atol=1e-6 * np.abs(grad).max()), (j, grad[j], fd)


@pytest.mark.parametrize("precision", ["fp32", "fp32_increment"])
def test_inv_and_logdet_vjp_fp32_vs_oracle(precision):
    # The inverse's and the log-det's cotangents scale as 1 / sigma and 1 / sigma^2, and
    # sigma^2 = c - b.w loses digits to cancellation in FP32 where the local Schur complement
    # is small, so the FP32 adjoints are compared on a well-conditioned field (3000 points,
    # spacing comparable to the length scale); the FP64 tests above cover the rest.
    g, og, cov, ocov, xi, w = _setup(3000, 3, 100, 12, 49)
    values = O.generate(og, ocov, xi).astype(np.float32)
    gvv, gvt = gp.generate_inv_vjp(g, cov, values, w.astype(np.float32), precision=precision)
    ogvv, ogvt = O.generate_inv_vjp(og, ocov, values, w.astype(np.float32))
    glt = gp.generate_logdet_vjp(g, cov, precision=precision)
    oglt = O.generate_logdet_vjp(og, ocov)
    ev = np.abs(cpu(gvv) - ogvv).max() / np.abs(ogvv).max()
    et = np.linalg.norm(cpu(gvt) - ogvt) / np.linalg.norm(ogvt)
    el = np.linalg.norm(cpu(glt) - oglt) / np.linalg.norm(oglt)
    print(f"FP32 adjoints vs FP64 oracle: values_bar {ev:.2e}, inverse table {et:.2e}, log-det table {el:.2e}",
          flush=True)
    assert ev < 1e-3
    assert et < 1e-2
    assert el < 1e-2


def test_autograd_inverse_and_logdet_likelihood():
    # the fitting caller: a Gaussian negative log-likelihood of observed values,
    # 0.5 |generate_inv(values)|^2 + generate_logdet, differentiated to the kernel
    # hyperparameters through the tables (autograd), against central differences
    rng = np.random.default_rng(51)
    n = 6000
    p = rng.uniform(0, 1, (n, 3)).astype(np.float32)
    g = gp.build_graph(p, n0=200, k=12)
    values = torch.as_tensor(rng.standard_normal(n), device="cuda")

    def nll(v, s, vals_in=values):
        cov = gp.extras.rbf_kernel(variance=v, scale=s, r_min=1e-5, r_max=2, n_bins=1000, jitter=1e-5)
        xi = gp.generate_inv(g, cov, vals_in, precision="fp64")
        return 0.5 * (xi * xi).sum() + gp.generate_logdet(g, cov, precision="fp64")

    var0, len0 = 1.3, 0.15
    v = torch.tensor(var0, dtype=torch.float64, device="cuda", requires_grad=True)
    s = torch.tensor(len0, dtype=torch.float64, device="cuda", requires_grad=True)
    vv = values.clone().requires_grad_(True)
    L = nll(v, s, vv)
    gv, gs, gval = torch.autograd.grad(L, [v, s, vv])
    with torch.no_grad():
        for name, g_, fd in (
            ("variance", gv, (nll(var0 * (1 + 1e-6), len0) - nll(var0 * (1 - 1e-6), len0)).item() / (2e-6 * var0)),
            ("scale", gs, (nll(var0, len0 * (1 + 1e-6)) - nll(var0, len0 * (1 - 1e-6))).item() / (2e-6 * len0)),
        ):
            assert abs(g_.item() - fd) <= 1e-5 * max(abs(fd), 1.0), (name, g_.item(), fd)
        # with respect to the observed values: the gradient of 0.5 |xi|^2 is J_inv^T xi
        xi = gp.generate_inv(g, gp.extras.rbf_kernel(variance=var0, scale=len0, r_min=1e-5, r_max=2, n_bins=1000,
                                                     jitter=1e-5), values, precision="fp64")
        ref, _ = gp.generate_inv_vjp(g, gp.extras.rbf_kernel(variance=var0, scale=len0, r_min=1e-5, r_max=2,
                                                              n_bins=1000, jitter=1e-5), values, xi, precision="fp64")
        assert torch.allclose(gval, ref, rtol=1e-10, atol=1e-12 * float(ref.abs().max()))


def test_jjt_approaches_dense_reference_configuration():
    # tests/test_refine.py:66-80: N = 1000 points (key 137), n0 = k = 200, matern p = 0,
    # graph order without the permutation.  J = d generate / d xi, built here row by row
    # from generate_vjp (row i = J^T e_i); J J^T must match the dense covariance within 0.02.
    # Rows: the first and last point and 46 sampled ones; (J J^T)[S, S] = J[S] J[S]^T, so
    # every entry checked is an entry of the reference's full product.
    import time

    from helpers import ref_points

    p = ref_points((1000, 3))
    g0 = gp.build_graph(p, n0=200, k=200)
    g = gp.Graph(g0.points, g0.neighbors, g0.offsets)
    cov = gp.extras.matern_kernel(p=0, variance=1.0, cutoff=1.0, r_min=1e-4, r_max=10, n_bins=1000, jitter=1e-5)
    n = 1000
    rows = np.unique(np.concatenate([[0, n - 1], np.random.default_rng(60).choice(n, 46, replace=False)]))
    zeros = torch.zeros(n, dtype=torch.float64, device="cuda")
    v0 = gp.generate(g, cov, zeros, precision="fp64")
    J = torch.empty((len(rows), n), dtype=torch.float64, device="cuda")
    t0 = time.time()
    for c, i in enumerate(rows):
        e = torch.zeros(n, dtype=torch.float64, device="cuda")
        e[i] = 1.0
        J[c] = gp.generate_vjp(g, cov, zeros, v0, e, precision="fp64")[0]
        if c % 8 == 0:
            torch.cuda.synchronize()
            print(f"J J^T: {c + 1} of {len(rows)} rows, {(time.time() - t0) / (c + 1):.3f} s per row", flush=True)
    dense = gp.compute_cov_matrix(cov, g.points, g.points)
    rs = torch.from_numpy(rows).cuda()
    err = float((J @ J.T - dense[rs][:, rs]).abs().max())
    print(f"max |J J^T - K| over {len(rows)}^2 entries = {err:.3e}", flush=True)
    assert err <= 0.02
    # the rows are the forward pass's Jacobian: J u = generate(u)
    u = torch.as_tensor(np.random.default_rng(61).standard_normal(n), device="cuda")
    assert torch.allclose(J @ u, gp.generate(g, cov, u, precision="fp64")[rs], rtol=1e-9, atol=1e-10)
    xi = torch.as_tensor(ref_points((1000,)), device="cuda")
    true_values = gp.generate_dense(g.points, cov, xi)
    assert float((true_values - gp.generate(g, cov, xi, precision="fp64")).abs().max()) <= 0.02
