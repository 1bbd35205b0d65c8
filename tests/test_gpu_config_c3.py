"""The differentiation run at its stated size: N = 4,194,304 points uniform in [0,1]^3,
n0 = 500, k = 20, RBF (variance 1, scale 0.1) on r = 0 plus 1000 log bins over [1e-5, 2],
jitter 1e-5, loss L = sum_i w_i values_i with w ~ N(0,1) -- the benchmark's inputs
(bench.py --workload differentiation, same seeds).

Gradients of generate (JAX autodiff of graphgp/refine.py:21-159 in the reference) from the
adjoint kernel: the dot-product identity <w, J u> = <J^T w, u> (generate is linear in xi),
central differences on 64 xi coordinates (h = 1e-3) and on 64 table entries, FP32 against
FP64; plus the graph's tree and sampled neighbours against the oracle, the log-det against
the oracle's and the inverse round trip."""
import numpy as np
import pytest
import torch

import graphgp_amd as gp
import oracle as O

pytestmark = pytest.mark.gpu

N, D, N0, K = 4_194_304, 3, 500, 20
KW = dict(variance=1.0, scale=0.1, r_min=1e-5, r_max=2.0, n_bins=1000, jitter=1e-5)


def cpu(t):
    return t.detach().cpu().numpy()


@pytest.fixture(scope="module")
def c3():
    print(f"[C3] N={N} d={D} n0={N0} k={K}", flush=True)
    rng = np.random.default_rng(3)
    pts = rng.uniform(0, 1, (N, D)).astype(np.float32)
    xi = rng.standard_normal(N)
    w = rng.standard_normal(N)
    g = gp.build_graph(torch.from_numpy(pts).cuda(), n0=N0, k=K)
    cov = gp.extras.rbf_kernel(**KW)
    xi_d, w_d = torch.from_numpy(xi).cuda(), torch.from_numpy(w).cuda()
    v = gp.generate(g, cov, xi_d, precision="fp64")
    gx, gv = gp.generate_vjp(g, cov, xi_d, v, w_d, precision="fp64")
    return pts, g, cov, xi_d, w_d, v, gx, gv


def test_c3_graph_vs_oracle(c3):
    pts, g, *_ = c3
    tp, sd, idx = gp.build_tree(torch.from_numpy(pts).cuda())
    op, osd, oidx = O.build_tree(pts)
    assert np.array_equal(cpu(idx), oidx) and np.array_equal(cpu(sd), osd) and np.array_equal(cpu(tp), op)
    nb = gp.query_preceding_neighbors(tp, sd, n0=N0, k=K)
    q = np.sort(np.random.default_rng(31).choice(np.arange(N0, N), 8_000, replace=False))
    assert np.array_equal(cpu(nb)[q - N0], O.query_neighbors(op, osd, q, q, k=K))
    dep = O.compute_depths(cpu(nb), n0=N0)
    po, io, no, do = O.order_by_depth(op, oidx, cpu(nb), dep)
    assert g.offsets == O.offsets_from_depths(do)
    assert np.array_equal(cpu(g.neighbors), no) and np.array_equal(cpu(g.indices), io)


def test_c3_dot_product_identity(c3):
    _, g, cov, xi, w, v, gx, _ = c3
    u = torch.from_numpy(np.random.default_rng(4).standard_normal(N)).cuda()
    lhs = float((w * gp.generate(g, cov, u, precision="fp64")).sum())
    rhs = float((gx * u).sum())
    print(f"[C3] <w, J u> = {lhs:.12e}, <J^T w, u> = {rhs:.12e}", flush=True)
    assert abs(lhs - rhs) <= 1e-10 * max(abs(lhs), 1.0)
    # J xi is the forward pass itself
    assert abs(float((w * v).sum()) - float((gx * xi).sum())) <= 1e-10 * float((w * v).abs().sum())


def test_c3_finite_differences_xi(c3):
    _, g, cov, xi, w, _, gx, _ = c3
    worst = 0.0
    gmax = float(gx.abs().max())
    for j in np.random.default_rng(5).choice(N, 64, replace=False):
        e = torch.zeros_like(xi)
        e[int(j)] = 1e-3
        vp = gp.generate(g, cov, xi + e, precision="fp64")
        vm = gp.generate(g, cov, xi - e, precision="fp64")
        fd = float((w * (vp - vm)).sum()) / 2e-3
        worst = max(worst, abs(fd - float(gx[int(j)])) / max(abs(float(gx[int(j)])), 1e-3 * gmax))
    print(f"[C3] FD on 64 xi coordinates (h = 1e-3): worst relative deviation {worst:.3e}", flush=True)
    assert worst <= 1e-6


def test_c3_finite_differences_table(c3):
    # central differences in the table values at h = 1e-8 |value|: at this length scale the
    # local Schur complements are tiny, so the loss is strongly curved in the table and
    # larger steps carry percent-level truncation error (h = 1e-4 |value| makes some local
    # systems indefinite)
    _, g, cov, xi, w, _, _, gv = c3
    bins, vals = cov
    nz = torch.nonzero(gv.abs() > 1e-3 * gv.abs().max()).flatten().cpu().numpy()
    devs = []
    for j in np.random.default_rng(6).choice(nz, min(64, len(nz)), replace=False):
        h = 1e-8 * max(abs(float(vals[int(j)])), 1e-6)
        e = torch.zeros_like(vals)
        e[int(j)] = h
        lp = float((w * gp.generate(g, (bins, vals + e), xi, precision="fp64")).sum())
        lm = float((w * gp.generate(g, (bins, vals - e), xi, precision="fp64")).sum())
        fd = (lp - lm) / (2 * h)
        devs.append(abs(fd - float(gv[int(j)])) / abs(float(gv[int(j)])))
    devs = np.array(devs)
    print(f"[C3] FD on {len(devs)} table entries: median {np.median(devs):.3e}, max {devs.max():.3e}", flush=True)
    assert np.median(devs) <= 1e-4
    assert devs.max() <= 5e-2


def test_c3_fp32_gradients_against_fp64(c3):
    _, g, cov, xi, w, _, gx, gv = c3
    v32 = gp.generate(g, cov, xi.float(), precision="fp32")
    gx32, gv32 = gp.generate_vjp(g, cov, xi.float(), v32, w.float(), precision="fp32")
    ex = float((gx32.double() - gx).abs().max()) / float(gx.abs().max())
    ev = float((gv32 - gv).abs().max()) / float(gv.abs().max())
    print(f"[C3] FP32 vs FP64: xi_bar rel {ex:.3e}, table_bar rel {ev:.3e}", flush=True)
    assert ex <= 1e-3
    assert ev <= 5e-2


def test_c3_logdet_and_inverse(c3):
    _, g, cov, xi, _, v, _, _ = c3
    og = O.OGraph(cpu(g.points), cpu(g.neighbors), g.offsets, cpu(g.indices))
    ref = O.generate_logdet(og, O.rbf_kernel(**KW))
    ld = float(gp.generate_logdet(g, cov, precision="fp64"))
    assert np.isclose(ld, ref, rtol=1e-9), (ld, ref)
    xb = gp.generate_inv(g, cov, v, precision="fp64")
    rt = float((xb - xi).abs().max())
    print(f"[C3] log-det {ld:.10e} (oracle {ref:.10e}); inverse round trip max |err| {rt:.3e}", flush=True)
    assert rt <= 1e-6
