"""The large-dynamic-range precision run at its stated size: N = 16,777,216 points in 3-D from
a mixture of 64 isotropic Gaussians (widths log-uniform in [1e-4, 1e-1]), n0 = 1000, k = 30,
Matern-3/2 (variance 1, scale 0.05) on r = 0 plus 2000 log bins over [1e-6, 2], jitter 1e-5,
xi ~ N(0,1) float32 -- the benchmark's inputs (bench.py, same seeds).

Against the oracle: the whole k-d tree bit for bit; 12,000 sampled preceding-kNN queries on
that tree bit for bit, and 32 against brute force; depths, depth order and offsets
recomputed by the oracle from the GPU's neighbours, bit for bit; FP64 values and the
inverse checked point by point against the FP64 local conditionals on 10,000 points; the
FP32-increment values within 1e-4 of FP64 (sigma = 1); the log-det against the oracle's."""
import numpy as np
import pytest
import torch

import graphgp_amd as gp
import oracle as O

pytestmark = pytest.mark.gpu

N, D, N0, K = 16_777_216, 3, 1000, 30
KW = dict(p=1, variance=1.0, cutoff=0.05, r_min=1e-6, r_max=2.0, n_bins=2001, jitter=1e-5)


def cpu(t):
    return t.detach().cpu().numpy()


@pytest.fixture(scope="module")
def c2():
    from bench import clustered

    print(f"[C2] N={N} d={D} n0={N0} k={K}: inputs", flush=True)
    pts = clustered(N, D, 64, 1e-4, 1e-1, seed=2024)
    xi = np.random.default_rng(7).standard_normal(N).astype(np.float32)
    return pts, xi


@pytest.fixture(scope="module")
def c2_tree(c2):
    pts, _ = c2
    tp, sd, idx = gp.build_tree(torch.from_numpy(pts).cuda())
    nb = gp.query_preceding_neighbors(tp, sd, n0=N0, k=K)
    out = cpu(tp), cpu(sd), cpu(idx), cpu(nb)
    del tp, sd, idx, nb
    torch.cuda.empty_cache()
    return out


@pytest.fixture(scope="module")
def c2_graph(c2):
    pts, _ = c2
    return gp.build_graph(torch.from_numpy(pts).cuda(), n0=N0, k=K)


def test_c2_tree_bit_exact(c2, c2_tree):
    pts, _ = c2
    tp, sd, idx, _ = c2_tree
    op, osd, oidx = O.build_tree(pts)
    assert np.array_equal(idx, oidx)
    assert np.array_equal(sd, osd)
    assert np.array_equal(tp, op)


def test_c2_sampled_knn_bit_exact(c2_tree):
    tp, sd, _, nb = c2_tree
    rng = np.random.default_rng(21)
    q = np.sort(rng.choice(np.arange(N0, N), 12_000, replace=False))
    q = np.concatenate([[N0, N0 + 1, N - 1], q])
    onb = O.query_neighbors(tp, sd, q, q, k=K)
    bad = np.nonzero(~np.all(nb[q - N0] == onb, axis=1))[0]
    assert bad.size == 0, f"{bad.size} of {len(q)} queries differ, first at tree position {q[bad[0]]}"
    # brute force over all preceding points: the same k smallest FP32 squared distances
    qb = np.sort(rng.choice(np.arange(N0, N), 32, replace=False))
    d2, _ = O.brute_preceding_knn(tp, qb, k=K)
    mine = np.sort(O.sqdist_f32(tp, nb[qb - N0], qb[:, None]), axis=1)
    assert np.array_equal(mine, d2)


def test_c2_depths_order_offsets_bit_exact(c2_tree, c2_graph):
    tp, _, idx, nb = c2_tree
    g = c2_graph
    dep = O.compute_depths(nb, n0=N0)
    po, io, no, do = O.order_by_depth(tp, idx, nb, dep)
    assert g.offsets == O.offsets_from_depths(do)
    assert np.array_equal(cpu(g.neighbors), no)
    assert np.array_equal(cpu(g.indices), io)
    assert np.array_equal(cpu(g.points), po)
    gp.check_graph(g)


@pytest.fixture(scope="module")
def c2_values(c2, c2_graph):
    _, xi = c2
    g = c2_graph
    cov = gp.extras.matern_kernel(**KW)
    xi_d = torch.from_numpy(xi).cuda()
    v64 = gp.generate(g, cov, xi_d, precision="fp64")
    v32 = gp.generate(g, cov, xi_d, precision="fp32_increment")
    return cov, v64, v32


def _graph_order(g, x):
    return x[cpu(g.indices)]


def _local(P, NB, r, ocov, vg):
    """The oracle's FP64 conditional mean and std of refined point n0 + r given its
    neighbours' values, on the (k + 1) points involved only."""
    rows = NB[r]
    local = np.concatenate([P[rows], P[N0 + r][None]]).astype(np.float64)
    return O.local_conditional(local, np.arange(K, dtype=np.int32), K, ocov, vg[rows])


def test_c2_fp64_values_match_local_conditionals(c2, c2_graph, c2_values):
    _, xi = c2
    g = c2_graph
    cov, v64, _ = c2_values
    ocov = O.matern_kernel(**KW)
    v = cpu(v64)
    assert np.all(np.isfinite(v))
    vg, xg = _graph_order(g, v), _graph_order(g, xi).astype(np.float64)
    P, NB = cpu(g.points), cpu(g.neighbors)
    # the seed block against the oracle's dense FP64 Cholesky
    seeds = O.generate_dense(P[:N0], ocov, xg[:N0])
    assert np.allclose(vg[:N0], seeds, rtol=1e-9, atol=1e-11)
    rng = np.random.default_rng(22)
    sample = rng.choice(N - N0, 10_000, replace=False)
    worst = 0.0
    for c, r in enumerate(sample):
        mean, std = _local(P, NB, r, ocov, vg)
        worst = max(worst, abs(mean + std * xg[N0 + r] - vg[N0 + r]))
        if c % 2500 == 0:
            print(f"[C2] local conditionals: {c} checked", flush=True)
    print(f"[C2] FP64 vs local conditionals, 10,000 points: max |err| = {worst:.3e}", flush=True)
    assert worst <= 1e-9


def test_c2_fp32_increment_within_bound(c2_values):
    _, v64, v32 = c2_values
    assert not torch.isnan(v32).any()
    err = float((v32.double() - v64).abs().max())
    print(f"[C2] FP32-increment vs FP64: max |err| = {err:.3e} (bound 1e-4 sigma, sigma = 1)", flush=True)
    assert err <= 1e-4


def test_c2_inverse(c2, c2_graph, c2_values):
    _, xi = c2
    g = c2_graph
    cov, v64, _ = c2_values
    ocov = O.matern_kernel(**KW)
    xb = gp.generate_inv(g, cov, v64, precision="fp64")
    rt = float((xb - torch.from_numpy(xi).cuda().double()).abs().max())
    print(f"[C2] inverse round trip (FP64): max |xi - inv(gen(xi))| = {rt:.3e}", flush=True)
    assert rt <= 1e-6
    # point by point: xi_i = (v_i - mean_i) / std_i from the oracle's local conditional
    vg = _graph_order(g, cpu(v64))
    xbg = _graph_order(g, cpu(xb))
    P, NB = cpu(g.points), cpu(g.neighbors)
    rng = np.random.default_rng(23)
    worst = 0.0
    for r in rng.choice(N - N0, 2_000, replace=False):
        mean, std = _local(P, NB, r, ocov, vg)
        worst = max(worst, abs((vg[N0 + r] - mean) / std - xbg[N0 + r]))
    assert worst <= 1e-6, worst


def test_c2_logdet(c2_graph, c2_values):
    g = c2_graph
    cov, _, _ = c2_values
    ocov = O.matern_kernel(**KW)
    og = O.OGraph(cpu(g.points), cpu(g.neighbors), g.offsets, cpu(g.indices))
    ref = O.generate_logdet(og, ocov)
    ld64 = float(gp.generate_logdet(g, cov, precision="fp64"))
    ld32 = float(gp.generate_logdet(g, cov, precision="fp32_increment"))
    print(f"[C2] log-det: oracle {ref:.10e}, FP64 {ld64:.10e}, FP32-increment {ld32:.10e} "
          f"(abs err {abs(ld32 - ref):.3e}, rel {abs(ld32 - ref) / abs(ref):.3e})", flush=True)
    assert np.isclose(ld64, ref, rtol=1e-9)
    assert abs(ld32 - ref) <= 1e-5 * abs(ref)
