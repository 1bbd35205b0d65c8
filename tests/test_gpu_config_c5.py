"""The billion-point field at its stated size: N = 1,000,000,000 points uniform in [0,1]^3,
n0 = 1000, k = 16, RBF (variance 1, scale 0.01) on r = 0 plus 4000 log bins over [1e-6, 2],
jitter 1e-5, xi ~ N(0,1) float32, int32 neighbour indices -- the benchmark's inputs
(bench.py --workload billion, same seeds), built stage by stage as the benchmark does.

Checks: 48 sampled preceding-kNN queries (plus the first and last) against brute force over
all preceding points (the same k smallest FP32 squared distances; tests/test_tree.py:44-59's
property); the graph passes check_graph's rules; the FP32 generate has no NaN and its values
match the FP64 local conditionals of the oracle, given the GPU's neighbour values, on 2,000
sampled points; the FP64 log-det is finite, agrees with the FP32-increment one, and on the
graph of the first 2^20 refined points equals the oracle's."""
import os

import numpy as np
import pytest
import torch

import graphgp_amd as gp
import oracle as O

pytestmark = pytest.mark.gpu

N, D, N0, K = 1_000_000_000, 3, 1000, 16
KW = dict(variance=1.0, scale=0.01, r_min=1e-6, r_max=2.0, n_bins=4001, jitter=1e-5)


def cpu(t):
    return t.detach().cpu().numpy()


def _need_memory():
    free, _ = torch.cuda.mem_get_info()
    return free < 200 * 2**30


def test_c5_billion_point_field():
    torch.cuda.empty_cache()
    if _need_memory():
        pytest.skip("the billion-point field needs about 200 GiB of free device memory")
    from graphgp_amd.graph import _offsets

    print(f"[C5] N={N} d={D} n0={N0} k={K}: tree", flush=True)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(2025)
    pts = torch.rand((N, D), generator=gen, device="cuda", dtype=torch.float32)
    tp, sd, idx = gp.build_tree(pts)
    del pts
    torch.cuda.empty_cache()
    print("[C5] kNN", flush=True)
    nb = gp.query_preceding_neighbors(tp, sd, n0=N0, k=K)
    del sd
    torch.cuda.empty_cache()
    # brute force over every preceding point for sampled queries
    rng = np.random.default_rng(51)
    q = np.sort(np.concatenate([[N0, N - 1], rng.choice(np.arange(N0 + 1, N - 1), 48, replace=False)]))
    nbq = cpu(nb[torch.from_numpy(q - N0).cuda()])
    tp_h = cpu(tp)
    print("[C5] brute-force kNN check", flush=True)
    d2, _ = O.brute_preceding_knn(tp_h, q, k=K)
    assert np.all(nbq < q[:, None]) and np.all(nbq >= 0)
    assert all(len(set(row)) == K for row in nbq)
    mine = np.sort(O.sqdist_f32(tp_h, nbq, q[:, None]), axis=1)
    assert np.array_equal(mine, d2)
    del tp_h
    print("[C5] depths, order, offsets", flush=True)
    dep = gp.compute_depths(nb, n0=N0)
    po, io, no, do = gp.order_by_depth(tp, idx, nb, dep)
    del tp, idx, nb, dep
    torch.cuda.empty_cache()
    offsets = _offsets(do, int(do[-1].item()))
    del do
    g = gp.Graph(po, no, offsets, io)
    assert offsets[0] == N0 and offsets[-1] == N
    gp.check_graph(g)
    print(f"[C5] {len(offsets) - 1} batches; generate", flush=True)
    cov = gp.extras.rbf_kernel(**KW)
    xi = torch.randn(N, generator=gen, device="cuda", dtype=torch.float32)
    v = gp.generate(g, cov, xi, precision="fp32")
    assert not torch.isnan(v).any()
    # sampled local conditionals (FP64 oracle) given the GPU's neighbour values
    ocov = O.rbf_kernel(**KW)
    sample = np.sort(rng.choice(N - N0, 2000, replace=False))
    rows = cpu(no[torch.from_numpy(sample).cuda()])
    ids = np.unique(np.concatenate([rows.ravel(), N0 + sample]))
    ids_d = torch.from_numpy(ids).cuda()
    P = {int(i): p for i, p in zip(ids, cpu(po[ids_d]).astype(np.float64))}
    perm = cpu(io[ids_d])
    V = dict(zip(ids.tolist(), cpu(v[torch.from_numpy(perm).cuda()]).astype(np.float64)))
    X = dict(zip(ids.tolist(), cpu(xi[torch.from_numpy(perm).cuda()]).astype(np.float64)))
    errs = []
    for r, row in zip(sample, rows):
        i = N0 + int(r)
        local = np.array([P[int(j)] for j in row] + [P[i]])
        mean, std = O.local_conditional(local, np.arange(K, dtype=np.int32), K, ocov,
                                        np.array([V[int(j)] for j in row]))
        errs.append(abs(mean + std * X[i] - V[i]))
    errs = np.array(errs)
    print(f"[C5] FP32 values vs FP64 local conditionals: max {errs.max():.3e}, median {np.median(errs):.3e}",
          flush=True)
    assert np.median(errs) <= 1e-4
    assert errs.max() <= 2e-2
    del v
    torch.cuda.empty_cache()
    print("[C5] log-det", flush=True)
    ld64 = float(gp.generate_logdet(g, cov, precision="fp64"))
    ld32i = float(gp.generate_logdet(g, cov, precision="fp32_increment"))
    ld32 = float(gp.generate_logdet(g, cov, precision="fp32"))
    print(f"[C5] log-det FP64 {ld64:.10e}; FP32-increment {ld32i:.10e} (rel {abs(ld32i - ld64) / abs(ld64):.2e}); "
          f"plain FP32 {ld32:.10e} (rel {abs(ld32 - ld64) / abs(ld64):.2e})", flush=True)
    assert np.isfinite(ld64)
    assert abs(ld32i - ld64) <= 1e-6 * abs(ld64)
    # the FP64 log-det kernel on this field's own neighbourhoods against the oracle: the first
    # 2^20 refined points in graph order with their seeds form a valid graph of their own
    # (every neighbour precedes its point), whose batch ends are the offsets below n0 + M
    M = 1 << 20
    sub_off = tuple(o for o in offsets if o < N0 + M) + (N0 + M,)
    sub = gp.Graph(po[: N0 + M], no[:M], sub_off, None)
    gp.check_graph(sub)
    og = O.OGraph(cpu(sub.points), cpu(sub.neighbors), sub_off, None)
    ref = O.generate_logdet(og, ocov)
    got = float(gp.generate_logdet(sub, cov, precision="fp64"))
    print(f"[C5] log-det of the first {M} refined points: FP64 {got:.10e}, oracle {ref:.10e}", flush=True)
    assert np.isclose(got, ref, rtol=1e-9)
