"""Serving layer: admission and residency logic on the CPU; on the GPU, a request's output
inside a concurrent mix equals its output when served alone, bit for bit."""
import numpy as np
import pytest

from graphgp_amd.serving import Admission, GraphCache, Request, graph_bytes, percentile, synthetic_mix


def test_cache_lru_and_pinning():
    c = GraphCache()
    c.put("a", "A", 100)
    c.put("b", "B", 200)
    c.put("c", "C", 300)
    assert c.bytes == 600 and len(c) == 3
    assert c.get("a") == "A"  # a becomes most recent
    c.pin("b")
    freed = c.evict_until(250)  # b is pinned: c (LRU after b) and then a go
    assert "b" in c and freed >= 250
    assert c.evictions >= 1
    c.unpin("b")
    assert c.evictable_bytes() == c.bytes


def test_admission_budget_and_eviction():
    c = GraphCache()
    adm = Admission(1000, c)
    c.put("g1", object(), 600)
    assert adm.try_admit(300)  # fits
    assert adm.used() == 900
    c.pin("g1")
    assert not adm.try_admit(200)  # would exceed; g1 pinned, cannot evict
    c.unpin("g1")
    assert adm.try_admit(200)  # evicts g1
    assert "g1" not in c and adm.used() == 500
    adm.release(300)
    adm.release(200)
    assert adm.used() == 0
    with pytest.raises(ValueError):
        adm.try_admit(5000)


def test_synthetic_mix_shape():
    specs = synthetic_mix(300, rate_per_s=100.0, seed=3)
    arr = [s["arrival"] for s in specs]
    assert arr == sorted(arr)
    keys = {s["graph"]["key"] for s in specs}
    assert 0.15 < len(keys) / len(specs) < 0.5  # about 30% new graphs
    for s in specs:
        g = s["graph"]
        assert 1e3 <= g["n"] <= 1e7 and g["k"] in (8, 16, 32) and g["d"] in (2, 3)
        assert g["n0"] == max(g["k"], min(1000, g["n"] // 100))
        assert 0.01 <= s["cov"]["scale"] <= 0.3


def test_percentile():
    assert percentile([3, 1, 2, 4], 50) == 2
    assert percentile(list(range(1, 101)), 99) == 99
    assert graph_bytes(100, 3, 4, 10) == 100 * 12 + 90 * 16 + 400


@pytest.mark.gpu
def test_result_alone_equals_result_in_mix():
    import torch

    import graphgp_amd as gp
    from graphgp_amd.serving import Server, materialize

    specs = synthetic_mix(24, rate_per_s=1e4, seed=11, n_min=2e3, n_max=2e5)
    reqs = materialize(specs)
    srv = Server(max_inflight=12)
    stats = srv.run(reqs, realtime=False)
    srv.close()
    assert stats["requests"] == len(reqs)
    for r in reqs:
        g = gp.build_graph(r.points, n0=r.n0, k=r.k, schedule=False)  # as the server builds
        fn = gp.generate if r.op == "generate" else gp.generate_inv
        alone = fn(g, r.covariance, r.data, precision=r.precision)
        assert alone.dtype == r.result.dtype == torch.float32
        assert torch.equal(alone.view(torch.int32), r.result.view(torch.int32)), (r.rid, r.op, r.n)
        assert not torch.isnan(r.result).any(), (r.rid, r.op, r.n)


@pytest.mark.gpu
def test_yielding_launches_give_the_same_results():
    # with the yield word set, workgroups beyond the kept CU share stop taking work at once:
    # generate (three precisions), the inverse and the kNN must not change
    import torch

    import graphgp_amd as gp
    from graphgp_amd.serving import make_points, make_table

    lib = gp._lib.lib()
    flag = torch.ones(1, dtype=torch.int32, pin_memory=True)
    pts = make_points(150_000, 3, True, 3)
    tab = make_table("matern32", 0.05, 1000)
    xi = torch.randn(150_000, device="cuda", generator=torch.Generator(device="cuda").manual_seed(4))
    ref = {}
    g = gp.build_graph(pts, n0=1000, k=16)
    for prec in ["fp64", "fp32", "fp32_increment"]:
        ref[prec] = (gp.generate(g, tab, xi, precision=prec), gp.generate_inv(g, tab, xi, precision=prec))
    for value in (1, 0):
        flag[0] = value
        gp._lib.check(lib.ggp_set_yield(flag.data_ptr(), 1), "set_yield")
        try:
            gy = gp.build_graph(pts, n0=1000, k=16)
            for prec in ["fp64", "fp32", "fp32_increment"]:
                v = gp.generate(gy, tab, xi, precision=prec)
                u = gp.generate_inv(gy, tab, xi, precision=prec)
                assert torch.equal(v, ref[prec][0]) and torch.equal(u, ref[prec][1]), (value, prec)
        finally:
            lib.ggp_set_yield(None, 0)
        assert torch.equal(gy.neighbors, g.neighbors) and torch.equal(gy.points, g.points)
        assert gy.offsets == g.offsets


@pytest.mark.gpu
def test_large_and_small_requests_alone_equal_in_mix():
    # a mix with large requests (their launches yield while small requests run)
    import torch

    import graphgp_amd as gp
    from graphgp_amd.serving import Server, materialize

    specs = synthetic_mix(16, rate_per_s=1e4, seed=12, n_min=2e4, n_max=3e6)
    reqs = materialize(specs)
    assert any(r.n > 1_000_000 for r in reqs) and any(r.n <= 1_000_000 for r in reqs)
    srv = Server(max_inflight=8)
    srv.run(reqs, realtime=False)
    srv.close()
    for r in reqs:
        g = gp.build_graph(r.points, n0=r.n0, k=r.k, schedule=False)  # as the server builds
        fn = gp.generate if r.op == "generate" else gp.generate_inv
        alone = fn(g, r.covariance, r.data, precision=r.precision)
        assert torch.equal(alone.view(torch.int32), r.result.view(torch.int32)), (r.rid, r.op, r.n)
