"""ggp_generate_logdet_noalloc (static device scratch, no allocation) against
ggp_generate_logdet with a caller workspace, both called through the C ABI."""
import ctypes

import numpy as np
import pytest
import torch

import graphgp_amd as gp
from graphgp_amd import _lib
from graphgp_amd._arrays import as_covariance
from helpers import clustered_points

pytestmark = pytest.mark.gpu

SCRATCH_BYTES = 64 << 20


def _args(g, cov, prec):
    pts = g.points
    nb = g.neighbors
    n, d = pts.shape
    n_ref, k = nb.shape
    bins, vals = as_covariance(cov)
    return (_lib.ptr(pts), _lib.dtype_code(pts), n, d, _lib.ptr(nb), n - n_ref, k, _lib.ptr(bins), _lib.ptr(vals),
            bins.numel())


def _with_workspace(g, cov, prec):
    L = _lib.lib()
    a = _args(g, cov, prec)
    out = torch.full((), float("nan"), dtype=torch.float64, device="cuda")
    n, d, n0, k, nbins = a[2], a[3], a[5], a[6], a[9]
    ws = _lib.workspace(L.ggp_refine_workspace(n, d, n0, k, nbins, _lib.PREC[prec], 1))
    _lib.check(L.ggp_generate_logdet(*a, _lib.ptr(out), _lib.PREC[prec], _lib.ptr(ws), ws.numel(), _lib.stream()),
               "generate_logdet")
    return out


def _noalloc(g, cov, prec):
    L = _lib.lib()
    out = torch.full((), float("nan"), dtype=torch.float64, device="cuda")
    _lib.check(L.ggp_generate_logdet_noalloc(*_args(g, cov, prec), _lib.ptr(out), _lib.PREC[prec], _lib.stream()),
               "generate_logdet_noalloc")
    return out


@pytest.fixture(scope="module")
def graphs():
    cov = gp.extras.matern_kernel(p=1, variance=1.0, cutoff=0.05, r_min=1e-6, r_max=2.0, n_bins=2001, jitter=1e-4)
    big = gp.build_graph(clustered_points(300_000, d=3, seed=71), n0=1000, k=30)
    small = gp.build_graph(clustered_points(20_000, d=2, seed=72), n0=200, k=10)
    return cov, big, small


@pytest.mark.parametrize("prec", ["fp64", "fp32", "fp32_increment"])
def test_noalloc_logdet_equals_workspace_entry(graphs, prec):
    cov, big, small = graphs
    # two calls in a row on the shared scratch, the second with a smaller problem
    for g in (big, small):
        print(f"noalloc logdet: n={g.points.shape[0]} k={g.neighbors.shape[1]} prec={prec}", flush=True)
        a = _noalloc(g, cov, prec)
        b = _with_workspace(g, cov, prec)
        torch.cuda.synchronize()
        assert torch.isfinite(a) and torch.isfinite(b)
        assert float(a) == float(b), (float(a), float(b))
        assert np.isclose(float(a), float(gp.generate_logdet(g, cov, precision=prec)), rtol=0, atol=0)
    # and the larger one again after the smaller (stale partial sums must not leak in)
    assert float(_noalloc(big, cov, prec)) == float(_with_workspace(big, cov, prec))


def test_noalloc_rejects_problems_beyond_scratch(graphs):
    cov, big, _ = graphs
    L = _lib.lib()
    need = L.ggp_refine_workspace(big.points.shape[0], 3, 1000, 30, 2001, 0, 1)
    assert need < SCRATCH_BYTES
    # a seed block of 3000 points needs 72 MB of scratch: rejected on the host, nothing launched
    rc = L.ggp_generate_logdet_noalloc(None, 0, 400_000, 3, None, 3000, 30, None, None, 2001, None, 0, None)
    assert rc == -1
    assert b"more than the" in L.ggp_last_error()
