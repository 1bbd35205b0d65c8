"""Plain method, small, at its stated size: N = 100,000 points in 2-D, i.i.d. standard normal
(seed 99), n0 = 100, k = 10, RBF (variance 1, scale 0.3) on r = 0 plus log bins over
[1e-4, 10], jitter 1e-4, xi ~ N(0,1) float32.  The whole graph bit for bit against the
oracle; values, inverse and log-det against the FP64 oracle (graphgp/graph.py:73-98,
graphgp/refine.py:21-56, 162-183, 234-240)."""
import numpy as np
import pytest
import torch

import graphgp_amd as gp
import oracle as O

pytestmark = pytest.mark.gpu

N, D, N0, K = 100_000, 2, 100, 10


def cpu(t):
    return t.detach().cpu().numpy()


@pytest.fixture(scope="module")
def c1():
    rng = np.random.default_rng(99)
    p = rng.standard_normal((N, D)).astype(np.float32)
    xi = rng.standard_normal(N).astype(np.float32)
    print(f"[C1] N={N} d={D} n0={N0} k={K}", flush=True)
    g = gp.build_graph(p, n0=N0, k=K)
    og = O.build_graph(p, n0=N0, k=K)
    return p, xi, g, og


def test_c1_graph_bit_exact(c1):
    _, _, g, og = c1
    assert g.offsets == og.offsets
    assert g.offsets[0] == N0 and g.offsets[-1] == N
    assert np.array_equal(cpu(g.neighbors), og.neighbors)
    assert np.array_equal(cpu(g.indices), og.indices)
    assert np.array_equal(cpu(g.points), og.points)
    gp.check_graph(g)


# n_bins 1001 is r = 0 plus 1,000 log bins; the reference README's example uses 1000
@pytest.mark.parametrize("n_bins", [1001, 1000])
def test_c1_values_inverse_logdet(c1, n_bins):
    _, xi, g, og = c1
    kw = dict(variance=1.0, scale=0.3, r_min=1e-4, r_max=10, n_bins=n_bins, jitter=1e-4)
    cov, ocov = gp.extras.rbf_kernel(**kw), O.rbf_kernel(**kw)
    ref = O.generate(og, ocov, xi)
    assert np.all(np.isfinite(ref))
    scale = float(np.max(np.abs(ref)))
    for prec, tol in (("fp64", 1e-10), ("fp32", 2e-3), ("fp32_increment", 2e-4)):
        v = cpu(gp.generate(g, cov, xi, precision=prec))
        err = float(np.max(np.abs(v - ref)))
        print(f"[C1] n_bins={n_bins} generate[{prec}] max|err| = {err:.3e}", flush=True)
        assert err <= tol * max(1.0, scale), (prec, err)
    oxi = O.generate_inv(og, ocov, ref)
    xb = cpu(gp.generate_inv(g, cov, ref, precision="fp64"))
    assert np.allclose(xb, oxi, rtol=1e-8, atol=1e-8)
    assert np.allclose(xb, xi, rtol=1e-6, atol=1e-6)
    ld = float(gp.generate_logdet(g, cov, precision="fp64"))
    assert np.isclose(ld, O.generate_logdet(og, ocov), rtol=1e-10)
