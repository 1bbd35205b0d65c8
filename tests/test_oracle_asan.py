"""The CPU oracle built with AddressSanitizer (`make -C oracle asan`) runs every entry point
on small inputs -- duplicates, ties, k = N - 1, d = 1 and 6, malformed graphs -- in a child
process with libasan preloaded; any out-of-bounds access or use after free ends that process
with ASan's report.  The oracle is the checker of every parity test, so a memory error in it
would silently corrupt the reference values."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SCRIPT = r"""
import numpy as np
import oracle as O

rng = np.random.default_rng(3)
cov = O.rbf_kernel(variance=1.0, scale=0.3, r_min=1e-4, r_max=10, n_bins=200, jitter=1e-4)
for n, d, n0, k, dup in [(600, 2, 20, 8, False), (300, 3, 10, 5, True), (200, 1, 4, 3, False),
                         (150, 6, 16, 12, False), (40, 3, 39, 39, False), (64, 2, 63, 63, False)]:
    p = rng.standard_normal((n, d)).astype(np.float32)
    if dup:
        p[n // 2:] = p[: n - n // 2]  # exact duplicates: ties everywhere
    tp, sd, idx = O.build_tree(p)
    tpl, sdl, idxl = O.build_tree(p, literal=True)
    assert np.array_equal(idx, idxl) and np.array_equal(sd, sdl)
    nb = O.query_preceding_neighbors(tp, sd, n0=n0, k=min(k, n0))
    q = np.arange(n0, n)
    O.query_neighbors(tp, sd, q, q, k=min(k, n0), return_visited=True)
    O.brute_preceding_knn(tp, q[:5], k=min(k, n0))
    g = O.build_graph(p, n0=n0, k=min(k, n0))
    dep = O.compute_depths(g.neighbors, n0=n0)
    O.compute_depths_jacobi(g.neighbors, n0=n0)
    if dup:
        continue  # duplicate points make the covariance singular: geometry only
    xi = rng.standard_normal(n)
    v = O.generate(g, cov, xi)
    O.generate_inv(g, cov, v)
    O.generate_logdet(g, cov)
    gv = rng.standard_normal(n)
    O.generate_vjp(g, cov, xi, gv)
    O.generate_inv_vjp(g, cov, v, gv)
    O.generate_logdet_vjp(g, cov)
    O.compute_cov_matrix(cov, p[:7].astype(np.float64), p[3:11].astype(np.float64))
# a malformed graph (a neighbour after its point): the depth routine reports it
bad = np.array([[0], [3], [1], [2]], dtype=np.int32)
try:
    O.compute_depths(bad, n0=1)
    raise AssertionError("malformed graph accepted")
except ValueError:
    pass
print("asan oracle ok")
"""


def test_oracle_under_address_sanitizer():
    asan = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    if not asan or not os.path.exists(asan):
        pytest.skip("libasan is not installed")
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "asan"], check=True)
    env = dict(os.environ)
    env.update(LD_PRELOAD=asan, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0:exitcode=23",
               GGP_ORACLE_LIB=os.path.join(ROOT, "oracle", "_build", "liboracle_asan.so"),
               OMP_NUM_THREADS="2", PYTHONPATH=ROOT)
    r = subprocess.run([sys.executable, "-c", SCRIPT], env=env, capture_output=True, text=True, timeout=600)
    assert "AddressSanitizer" not in r.stderr, r.stderr[-4000:]
    assert r.returncode == 0 and "asan oracle ok" in r.stdout, (r.returncode, r.stdout[-2000:], r.stderr[-4000:])
