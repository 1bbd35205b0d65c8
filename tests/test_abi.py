"""The C-ABI library loads and exports every entry point that include/graphgp_amd.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "graphgp_amd.h")
LIB = os.path.join(ROOT, "graphgp_amd", "libgraphgp_amd.so")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ggp_[a-z0-9_]+)\s*\(", text)))


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    names = declared()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_python_binding_covers_header():
    from graphgp_amd import _lib

    assert set(declared()) == set(_lib.exported_symbols())


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_workspace_queries_and_errors():
    from graphgp_amd import _lib

    L = _lib.lib()
    assert L.ggp_abi_version() == 1
    assert L.ggp_build_tree_workspace(1000, 3, 0) > 1000 * 4
    assert L.ggp_refine_workspace(10000, 3, 100, 10, 1001, 1, 1) > 0
    # argument errors are reported without touching the GPU
    rc = L.ggp_query_preceding_neighbors(None, None, 100, 3, 0, 5, 10, None, None, 0, None)
    assert rc == -1 and b"n0 must be at least k" in L.ggp_last_error()
    rc = L.ggp_generate(None, 0, 100, 3, None, 10, 5, None, None, None, 1001, None, 0, None, 0, 7, None, 0, None)
    assert rc == -1 and b"precision" in L.ggp_last_error()


def test_python_api_matches_reference_names():
    import graphgp_amd as gp

    ref = ["build_tree", "query_preceding_neighbors", "Graph", "check_graph", "build_graph", "compute_depths",
           "order_by_depth", "generate", "generate_inv", "generate_logdet", "generate_dense", "generate_dense_inv",
           "generate_dense_logdet", "refine", "refine_inv", "refine_logdet", "compute_cov_matrix", "extras"]
    for name in ref:
        assert hasattr(gp, name), name
    for name in ["rbf_kernel", "matern_kernel", "make_cov_bins"]:
        assert hasattr(gp.extras, name)


def test_extras_match_oracle_tables():
    import numpy as np

    import graphgp_amd as gp
    import oracle as O

    a = gp.extras.matern_kernel(p=1, variance=1.0, cutoff=0.05, r_min=1e-6, r_max=2, n_bins=2001, jitter=1e-5)
    b = O.matern_kernel(p=1, variance=1.0, cutoff=0.05, r_min=1e-6, r_max=2, n_bins=2001, jitter=1e-5)
    assert np.allclose(a[0].cpu().numpy(), b[0], rtol=1e-15) and np.allclose(a[1].cpu().numpy(), b[1], rtol=1e-14)
    a = gp.extras.rbf_kernel(variance=1.0, scale=0.3, r_min=1e-4, r_max=10, n_bins=1001, jitter=1e-4)
    b = O.rbf_kernel(variance=1.0, scale=0.3, r_min=1e-4, r_max=10, n_bins=1001, jitter=1e-4)
    assert np.allclose(a[1].cpu().numpy(), b[1], rtol=1e-14)
