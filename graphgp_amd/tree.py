"""k-d tree build and preceding-neighbour queries (reference: graphgp/tree.py).

Same names, arguments and errors as the reference; the work runs in the gfx950 kernels of
``csrc/tree.hip`` and ``csrc/knn.hip``.  ``cuda`` is accepted for drop-in compatibility:
there is only one (GPU) code path.
"""
from __future__ import annotations

import torch

from . import _lib
from ._arrays import as_index, as_points


def build_tree(points, cuda: bool = False):
    """Build the k-d tree in the reference's special order (tree.py:18-38).

    Returns ``(points, split_dims, indices)``: points in tree order, the split dimension of
    every node and the original index of every tree position (int32).
    """
    p = as_points(points)
    n, d = p.shape
    L = _lib.lib()
    dt = _lib.dtype_code(p)
    out = torch.empty_like(p)
    split = torch.empty(n, dtype=torch.int32, device=p.device)
    idx = torch.empty(n, dtype=torch.int32, device=p.device)
    ws = _lib.workspace(L.ggp_build_tree_workspace(n, d, dt))
    _lib.check(L.ggp_build_tree(_lib.ptr(p), n, d, dt, _lib.ptr(out), _lib.ptr(split), _lib.ptr(idx), _lib.ptr(ws),
                                ws.numel(), _lib.stream()), "build_tree")
    return out, split, idx


def query_preceding_neighbors(points, split_dims, *, n0: int, k: int, cuda: bool = False):
    """k nearest neighbours of every point i >= n0 among the points before it (tree.py:41-63).

    Returns ``neighbors`` of shape ``(N - n0, k)`` (int32), sorted by distance then index.
    """
    if n0 < k:
        raise ValueError(f"n0 must be at least k. Got n0={n0}, k={k}.")
    p = as_points(points)
    s = as_index(split_dims)
    n, d = p.shape
    L = _lib.lib()
    dt = _lib.dtype_code(p)
    out = torch.empty((n - n0, k), dtype=torch.int32, device=p.device)
    ws = _lib.workspace(L.ggp_query_neighbors_workspace(n, d, n - n0, k, dt))
    _lib.check(L.ggp_query_preceding_neighbors(_lib.ptr(p), _lib.ptr(s), n, d, dt, n0, k, _lib.ptr(out),
                                               _lib.ptr(ws), ws.numel(), _lib.stream()), "query_preceding_neighbors")
    return out


def query_neighbors(points, split_dims, query_indices, max_indices, *, k: int, cuda: bool = False):
    """k nearest tree nodes with index < max_indices[i] of point query_indices[i] (tree.py:66-76)."""
    p = as_points(points)
    s = as_index(split_dims)
    q = as_index(query_indices)
    m = as_index(max_indices)
    n, d = p.shape
    L = _lib.lib()
    dt = _lib.dtype_code(p)
    out = torch.empty((q.numel(), k), dtype=torch.int32, device=p.device)
    ws = _lib.workspace(L.ggp_query_neighbors_workspace(n, d, q.numel(), k, dt))
    _lib.check(L.ggp_query_neighbors(_lib.ptr(p), _lib.ptr(s), n, d, dt, _lib.ptr(q), _lib.ptr(m), q.numel(), k,
                                     _lib.ptr(out), _lib.ptr(ws), ws.numel(), _lib.stream()), "query_neighbors")
    return out
