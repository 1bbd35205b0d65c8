"""Neighbour graph: depths, depth order and batch offsets (reference: graphgp/graph.py).

``build_graph`` runs the whole device pipeline (``ggp_build_graph``: tree, preceding
kNN, depths, stable depth order) and reads back the largest depth once to size the
static ``offsets`` tuple, which the reference also materialises on the host
(graph.py:95-96).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Optional, Tuple

import ctypes

import torch

from . import _lib
from ._arrays import as_index, as_points


@dataclass
class Graph:
    """Nearest-neighbour dependency graph (graph.py:20-41).

    Fields:
        points: modeled points in graph order, ``(N, d)``.
        neighbors: indices of the neighbours, ``(N - offsets[0], k)`` (int32).
        offsets: tuple with the end index of every batch (static, host side).
        indices: original index of every graph position, ``(N,)`` (int32), or ``None``.
    """

    points: torch.Tensor
    neighbors: torch.Tensor
    offsets: Tuple[int, ...] = field(metadata=dict(static=True))
    indices: Optional[torch.Tensor] = None


def check_graph(graph: Graph):
    """Verify that ``graph`` is valid for GraphGP (graph.py:44-70); raises AssertionError."""
    points, neighbors, offsets = graph.points, graph.neighbors, tuple(int(x) for x in graph.offsets)
    n, n_ref = points.shape[0], neighbors.shape[0]
    assert len(offsets) >= 1 and offsets[0] == n - n_ref, "Neighbors should start from first offset"
    assert all(b >= a for a, b in zip(offsets[:-1], offsets[1:])), "Offsets must be non-decreasing"
    assert offsets[-1] <= n, "Last offset must be less than or equal to the number of points"
    if n_ref == 0:
        return
    nb = as_index(neighbors)
    off = torch.tensor(offsets, dtype=torch.int64, device=nb.device)
    status = torch.zeros(1, dtype=torch.int32, device=nb.device)
    L = _lib.lib()
    _lib.check(L.ggp_check_graph(_lib.ptr(nb), n_ref, nb.shape[1], n, _lib.ptr(off), len(offsets), _lib.ptr(status),
                                 _lib.stream()), "check_graph")
    s = int(_lib.to_host(status)[0])
    assert not (s & 2), "Points are not in topological order"
    assert not (s & 4), "Neighbors must not be in the same batch"


def _offsets(depths_sorted: torch.Tensor, max_depth: int) -> Tuple[int, ...]:
    n_off = int(max_depth) + 1
    off = torch.empty(n_off, dtype=torch.int64, device=depths_sorted.device)
    _lib.check(_lib.lib().ggp_depth_offsets(_lib.ptr(depths_sorted), depths_sorted.numel(), n_off, _lib.ptr(off),
                                            _lib.stream()), "offsets")
    return tuple(int(x) for x in _lib.to_host(off).tolist())


# Graphs from this many refined points on get a dataflow schedule (ggp_graph_schedule):
# ordering every batch by space makes a tile's neighbour values and coordinates share
# cache lines.  It costs one sort, an (n - n0, k) int32 array and a copy of the points per
# graph, kept on the Graph.  FP32-increment generate of the clustered 3-D mixture, k = 30,
# with and without it (ms): 150k points 1.60 / 2.19, 300k 1.77 / 2.42, 600k 2.00 / 2.73,
# 1.05M 2.41 / 3.25.
SCHEDULE_MIN_REFINED = 1 << 17


def _schedule(graph: Graph, pts, nb):
    """The graph's dataflow schedule, computed on first use and kept on the graph; None for
    small graphs or more than 65535 batches.  The cache key holds the tensors' addresses
    and version counters, so a graph whose points or neighbours were changed in place
    gets a fresh schedule instead of stale slot-order copies."""
    n, d = pts.shape
    n_ref, k = nb.shape
    if n_ref < SCHEDULE_MIN_REFINED or len(graph.offsets) > 65536:
        return None
    key = (pts.data_ptr(), pts._version, nb.data_ptr(), nb._version, tuple(graph.offsets))
    cached = getattr(graph, "_ggp_schedule", None)
    if cached is not None and cached[0] in (key, "off"):
        return cached[1]
    n0 = n - n_ref
    L = _lib.lib()
    kept = n_ref * (k + 1) * 4 + pts.numel() * pts.element_size()
    ws_bytes = L.ggp_graph_schedule_workspace(n, n0)
    free, _ = torch.cuda.mem_get_info(pts.device)
    if kept + ws_bytes > free // 2:  # a large graph near the memory limit runs in graph order
        return None
    # pinned and non-blocking: a serving thread must not wait here for its stream's work
    off = torch.tensor(graph.offsets, dtype=torch.int64).pin_memory().to(pts.device, non_blocking=True)
    order = torch.empty(n_ref, dtype=torch.int32, device=pts.device)
    nbv = torch.empty((n_ref, k), dtype=torch.int32, device=pts.device)
    pts_s = torch.empty_like(pts)
    ws = _lib.workspace(ws_bytes)
    _lib.check(L.ggp_graph_schedule(_lib.ptr(pts), _lib.dtype_code(pts), n, d, _lib.ptr(nb), n0, k, _lib.ptr(off),
                                    off.numel(), _lib.ptr(order), _lib.ptr(nbv), _lib.ptr(pts_s), _lib.ptr(ws),
                                    ws.numel(), _lib.stream()), "graph schedule")
    graph._ggp_schedule = (key, (order, nbv, pts_s))
    return order, nbv, pts_s


def build_graph(points, *, n0: int, k: int, cuda: bool = False, return_depths: bool = False,
                schedule: bool = True) -> Graph:
    """Graph of each point's k nearest preceding neighbours in k-d tree order (graph.py:73-98).

    ``schedule`` (not in the reference): for large graphs, also compute the dataflow
    schedule that ``generate`` uses (an order of evaluation; results are identical without
    it).  ``False`` skips it for this graph, saving its memory."""
    p = as_points(points)
    n, d = p.shape
    if n0 < k:
        raise ValueError(f"n0 must be at least k. Got n0={n0}, k={k}.")
    if n0 > n:
        raise ValueError(f"n0 must not exceed the number of points. Got n0={n0}, N={n}.")
    L = _lib.lib()
    dt = _lib.dtype_code(p)
    dev = p.device
    pts = torch.empty_like(p)
    idx = torch.empty(n, dtype=torch.int32, device=dev)
    nbrs = torch.empty((n - n0, k), dtype=torch.int32, device=dev)
    depths = torch.empty(n, dtype=torch.int32, device=dev)
    ws = _lib.workspace(L.ggp_build_graph_workspace(n, d, dt, n0, k))
    mx = ctypes.c_int32(0)
    _lib.check(L.ggp_build_graph(_lib.ptr(p), n, d, dt, n0, k, _lib.ptr(pts), _lib.ptr(idx), _lib.ptr(nbrs),
                                 _lib.ptr(depths), ctypes.byref(mx), _lib.ptr(ws), ws.numel(), _lib.stream()),
               "build_graph")
    del ws
    g = Graph(pts, nbrs, _offsets(depths, mx.value), idx)
    if schedule:
        _schedule(g, pts, nbrs)  # large graphs: the dataflow schedule, while the build has the stream
    else:
        g._ggp_schedule = ("off", None)
    return (g, depths) if return_depths else g


DEPTH_METHODS = {"flow": 0, "rounds": 1, "sweeps": 2}


def compute_depths(neighbors, *, n0: int, cuda: bool = False, method: str = "flow"):
    """Longest-path depth of every point (graph.py:101-120); seeds have depth 0.

    ``method="flow"`` (default) computes the depths as a dataflow in index order, in one
    launch; ``"rounds"`` runs Kahn rounds over the transposed graph; ``"sweeps"`` runs
    in-place sweeps without the transpose (for graphs whose transpose would not fit in
    device memory).  All give the reference's fixed point.
    """
    if method not in DEPTH_METHODS:
        raise ValueError(f"method must be one of {tuple(DEPTH_METHODS)}, got {method!r}")
    nb = as_index(neighbors)
    n_ref, k = nb.shape
    L = _lib.lib()
    out = torch.empty(n0 + n_ref, dtype=torch.int32, device=nb.device)
    ws = _lib.workspace(L.ggp_compute_depths_workspace(n_ref, k) if method == "rounds" or k > 64 else 64)
    _lib.check(L.ggp_compute_depths_method(_lib.ptr(nb), n_ref, k, n0, _lib.ptr(out), DEPTH_METHODS[method],
                                           _lib.ptr(ws), ws.numel(), _lib.stream()), "compute_depths")
    status = ctypes.c_uint32(0)
    _lib.check(L.ggp_compute_depths_status(_lib.ptr(ws), ctypes.byref(status), _lib.stream()), "compute_depths")
    if status.value:
        raise ValueError("neighbors are not in topological order (a neighbour does not precede its point)")
    return out


def order_by_depth(points, indices, neighbors, depths, *, cuda: bool = False):
    """Stable reorder by depth with the neighbour remap (graph.py:123-137)."""
    p = as_points(points)
    n, d = p.shape
    idx = as_index(indices) if indices is not None else None
    nb = as_index(neighbors)
    dep = as_index(depths)
    n_ref, k = nb.shape
    L = _lib.lib()
    dt = _lib.dtype_code(p)
    max_depth = int(_lib.to_host(dep.max())) if n else 0
    po = torch.empty_like(p)
    io = torch.empty(n, dtype=torch.int32, device=p.device)
    no = torch.empty_like(nb)
    do = torch.empty_like(dep)
    ws = _lib.workspace(L.ggp_order_by_depth_workspace(n))
    _lib.check(L.ggp_order_by_depth(_lib.ptr(p), n, d, dt, _lib.ptr(idx), _lib.ptr(nb), n_ref, k, _lib.ptr(dep),
                                    max_depth, _lib.ptr(po), _lib.ptr(io), _lib.ptr(no), _lib.ptr(do), _lib.ptr(ws),
                                    ws.numel(), _lib.stream()), "order_by_depth")
    return po, io, no, do
