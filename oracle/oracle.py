"""ctypes front end of the CPU oracle (oracle/graphgp_oracle.cpp) -- TEST INFRASTRUCTURE ONLY.

Also holds NumPy restatements of the reference's table builders (graphgp/extras.py:14-68)
and of the small host steps of graph.py, used to produce oracle inputs and to pin the
product's own table builders against closed forms.
"""
from __future__ import annotations

import ctypes
import math
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# GGP_ORACLE_LIB selects another build of the same source (the AddressSanitizer build)
_LIB_PATH = os.environ.get("GGP_ORACLE_LIB") or os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

_i64 = ctypes.c_int64
_i32 = ctypes.c_int32
_dbl = ctypes.c_double
_p = ctypes.c_void_p


def build():
    """Compile the oracle with its Makefile (gcc), in place."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _lib.ora_interp.restype = _dbl
        _lib.ora_generate_dense_logdet.restype = _dbl
        _lib.ora_refine_logdet.restype = _dbl
        _lib.ora_offsets.restype = _i64
        _lib.ora_compute_depths.restype = ctypes.c_int
        _lib.ora_num_threads.restype = ctypes.c_int
    return _lib


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def num_threads():
    return lib().ora_num_threads()


def set_threads(n):
    lib().ora_set_threads(int(n))


# ----------------------------------------------------------------- extras.py
def make_cov_bins(*, r_min, r_max, n_bins):
    """extras.py:65-68: [0, logspace(log10 r_min, log10 r_max, n_bins - 1)]."""
    b = np.logspace(np.log10(r_min), np.log10(r_max), n_bins - 1)
    return np.concatenate([[0.0], b])


def rbf_kernel(*, variance, scale, r_min, r_max, n_bins, jitter=0.0):
    """extras.py:14-31."""
    r = make_cov_bins(r_min=r_min, r_max=r_max, n_bins=n_bins)
    cov = variance * np.exp(-0.5 * (r / scale) ** 2)
    cov = np.where(r == 0.0, cov[0] * (1.0 + jitter), cov)
    return r, cov


def matern_kernel(*, p, variance, cutoff, r_min, r_max, n_bins, jitter=0.0):
    """extras.py:34-58: nu = p + 1/2, polyval of p!(p+i)!/(i!(p-i)!(2p)!) at 2x."""
    r = make_cov_bins(r_min=r_min, r_max=r_max, n_bins=n_bins)
    x = math.sqrt(2 * p + 1) * r / cutoff
    lf = lambda v: math.lgamma(v + 1)  # noqa: E731
    coeff = [math.exp(lf(p) + lf(p + i) - lf(i) - lf(p - i) - lf(2 * p)) for i in range(p + 1)]
    cov = variance * np.exp(-x) * np.polyval(coeff, 2 * x)
    cov = np.where(r == 0.0, cov[0] * (1.0 + jitter), cov)
    return r, cov


# ------------------------------------------------------------------- tree.py
def build_tree(points, literal=False):
    """tree.py:153-184.  Returns (points, split_dims int32, indices int64).  ``literal`` runs
    the whole-array sort per level as the reference writes it (slow; for pinning the
    default, which sorts each node's contiguous run of members independently)."""
    pts = np.ascontiguousarray(points)
    if pts.dtype not in (np.float32, np.float64):
        pts = pts.astype(np.float64)
    n, d = pts.shape
    po = np.empty_like(pts)
    sd = np.empty(n, np.int32)
    idx = np.empty(n, np.int64)
    if literal:
        fn = lib().ora_build_tree_literal_f32 if pts.dtype == np.float32 else lib().ora_build_tree_literal_f64
    else:
        fn = lib().ora_build_tree_f32 if pts.dtype == np.float32 else lib().ora_build_tree_f64
    fn(_ptr(pts), _i64(n), ctypes.c_int(d), _ptr(po), _ptr(sd), _ptr(idx))
    return po, sd, idx


def query_neighbors(points, split_dims, query_indices, max_indices, *, k, return_visited=False):
    """tree.py:66-105."""
    pts = np.ascontiguousarray(points)
    sd = np.ascontiguousarray(split_dims, dtype=np.int32)
    q = np.ascontiguousarray(query_indices, dtype=np.int64)
    m = np.ascontiguousarray(max_indices, dtype=np.int64)
    n, d = pts.shape
    out = np.empty((len(q), k), np.int32)
    vis = np.empty(len(q), np.int64)
    fn = lib().ora_query_neighbors_f32 if pts.dtype == np.float32 else lib().ora_query_neighbors_f64
    fn(_ptr(pts), _ptr(sd), _i64(n), ctypes.c_int(d), _ptr(q), _ptr(m), _i64(len(q)), ctypes.c_int(k),
       _ptr(out), _ptr(vis))
    return (out, vis) if return_visited else out


def query_preceding_neighbors(points, split_dims, *, n0, k):
    """tree.py:41-63."""
    if n0 < k:
        raise ValueError(f"n0 must be at least k. Got n0={n0}, k={k}.")
    q = np.arange(n0, len(points), dtype=np.int64)
    return query_neighbors(points, split_dims, q, q, k=k)


def brute_preceding_knn(points, queries, *, k):
    """The k smallest FP32 squared distances (and their tree positions) among the points
    preceding each query position, ordered by (distance^2, index): what the preceding kNN
    (tree.py:41-63) must find, computed by brute force in one pass over the points."""
    pts = np.ascontiguousarray(points, dtype=np.float32)
    q = np.ascontiguousarray(queries, dtype=np.int64)
    n, d = pts.shape
    d2 = np.empty((len(q), k), np.float32)
    idx = np.empty((len(q), k), np.int64)
    lib().ora_brute_preceding_knn_f32(_ptr(pts), _i64(n), ctypes.c_int(d), _ptr(q), _i64(len(q)), ctypes.c_int(k),
                                      _ptr(d2), _ptr(idx))
    return d2, idx


def sqdist_f32(points, a, b):
    """FP32 squared distances between rows a and b of points, summed left to right with
    separate roundings (the kNN's arithmetic)."""
    pts = np.asarray(points, dtype=np.float32)
    diff = pts[a] - pts[b]
    s = diff[..., 0] * diff[..., 0]
    for j in range(1, pts.shape[1]):
        s = s + diff[..., j] * diff[..., j]
    return s


# ------------------------------------------------------------------ graph.py
def compute_depths(neighbors, *, n0):
    """graph.py:101-120 (forward pass == the reference's fixed point for valid graphs)."""
    nb = np.ascontiguousarray(neighbors, dtype=np.int32)
    n_ref, k = nb.shape
    out = np.empty(n0 + n_ref, np.int32)
    rc = lib().ora_compute_depths(_ptr(nb), _i64(n_ref), ctypes.c_int(k), _i64(n0), _ptr(out))
    if rc != 0:
        raise ValueError("neighbors are not in topological order")
    return out


def compute_depths_jacobi(neighbors, *, n0):
    """graph.py:110-119 literally: Jacobi sweeps until nothing changes (small inputs)."""
    nb = np.asarray(neighbors)
    depths = np.zeros(n0 + len(nb), np.int32)
    old = depths - 1
    while np.any(old != depths):
        new = depths.copy()
        new[n0:] = 1 + np.max(depths[nb], axis=1)
        old, depths = depths, new
    return depths


def order_by_depth(points, indices, neighbors, depths):
    """graph.py:123-137."""
    pts = np.ascontiguousarray(points, dtype=np.float64)
    n, d = pts.shape
    nb = np.ascontiguousarray(neighbors, dtype=np.int32)
    n0 = n - len(nb)
    k = nb.shape[1]
    idx = np.ascontiguousarray(indices, dtype=np.int64)
    dep = np.ascontiguousarray(depths, dtype=np.int32)
    po = np.empty_like(pts)
    io = np.empty(n, np.int64)
    no = np.empty_like(nb)
    do = np.empty(n, np.int32)
    lib().ora_order_by_depth(_ptr(pts), _i64(n), ctypes.c_int(d), _ptr(idx), _ptr(nb), ctypes.c_int(k), _ptr(dep),
                             _ptr(po), _ptr(io), _ptr(no), _ptr(do), _i64(n0))
    return po.astype(points.dtype, copy=False), io, no, do


def offsets_from_depths(depths_sorted):
    """graph.py:95-96."""
    dep = np.ascontiguousarray(depths_sorted, dtype=np.int32)
    cap = int(dep.max()) + 1 if len(dep) else 1
    out = np.empty(cap, np.int64)
    cnt = lib().ora_offsets(_ptr(dep), _i64(len(dep)), _ptr(out), _i64(cap))
    return tuple(int(x) for x in out[:cnt])


class OGraph:
    """graph.py:20-41 fields."""

    def __init__(self, points, neighbors, offsets, indices=None):
        self.points, self.neighbors, self.offsets, self.indices = points, neighbors, offsets, indices


def build_graph(points, *, n0, k):
    """graph.py:73-98."""
    p, sd, idx = build_tree(points)
    nb = query_preceding_neighbors(p, sd, n0=n0, k=k)
    dep = compute_depths(nb, n0=n0)
    p2, idx2, nb2, dep2 = order_by_depth(p, idx, nb, dep)
    return OGraph(p2.astype(p.dtype), nb2, offsets_from_depths(dep2), idx2)


# ----------------------------------------------------------------- refine.py
def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _cov(covariance):
    bins, vals = covariance
    return _f64(bins), _f64(vals)


def interp(x, bins, vals):
    bins, vals = _f64(bins), _f64(vals)
    return np.array([lib().ora_interp(_dbl(float(v)), _ptr(bins), _ptr(vals), ctypes.c_int(len(bins)))
                     for v in np.ravel(x)]).reshape(np.shape(x))


def compute_cov_matrix(covariance, a, b):
    """refine.py:288-295."""
    bins, vals = _cov(covariance)
    a, b = _f64(a), _f64(b)
    out = np.empty((len(a), len(b)))
    lib().ora_cov_matrix(_ptr(a), _i64(len(a)), _ptr(b), _i64(len(b)), ctypes.c_int(a.shape[1]), _ptr(bins),
                         _ptr(vals), ctypes.c_int(len(bins)), _ptr(out))
    return out


def generate_dense(points, covariance, xi, return_L=False):
    """refine.py:59-76."""
    bins, vals = _cov(covariance)
    p, x = _f64(points), _f64(xi)
    n, d = p.shape
    out = np.empty(n)
    L = np.empty((n, n)) if return_L else None
    lib().ora_generate_dense(_ptr(p), _i64(n), ctypes.c_int(d), _ptr(bins), _ptr(vals), ctypes.c_int(len(bins)),
                             _ptr(x), _ptr(out), _ptr(L) if return_L else None)
    return (out, L) if return_L else out


def generate_dense_inv(points, covariance, values):
    """refine.py:186-195."""
    bins, vals = _cov(covariance)
    p, v = _f64(points), _f64(values)
    n, d = p.shape
    out = np.empty(n)
    lib().ora_generate_dense_inv(_ptr(p), _i64(n), ctypes.c_int(d), _ptr(bins), _ptr(vals), ctypes.c_int(len(bins)),
                                 _ptr(v), _ptr(out))
    return out


def generate_dense_logdet(points, covariance):
    """refine.py:243-248."""
    bins, vals = _cov(covariance)
    p = _f64(points)
    n, d = p.shape
    return lib().ora_generate_dense_logdet(_ptr(p), _i64(n), ctypes.c_int(d), _ptr(bins), _ptr(vals),
                                           ctypes.c_int(len(bins)))


def _graph_args(points, neighbors):
    p = _f64(points)
    nb = np.ascontiguousarray(neighbors, dtype=np.int32)
    n, d = p.shape
    return p, nb, n, d, nb.shape[1], n - len(nb)


def refine_weights(points, neighbors, covariance):
    """refine.py:128-135: (mean_vec, std) for every refined point."""
    bins, vals = _cov(covariance)
    p, nb, n, d, k, n0 = _graph_args(points, neighbors)
    mv = np.empty((n - n0, k))
    sd = np.empty(n - n0)
    lib().ora_refine_weights(_ptr(p), _i64(n), ctypes.c_int(d), _ptr(nb), ctypes.c_int(k), _i64(n0), _ptr(bins),
                             _ptr(vals), ctypes.c_int(len(bins)), _ptr(mv), _ptr(sd))
    return mv, sd


def refine(points, neighbors, offsets, covariance, initial_values, xi, *, loop=False):
    """refine.py:79-159 (fast_jit form, or the loop form with ``loop=True``)."""
    bins, vals = _cov(covariance)
    p, nb, n, d, k, n0 = _graph_args(points, neighbors)
    iv, x = _f64(initial_values), _f64(xi)
    out = np.empty(n)
    if loop:
        off = np.ascontiguousarray(offsets, dtype=np.int64)
        lib().ora_refine_loop(_ptr(p), _i64(n), ctypes.c_int(d), _ptr(nb), ctypes.c_int(k), _i64(n0), _ptr(off),
                              _i64(len(off)), _ptr(bins), _ptr(vals), ctypes.c_int(len(bins)), _ptr(iv), _ptr(x),
                              _ptr(out))
    else:
        lib().ora_refine(_ptr(p), _i64(n), ctypes.c_int(d), _ptr(nb), ctypes.c_int(k), _i64(n0), _ptr(bins),
                         _ptr(vals), ctypes.c_int(len(bins)), _ptr(iv), _ptr(x), _ptr(out))
    return out


def refine_inv(points, neighbors, offsets, covariance, values):
    """refine.py:198-231."""
    bins, vals = _cov(covariance)
    p, nb, n, d, k, n0 = _graph_args(points, neighbors)
    v = _f64(values)
    xi = np.empty(n - n0)
    lib().ora_refine_inv(_ptr(p), _i64(n), ctypes.c_int(d), _ptr(nb), ctypes.c_int(k), _i64(n0), _ptr(bins),
                         _ptr(vals), ctypes.c_int(len(bins)), _ptr(v), _ptr(xi))
    return v[:n0].copy(), xi


def refine_logdet(points, neighbors, offsets, covariance):
    """refine.py:251-275."""
    bins, vals = _cov(covariance)
    p, nb, n, d, k, n0 = _graph_args(points, neighbors)
    return lib().ora_refine_logdet(_ptr(p), _i64(n), ctypes.c_int(d), _ptr(nb), ctypes.c_int(k), _i64(n0),
                                   _ptr(bins), _ptr(vals), ctypes.c_int(len(bins)))


def local_conditional(points, neighbors_row, fine_index, covariance, neighbor_values):
    """Conditional mean and std of one refined point given its neighbours' values."""
    bins, vals = _cov(covariance)
    rows = np.asarray(neighbors_row, dtype=np.int64)
    # only the k + 1 points involved are converted (the point set may be large)
    p = _f64(np.concatenate([np.asarray(points)[rows], np.asarray(points)[int(fine_index)][None]]))
    nb = np.arange(len(rows), dtype=np.int32)
    fine_index = len(rows)
    nv = _f64(neighbor_values)
    mean = ctypes.c_double()
    std = ctypes.c_double()
    lib().ora_local_conditional(_ptr(p), ctypes.c_int(p.shape[1]), _ptr(nb), ctypes.c_int(len(nb)),
                                _i64(fine_index), _ptr(bins), _ptr(vals), ctypes.c_int(len(bins)), _ptr(nv),
                                ctypes.byref(mean), ctypes.byref(std))
    return mean.value, std.value


def generate(graph, covariance, xi):
    """refine.py:21-56."""
    n0 = len(graph.points) - len(graph.neighbors)
    xi = _f64(xi)
    if graph.indices is not None:
        xi = xi[graph.indices]
    init = generate_dense(graph.points[:n0], covariance, xi[:n0])
    v = refine(graph.points, graph.neighbors, graph.offsets, covariance, init, xi[n0:])
    if graph.indices is not None:
        out = np.empty_like(v)
        out[graph.indices] = v
        v = out
    if np.any(np.isnan(v)):
        v = np.full_like(v, np.nan)
    return v


def generate_inv(graph, covariance, values):
    """refine.py:162-183."""
    n0 = len(graph.points) - len(graph.neighbors)
    values = _f64(values)
    if graph.indices is not None:
        values = values[graph.indices]
    init, xi = refine_inv(graph.points, graph.neighbors, graph.offsets, covariance, values)
    xi0 = generate_dense_inv(graph.points[:n0], covariance, init)
    xi = np.concatenate([xi0, xi])
    if graph.indices is not None:
        out = np.empty_like(xi)
        out[graph.indices] = xi
        xi = out
    if np.any(np.isnan(xi)):
        xi = np.full_like(xi, np.nan)
    return xi


def generate_logdet(graph, covariance):
    """refine.py:234-240."""
    n0 = len(graph.points) - len(graph.neighbors)
    return generate_dense_logdet(graph.points[:n0], covariance) + refine_logdet(
        graph.points, graph.neighbors, graph.offsets, covariance)


def refine_vjp(points, neighbors, covariance, values, xi, gvalues):
    """Hand-written reverse mode of refine for fixed seed values (see the C source)."""
    bins, vals = _cov(covariance)
    p, nb, n, d, k, n0 = _graph_args(points, neighbors)
    v, x, g = _f64(values), _f64(xi), _f64(gvalues)
    gxi = np.empty(n - n0)
    ginit = np.empty(n0)
    gvals = np.empty(len(bins))
    lib().ora_refine_vjp(_ptr(p), _i64(n), ctypes.c_int(d), _ptr(nb), ctypes.c_int(k), _i64(n0), _ptr(bins),
                         _ptr(vals), ctypes.c_int(len(bins)), _ptr(v), _ptr(x), _ptr(g), _ptr(gxi), _ptr(ginit),
                         _ptr(gvals))
    return gxi, ginit, gvals


def refine_inv_vjp(points, neighbors, covariance, values, xi_bar):
    """Hand-written reverse mode of refine_inv's refined points (refine.py:198-231) for
    fixed tables: (values_bar (N,), cov_vals_bar) for the cotangent xi_bar (N - n0,)."""
    bins, vals = _cov(covariance)
    p, nb, n, d, k, n0 = _graph_args(points, neighbors)
    v, t = _f64(values), _f64(xi_bar)
    gv = np.empty(n)
    gvals = np.empty(len(bins))
    lib().ora_refine_local_vjp(_ptr(p), _i64(n), ctypes.c_int(d), _ptr(nb), ctypes.c_int(k), _i64(n0), _ptr(bins),
                               _ptr(vals), ctypes.c_int(len(bins)), ctypes.c_int(1), _ptr(v), _ptr(t), _ptr(gv),
                               _ptr(gvals))
    return gv, gvals


def refine_logdet_vjp(points, neighbors, covariance, logdet_bar=1.0):
    """Hand-written reverse mode of refine_logdet (refine.py:251-275): cov_vals_bar."""
    bins, vals = _cov(covariance)
    p, nb, n, d, k, n0 = _graph_args(points, neighbors)
    lb = np.array([float(logdet_bar)])
    gvals = np.empty(len(bins))
    lib().ora_refine_local_vjp(_ptr(p), _i64(n), ctypes.c_int(d), _ptr(nb), ctypes.c_int(k), _i64(n0), _ptr(bins),
                               _ptr(vals), ctypes.c_int(len(bins)), ctypes.c_int(2), None, _ptr(lb), None,
                               _ptr(gvals))
    return gvals


def _chol_adjoint(L, Lbar):
    """Cotangent of the symmetric K for the cotangent Lbar of L = chol(K) (Murray 2016, eq. 10):
    1/2 L^{-T} Phi(L^T Lbar) L^{-1}, symmetrised."""
    import scipy.linalg as sla

    P = L.T @ Lbar
    P = np.tril(P) - 0.5 * np.diag(np.diag(P))
    S = sla.solve_triangular(L.T, sla.solve_triangular(L.T, P, lower=False).T, lower=False).T
    return 0.5 * (S + S.T)


def _pair_r(p):
    return np.sqrt(((p[:, None, :] - p[None, :, :]) ** 2).sum(-1))


def dense_inv_vjp(points, covariance, values, xi_bar):
    """Reverse mode of generate_dense_inv (xi = L^{-1} values, refine.py:186-195) in NumPy
    FP64: (values_bar, cov_vals_bar)."""
    import scipy.linalg as sla

    bins, _ = _cov(covariance)
    p = _f64(points)
    n = len(p)
    _, L = generate_dense(p, covariance, np.zeros(n), return_L=True)
    v, t = _f64(values), _f64(xi_bar)
    xs = sla.solve_triangular(L, v, lower=True)
    a = sla.solve_triangular(L.T, t, lower=False)
    Kbar = _chol_adjoint(L, -np.tril(np.outer(a, xs)))
    return a, _lookup_grad_all(bins, _pair_r(p).ravel(), Kbar.ravel())


def dense_logdet_vjp(points, covariance, logdet_bar=1.0):
    """Reverse mode of generate_dense_logdet (slogdet(K)[1] / 2, refine.py:243-248):
    K_bar = logdet_bar K^{-1} / 2 through every pair's lookup."""
    bins, _ = _cov(covariance)
    p = _f64(points)
    K = compute_cov_matrix(covariance, p, p)
    Kbar = 0.5 * float(logdet_bar) * np.linalg.inv(K)
    return _lookup_grad_all(bins, _pair_r(p).ravel(), Kbar.ravel())


def generate_inv_vjp(graph, covariance, values, xi_bar):
    """Reverse mode of generate_inv (refine.py:162-183): caller order in and out; returns
    (values_bar (N,), cov_vals_bar (n_bins,))."""
    n = len(graph.points)
    n0 = n - len(graph.neighbors)
    v, t = _f64(values), _f64(xi_bar)
    if graph.indices is not None:
        v, t = v[graph.indices], t[graph.indices]
    gv, gtab = refine_inv_vjp(graph.points, graph.neighbors, covariance, v, t[n0:])
    ga, gtd = dense_inv_vjp(graph.points[:n0], covariance, v[:n0], t[:n0])
    gv[:n0] += ga
    if graph.indices is not None:
        out = np.empty_like(gv)
        out[graph.indices] = gv
        gv = out
    return gv, gtab + gtd


def generate_logdet_vjp(graph, covariance, logdet_bar=1.0):
    """Reverse mode of generate_logdet (refine.py:234-240): cov_vals_bar (n_bins,)."""
    n0 = len(graph.points) - len(graph.neighbors)
    return (refine_logdet_vjp(graph.points, graph.neighbors, covariance, logdet_bar)
            + dense_logdet_vjp(graph.points[:n0], covariance, logdet_bar))


def dense_vjp(points, covariance, xi, gvalues):
    """Reverse mode of generate_dense (values = chol(K) @ xi) in NumPy FP64.

    dxi = L^T g; with Lbar = tril(g xi^T), the Cholesky adjoint (Murray 2016, eq. 10)
    Kbar = 1/2 L^{-T} Phi(L^T Lbar) L^{-1} symmetrised; each off-diagonal table lookup
    feeds K_ij and K_ji, so its gradient is Kbar_ij + Kbar_ji.
    """
    import scipy.linalg as sla

    bins, vals = _cov(covariance)
    p = _f64(points)
    n = len(p)
    _, L = generate_dense(p, covariance, np.zeros(n), return_L=True)
    g = _f64(gvalues)
    x = _f64(xi)
    gxi = L.T @ g
    Lbar = np.tril(np.outer(g, x))
    P = L.T @ Lbar
    P = np.tril(P) - 0.5 * np.diag(np.diag(P))
    # S = L^{-T} P L^{-1} by two triangular solves
    S = sla.solve_triangular(L.T, sla.solve_triangular(L.T, P, lower=False).T, lower=False).T
    Kbar = 0.5 * (S + S.T)
    r = np.sqrt(((p[:, None, :] - p[None, :, :]) ** 2).sum(-1))
    return gxi, _lookup_grad_all(bins, r.ravel(), Kbar.ravel())


def generate_vjp(graph, covariance, xi, gvalues):
    """Reverse mode of generate (refine.py:21-56): refine_vjp over the refined points, then
    dense_vjp of the seed block with the seeds' accumulated cotangents.  Caller order in and
    out; returns (xi_bar (N,), cov_vals_bar (n_bins,))."""
    n = len(graph.points)
    n0 = n - len(graph.neighbors)
    x = _f64(xi)
    g = _f64(gvalues)
    if graph.indices is not None:
        x = x[graph.indices]
        g = g[graph.indices]
    init = generate_dense(graph.points[:n0], covariance, x[:n0])
    v = refine(graph.points, graph.neighbors, graph.offsets, covariance, init, x[n0:])
    gxr, ginit, gvr = refine_vjp(graph.points, graph.neighbors, covariance, v, x[n0:], g)
    gxs, gvd = dense_vjp(graph.points[:n0], covariance, x[:n0], ginit)
    gx = np.concatenate([gxs, gxr])
    if graph.indices is not None:
        out = np.empty_like(gx)
        out[graph.indices] = gx
        gx = out
    return gx, gvr + gvd


def _lookup_grad_all(bins, r, g):
    """Cotangent of interp(r, bins, vals) w.r.t. vals, summed over arrays r with weights g:
    (1 - f) to vals[i-1] and f to vals[i]; the clamps feed the end values."""
    nb = len(bins)
    gvals = np.zeros(nb)
    lo, hi = r < bins[0], r > bins[-1]
    gvals[0] += g[lo].sum()
    gvals[-1] += g[hi].sum()
    mid = ~(lo | hi)
    rm, gm = r[mid], g[mid]
    i = np.clip(np.searchsorted(bins, rm, side="right"), 1, nb - 1)
    f = (rm - bins[i - 1]) / (bins[i] - bins[i - 1])
    np.add.at(gvals, i - 1, gm * (1 - f))
    np.add.at(gvals, i, gm * f)
    return gvals
