"""Generation, inverse and log-determinant (reference: graphgp/refine.py).

Every function keeps the reference's name, arguments and argument checks and runs on the
gfx950 kernels (``csrc/refine*.hip`` for the refined points, ``csrc/dense.hip`` for the
n0 x n0 seed block, always FP64).  ``cuda`` and ``fast_jit`` are accepted for drop-in
compatibility; there is one code path.

``precision`` selects the arithmetic of the local systems:
  * ``"fp64"``           -- FP64, the reference formulation;
  * ``"fp32"``           -- FP32, the reference formulation;
  * ``"fp32_increment"`` -- FP32 in increments relative to the nearest neighbour (keeps the
                            relative precision of the covariance entries when points cluster);
                            points whose FP32 pivots or Schur complement fall below 1e-3 of
                            their diagonal are recomputed in FP64 (generate: by a separate
                            pass over the list of such points, and if they are more than a
                            quarter of all points, by the FP64 refinement of every point;
                            inverse and log-det: inside the same kernel).
The default is ``"fp64"`` when the floating-point input (xi / values / points) is float64,
else ``"fp32"``.  Outputs have the compute dtype.
"""
from __future__ import annotations

from typing import Tuple

import torch

from . import _lib
from ._arrays import as_covariance, as_device, as_index, as_points
from .graph import Graph, _schedule

PRECISIONS = tuple(_lib.PREC)


def _precision(precision, *tensors):
    if precision is None:
        precision = "fp64" if any(t is not None and t.dtype == torch.float64 for t in tensors) else "fp32"
    if precision not in _lib.PREC:
        raise ValueError(f"precision must be one of {PRECISIONS}, got {precision!r}")
    return precision


def _out_dtype(precision):
    return torch.float64 if precision == "fp64" else torch.float32


def _counters(ws):
    """Counters of the refinement flow that used workspace ``ws`` (synchronises)."""
    import ctypes

    buf = (ctypes.c_uint32 * 4)()
    _lib.check(_lib.lib().ggp_refine_counters(_lib.ptr(ws), buf, _lib.stream()), "refine counters")
    # error word bits: 1 = a wait gave up (malformed graph), 2 = more FP32-increment points
    # needed the FP64 recomputation than its list holds (a quarter of the points), and the
    # generate returned the FP64 refinement's values instead
    return {"tiles": int(buf[0]), "nan": int(buf[1]), "malformed": int(buf[2] & 1),
            "fp64_list_overflow": int((buf[2] >> 1) & 1), "fp64_recomputed": int(buf[3])}


def generate(graph: Graph, covariance, xi, *, cuda: bool = False, fast_jit: bool = True, precision=None,
             return_counters: bool = False):
    """GP realisation: dense Cholesky of the seed block, then refinement (refine.py:21-56).

    ``xi`` (N,) in the caller's (original) order; returns values (N,) in that order.  A
    non-positive-definite local covariance makes every value NaN, as in the reference.
    Differentiable with respect to ``xi`` and the covariance values (``covariance[1]``)
    when either requires grad: the backward pass is the adjoint kernel (``generate_vjp``).
    """
    pts = as_points(graph.points)
    x = as_device(xi)
    n, d = pts.shape
    if x.numel() != n:
        raise ValueError("Length of xi must match number of points in graph.")
    nb = as_index(graph.neighbors)
    bins, vals = as_covariance(covariance)
    prec = _precision(precision, x, pts)
    perm = as_index(graph.indices) if graph.indices is not None else None
    if torch.is_grad_enabled() and (x.requires_grad or vals.requires_grad):
        if return_counters:
            raise ValueError("return_counters is not available on the differentiable path")
        from .autograd import GenerateFunction

        return GenerateFunction.apply(x, vals, pts, nb, perm, bins, prec)
    sched = _schedule(graph, pts, nb)
    return _generate_raw(pts, nb, perm, bins, vals, x, prec, return_counters, sched)


def _generate_raw(pts, nb, perm, bins, vals, x, prec, return_counters=False, sched=None):
    n, d = pts.shape
    n_ref, k = nb.shape
    n0 = n - n_ref
    out = torch.empty(n, dtype=_out_dtype(prec), device=pts.device)
    L = _lib.lib()
    ws = _lib.workspace(L.ggp_refine_workspace(n, d, n0, k, bins.numel(), _lib.PREC[prec], 1))
    args = (_lib.ptr(pts), _lib.dtype_code(pts), n, d, _lib.ptr(nb), n0, k, _lib.ptr(perm), _lib.ptr(bins),
            _lib.ptr(vals), bins.numel(), _lib.ptr(x), _lib.dtype_code(x), _lib.ptr(out), _lib.dtype_code(out),
            _lib.PREC[prec])
    if sched is not None:
        _lib.check(L.ggp_generate_sched(*args, _lib.ptr(sched[0]), _lib.ptr(sched[1]), _lib.ptr(sched[2]),
                                        _lib.ptr(ws), ws.numel(), _lib.stream()), "generate")
    else:
        _lib.check(L.ggp_generate(*args, _lib.ptr(ws), ws.numel(), _lib.stream()), "generate")
    return (out, _counters(ws)) if return_counters else out


def generate_vjp(graph: Graph, covariance, xi, values, values_bar, *, precision=None):
    """Reverse mode of ``generate``: what JAX autodiff of the reference's ``generate``
    (refine.py:21-56, 79-159) returns for the cotangent ``values_bar``.

    ``xi`` and ``values`` are the forward pass's input and output (caller order).  Returns
    ``(xi_bar, cov_vals_bar)``: xi_bar (N,) in the compute dtype, cov_vals_bar (n_bins,)
    FP64.  The local factors are recomputed, not stored.
    """
    pts = as_points(graph.points)
    x = as_device(xi)
    v = as_device(values)
    g = as_device(values_bar)
    n, d = pts.shape
    if x.numel() != n or v.numel() != n or g.numel() != n:
        raise ValueError("Length of xi, values and values_bar must match number of points in graph.")
    nb = as_index(graph.neighbors)
    bins, vals = as_covariance(covariance)
    prec = _precision(precision, x, pts)
    perm = as_index(graph.indices) if graph.indices is not None else None
    return _generate_vjp_raw(pts, nb, perm, bins, vals, x, v, g, prec)


def _generate_vjp_raw(pts, nb, perm, bins, vals, x, v, g, prec):
    n, d = pts.shape
    n_ref, k = nb.shape
    n0 = n - n_ref
    gxi = torch.empty(n, dtype=_out_dtype(prec), device=pts.device)
    gvals = torch.empty(bins.numel(), dtype=torch.float64, device=pts.device)
    L = _lib.lib()
    ws = _lib.workspace(L.ggp_generate_vjp_workspace(n, d, n0, k, bins.numel(), _lib.PREC[prec]))
    _lib.check(L.ggp_generate_vjp(_lib.ptr(pts), _lib.dtype_code(pts), n, d, _lib.ptr(nb), n0, k, _lib.ptr(perm),
                                  _lib.ptr(bins), _lib.ptr(vals), bins.numel(), _lib.ptr(x), _lib.dtype_code(x),
                                  _lib.ptr(v), _lib.dtype_code(v), _lib.ptr(g), _lib.dtype_code(g), _lib.ptr(gxi),
                                  _lib.dtype_code(gxi), _lib.ptr(gvals), _lib.PREC[prec], _lib.ptr(ws), ws.numel(),
                                  _lib.stream()), "generate_vjp")
    return gxi, gvals


def generate_inv_vjp(graph: Graph, covariance, values, xi_bar, *, precision=None):
    """Reverse mode of ``generate_inv``: what JAX autodiff of the reference's ``generate_inv``
    (refine.py:162-183, 186-195, 198-231) returns for the cotangent ``xi_bar``.

    ``values`` is the inverse's input (caller order).  Returns ``(values_bar, cov_vals_bar)``:
    values_bar (N,) in the compute dtype, cov_vals_bar (n_bins,) FP64.
    """
    pts = as_points(graph.points)
    v = as_device(values)
    t = as_device(xi_bar)
    n, d = pts.shape
    if v.numel() != n or t.numel() != n:
        raise ValueError("Length of values and xi_bar must match number of points in graph.")
    nb = as_index(graph.neighbors)
    bins, vals = as_covariance(covariance)
    prec = _precision(precision, v, pts)
    perm = as_index(graph.indices) if graph.indices is not None else None
    return _generate_inv_vjp_raw(pts, nb, perm, bins, vals, v, t, prec)


def _generate_inv_vjp_raw(pts, nb, perm, bins, vals, v, t, prec):
    n, d = pts.shape
    n_ref, k = nb.shape
    n0 = n - n_ref
    gv = torch.empty(n, dtype=_out_dtype(prec), device=pts.device)
    gtab = torch.empty(bins.numel(), dtype=torch.float64, device=pts.device)
    L = _lib.lib()
    ws = _lib.workspace(L.ggp_generate_vjp_workspace(n, d, n0, k, bins.numel(), _lib.PREC[prec]))
    _lib.check(L.ggp_generate_inv_vjp(_lib.ptr(pts), _lib.dtype_code(pts), n, d, _lib.ptr(nb), n0, k, _lib.ptr(perm),
                                      _lib.ptr(bins), _lib.ptr(vals), bins.numel(), _lib.ptr(v), _lib.dtype_code(v),
                                      _lib.ptr(t), _lib.dtype_code(t), _lib.ptr(gv), _lib.dtype_code(gv),
                                      _lib.ptr(gtab), _lib.PREC[prec], _lib.ptr(ws), ws.numel(), _lib.stream()),
               "generate_inv_vjp")
    return gv, gtab


def generate_logdet_vjp(graph: Graph, covariance, logdet_bar=1.0, *, precision=None):
    """Reverse mode of ``generate_logdet``: the cotangent of the covariance table values
    (n_bins,) FP64 for the log-det's cotangent ``logdet_bar`` (refine.py:234-275)."""
    pts = as_points(graph.points)
    nb = as_index(graph.neighbors)
    bins, vals = as_covariance(covariance)
    prec = _precision(precision, pts)
    lb = torch.as_tensor(logdet_bar, dtype=torch.float64, device=pts.device).reshape(())
    return _generate_logdet_vjp_raw(pts, nb, bins, vals, lb, prec)


def _generate_logdet_vjp_raw(pts, nb, bins, vals, lb, prec):
    n, d = pts.shape
    n_ref, k = nb.shape
    n0 = n - n_ref
    gtab = torch.empty(bins.numel(), dtype=torch.float64, device=pts.device)
    lb = lb.to(torch.float64).contiguous()
    L = _lib.lib()
    ws = _lib.workspace(L.ggp_generate_vjp_workspace(n, d, n0, k, bins.numel(), _lib.PREC[prec]))
    _lib.check(L.ggp_generate_logdet_vjp(_lib.ptr(pts), _lib.dtype_code(pts), n, d, _lib.ptr(nb), n0, k,
                                         _lib.ptr(bins), _lib.ptr(vals), bins.numel(), _lib.ptr(lb), _lib.ptr(gtab),
                                         _lib.PREC[prec], _lib.ptr(ws), ws.numel(), _lib.stream()),
               "generate_logdet_vjp")
    return gtab


def generate_inv(graph: Graph, covariance, values, *, cuda: bool = False, precision=None,
                 return_counters: bool = False):
    """Inverse of ``generate`` (refine.py:162-183): values (N,) -> xi (N,), caller order.

    Differentiable with respect to ``values`` and the covariance values when either
    requires grad (backward: ``generate_inv_vjp``)."""
    pts = as_points(graph.points)
    v = as_device(values)
    n, d = pts.shape
    if v.numel() != n:
        raise ValueError("Length of values must match number of points in graph.")
    nb = as_index(graph.neighbors)
    n_ref, k = nb.shape
    n0 = n - n_ref
    bins, vals = as_covariance(covariance)
    prec = _precision(precision, v, pts)
    perm = as_index(graph.indices) if graph.indices is not None else None
    if torch.is_grad_enabled() and (v.requires_grad or vals.requires_grad):
        if return_counters:
            raise ValueError("return_counters is not available on the differentiable path")
        from .autograd import GenerateInvFunction

        return GenerateInvFunction.apply(v, vals, pts, nb, perm, bins, prec)
    return _generate_inv_raw(pts, nb, perm, bins, vals, v, prec, return_counters)


def _generate_inv_raw(pts, nb, perm, bins, vals, v, prec, return_counters=False):
    n, d = pts.shape
    n_ref, k = nb.shape
    n0 = n - n_ref
    out = torch.empty(n, dtype=_out_dtype(prec), device=pts.device)
    L = _lib.lib()
    ws = _lib.workspace(L.ggp_refine_workspace(n, d, n0, k, bins.numel(), _lib.PREC[prec], 1))
    _lib.check(L.ggp_generate_inv(_lib.ptr(pts), _lib.dtype_code(pts), n, d, _lib.ptr(nb), n0, k, _lib.ptr(perm),
                                  _lib.ptr(bins), _lib.ptr(vals), bins.numel(), _lib.ptr(v), _lib.dtype_code(v),
                                  _lib.ptr(out), _lib.dtype_code(out), _lib.PREC[prec], _lib.ptr(ws), ws.numel(),
                                  _lib.stream()), "generate_inv")
    return (out, _counters(ws)) if return_counters else out


def generate_logdet(graph: Graph, covariance, *, cuda: bool = False, precision=None):
    """Log-determinant of ``generate`` (refine.py:234-240), summed in FP64; a 0-d tensor.

    Differentiable with respect to the covariance values when they require grad
    (backward: ``generate_logdet_vjp``)."""
    pts = as_points(graph.points)
    nb = as_index(graph.neighbors)
    bins, vals = as_covariance(covariance)
    prec = _precision(precision, pts)
    if torch.is_grad_enabled() and vals.requires_grad:
        from .autograd import GenerateLogdetFunction

        return GenerateLogdetFunction.apply(vals, pts, nb, bins, prec)
    return _generate_logdet_raw(pts, nb, bins, vals, prec)


def _generate_logdet_raw(pts, nb, bins, vals, prec):
    n, d = pts.shape
    n_ref, k = nb.shape
    n0 = n - n_ref
    out = torch.empty((), dtype=torch.float64, device=pts.device)
    L = _lib.lib()
    ws = _lib.workspace(L.ggp_refine_workspace(n, d, n0, k, bins.numel(), _lib.PREC[prec], 1))
    _lib.check(L.ggp_generate_logdet(_lib.ptr(pts), _lib.dtype_code(pts), n, d, _lib.ptr(nb), n0, k, _lib.ptr(bins),
                                     _lib.ptr(vals), bins.numel(), _lib.ptr(out), _lib.PREC[prec], _lib.ptr(ws),
                                     ws.numel(), _lib.stream()), "generate_logdet")
    return out


def generate_dense(points, covariance, xi):
    """values = chol(K) @ xi in FP64 (refine.py:59-76)."""
    p = as_points(points)
    x = as_device(xi)
    n, d = p.shape
    if x.numel() != n:
        raise ValueError("Length of xi must match number of points.")
    bins, vals = as_covariance(covariance)
    out = torch.empty(n, dtype=torch.float64, device=p.device)
    L = _lib.lib()
    ws = _lib.workspace(L.ggp_dense_workspace(n))
    _lib.check(L.ggp_generate_dense(_lib.ptr(p), _lib.dtype_code(p), n, d, _lib.ptr(bins), _lib.ptr(vals),
                                    bins.numel(), _lib.ptr(x), _lib.dtype_code(x), _lib.ptr(out), _lib.F64,
                                    _lib.ptr(ws), ws.numel(), _lib.stream()), "generate_dense")
    return out


def generate_dense_inv(points, covariance, values):
    """Inverse of ``generate_dense`` (refine.py:186-195)."""
    p = as_points(points)
    v = as_device(values)
    n, d = p.shape
    if v.numel() != n:
        raise ValueError("Length of values must match number of points.")
    bins, vals = as_covariance(covariance)
    out = torch.empty(n, dtype=torch.float64, device=p.device)
    L = _lib.lib()
    ws = _lib.workspace(L.ggp_dense_workspace(n))
    _lib.check(L.ggp_generate_dense_inv(_lib.ptr(p), _lib.dtype_code(p), n, d, _lib.ptr(bins), _lib.ptr(vals),
                                        bins.numel(), _lib.ptr(v), _lib.dtype_code(v), _lib.ptr(out), _lib.F64,
                                        _lib.ptr(ws), ws.numel(), _lib.stream()), "generate_dense_inv")
    return out


def generate_dense_logdet(points, covariance):
    """slogdet(K)[1] / 2 (refine.py:243-248), via the FP64 Cholesky."""
    p = as_points(points)
    n, d = p.shape
    bins, vals = as_covariance(covariance)
    out = torch.empty((), dtype=torch.float64, device=p.device)
    L = _lib.lib()
    ws = _lib.workspace(L.ggp_dense_workspace(n))
    _lib.check(L.ggp_generate_dense_logdet(_lib.ptr(p), _lib.dtype_code(p), n, d, _lib.ptr(bins), _lib.ptr(vals),
                                           bins.numel(), _lib.ptr(out), _lib.ptr(ws), ws.numel(), _lib.stream()),
               "generate_dense_logdet")
    return out


def _raw_args(points, neighbors, covariance):
    pts = as_points(points)
    nb = as_index(neighbors)
    n, d = pts.shape
    n_ref, k = nb.shape
    bins, vals = as_covariance(covariance)
    return pts, nb, n, d, n - n_ref, k, bins, vals


def refine(points, neighbors, offsets: Tuple[int, ...], covariance, initial_values, xi, *, cuda: bool = False,
           fast_jit: bool = True, precision=None):
    """Conditional generation from given seed values (refine.py:79-159); graph order, no permutation.

    ``offsets`` is accepted for API parity: the kernel follows each point's own
    neighbours, which for a valid graph (``check_graph``) gives the batch result.
    """
    pts, nb, n, d, n0, k, bins, vals = _raw_args(points, neighbors, covariance)
    iv = as_device(initial_values)
    x = as_device(xi)
    if iv.numel() != n0:
        raise ValueError("Length of initial_values must match number of initial points.")
    if x.numel() != n - n0:
        raise ValueError("Length of xi must match number of refined points.")
    prec = _precision(precision, x, iv, pts)
    out = torch.empty(n, dtype=_out_dtype(prec), device=pts.device)
    L = _lib.lib()
    ws = _lib.workspace(L.ggp_refine_workspace(n, d, n0, k, bins.numel(), _lib.PREC[prec], 0))
    _lib.check(L.ggp_refine(_lib.ptr(pts), _lib.dtype_code(pts), n, d, _lib.ptr(nb), n0, k, _lib.ptr(bins),
                            _lib.ptr(vals), bins.numel(), _lib.ptr(iv), _lib.dtype_code(iv), _lib.ptr(x),
                            _lib.dtype_code(x), _lib.ptr(out), _lib.PREC[prec], _lib.ptr(ws), ws.numel(),
                            _lib.stream()), "refine")
    return out


def refine_inv(points, neighbors, offsets, covariance, values, *, cuda: bool = False, precision=None):
    """Inverse of ``refine`` (refine.py:198-231): returns (initial_values, xi)."""
    pts, nb, n, d, n0, k, bins, vals = _raw_args(points, neighbors, covariance)
    v = as_device(values)
    if v.numel() != n:
        raise ValueError("Length of values must match number of points.")
    prec = _precision(precision, v, pts)
    init = torch.empty(n0, dtype=torch.float64, device=pts.device)
    xi = torch.empty(n - n0, dtype=_out_dtype(prec), device=pts.device)
    L = _lib.lib()
    ws = _lib.workspace(L.ggp_refine_workspace(n, d, n0, k, bins.numel(), _lib.PREC[prec], 0))
    _lib.check(L.ggp_refine_inv(_lib.ptr(pts), _lib.dtype_code(pts), n, d, _lib.ptr(nb), n0, k, _lib.ptr(bins),
                                _lib.ptr(vals), bins.numel(), _lib.ptr(v), _lib.dtype_code(v), _lib.ptr(init),
                                _lib.ptr(xi), _lib.dtype_code(xi), _lib.PREC[prec], _lib.ptr(ws), ws.numel(),
                                _lib.stream()), "refine_inv")
    return init.to(_out_dtype(prec)), xi


def refine_logdet(points, neighbors, offsets, covariance, *, cuda: bool = False, precision=None):
    """Sum of log conditional standard deviations (refine.py:251-275), FP64 sum."""
    pts, nb, n, d, n0, k, bins, vals = _raw_args(points, neighbors, covariance)
    prec = _precision(precision, pts)
    out = torch.empty((), dtype=torch.float64, device=pts.device)
    L = _lib.lib()
    ws = _lib.workspace(L.ggp_refine_workspace(n, d, n0, k, bins.numel(), _lib.PREC[prec], 0))
    _lib.check(L.ggp_refine_logdet(_lib.ptr(pts), _lib.dtype_code(pts), n, d, _lib.ptr(nb), n0, k, _lib.ptr(bins),
                                   _lib.ptr(vals), bins.numel(), _lib.ptr(out), _lib.PREC[prec], _lib.ptr(ws),
                                   ws.numel(), _lib.stream()), "refine_logdet")
    return out


def compute_cov_matrix(covariance, points_a, points_b):
    """Covariance between two point sets by table interpolation (refine.py:288-295), FP64."""
    bins, vals = as_covariance(covariance)
    a = as_points(points_a)
    b = as_points(points_b)
    if a.dtype != b.dtype:
        b = b.to(a.dtype)
    out = torch.empty((a.shape[0], b.shape[0]), dtype=torch.float64, device=a.device)
    _lib.check(_lib.lib().ggp_cov_matrix(_lib.ptr(a), a.shape[0], _lib.ptr(b), b.shape[0], a.shape[1],
                                         _lib.dtype_code(a), _lib.ptr(bins), _lib.ptr(vals), bins.numel(),
                                         _lib.ptr(out), _lib.stream()), "compute_cov_matrix")
    return out
