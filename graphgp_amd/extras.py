"""Discretised covariance tables (reference: graphgp/extras.py:14-68).

Tables are ``(cov_bins, cov_vals)`` FP64 tensors on the device: r = 0 followed by
``n_bins - 1`` log-spaced radii, with the r = 0 value inflated by ``(1 + jitter)``.

The values are computed with torch operations on the device, so the kernel
hyperparameters may be tensors that require gradients: ``generate``'s gradient with
respect to the table values (the adjoint kernel) then reaches ``variance`` and ``scale``
(or ``cutoff``) by the chain rule, as a JAX caller gets it through ``custom_vjp`` of the
refinement and autodiff of the table formula.  Plain floats give the same tables.
"""
from __future__ import annotations

import math

import numpy as np
import torch


def _device():
    return torch.device("cuda") if torch.cuda.is_available() else torch.device("cpu")


def _param(x, dev):
    if isinstance(x, torch.Tensor):
        return x.to(device=dev, dtype=torch.float64)
    return torch.tensor(float(x), dtype=torch.float64, device=dev)


def make_cov_bins(*, r_min: float, r_max: float, n_bins: int):
    """[0, logspace(log10 r_min, log10 r_max, n_bins - 1)] (extras.py:65-68)."""
    r = np.concatenate([[0.0], np.logspace(np.log10(r_min), np.log10(r_max), n_bins - 1)])
    return torch.as_tensor(r, dtype=torch.float64, device=_device())


def _with_jitter(r, cov, jitter: float):
    # vals[0] *= (1 + jitter): the r = 0 entry (extras.py:30, 57)
    return torch.where(r == 0.0, cov * (1.0 + jitter), cov)


def rbf_kernel(*, variance, scale, r_min: float, r_max: float, n_bins: int, jitter: float = 0.0):
    """Squared exponential variance * exp(-(r / scale)^2 / 2) (extras.py:14-31).  ``variance``
    and ``scale`` may be tensors that require gradients."""
    r = make_cov_bins(r_min=r_min, r_max=r_max, n_bins=n_bins)
    v, s = _param(variance, r.device), _param(scale, r.device)
    cov = v * torch.exp(-0.5 * (r / s) ** 2)
    return r, _with_jitter(r, cov, jitter)


def matern_kernel(*, p: int, variance, cutoff, r_min: float, r_max: float, n_bins: int, jitter: float = 0.0):
    """Matern covariance with nu = p + 1/2 (extras.py:34-58).

    x = sqrt(2p + 1) r / cutoff; cov = variance * exp(-x) * sum_i c_i (2x)^(p - i) with
    c_i = p! (p + i)! / (i! (p - i)! (2p)!).  ``variance`` and ``cutoff`` may be tensors that
    require gradients.
    """
    r = make_cov_bins(r_min=r_min, r_max=r_max, n_bins=n_bins)
    v, c = _param(variance, r.device), _param(cutoff, r.device)
    x = math.sqrt(2 * p + 1) * r / c
    lf = lambda t: math.lgamma(t + 1)  # noqa: E731
    coeff = [math.exp(lf(p) + lf(p + i) - lf(i) - lf(p - i) - lf(2 * p)) for i in range(p + 1)]
    poly = torch.zeros_like(x)
    for ci in coeff:
        poly = poly * (2 * x) + ci
    cov = v * torch.exp(-x) * poly
    return r, _with_jitter(r, cov, jitter)
