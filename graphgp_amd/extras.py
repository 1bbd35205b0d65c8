"""Discretised covariance tables (reference: graphgp/extras.py:14-68).

Tables are computed on the host in FP64 and returned as FP64 device tensors
``(cov_bins, cov_vals)``: r = 0 followed by ``n_bins - 1`` log-spaced radii, with the r = 0
value inflated by ``(1 + jitter)``.
"""
from __future__ import annotations

import math

import numpy as np
import torch


def _dev(a):
    if torch.cuda.is_available():
        return torch.as_tensor(a, dtype=torch.float64, device="cuda")
    return torch.as_tensor(a, dtype=torch.float64)


def make_cov_bins(*, r_min: float, r_max: float, n_bins: int):
    """[0, logspace(log10 r_min, log10 r_max, n_bins - 1)] (extras.py:65-68)."""
    r = np.concatenate([[0.0], np.logspace(np.log10(r_min), np.log10(r_max), n_bins - 1)])
    return _dev(r)


def _table(r, cov, jitter):
    cov = np.where(r == 0.0, cov[0] * (1.0 + jitter), cov)
    return _dev(r), _dev(cov)


def rbf_kernel(*, variance: float, scale: float, r_min: float, r_max: float, n_bins: int, jitter: float = 0.0):
    """Squared exponential variance * exp(-(r / scale)^2 / 2) (extras.py:14-31)."""
    r = make_cov_bins(r_min=r_min, r_max=r_max, n_bins=n_bins).cpu().numpy()
    cov = variance * np.exp(-0.5 * (r / scale) ** 2)
    return _table(r, cov, jitter)


def matern_kernel(*, p: int, variance: float, cutoff: float, r_min: float, r_max: float, n_bins: int,
                  jitter: float = 0.0):
    """Matern covariance with nu = p + 1/2 (extras.py:34-58).

    x = sqrt(2p + 1) r / cutoff; cov = variance * exp(-x) * sum_i c_i (2x)^(p - i) with
    c_i = p! (p + i)! / (i! (p - i)! (2p)!).
    """
    r = make_cov_bins(r_min=r_min, r_max=r_max, n_bins=n_bins).cpu().numpy()
    x = math.sqrt(2 * p + 1) * r / cutoff
    lf = lambda v: math.lgamma(v + 1)  # noqa: E731
    coeff = [math.exp(lf(p) + lf(p + i) - lf(i) - lf(p - i) - lf(2 * p)) for i in range(p + 1)]
    poly = np.zeros_like(x)
    for c in coeff:
        poly = poly * (2 * x) + c
    cov = variance * np.exp(-x) * poly
    return _table(r, cov, jitter)
