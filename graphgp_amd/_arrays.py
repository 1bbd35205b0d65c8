"""Array plumbing: everything the kernels touch is a contiguous torch tensor on the GPU."""
from __future__ import annotations

import numpy as np
import torch

from . import _lib


def as_device(x, dtype=None):
    """Contiguous GPU tensor (float inputs keep float32/float64; ints become int32)."""
    dev = _lib.device()
    if isinstance(x, torch.Tensor):
        t = x
    else:
        t = torch.as_tensor(np.asarray(x))
    if dtype is not None:
        t = t.to(dtype)
    elif t.dtype in (torch.float16, torch.bfloat16):
        t = t.to(torch.float32)
    elif not t.is_floating_point():
        t = t.to(torch.float64)
    return t.to(dev).contiguous()


def as_points(x):
    t = as_device(x)
    if t.ndim != 2:
        raise ValueError(f"Points must have shape (N, d). Got shape {tuple(t.shape)}.")
    return t


def as_index(x):
    return as_device(x, torch.int32)


def as_covariance(covariance):
    """(cov_bins, cov_vals) as FP64 device tensors (refine.py:288-295 requires a pair)."""
    if not (isinstance(covariance, tuple) and len(covariance) == 2):
        raise ValueError("Invalid covariance specification.")
    bins, vals = covariance
    bins = as_device(bins, torch.float64)
    vals = as_device(vals, torch.float64)
    if bins.ndim != 1 or bins.shape != vals.shape or bins.numel() < 2:
        raise ValueError("Invalid covariance specification.")
    return bins, vals
