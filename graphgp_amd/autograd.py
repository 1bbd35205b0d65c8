"""``torch.autograd`` binding of the adjoint kernel.

The reference makes ``generate`` differentiable through JAX autodiff (refine.py:21-159;
README.md:36 notes the custom primitives of its CUDA extension).  Here the forward pass is
the dataflow refinement kernel and the backward pass is ``generate_vjp`` (one reverse
dataflow launch that recomputes every local Cholesky factor, plus the FP64 seed-block
adjoint), so ``torch.autograd.grad`` / ``.backward()`` of any loss of ``generate``'s values
gives the gradient with respect to xi and the covariance table values.
"""
from __future__ import annotations

import torch


class GenerateFunction(torch.autograd.Function):
    """values = generate(graph, (bins, cov_vals), xi); differentiable in xi and cov_vals."""

    @staticmethod
    def forward(ctx, xi, cov_vals, pts, nb, perm, bins, precision):
        from .refine import _generate_raw

        values = _generate_raw(pts, nb, perm, bins, cov_vals.detach(), xi.detach(), precision)
        ctx.precision = precision
        ctx.perm = perm
        ctx.save_for_backward(xi, values, cov_vals, pts, nb, bins)
        return values

    @staticmethod
    def backward(ctx, values_bar):
        from .refine import _generate_vjp_raw

        xi, values, cov_vals, pts, nb, bins = ctx.saved_tensors
        gxi, gvals = _generate_vjp_raw(pts, nb, ctx.perm, bins, cov_vals.detach(), xi.detach(), values,
                                       values_bar.contiguous(), ctx.precision)
        gxi = gxi.to(xi.dtype) if ctx.needs_input_grad[0] else None
        gvals = gvals.to(cov_vals.dtype) if ctx.needs_input_grad[1] else None
        return gxi, gvals, None, None, None, None, None
