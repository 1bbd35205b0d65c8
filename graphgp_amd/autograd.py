"""``torch.autograd`` binding of the adjoint kernels.

The reference makes ``generate``, ``generate_inv`` and ``generate_logdet`` differentiable
through JAX autodiff (refine.py:21-275; README.md:36 notes the custom primitives of its
CUDA extension).  Here each forward pass is the refinement kernel and each backward pass
its adjoint (``generate_vjp``: one reverse dataflow launch; ``generate_inv_vjp`` and
``generate_logdet_vjp``: one launch without a chain; all recompute every local Cholesky
factor, plus the FP64 seed-block adjoint), so ``torch.autograd.grad`` / ``.backward()`` of
any loss gives the gradient with respect to xi / the values and the covariance table values
(and, through ``extras``, the kernel hyperparameters).
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


class GenerateInvFunction(torch.autograd.Function):
    """xi = generate_inv(graph, (bins, cov_vals), values); differentiable in values and cov_vals."""

    @staticmethod
    def forward(ctx, values, cov_vals, pts, nb, perm, bins, precision):
        from .refine import _generate_inv_raw

        xi = _generate_inv_raw(pts, nb, perm, bins, cov_vals.detach(), values.detach(), precision)
        ctx.precision = precision
        ctx.perm = perm
        ctx.save_for_backward(values, cov_vals, pts, nb, bins)
        return xi

    @staticmethod
    def backward(ctx, xi_bar):
        from .refine import _generate_inv_vjp_raw

        values, cov_vals, pts, nb, bins = ctx.saved_tensors
        gv, gvals = _generate_inv_vjp_raw(pts, nb, ctx.perm, bins, cov_vals.detach(), values.detach(),
                                          xi_bar.contiguous(), ctx.precision)
        gv = gv.to(values.dtype) if ctx.needs_input_grad[0] else None
        gvals = gvals.to(cov_vals.dtype) if ctx.needs_input_grad[1] else None
        return gv, gvals, None, None, None, None, None


class GenerateLogdetFunction(torch.autograd.Function):
    """logdet = generate_logdet(graph, (bins, cov_vals)); differentiable in cov_vals."""

    @staticmethod
    def forward(ctx, cov_vals, pts, nb, bins, precision):
        from .refine import _generate_logdet_raw

        ld = _generate_logdet_raw(pts, nb, bins, cov_vals.detach(), precision)
        ctx.precision = precision
        ctx.save_for_backward(cov_vals, pts, nb, bins)
        return ld

    @staticmethod
    def backward(ctx, ld_bar):
        from .refine import _generate_logdet_vjp_raw

        cov_vals, pts, nb, bins = ctx.saved_tensors
        gvals = _generate_logdet_vjp_raw(pts, nb, bins, cov_vals.detach(), ld_bar.detach(), ctx.precision)
        return gvals.to(cov_vals.dtype), None, None, None, None
