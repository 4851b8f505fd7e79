"""torch.autograd binding of the MLS transfer: y = C(t) x.

Forward builds the operator for the given target coordinates (double
geometry, so the operator is differentiable in t) and applies it; backward
returns dL/dx = C^T ybar (the transpose apply) and dL/dt (the adjoint kernel
through the weights, the radius and the basis rows, neighbour sets frozen at
the forward's). The reference MovingLeastSquaresOperator has neither; the
forward matches its constructor + apply (DTK_MovingLeastSquaresOperator_def.hpp
:24-141).
"""
from __future__ import annotations

import torch

from .operators import MovingLeastSquaresOperator


class MLSTransfer(torch.autograd.Function):
    @staticmethod
    def forward(ctx, source_values, target_points, source_points, neighbors,
                options):
        opts = dict(options)
        opts["geometry"] = "double"
        op = MovingLeastSquaresOperator(source_points,
                                        target_points.detach().contiguous(),
                                        neighbors=neighbors, **opts)
        x = source_values.detach().contiguous()
        y = torch.empty((x.shape[0], target_points.shape[0]), dtype=x.dtype,
                        device=x.device)
        op.apply(x, y)
        ctx.op = op
        ctx.save_for_backward(x)
        return y

    @staticmethod
    def backward(ctx, ybar):
        (x,) = ctx.saved_tensors
        ybar = ybar.contiguous()
        gx = gt = None
        if ctx.needs_input_grad[0]:
            gx = torch.empty_like(x)
            ctx.op.apply_transpose(ybar, gx)
        if ctx.needs_input_grad[1]:
            gt = ctx.op.target_gradient(x, ybar)
        return gx, gt, None, None, None


def mls_transfer(source_points, target_points, source_values, *, neighbors=None,
                 **options):
    """Differentiable MLS transfer of source_values ([F, N_s] or [N_s],
    float64 on the GPU) to target_points ([N_t, 3] float64). Options are
    those of MovingLeastSquaresOperator (geometry is always double);
    neighbors ([N_t, k] int32, optional) freezes the neighbour sets."""
    single = source_values.dim() == 1
    x = source_values[None] if single else source_values
    y = MLSTransfer.apply(x, target_points, source_points, neighbors, options)
    return y[0] if single else y
