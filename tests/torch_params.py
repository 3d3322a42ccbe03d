"""Test-only torch.autograd.Function over mm_set_site_params and
mm_param_gradients.

forward(theta, engine, sys, f_ref, site_type) sets the system's per-site
charges, sigma and epsilon from the per-type parameters theta (ntype x 3:
q, sigma, epsilon) through the C ABI, evaluates the force-matching loss
L = sum |F - f_ref|^2 with the adjoint kernels and keeps dL/dtheta for
backward.  The site types must stay distinct (their numbering is that of the
engine's first appearance, as site_type gives it).
"""
import numpy as np
import torch


class TypeParamLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta, engine, sys, f_ref, site_type):
        th = theta.detach().cpu().numpy().astype(np.float64)
        per_site = th[np.asarray(site_type)]
        engine.set_site_params(sys, per_site[:, 0], per_site[:, 1], per_site[:, 2])
        g = engine.param_gradients(sys, f_ref=f_ref)
        if g["ntype"] != th.shape[0] or not np.array_equal(g["site_type"], site_type):
            raise ValueError("theta merged or reordered the engine's site types")
        ctx.save_for_backward(torch.as_tensor(g["dL"], dtype=theta.dtype))
        return torch.tensor(g["loss"], dtype=theta.dtype)

    @staticmethod
    def backward(ctx, grad_out):
        (dL,) = ctx.saved_tensors
        return grad_out * dL, None, None, None, None
