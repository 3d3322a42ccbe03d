"""Test-only torch.autograd.Function over mm_param_gradients.

forward(theta, engine, sys, f_ref) returns the force-matching loss at the
engine's current parameters (theta must equal them: the engine does not
take parameters from torch; this wrapper demonstrates the gradient path);
backward returns dL/dtheta from the device basis evaluation.
"""
import numpy as np
import torch


class TypeParamLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta, engine, sys, f_ref):
        g = engine.param_gradients(sys, f_ref=f_ref)
        if not np.allclose(theta.detach().cpu().numpy(), g["params"], rtol=1e-15, atol=0):
            raise ValueError("theta differs from the engine's parameters")
        ctx.save_for_backward(torch.as_tensor(g["dL"], dtype=theta.dtype))
        return torch.tensor(g["loss"], dtype=theta.dtype)

    @staticmethod
    def backward(ctx, grad_out):
        (dL,) = ctx.saved_tensors
        return grad_out * dL, None, None, None
