"""torch.autograd.Function over the GPU log-likelihood.

The forward pass runs the batched pipeline once and gets LL and its gradient
together (the gradient is the adjoint kernel's contraction of
W = x x'/sigma2 - R^-1 with dR/dtheta recomputed from X, all components in one
pass); backward only scales that gradient by the incoming cotangent.  theta
may be a vector (one evaluation) or a (batch, d) matrix (one batched launch).
There is no counterpart in the reference, which has no torch surface.
"""
import numpy as np
import torch

from .kriging import Kriging


class KrigingLogLikelihood(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta, model):
        if not isinstance(model, Kriging):
            raise TypeError("model must be a libkriging_amd.Kriging")
        th = theta.detach().to("cpu", torch.float64).numpy()
        single = th.ndim == 1
        thb = th.reshape(1, -1) if single else th
        ll, grad = model.logLikelihoodFunBatch(thb, return_grad=True)
        ll_t = torch.from_numpy(np.ascontiguousarray(ll)).to(theta.device, torch.float64)
        g_t = torch.from_numpy(np.ascontiguousarray(grad)).to(theta.device, torch.float64)
        if single:
            ll_t = ll_t.reshape(())
            g_t = g_t.reshape(-1)
        ctx.save_for_backward(g_t)
        return ll_t

    @staticmethod
    def backward(ctx, g_out):
        (grad,) = ctx.saved_tensors
        if grad.ndim == 1:
            return g_out * grad, None
        return g_out.unsqueeze(-1) * grad, None


def log_likelihood(theta, model):
    """LL(theta) of a fitted-or-not Kriging model, differentiable in theta."""
    return KrigingLogLikelihood.apply(theta, model)
