"""torch.autograd.Function over the C ABI: (eps_r, sigma) -> traces.

forward  = gjw_forward of every shot of the survey (checkpoints kept);
backward = gjw_gradient in EXACT mode with the incoming trace gradient as the
           adjoint source, i.e. the transpose of the discrete forward map
           (src/w/w_grad.m is the reference's counterpart).

Example::

    sv = Survey(grid); sv.set_shots(src, wavelets, receivers)
    eps = torch.tensor(eps0, dtype=torch.float64, requires_grad=True)
    sig = torch.tensor(sig0, dtype=torch.float64, requires_grad=True)
    d = TracesFn.apply(eps, sig, sv)          # (ntraces, nt) float32 on cuda
    loss = 0.5 * ((d - d_obs) ** 2).sum()
    loss.backward()                            # eps.grad, sig.grad: FP64 (m, n)
"""
from __future__ import annotations

import numpy as np
import torch


class TracesFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, eps, sigma, survey):
        survey.set_model(eps.detach().cpu().numpy(), sigma.detach().cpu().numpy())
        dev = torch.device("cuda", survey.ctx.device)
        d = torch.empty(survey.traces_shape(), dtype=torch.float32, device=dev)
        torch.cuda.synchronize(dev)
        survey.forward(keep=True, out=d)
        ctx.survey = survey
        ctx.shape = tuple(eps.shape)
        ctx.in_device = eps.device
        return d

    @staticmethod
    def backward(ctx, grad_d):
        sv = ctx.survey
        m, n = ctx.shape
        dev = grad_d.device
        adj = grad_d.detach().to(torch.float32).contiguous()
        ge = torch.zeros((n, m), dtype=torch.float64, device=dev)  # column-major (m, n)
        gs = torch.zeros((n, m), dtype=torch.float64, device=dev)
        sv.gradient(adj=adj, g_eps=ge, g_sig=gs)
        return ge.t().to(ctx.in_device), gs.t().to(ctx.in_device), None


def traces(eps, sigma, survey):
    """Differentiable forward: (m, n) FP64 tensors -> (ntraces, nt) FP32 traces."""
    return TracesFn.apply(eps, sigma, survey)
