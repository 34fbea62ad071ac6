"""torch.autograd.Function over the C ABI: (eps_r, sigma) -> traces.

forward  = gjw_forward of every shot of the survey (checkpoints kept);
backward = gjw_gradient in EXACT mode with the incoming trace gradient as the
           adjoint source, i.e. the transpose of the discrete forward map
           (src/w/w_grad.m is the reference's counterpart).

The survey is mutable: a later ``set_model``, ``set_shots``, ``forward`` or
``fwi_gradient`` on it replaces the checkpoints a backward needs.  Every such
call bumps ``Survey.generation``; forward records it with a copy of its
model, and a backward that finds the survey moved on first restores that
model and re-runs the forward with checkpoints, so each graph node always
gets the gradient of its own model.

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
        e = np.array(eps.detach().cpu().numpy(), dtype=np.float64)
        s = np.array(sigma.detach().cpu().numpy(), dtype=np.float64)
        survey.set_model(e, s)
        dev = torch.device("cuda", survey.ctx.device)
        d = torch.empty(survey.traces_shape(), dtype=torch.float32, device=dev)
        torch.cuda.synchronize(dev)
        survey.forward(keep=True, out=d)
        ctx.survey = survey
        ctx.model = (e, s)
        ctx.generation = survey.generation
        ctx.shape = tuple(eps.shape)
        ctx.in_device = eps.device
        return d

    @staticmethod
    def backward(ctx, grad_d):
        sv = ctx.survey
        m, n = ctx.shape
        dev = grad_d.device
        if sv.generation != ctx.generation:
            # the survey ran another model since this node's forward: rebuild
            # this node's checkpoints before the adjoint
            sv.set_model(*ctx.model)
            scratch = torch.empty(sv.traces_shape(), dtype=torch.float32, device=dev)
            torch.cuda.synchronize(dev)
            sv.forward(keep=True, out=scratch)
            ctx.generation = sv.generation
        adj = grad_d.detach().to(torch.float32).contiguous()
        ge = torch.zeros((n, m), dtype=torch.float64, device=dev)  # column-major (m, n)
        gs = torch.zeros((n, m), dtype=torch.float64, device=dev)
        sv.gradient(adj=adj, g_eps=ge, g_sig=gs)
        return ge.t().to(ctx.in_device), gs.t().to(ctx.in_device), None


def traces(eps, sigma, survey):
    """Differentiable forward: (m, n) FP64 tensors -> (ntraces, nt) FP32 traces."""
    return TracesFn.apply(eps, sigma, survey)
