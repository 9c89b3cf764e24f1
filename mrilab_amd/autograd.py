"""torch.autograd entry point of the adjoint Bloch kernels.

``excitation_loss(adj, rf, T1, T2, dB0)`` evaluates the rf-design loss of an
engine.Adjoint problem on float32 CUDA tensors and differentiates it with the
hand-written backward-in-time kernel (mrs_adjoint_backward), not with
autograd's tape: the forward pass keeps a checkpoint every K steps on the
device and the backward pass recomputes in between.
"""
from __future__ import annotations

import torch


class BlochExcitationLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, rf, T1, T2, dB0, adj):
        loss = adj.forward(rf, T1, T2, dB0)
        ctx.adj = adj
        ctx.shapes = (rf.shape, T1.shape, T2.shape, dB0.shape)
        ctx.device = rf.device
        return torch.tensor(loss, dtype=torch.float64, device=rf.device)

    @staticmethod
    def backward(ctx, grad_out):
        dev = ctx.device
        g = [torch.empty(s, dtype=torch.float32, device=dev) for s in ctx.shapes]
        torch.cuda.synchronize(dev)
        ctx.adj.backward(out=g)
        scale = grad_out.to(torch.float32)
        return g[0] * scale, g[1] * scale, g[2] * scale, g[3] * scale, None


def excitation_loss(adj, rf, T1, T2, dB0):
    """Differentiable rf-design loss (float64 scalar tensor)."""
    return BlochExcitationLoss.apply(rf, T1, T2, dB0, adj)
