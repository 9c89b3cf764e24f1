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
    """The adjoint handle differentiates the forward pass it ran last. Each
    forward records the handle's generation; a backward whose forward is no
    longer the last one (two losses on one handle, or a forward call in
    between) first runs its forward again on the saved inputs."""

    @staticmethod
    def forward(ctx, rf, T1, T2, dB0, adj):
        loss = adj.forward(rf, T1, T2, dB0)
        ctx.adj = adj
        ctx.generation = adj.generation
        ctx.save_for_backward(rf, T1, T2, dB0)
        ctx.shapes = (rf.shape, T1.shape, T2.shape, dB0.shape)
        ctx.device = rf.device
        return torch.tensor(loss, dtype=torch.float64, device=rf.device)

    @staticmethod
    def backward(ctx, grad_out):
        dev = ctx.device
        adj = ctx.adj
        if adj.generation != ctx.generation:
            adj.forward(*ctx.saved_tensors)
        g = [torch.empty(s, dtype=torch.float32, device=dev) for s in ctx.shapes]
        torch.cuda.synchronize(dev)
        adj.backward(out=g)
        scale = grad_out.to(torch.float32)
        return g[0] * scale, g[1] * scale, g[2] * scale, g[3] * scale, None


def excitation_loss(adj, rf, T1, T2, dB0):
    """Differentiable rf-design loss (float64 scalar tensor)."""
    return BlochExcitationLoss.apply(rf, T1, T2, dB0, adj)
