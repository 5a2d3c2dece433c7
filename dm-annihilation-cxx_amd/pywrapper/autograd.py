"""torch.autograd.Function around the projector and its adjoint kernel.

    proj = Projector(params, max_parts)           # one map (nside, observer)
    m = project_map(parts, proj)                  # parts: (n, 12) float32 CUDA tensor
    loss = ((m - target) ** 2 / target).sum()
    loss.backward()                               # parts.grad[:, :6]: d/d(mass, dens, h, x, y, z)

The forward renders with the projector's precision; the backward runs the
FP64 adjoint kernel (dmann_adjoint_device), which recomputes each particle's
footprint and gathers dL/dmap over it.  Footprint membership is piecewise
constant, so its jumps are not part of the gradient.
"""
from __future__ import annotations

import torch


class ProjectMap(torch.autograd.Function):
    @staticmethod
    def forward(ctx, parts: torch.Tensor, projector):
        if not (parts.is_cuda and parts.dtype == torch.float32 and parts.dim() == 2 and parts.shape[1] == 12):
            raise ValueError("parts must be an (n, 12) float32 CUDA tensor")
        parts_c = parts.detach().contiguous()
        out = torch.empty(projector.npix, dtype=torch.float64, device=parts.device)
        projector.reset()
        projector.render(parts_c)
        projector.get_map_device(out)
        ctx.projector = projector
        ctx.save_for_backward(parts_c)
        return out

    @staticmethod
    def backward(ctx, grad_map):
        (parts,) = ctx.saved_tensors
        g6 = ctx.projector.adjoint(parts, grad_map.to(torch.float64).contiguous())
        grad = torch.zeros_like(parts)
        grad[:, :6] = g6.to(parts.dtype)
        return grad, None


def project_map(parts: torch.Tensor, projector) -> torch.Tensor:
    """FP64 RING map (divided by dOmega) of `parts`, differentiable."""
    return ProjectMap.apply(parts, projector)
