"""torch.autograd.Function over the library's porous-drag adjoint
(simlab_cfd_adjoint_forward / _backward), for gradient checks against
central finite differences. The lattice does the work; torch only carries
the solid-fraction tensor and its gradient."""
from __future__ import annotations

import numpy as np
import torch


class PorousDrag(torch.autograd.Function):
    """J(phi) = mean drag of the porous solid over steps [avg_from, nsteps)."""

    @staticmethod
    def forward(ctx, phi, lat, nsteps, ckpt_every, avg_from, axis):
        lat.set_solid_fraction(phi.detach().cpu().numpy().astype(np.float32))
        j, _ = lat.adjoint_forward(nsteps, ckpt_every, avg_from, axis)
        ctx.save_for_backward(phi.detach().clone())
        ctx.args = (lat, nsteps, ckpt_every, avg_from, axis)
        return torch.tensor(j, dtype=torch.float64)

    @staticmethod
    def backward(ctx, grad_j):
        (phi,) = ctx.saved_tensors
        lat, nsteps, ckpt_every, avg_from, axis = ctx.args
        # the lattice keeps the checkpoints of its last forward only: rerun it
        lat.set_solid_fraction(phi.cpu().numpy().astype(np.float32))
        lat.adjoint_forward(nsteps, ckpt_every, avg_from, axis)
        g = torch.from_numpy(lat.adjoint_backward()).reshape(phi.shape)
        return (grad_j * g).to(phi.dtype), None, None, None, None, None


def porous_drag(phi: torch.Tensor, lat, nsteps: int, ckpt_every: int = 100, avg_from: int = 0, axis: int = 0):
    return PorousDrag.apply(phi, lat, nsteps, ckpt_every, avg_from, axis)
