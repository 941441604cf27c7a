"""torch.autograd.Function over the differentiable timestep.

``simulate(diff, phase, E, nsteps)`` runs ``nsteps`` of the differentiable
mode with per-voxel actuation phase (cycles) and Young's modulus (Pa) taken
from torch tensors and returns the final absolute voxel positions (V x 3,
float64).  Its backward calls the hand-written adjoint (vx3mi_diff_backward).
PyTorch is used only here; the simulation itself runs behind the C ABI.
"""
from __future__ import annotations

import numpy as np
import torch

from . import Diff


class _Simulate(torch.autograd.Function):
    @staticmethod
    def forward(ctx, phase: torch.Tensor, E: torch.Tensor, diff: Diff, nsteps: int):
        diff.set_params(phase.detach().to("cpu", torch.float64).numpy(), E.detach().to("cpu", torch.float64).numpy())
        diff.forward(nsteps)
        ctx.diff = diff
        ctx.devices = (phase.device, E.device)
        ctx.dtypes = (phase.dtype, E.dtype)
        return torch.from_numpy(diff.state()["pos"]).to(phase.device)

    @staticmethod
    def backward(ctx, grad_pos: torch.Tensor):
        g = np.ascontiguousarray(grad_pos.detach().to("cpu", torch.float64).numpy())
        gphase, gE = ctx.diff.backward(grad_pos=g)
        return (torch.from_numpy(gphase).to(ctx.devices[0], ctx.dtypes[0]),
                torch.from_numpy(gE).to(ctx.devices[1], ctx.dtypes[1]), None, None)


def simulate(diff: Diff, phase: torch.Tensor, E: torch.Tensor, nsteps: int) -> torch.Tensor:
    """Final absolute positions after ``nsteps``; differentiable in phase and E."""
    return _Simulate.apply(phase, E, diff, int(nsteps))


def com_x_displacement(diff: Diff, pos: torch.Tensor, masses: torch.Tensor, x0: torch.Tensor) -> torch.Tensor:
    """Mass-weighted centre-of-mass x displacement per robot from final positions."""
    out = []
    for r in range(diff.nr):
        a, b = int(diff.voxel_offsets[r]), int(diff.voxel_offsets[r + 1])
        m = masses[a:b]
        out.append((pos[a:b, 0] * m).sum() / m.sum() - x0[r])
    return torch.stack(out)
