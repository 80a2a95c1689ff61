"""torch.autograd.Function over the library's differentiable batched rollout.

`rollout(batch, state0, friction, restitution, frames)` returns the packed
state (pos 3N | quat 4N | vel 3N | omega 3N) after `frames` frames of every
scene in `batch`; its backward calls sv_batch_diff_backward, which
recomputes each frame from the checkpoint the forward saved and runs the
adjoint kernels (collision detection held fixed per substep).  Gradients
flow to the initial packed state and to the per-geom friction and
restitution; `per_material` sums per-geom gradients over each material's
geoms, since the library stores the coefficients per geom (as the
reference does, src/data.zig:474-475).

Tensors may live on the GPU or the CPU (values are staged through host
memory into the library's buffers).  One rollout per batch may be in flight: a
second forward overwrites the checkpoints the first backward needs.
"""
from __future__ import annotations

import torch

from .api import Batch


class Rollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, batch: Batch, frames: int, substeps: int, dt: float, state0: torch.Tensor,
                friction: torch.Tensor, restitution: torch.Tensor) -> torch.Tensor:
        n = batch.num_bodies
        s0 = state0.detach().to(torch.float32).contiguous()
        fr = friction.detach().to(torch.float32).contiguous()
        re = restitution.detach().to(torch.float32).contiguous()
        assert s0.numel() == 13 * n and fr.numel() == n and re.numel() == n
        batch.set_packed_state(s0.cpu().numpy())
        batch.ctx.upload("GEOM_FRICTION", fr.cpu().numpy())
        batch.ctx.upload("GEOM_RESTITUTION", re.cpu().numpy())
        batch.diff_forward(frames, substeps, dt)
        out = torch.from_numpy(batch.packed_state()).to(state0.device)
        ctx.batch, ctx.frames, ctx.substeps, ctx.dt = batch, frames, substeps, dt
        ctx.device = state0.device
        return out

    @staticmethod
    def backward(ctx, grad_out: torch.Tensor):
        b = ctx.batch
        g = grad_out.detach().to(torch.float32).contiguous()
        so, fo, ro = b.diff_backward(ctx.frames, g.cpu().numpy(), ctx.substeps, ctx.dt)
        dev = ctx.device
        return (None, None, None, None, torch.from_numpy(so).to(dev), torch.from_numpy(fo).to(dev),
                torch.from_numpy(ro).to(dev))


def rollout(batch: Batch, state0: torch.Tensor, friction: torch.Tensor, restitution: torch.Tensor,
            frames: int, substeps: int = 10, dt: float = 1.0 / 60.0) -> torch.Tensor:
    return Rollout.apply(batch, frames, substeps, dt, state0, friction, restitution)


def per_material(geom_grad: torch.Tensor, geom_material: torch.Tensor, num_materials: int) -> torch.Tensor:
    """Sum of per-geom gradients over each material's geoms."""
    out = torch.zeros(num_materials, dtype=geom_grad.dtype, device=geom_grad.device)
    return out.index_add_(0, geom_material.to(geom_grad.device).long(), geom_grad)
