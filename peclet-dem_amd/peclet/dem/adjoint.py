"""Differentiable rollouts of the XPBD substep (an extension: the reference has no adjoint).

``rollout`` runs ``substeps`` substeps of a prepared ``Simulation`` while recording them, and
returns the end state as a torch tensor; its backward pass calls the hand-written adjoint of the
recorded substeps (``Simulation.tape_adjoint``), which replays checkpointed segments and walks
every substep backwards — position sweeps, re-integration, velocity sweeps in reverse colour
order with the friction cone's active set as the forward pass left it, contacts, ghosts,
predict.  Gradients flow to the initial velocities and angular velocities, the initial
positions, the friction coefficient and the normal restitution.

The state of a body is 13 numbers: position (3), velocity (3), angular velocity (3),
orientation quaternion (x, y, z, w).
"""
from __future__ import annotations

import numpy as np
import torch

__all__ = ["rollout", "bed_height", "packing_fraction_below", "STATE_COLUMNS"]

STATE_COLUMNS = {"position": slice(0, 3), "velocity": slice(3, 6),
                 "angular_velocity": slice(6, 9), "quaternion": slice(9, 13)}


class _Rollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vel0, angvel0, friction, restitution, sim, dt, substeps, every, rest_t):
        sim.set_velocities(vel0.detach().cpu().double().numpy())
        sim.set_angular_velocities(angvel0.detach().cpu().double().numpy())
        sim.set_material_params(float(restitution), float(rest_t), float(friction))
        sim.tape_begin(every)
        try:
            sim.step_n(dt, substeps)
        finally:
            sim.tape_end()
        end = np.concatenate([sim.get_positions(), sim.get_velocities(),
                              sim.get_angular_velocities(), sim.get_quaternions()], axis=1)
        ctx.sim = sim
        ctx.generation = sim.tape_generation()
        ctx.device = vel0.device
        return torch.from_numpy(end).to(vel0.device)

    @staticmethod
    def backward(ctx, grad_end):
        if ctx.sim.tape_generation() != ctx.generation:
            raise RuntimeError("rollout backward: the simulation has recorded another rollout "
                               "since this one (its recording was replaced); run backward before "
                               "the next rollout or tape_begin on the same simulation")
        g = grad_end.detach().cpu().double().numpy()
        g0, gp = ctx.sim.tape_adjoint(g)
        dev = ctx.device
        return (torch.from_numpy(np.ascontiguousarray(g0[:, 3:6])).to(dev),
                torch.from_numpy(np.ascontiguousarray(g0[:, 6:9])).to(dev),
                torch.tensor(gp[0], dtype=torch.float64, device=dev),
                torch.tensor(gp[1], dtype=torch.float64, device=dev),
                None, None, None, None, None)


def rollout(sim, vel0, friction, restitution, dt, substeps, angvel0=None, checkpoint_every=16,
            restitution_tangent=0.0):
    """End state (n, 13) after ``substeps`` substeps of ``dt`` from the simulation's current
    positions and orientations with initial velocities ``vel0`` (n, 3) and angular velocities
    ``angvel0`` (default: the simulation's), friction ``friction`` and normal restitution
    ``restitution`` (tensors that may require gradients)."""
    if angvel0 is None:
        angvel0 = torch.from_numpy(sim.get_angular_velocities()).to(vel0.device)
    fr = friction if torch.is_tensor(friction) else torch.tensor(float(friction), dtype=torch.float64)
    re = (restitution if torch.is_tensor(restitution)
          else torch.tensor(float(restitution), dtype=torch.float64))
    return _Rollout.apply(vel0, angvel0, fr, re, sim, float(dt), int(substeps),
                          int(checkpoint_every), float(restitution_tangent))


def bed_height(pos):
    """Mean height of the particle centres."""
    return pos[:, 2].mean()


def packing_fraction_below(pos, radii, area, fraction=0.6):
    """Solid fraction of the layer z < fraction * H, H = 2 x the mean centre height (the height
    of a uniform bed whose centres fill it): the volume of the spheres whose centres lie in the
    layer over the layer's volume.  Which spheres count is piecewise constant in the positions,
    so the gradient flows through H."""
    H = 2.0 * pos[:, 2].mean()
    top = fraction * H
    inside = (pos[:, 2] < top).to(pos.dtype)
    vol = (4.0 / 3.0) * np.pi * torch.as_tensor(radii, dtype=pos.dtype, device=pos.device) ** 3
    return (inside * vol).sum() / (area * top)
