"""torch.autograd surface of the differentiable batched form.

``Rollout`` holds a batch of scenes and a horizon; ``rollout(vel0, inv_mass,
r)`` returns the final body positions as a tensor whose backward runs the
gfx950 reverse sweep (dy_diff_backward): gradients with respect to the
initial linear velocities and the inverse masses.  Static bodies (inverse
mass 0) stay static and receive zero gradient.

PyTorch is only the caller's side here: the tensors cross the boundary as
host arrays, and the library's own stream does the work.  One forward then
one backward per Rollout at a time (the forward's checkpoints are what the
backward reverses).
"""
from __future__ import annotations

import numpy as np
import torch

from . import capi


class Rollout:
    def __init__(self, scene_list, frames: int, checkpoint_every: int = 10, n_substeps: int = 1, **tape):
        """tape: tape_contacts_per_body / tape_bytes, as capi.Diff takes them."""
        self.batch = capi.Batch(scene_list)
        self.diff = capi.Diff(self.batch, frames, checkpoint_every, n_substeps=n_substeps, **tape)
        self.frames = frames
        self.init = self.batch.bodies.host_arrays(self.batch.merged)
        self.static = self.init["inv_mass"] == 0

    @property
    def n_bodies(self) -> int:
        return self.batch.n_bodies

    def initial(self) -> tuple[np.ndarray, np.ndarray]:
        v = np.stack([self.init["vel_x"], self.init["vel_y"], self.init["vel_z"]], 1)
        return v.astype(np.float32), self.init["inv_mass"].astype(np.float32)

    def run(self, vel0: np.ndarray, inv_mass: np.ndarray) -> np.ndarray:
        if vel0.shape != (self.n_bodies, 3) or inv_mass.shape != (self.n_bodies,):
            raise ValueError("vel0 must be (n, 3) and inv_mass (n,)")
        if np.any(inv_mass[self.static] != 0) or np.any(inv_mass[~self.static] == 0):
            raise ValueError("inverse masses must stay zero exactly for the static bodies")
        state = dict(self.init)
        state["vel_x"], state["vel_y"], state["vel_z"] = (np.ascontiguousarray(vel0[:, k], np.float32)
                                                          for k in range(3))
        state["inv_mass"] = np.ascontiguousarray(inv_mass, np.float32)
        self.batch.bodies.upload_state(state, self.batch.s)
        self.diff.forward(self.frames)
        st = self.batch.state()
        return np.stack([st["pos_x"], st["pos_y"], st["pos_z"]], 1)

    def grads(self, grad_pos: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
        g = np.asarray(grad_pos, np.float32)
        adj = self.diff.backward({"pos_x": g[:, 0], "pos_y": g[:, 1], "pos_z": g[:, 2]})
        gv = np.stack([adj["vel_x"], adj["vel_y"], adj["vel_z"]], 1)
        return gv, adj["inv_mass"]


class RolloutFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vel0: torch.Tensor, inv_mass: torch.Tensor, r: Rollout) -> torch.Tensor:
        ctx.rollout = r
        x = r.run(vel0.detach().to("cpu", torch.float32).numpy(),
                  inv_mass.detach().to("cpu", torch.float32).numpy())
        return torch.as_tensor(x, dtype=vel0.dtype, device=vel0.device)

    @staticmethod
    def backward(ctx, grad_x: torch.Tensor):
        gv, gim = ctx.rollout.grads(grad_x.detach().to("cpu", torch.float32).numpy())
        return (torch.as_tensor(gv, dtype=grad_x.dtype, device=grad_x.device),
                torch.as_tensor(gim, dtype=grad_x.dtype, device=grad_x.device), None)


def rollout(vel0: torch.Tensor, inv_mass: torch.Tensor, r: Rollout) -> torch.Tensor:
    """Final positions (n, 3) after r.frames frames from initial velocities
    vel0 (n, 3) and inverse masses inv_mass (n,); differentiable in both."""
    return RolloutFunction.apply(vel0, inv_mass, r)
