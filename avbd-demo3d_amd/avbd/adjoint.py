"""Differentiable rollouts of the resident AVBD step as a torch.autograd.Function.

A thin layer over the C ABI's adjoint (avbd_adjoint_forward / avbd_adjoint_backward in
include/avbd_hip.h): the forward uploads a state of boxes, contacts, joints and springs whose initial velocities, angular
velocities, masses (moments scale with the mass) and friction coefficients come from tensors,
steps it on the GPU keeping checkpoints, and returns the final positions; the backward runs
the hand-written reverse sweep on the GPU and returns the gradients of those inputs. There is
no CPU path: without the HIP library or a GPU the rollout raises.

    sim = Rollout(state, params, steps=120, checkpoint_every=10)
    pos = sim(vel, ang_vel, mass, friction)        # (bodies, 3) float64, differentiable
    loss = ((com(pos, mass) - target) ** 2).sum(); loss.backward()
"""
from __future__ import annotations

import numpy as np

from . import capi
from .state import State


class _RolloutFn:
    """Holder of the torch.autograd.Function subclass, built on first use so that importing this
    module does not import torch."""
    fn = None

    @classmethod
    def get(cls):
        if cls.fn is None:
            import torch

            class RolloutFunction(torch.autograd.Function):
                @staticmethod
                def forward(ctx, vel, ang_vel, mass, friction, sim):
                    pos = sim._forward(vel.detach().cpu().numpy(), ang_vel.detach().cpu().numpy(),
                                       mass.detach().cpu().numpy(), friction.detach().cpu().numpy())
                    ctx.sim = sim
                    ctx.device = vel.device
                    return torch.from_numpy(pos).to(vel.device)

                @staticmethod
                def backward(ctx, grad_pos):
                    g = ctx.sim._backward(grad_pos.detach().cpu().numpy())
                    dev = ctx.device
                    t = lambda a: torch.from_numpy(a).to(dev)
                    return t(g["vel"]), t(g["ang_vel"]), t(g["mass"]), t(g["friction"]), None

            cls.fn = RolloutFunction
        return cls.fn


class Rollout:
    """`steps` GPU steps of a world of boxes (contacts, joints, springs) as a differentiable function of the initial
    velocities, angular velocities, masses and friction coefficients (all bodies, list order).
    checkpoint_every: checkpoints every that many steps, each segment recomputed by the backward;
    0 tapes every step in the forward (more HBM, no recompute)."""

    def __init__(self, state: State, params, steps: int, checkpoint_every: int = 10, **world_kw):
        self.state = state.copy()
        self.params = params
        self.steps = int(steps)
        self.every = int(checkpoint_every)
        self.world = capi.HipWorld.for_state(self.state, params, **world_kw)
        dyn = self.state.bodies["mass"] > 0
        self._ratio = np.where(dyn[:, None], self.state.bodies["moment"].astype(np.float64) /
                               np.where(dyn, self.state.bodies["mass"], 1.0)[:, None].astype(np.float64), 0.0)
        self._dyn = dyn
        self.final = None

    def __call__(self, vel, ang_vel, mass, friction):
        return _RolloutFn.get().apply(vel, ang_vel, mass, friction, self)

    def initial_state(self, vel, ang_vel, mass, friction) -> State:
        st = self.state.copy()
        b = st.bodies
        b["vel"] = np.asarray(vel, np.float64).reshape(-1, 3).astype(np.float32)
        b["ang_vel"] = np.asarray(ang_vel, np.float64).reshape(-1, 3).astype(np.float32)
        m = np.asarray(mass, np.float64).reshape(-1)
        b["mass"] = np.where(self._dyn, m, 0.0).astype(np.float32)
        b["moment"] = (self._ratio * np.where(self._dyn, m, 0.0)[:, None]).astype(np.float32)
        b["friction"] = np.asarray(friction, np.float64).reshape(-1).astype(np.float32)
        return st

    def _forward(self, vel, ang_vel, mass, friction):
        self.world.upload(self.initial_state(vel, ang_vel, mass, friction))
        self.world.adjoint_forward(self.steps, self.every)
        self.final = self.world.state()
        return self.final.bodies["pos"].astype(np.float64)

    def _backward(self, grad_pos):
        return self.world.adjoint_backward(d_pos=grad_pos)

    def close(self):
        self.world.close()
