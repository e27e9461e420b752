"""Differentiable rollouts of the resident AVBD step as a torch.autograd.Function.

A thin layer over the C ABI's adjoint (avbd_adjoint_forward / avbd_adjoint_backward in
include/avbd_hip.h): the forward sets a state of boxes and round shapes with contacts, joints,
springs and IgnoreCollision pairs whose initial velocities, angular velocities, masses (moments scale with
the mass) and friction coefficients come from tensors, steps it on the GPU keeping checkpoints,
and returns the final positions; the backward runs the hand-written reverse sweep on the GPU and
returns the gradients of those inputs. With CUDA (ROCm) tensors nothing goes through host
memory: the inputs are read from the tensors' device memory (avbd_set_body_inputs) into the
state uploaded once at construction (each later forward first rewinds to it with
avbd_adjoint_rewind), and positions and gradients are written into device tensors. With CPU
tensors each forward uploads the state from the host. The two give the same bits. There is no
CPU compute path: without the HIP library or a GPU the rollout raises.

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
                    ctx.sim = sim
                    ctx.device = vel.device
                    if vel.is_cuda:
                        return sim._forward_device(vel, ang_vel, mass, friction)
                    pos = sim._forward(vel.detach().numpy(), ang_vel.detach().numpy(),
                                       mass.detach().numpy(), friction.detach().numpy())
                    return torch.from_numpy(pos)

                @staticmethod
                def backward(ctx, grad_pos):
                    if grad_pos.is_cuda:
                        g = ctx.sim._backward_device(grad_pos)
                        return g["vel"], g["ang_vel"], g["mass"], g["friction"], None
                    g = ctx.sim._backward(grad_pos.detach().cpu().numpy())
                    t = lambda a: torch.from_numpy(a).to(ctx.device)
                    return t(g["vel"]), t(g["ang_vel"]), t(g["mass"]), t(g["friction"]), None

            cls.fn = RolloutFunction
        return cls.fn


class Rollout:
    """`steps` GPU steps of a world (boxes and round shapes; contacts, joints, springs) as a differentiable function of the initial
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
        self._final = None
        # the handle holds self.state as uploaded (no inputs set on it yet); after a device-path
        # forward it holds that forward's record, which avbd_adjoint_rewind returns to
        self._original = True
        self._dev_recorded = False

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

    @property
    def final(self) -> State:
        """The state after the last forward (downloaded on first use after a device-path forward)."""
        if self._final is None:
            self._final = self.world.state()
        return self._final

    def _forward(self, vel, ang_vel, mass, friction):
        self.world.upload(self.initial_state(vel, ang_vel, mass, friction))
        self._original = self._dev_recorded = False
        self.world.adjoint_forward(self.steps, self.every)
        self._final = self.world.state()
        return self._final.bodies["pos"].astype(np.float64)

    def _forward_device(self, vel, ang_vel, mass, friction):
        import torch
        nb = len(self.state.bodies)
        ins = [x.detach().to(torch.float64).contiguous() for x in (vel, ang_vel, mass, friction)]
        for x, shape in zip(ins, ((nb, 3), (nb, 3), (nb,), (nb,))):
            if x.numel() != int(np.prod(shape)):
                raise ValueError(f"rollout input of {x.numel()} values, expected {shape}")
        pos = torch.empty((nb, 3), dtype=torch.float64, device=vel.device)
        torch.cuda.current_stream(vel.device).synchronize()
        if self._dev_recorded:
            self.world.adjoint_rewind()
        elif not self._original:
            self.world.upload(self.state)
            self._original = True
        self.world.set_body_inputs_ptr(*[x.data_ptr() for x in ins])
        self.world.adjoint_forward(self.steps, self.every)
        self._dev_recorded = True
        self._final = None
        self.world.get_body_state_f64_ptr(pos=pos.data_ptr())
        return pos

    def _backward(self, grad_pos):
        return self.world.adjoint_backward(d_pos=grad_pos)

    def _backward_device(self, grad_pos):
        import torch
        nb = len(self.state.bodies)
        d = grad_pos.detach().to(torch.float64).contiguous()
        f = lambda *shape: torch.empty(shape, dtype=torch.float64, device=d.device)
        g = {"vel": f(nb, 3), "ang_vel": f(nb, 3), "mass": f(nb), "friction": f(nb)}
        torch.cuda.current_stream(d.device).synchronize()
        self.world.adjoint_backward_ptr(d_pos=d.data_ptr(), g_vel=g["vel"].data_ptr(),
                                        g_ang_vel=g["ang_vel"].data_ptr(), g_mass=g["mass"].data_ptr(),
                                        g_friction=g["friction"].data_ptr())
        return g

    def close(self):
        self.world.close()
