"""Differentiable VBD rollouts as a ``torch.autograd.Function`` (no reference counterpart).

``rollout(integrator, log_y, x0, v0, dt, iterations, steps)`` sets the
per-tet log Young's moduli and the initial state, runs ``steps`` VBD steps
on the device and returns ``(x_T, v_T)`` (each 3 x nV). Its backward pass is
the library's implicit-function-theorem adjoint (``Integrator.rollout_backward``):
one linear solve with the step Hessian per step, walking checkpoints kept
every ``checkpoint_every`` steps and recomputing the steps between them.

Torch is only the autograd front end here; states cross the C ABI as host
arrays and all simulation and adjoint work runs in the HIP kernels.
"""
from __future__ import annotations

import numpy as np

from .vbd import Integrator


def _torch():
    import torch
    return torch


def _make_function():
    torch = _torch()

    class VbdRollout(torch.autograd.Function):
        @staticmethod
        def forward(ctx, log_y, x0, v0, integrator, dt, iterations, steps, checkpoint_every, solver):
            integrator.set_log_young_modulus(log_y.detach().cpu().numpy())
            integrator.x = x0.detach().cpu().numpy()
            integrator.v = v0.detach().cpu().numpy()
            integrator.rollout_forward(dt, iterations, steps, checkpoint_every)
            ctx.integrator = integrator
            ctx.generation = integrator.rollout_generation()[0]
            ctx.solver = dict(solver or {})
            ctx.like = (log_y, x0, v0)
            xT = torch.from_numpy(np.ascontiguousarray(integrator.x)).to(dtype=x0.dtype, device=x0.device)
            vT = torch.from_numpy(np.ascontiguousarray(integrator.v)).to(dtype=v0.dtype, device=v0.device)
            return xT, vT

        @staticmethod
        def backward(ctx, gx, gv):
            log_y, x0, v0 = ctx.like
            generation, held = ctx.integrator.rollout_generation()
            if not held or generation != ctx.generation:
                # the integrator holds one rollout: another rollout, or a change of its
                # state, moduli or settings, replaced the one this graph recorded
                raise RuntimeError(
                    "the integrator's rollout changed after this graph's forward pass (another "
                    "rollout, or a new state, moduli or solver setting); its gradients would "
                    "belong to that rollout. Use one integrator per pending backward pass.")
            gxa = np.zeros(tuple(x0.shape), dtype=np.float32) if gx is None else gx.detach().cpu().numpy()
            gva = None if gv is None else gv.detach().cpu().numpy()
            r = ctx.integrator.rollout_backward(gxa, gva, **ctx.solver)
            ctx.stats = r["stats"]
            gl = torch.from_numpy(r["log_young_modulus"]).to(dtype=log_y.dtype, device=log_y.device)
            g0 = torch.from_numpy(np.ascontiguousarray(r["x0"])).to(dtype=x0.dtype, device=x0.device)
            gv0 = torch.from_numpy(np.ascontiguousarray(r["v0"])).to(dtype=v0.dtype, device=v0.device)
            return gl, g0, gv0, None, None, None, None, None, None

    return VbdRollout


_FUNCTION = None


def rollout(integrator: Integrator, log_y, x0, v0, dt: float = 0.01, iterations: int = 20, steps: int = 1,
            checkpoint_every: int = 10, solver: dict | None = None):
    """(x_T, v_T) after `steps` VBD steps; differentiable in log_y (nT), x0 and v0 (3 x nV)."""
    global _FUNCTION
    if _FUNCTION is None:
        _FUNCTION = _make_function()
    return _FUNCTION.apply(log_y, x0, v0, integrator, float(dt), int(iterations), int(steps),
                           int(checkpoint_every), solver)
