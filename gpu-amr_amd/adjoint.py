"""torch.autograd face of the discrete adjoint (amr_tape_* in include/amr_hip.h).

`AdvectionSolve.apply(u0, velocity, tape)` runs the recorded solve (topology
and dt sequence frozen at recording) from initial interiors `u0` (a float64
CUDA tensor, initial_leaves * cells_per_leaf values) with `velocity` (a
float64 tensor of `rank` values) and returns u(T); its backward runs the HIP
adjoint kernels with checkpoint recomputation and returns the gradients with
respect to u0 and velocity (dt held fixed).  Nothing here computes on the
CPU: forward and backward are single calls into libamr_hip.so.
"""
from __future__ import annotations

import numpy as np


def _torch():
    import torch

    return torch


class _Generation:
    """The tape differentiates its most recent forward pass; this counter lets
    a backward detect that the tape was replayed in between."""

    def __init__(self):
        self.value = 0


def _generation(tape):
    if not hasattr(tape, "_generation"):
        tape._generation = _Generation()
    return tape._generation


def _make_function():
    torch = _torch()

    class AdvectionSolve(torch.autograd.Function):
        @staticmethod
        def forward(ctx, u0, velocity, tape):
            info = tape.info()
            if u0.dtype != torch.float64 or not u0.is_cuda:
                raise TypeError("u0 must be a float64 CUDA tensor")
            if u0.numel() != info["initial_leaves"] * info["cells_per_leaf"]:
                raise ValueError("u0 must hold initial_leaves * cells_per_leaf values")
            u0c = u0.detach().contiguous()
            v = [float(x) for x in velocity.detach().cpu().tolist()]
            uT = torch.empty(info["final_leaves"] * info["cells_per_leaf"], dtype=torch.float64, device=u0.device)
            tape.forward(u0c.data_ptr(), v, uT.data_ptr())
            gen = _generation(tape)
            gen.value += 1
            ctx.tape, ctx.gen, ctx.rank = tape, gen.value, len(v)
            ctx.u0_shape, ctx.v_dev = u0.shape, velocity.device
            return uT

        @staticmethod
        def backward(ctx, grad_uT):
            tape = ctx.tape
            if _generation(tape).value != ctx.gen:
                raise RuntimeError("the tape was replayed after this forward pass; run the forward again")
            g = grad_uT.detach().to(torch.float64).contiguous()
            grad_u0 = torch.empty(int(np.prod(ctx.u0_shape)), dtype=torch.float64, device=g.device)
            gv = tape.adjoint(g.data_ptr(), grad_u0.data_ptr())
            grad_v = torch.tensor(gv[: ctx.rank], dtype=torch.float64, device=ctx.v_dev)
            return grad_u0.reshape(ctx.u0_shape), grad_v, None

    return AdvectionSolve


_fn = None


def advection_solve(u0, velocity, tape):
    """u(T) of the recorded solve, differentiable in u0 and velocity."""
    global _fn
    if _fn is None:
        _fn = _make_function()
    return _fn.apply(u0, velocity, tape)


def misfit(u0, velocity, tape, u_target):
    """J = 0.5 * sum over leaf cells of (u(T) - u_target)^2 (unweighted)."""
    uT = advection_solve(u0, velocity, tape)
    return 0.5 * ((uT - u_target) ** 2).sum()
