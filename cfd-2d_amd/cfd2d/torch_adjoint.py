"""torch.autograd bridge to the discrete adjoint (cfd2d_adjoint).

Used by the gradient test harness: the wall force J = sum over the masked
edges of p(c1) * l * n_y after nsteps steps, as a differentiable function of
the initial state U0 (4, nc) and the boundary parameters (n_bc, 4).  The
forward call runs cfd2d_adjoint once (the steps and the reverse sweep) and
keeps dJ/dU0 and dJ/dpar for backward(); nothing is differentiated by torch
itself.  Tensors may live on any device; they cross to the host buffers of
the C ABI.
"""
from __future__ import annotations

import torch


class WallForceY(torch.autograd.Function):
    @staticmethod
    def forward(ctx, U0, bc_par, solver, nsteps, edge_mask, ckpt_every):
        solver.set_bc_par(bc_par.detach().to("cpu", torch.float64).numpy())
        solver.set_state(U0.detach().to("cpu", torch.float64).numpy())
        J, gU, gb = solver.adjoint(nsteps, edge_mask, ckpt_every)
        ctx.save_for_backward(torch.from_numpy(gU).to(U0.device, U0.dtype),
                              torch.from_numpy(gb).to(bc_par.device, bc_par.dtype))
        return torch.tensor(J, dtype=U0.dtype, device=U0.device)

    @staticmethod
    def backward(ctx, gJ):
        gU, gb = ctx.saved_tensors
        return gJ * gU, gJ * gb, None, None, None, None


def wall_force_y(U0, bc_par, solver, nsteps, edge_mask, ckpt_every=0):
    """J after nsteps steps of `solver` (a cfd2d.native.GpuSolver) from U0 with
    boundary parameters bc_par; differentiable in U0 and bc_par."""
    return WallForceY.apply(U0, bc_par, solver, nsteps, edge_mask, ckpt_every)
