"""torch.autograd binding of the 2D adjoint.

``step_loss(solver, ux0, uy0, dts, every)`` runs ``len(dts)`` fused time steps
of ``solver`` from the initial velocity (ux0, uy0) (padded arrays, flat or
shaped like ``Solver.get``) with the frozen dt sequence ``dts`` and returns
the loss (droplet centroid x + interface length) as a 0-d float64 tensor.
Its backward is the hand-written adjoint sweep of the library
(vof_adjoint_gradient): one forward with checkpoints every ``every`` steps,
recomputation per segment and the reverse stages.  The solver's other state
(pressure, vof) is the one it holds when called and is left unchanged.
"""
from __future__ import annotations

import numpy as np
import torch


class _StepLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, ux0, uy0, solver, dts, every):
        shape_x, shape_y = ux0.shape, uy0.shape
        saved_x, saved_y = solver.get("ux"), solver.get("uy")
        solver.set("ux", ux0.detach().cpu().numpy().astype(np.float64).ravel())
        solver.set("uy", uy0.detach().cpu().numpy().astype(np.float64).ravel())
        solver.boundaries("ux")
        solver.boundaries("uy")
        try:
            loss, gux, guy = solver.adjoint_gradient(list(dts), every=every)
        finally:
            solver.set("ux", saved_x)
            solver.set("uy", saved_y)
        ctx.save_for_backward(torch.from_numpy(gux).reshape(shape_x).to(ux0.device),
                              torch.from_numpy(guy).reshape(shape_y).to(uy0.device))
        return torch.tensor(loss, dtype=torch.float64, device=ux0.device)

    @staticmethod
    def backward(ctx, gout):
        gux, guy = ctx.saved_tensors
        return gout * gux, gout * guy, None, None, None


def step_loss(solver, ux0: torch.Tensor, uy0: torch.Tensor, dts, every: int = 10) -> torch.Tensor:
    """Loss after len(dts) steps from (ux0, uy0), differentiable in both."""
    if solver.ndims != 2:
        raise ValueError("the adjoint is implemented for 2D")
    return _StepLoss.apply(ux0, uy0, solver, tuple(float(d) for d in dts), int(every))
