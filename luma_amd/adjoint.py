"""torch.autograd binding of the discrete adjoint (lumagpu_adj_run), for testing gradients
from PyTorch.

``DragObjective.apply(omega, inlet_scale, positions)`` builds a grid from ``make_grid``
(a callable taking omega, inlet_scale and an (n, 3) numpy array of marker positions and
returning a ready GpuGrid), runs T steps forward with checkpoints, and returns the
time-averaged drag J as a 0-d float64 tensor; ``backward`` supplies dJ/domega, dJ/dinlet_scale
and dJ/dpositions from the adjoint run the forward already performed.  The tensors live on
the CPU: they carry parameters and gradients, the flow itself stays in HBM behind the C ABI.
"""
from __future__ import annotations

import numpy as np
import torch


class DragObjective(torch.autograd.Function):
    @staticmethod
    def forward(ctx, omega, inlet_scale, positions, make_grid, nsteps, checkpoint_every, direction=(1.0, 0.0, 0.0)):
        pos = positions.detach().cpu().numpy().astype(np.float64)
        with make_grid(float(omega), float(inlet_scale), pos) as g:
            out = g.adj_run(int(nsteps), int(checkpoint_every), direction=direction, backward=True, positions=True)
        ctx.grads = (out["dJ_domega"], out["dJ_dinlet"], out.get("dJ_dpos", np.zeros_like(pos)))
        return torch.tensor(out["J"], dtype=torch.float64)

    @staticmethod
    def backward(ctx, gJ):
        d_om, d_in, d_pos = ctx.grads
        g = gJ.to(torch.float64)
        return (g * d_om, g * d_in, g * torch.as_tensor(d_pos, dtype=torch.float64), None, None, None, None)
