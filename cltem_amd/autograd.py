"""torch.autograd binding of the adjoint slice loop.

`exit_intensity(positions, engine, Z)` is the CTEM plane-wave exit intensity |psi_exit|^2
as a differentiable function of the atom positions: the forward runs the engine's slice
loop keeping every state, or sqrt(N) checkpoints beyond the engine's checkpoint budget
(mse_forward_checkpointed), and the backward runs the adjoint
slice loop (mse_adjoint_positions) with the incoming dL/dI.  Positions are float64 (n, 3)
tensors on the CPU: they are host data that the engine bins on the GPU.
"""
from __future__ import annotations

import numpy as np
import torch


class ExitIntensity(torch.autograd.Function):
    @staticmethod
    def forward(ctx, positions: torch.Tensor, engine, Z, config: int = 0, plane_value: float = 1.0):
        pos = positions.detach().cpu().numpy().astype(np.float64)
        if pos.ndim != 2 or pos.shape[1] != 3:
            raise ValueError("positions must have shape (n, 3)")
        engine.sort_atoms(config, pos[:, 0], pos[:, 1], pos[:, 2], Z, want_ids=False)
        engine.build_transmission(config)
        engine.set_wave_count(1)
        engine.set_wave_configs([config])
        engine.init_plane_wave(plane_value)
        intensity = engine.forward_checkpointed(config)[0]
        ctx.engine = engine
        ctx.n_atoms = len(pos)
        return torch.from_numpy(intensity).to(positions.dtype)

    @staticmethod
    def backward(ctx, grad_intensity: torch.Tensor):
        g = grad_intensity.detach().cpu().numpy().astype(np.float64)[None]
        grad = ctx.engine.adjoint_positions(g, ctx.n_atoms)
        return torch.from_numpy(grad), None, None, None, None


def exit_intensity(positions: torch.Tensor, engine, Z, config: int = 0, plane_value: float = 1.0):
    """|psi_exit|^2 (res, res) of a plane wave through the atoms at `positions`."""
    return ExitIntensity.apply(positions, engine, Z, config, plane_value)
