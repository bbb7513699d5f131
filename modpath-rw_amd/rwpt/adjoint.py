"""Calibration objective with a pathwise gradient for PyTorch.

``ZonedMisfit.apply(theta, problem)`` evaluates the observation misfit of
rwpt_adjoint (Gaussian-smoothed concentrations of the active particles at
observation points, squared differences to the targets) for the zoned
parameter vector

    theta = (aL_1 .. aL_Z, aT_1 .. aT_Z, porosity scale, Dm)

and returns it as a 0-d float64 tensor whose backward is the adjoint's
gradient (common random numbers, every particle's branches held fixed).
All tracking and the reverse sweep run in librwpt on the GPU; torch only
carries the parameter vector and the gradient.
"""
from __future__ import annotations

import dataclasses

import numpy as np
import torch

from . import problems
from .engine import Engine


@dataclasses.dataclass
class CalibrationProblem:
    problem: problems.Problem       # flow field, porosity at scale 1, dispersion model
    zones: np.ndarray               # 0-based zone of every cell
    release: tuple                  # (cell, lx, ly, lz) of the particles
    windows: tuple                  # output / checkpoint times
    obs_xyz: np.ndarray             # (n_obs, 3) global coordinates
    target: np.ndarray              # (n_obs,) target concentrations
    bandwidth: tuple = (2.0, 2.0, 1.0)
    precision: int = 0

    @property
    def nzones(self) -> int:
        return int(self.zones.max()) + 1

    def evaluate(self, theta: np.ndarray, gradient: bool):
        """Misfit and (with gradient) d misfit / d theta for one parameter vector."""
        Z = self.nzones
        aL, aT = theta[:Z], theta[Z:2 * Z]
        scale, dm = float(theta[2 * Z]), float(theta[2 * Z + 1])
        Q = problems.zoned_dispersion(self.problem, self.zones, aL, aT, dm)
        Q.porosity = self.problem.porosity * scale
        E = Engine(Q, precision=self.precision)
        try:
            E.load_particles(*self.release)
            out = E.adjoint(self.windows, self.obs_xyz, self.target, self.bandwidth,
                            gradient=gradient)
        finally:
            E.close()
        if not gradient:
            return out["misfit"], None
        g = np.zeros(2 * Z + 2)
        for z in range(Z):
            g[z] = out["grad_al"][self.zones == z].sum()
            g[Z + z] = out["grad_at"][self.zones == z].sum()
        # porosity enters as porosity0 * scale: d/d scale at the given scale
        g[2 * Z] = out["grad_porosity_scale"] / scale
        g[2 * Z + 1] = out["grad_dm"].sum()
        return out["misfit"], g


class ZonedMisfit(torch.autograd.Function):
    """J(theta) with the pathwise adjoint as its backward."""

    @staticmethod
    def forward(ctx, theta: torch.Tensor, problem: CalibrationProblem) -> torch.Tensor:
        th = theta.detach().to("cpu", torch.float64).numpy()
        J, g = problem.evaluate(th, gradient=theta.requires_grad)
        if g is not None:
            ctx.save_for_backward(torch.from_numpy(g).to(theta.device, theta.dtype))
        return torch.tensor(J, dtype=theta.dtype, device=theta.device)

    @staticmethod
    def backward(ctx, grad_out: torch.Tensor):
        (g,) = ctx.saved_tensors
        return grad_out * g, None
