"""torch.autograd.Function over the C-ABI discrete adjoint.

`InletObjective.apply(profile, inv_tau, problem)` returns the objective
J = sum_{t in window} sum_x w(x) u_x(x, t) of a run whose EQU inlet cells take
their x-velocity from `profile` (one value per inlet cell, in C order of the
inlet mask), with gradients dJ/dprofile and dJ/dinvTau from culbm_adjoint.
The initial populations are constants of the problem (rest by default, as the
reference initialises them), so the profile enters through the EQU cells only.

The forward call runs the adjoint as well (one pass yields J and every
gradient); backward scales the stored gradients by the incoming one.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Optional, Tuple

import numpy as np
import torch

from . import cases
from .lattice import Lattice, LatticeConfig


@dataclass
class InletProblem:
    """A run whose inlet x-velocity profile is the parameter.

    state: flags and fields of the run (cases.State); inlet: boolean mask
    (nz, ny, nx) of the cells whose vx is the profile (EQU cells); weight: the
    objective's w (nz, ny, nx); nsteps / window: as Lattice.adjoint."""

    state: cases.State
    inlet: np.ndarray
    weight: np.ndarray
    nsteps: int
    window: Tuple[int, int]
    storage: str = "fp32"
    init: str = "rest"
    checkpoint_every: int = 0
    mem_cap: int = 0
    max_slots: int = 0
    device: int = 0

    def __post_init__(self):
        self.inlet = np.asarray(self.inlet, dtype=bool)
        if self.inlet.shape != self.state.shape:
            raise ValueError("inlet mask must have the state's shape")
        if not np.all(self.state.flag[self.inlet] & cases.EQU):
            raise ValueError("every inlet cell must carry the EQU flag")
        self._lat: Optional[Lattice] = None
        self._tau: Optional[float] = None
        self.last_stats: dict = {}

    @property
    def size(self) -> int:
        return int(self.inlet.sum())

    def default_profile(self) -> np.ndarray:
        return self.state.vx[self.inlet].astype(np.float64)

    def _lattice(self, inv_tau: float) -> Lattice:
        if self._lat is None or self._tau != inv_tau:
            if self._lat is not None:
                self._lat.close()
            self._lat = Lattice(LatticeConfig(domain=self.state.domain, inv_tau=inv_tau,
                                              storage=self.storage, device=self.device))
            self._tau = inv_tau
        return self._lat

    def evaluate(self, profile: np.ndarray, inv_tau: float):
        """(J, dJ/dprofile, dJ/dinvTau) for one profile (float64 values)."""
        profile = np.asarray(profile, dtype=np.float64).reshape(-1)
        if profile.size != self.size:
            raise ValueError(f"profile has {profile.size} values, the inlet has {self.size} cells")
        lat = self._lattice(float(np.float32(inv_tau)))
        vx = self.state.vx.copy()
        vx[self.inlet] = profile.astype(np.float32)
        lat.set_state(self.state.flag, self.state.rho, vx, self.state.vy, self.state.vz)
        lat.init_ddf(self.init)
        r = lat.adjoint(self.weight, self.nsteps, self.window, checkpoint_every=self.checkpoint_every,
                        mem_cap=self.mem_cap, max_slots=self.max_slots)
        self.last_stats = r.stats
        return r.objective, r.grad_ux[self.inlet].copy(), r.grad_inv_tau

    def close(self) -> None:
        if self._lat is not None:
            self._lat.close()
            self._lat = None


class InletObjective(torch.autograd.Function):
    @staticmethod
    def forward(ctx, profile: torch.Tensor, inv_tau: torch.Tensor, problem: InletProblem):
        J, gp, gt = problem.evaluate(profile.detach().cpu().double().numpy(),
                                     float(inv_tau.detach().cpu()))
        ctx.save_for_backward(torch.from_numpy(gp).to(profile.device, profile.dtype),
                              torch.tensor(gt, dtype=inv_tau.dtype, device=inv_tau.device))
        return torch.tensor(J, dtype=torch.float64, device=profile.device)

    @staticmethod
    def backward(ctx, grad_out: torch.Tensor):
        gp, gt = ctx.saved_tensors
        g = grad_out.to(gp.dtype)
        return g * gp, grad_out.to(gt.dtype) * gt, None


def inlet_objective(profile: torch.Tensor, inv_tau: torch.Tensor, problem: InletProblem) -> torch.Tensor:
    """Differentiable J(profile, inv_tau) for `problem`."""
    return InletObjective.apply(profile, inv_tau, problem)
