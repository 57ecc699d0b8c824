"""NavierStokesVelVortPerturbationInstationary (reference
navier_stokes_perturbation_instationary.py:29-108): the nonlinear
perturbation step about a time-dependent base flow U(y, t) between walls
moving with +-g_w(t), g_w = 1 - exp(-kappa t) (``accelerate``) or
exp(-kappa t).

The profile (vel_base_fn, :39-80) is the series solution of
U_t = U_yy / Re started from rest (accelerating) or from Couette flow:

    U = g_w y + Re/6 g_w'(t) (y^3 - y)
        + sum_{n=1}^{n_max-1} [(e^{-kappa t} - e^{-a_n t}) p_n kappa^2 s / (a_n - kappa)
                               + e^{-a_n t} p_n g_w'(0)] sin(n pi y),

a_n = (n pi)^2 / Re, p_n = -2 Re (-1)^n / (pi n)^3, s = -1 (accelerating) or
+1.  It is evaluated on the host at t = time_step * dt and handed to the
HIP step (jsd_set_base_flow) before each step, the same for the step's
three substeps (navier_stokes.py:1695); solve_scan steps through a device
table of all steps' profiles built once (base_flow_table), so the loop
never waits for the GPU.  The nonlinear term is the full
(linearise = False) skew-symmetric perturbation term with
coupling_term = False (:95-108): h_g gains d_z hat(U' v).  No pressure
gradient (:83-89).
"""

from __future__ import annotations

import math
from typing import Any, List, Optional, Tuple

import numpy as np
import torch

from .field import FourierField, VectorField
from .navier_stokes_perturbation import NavierStokesVelVortPerturbation


def instationary_base_profile(
    y: np.ndarray, t: float, re: float, kappa: float = 1.0, accelerate: bool = True, n_max: int = 100
) -> np.ndarray:
    """U(y, t) of vel_base_fn (navier_stokes_perturbation_instationary.py:39-80);
    g_w' is the exact derivative (the reference takes it with jax.grad)."""
    gw = (1.0 - math.exp(-kappa * t)) if accelerate else math.exp(-kappa * t)
    dgw_t = (kappa if accelerate else -kappa) * math.exp(-kappa * t)
    dgw_0 = kappa if accelerate else -kappa
    s = -1.0 if accelerate else 1.0
    n = np.arange(1, n_max)
    p = -2.0 * re * (-1.0) ** n / (math.pi * n) ** 3
    a = (math.pi * n) ** 2 / re
    coef = (math.exp(-kappa * t) - np.exp(-a * t)) * p * (kappa**2 * s / (a - kappa)) + np.exp(-a * t) * p * dgw_0
    return gw * y + re / 6.0 * dgw_t * (y**3 - y) + np.sin(np.pi * np.outer(y, n)) @ coef


class NavierStokesVelVortPerturbationInstationary(NavierStokesVelVortPerturbation):
    name = "Navier Stokes equation (velocity-vorticity formulation) for perturbations on top of a time-dependent base flow."

    def __init__(self, velocity_field: VectorField, **params: Any):
        params = dict(params)
        params["linearise"] = False  # set_linearise installs the full term (:95-108)
        params["coupling_term"] = False
        super().__init__(velocity_field, **params)
        self.kappa = float(params.get("kappa", 1.0))
        self.accelerate = bool(params.get("accelerate", True))
        self.n_max = int(params.get("n_max", 100))
        self.dPdx = 0.0

    def vel_base_profile(self, time_step: int) -> np.ndarray:
        """U(y) at t = time_step * dt (time_steps = arange(0, end_time, dt))."""
        y = self.physical_domain.grid_math.y
        return instationary_base_profile(y, time_step * self.get_dt(), self.Re_tau, self.kappa, self.accelerate, self.n_max)

    def base_flow_table(self) -> torch.Tensor:
        """(U, U') of every time step 0 .. n as one device table (n + 1, 2, Ny),
        built once on the host (the same values set_base_flow derives from
        vel_base_profile) and uploaded once: stepping through it queues a
        device-to-device copy before each step instead of a host copy and a
        wait (jsd_set_base_flow_device)."""
        n = self.number_of_time_steps()
        tab = getattr(self, "_base_table", None)
        if tab is None or tab.shape[0] != n + 1:
            D = self.physical_domain.grid_math.D
            rows = []
            for k in range(n + 1):
                u = np.ascontiguousarray(self.vel_base_profile(k), dtype=np.float64)
                rows.append(np.stack([u, np.ascontiguousarray(D @ u, dtype=np.float64)]))
            tab = torch.from_numpy(np.stack(rows)).to(self._state(None).device)
            self._base_table = tab
        return tab

    def perform_time_step(
        self, vel_hat_data: Optional[torch.Tensor] = None, dPdx: Optional[float] = None, time_step: Optional[int] = None
    ) -> Tuple[torch.Tensor, float]:
        """One step with the base flow of ``time_step`` (default: the index of
        the latest field in the history)."""
        if time_step is None:
            time_step = self.get_latest_field("velocity_hat").get_time_step()
        self.plan().set_base_flow(self.vel_base_profile(int(time_step)))
        return super().perform_time_step(vel_hat_data, 0.0, time_step)

    def solve_scan(self) -> Tuple[torch.Tensor, List[float], int]:
        """All int(T/dt) steps, each with its own base flow; returns the final
        state (navier_stokes.py:1826-2013 with a per-step base flow)."""
        n = self.number_of_time_steps()
        u = self._state(None).clone()
        plan = self.plan()
        table = self.base_flow_table()
        traj = [u.clone()]
        for k in range(n):
            plan.set_base_flow(table[k, 0], table[k, 1])
            plan.step(u, 0.0, 1)
            if self.write_intermediate_output or self.write_entire_output:
                traj.append(u.clone())
        final = VectorField.FromData(FourierField, self.physical_domain, u.clone(), "velocity_hat")
        final.set_time_step(n)
        self.append_field("velocity_hat", final, in_place=False)
        if self.write_intermediate_output or self.write_entire_output:
            return torch.stack(traj), [0.0] * n, n
        return u.unsqueeze(0), [0.0], n
