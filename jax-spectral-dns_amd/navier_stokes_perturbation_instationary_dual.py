"""NavierStokesVelVortPerturbationInstationaryDual (reference
navier_stokes_perturbation_instationary_dual.py:70-119): the dual of the
perturbation equation about the time-dependent base flow U(y, t).

Dual step t (counted from the final time) uses the forward's base flow at
forward time step n - t (vel_base_fn, :96-101) and the forward state of that
time step (navier_stokes_perturbation_dual.py:591-602), with the full
(linearise = False) dual term and coupling_term = False: h_v gains
(d_x^2 + d_z^2) hat(U' v_x) (set_linearise, :103-119;
navier_stokes_perturbation_dual.py:269-278).  No pressure gradient (:87-92).
The forward run and its checkpointed recomputation set the forward plan's
base flow before every forward step, as the forward equation does; every
profile comes from the forward equation's device table (base_flow_table),
so neither direction waits for the GPU between steps.
"""

from __future__ import annotations

from typing import Any, Optional

import torch

from .navier_stokes_perturbation_dual import NavierStokesVelVortPerturbationDual
from .navier_stokes_perturbation_instationary import NavierStokesVelVortPerturbationInstationary


class NavierStokesVelVortPerturbationInstationaryDual(NavierStokesVelVortPerturbationDual):
    name = (
        "Dual Navier Stokes equation (velocity-vorticity formulation) for perturbations "
        "on top of a time-dependent base flow."
    )

    def __init__(self, forward_equation: NavierStokesVelVortPerturbationInstationary, **params: Any):
        if not isinstance(forward_equation, NavierStokesVelVortPerturbationInstationary):
            raise TypeError("the forward equation must be a NavierStokesVelVortPerturbationInstationary")
        params = dict(params)
        params["coupling_term"] = False
        super().__init__(forward_equation, **params)
        self.kappa = forward_equation.kappa
        self.accelerate = forward_equation.accelerate
        self.n_max = forward_equation.n_max

    def update_pressure_gradient(self, vel_new_field_hat=None, dPdx: Optional[float] = None) -> float:
        return 0.0

    def vel_base_profile(self, time_step: int):
        """U(y) of dual time step time_step: the forward's at step n - time_step."""
        return self.forward_equation.vel_base_profile(self.number_of_time_steps - time_step)

    def _advance_forward(self, u: torch.Tensor, k0: int, nsteps: int) -> None:
        fplan = self.forward_equation.plan()
        table = self.forward_equation.base_flow_table()
        for k in range(k0, k0 + nsteps):
            fplan.set_base_flow(table[k, 0], table[k, 1])
            fplan.step(u, 0.0, 1)

    def _dual_step(self, v: torch.Tensor, t: int, fwd_state: Optional[torch.Tensor]) -> None:
        plan = self.plan_for_step(t)
        table = self.forward_equation.base_flow_table()
        k = self.number_of_time_steps - t  # the forward's step n - t
        plan.set_base_flow(table[k, 0], table[k, 1])
        plan.dual_step(v, fwd_state)
