"""NavierStokesVelVort: the reference's equation class (navier_stokes.py:310-2013)
running its time step through the HIP library.

Constructor parameters, field history and entry points follow the reference:
``NavierStokesVelVort(velocity_field, Re=..., dt=..., end_time=...)``,
``perform_time_step(vel_hat_data, dPdx, time_step)``, ``solve()``,
``solve_scan()`` and the classmethods ``FromVelocityField`` / ``FromDomain``.
The state is a complex128 tensor (3, Nx_modes, Ny, Nz_modes) on the GPU.
"""

from __future__ import annotations

from functools import reduce
from typing import Any, Callable, List, Optional, Tuple

import numpy as np
import torch

from . import _lib
from .domain import PhysicalDomain
from .equation import Equation
from .field import FourierField, PhysicalField, VectorField
from .plan import StepPlan


def median_factor(n: int) -> int:
    """Median integer factor (navier_stokes.py:1853-1861)."""
    factors = reduce(list.__add__, ([i, n // i] for i in range(1, int(n**0.5) + 1) if n % i == 0))
    factors.sort()
    return factors[len(factors) // 2]


class NavierStokesVelVort(Equation):
    name = "Navier Stokes equation (velocity-vorticity formulation)"
    kind = _lib.KIND_NS

    def __init__(self, velocity_field: VectorField, **params: Any):
        pd: PhysicalDomain = params.get("physical_domain", velocity_field[0].physical_domain)
        super().__init__(pd.hat(), velocity_field, **params)
        self.physical_domain = pd
        self.u_max_over_u_tau = params.get("u_max_over_u_tau", 1.0)
        self.Re_tau = self._get_Re_tau(**params)
        self.max_cfl = params.get("max_cfl", 0.7)
        if params.get("constant_mass_flux", False):
            raise NotImplementedError("constant mass flux is not part of the MI355X step")
        self.constant_mass_flux = False
        self.dPdx = -1.0  # navier_stokes.py:371-375
        self.linearise_flag = False
        self._plan: Optional[StepPlan] = None
        self._plan_params = params

    @staticmethod
    def _get_Re_tau(**params: Any) -> float:
        """navier_stokes.py:661-674."""
        if "Re_tau" in params:
            return float(params["Re_tau"])
        if "Re" in params:
            return float(params["Re"]) / params.get("u_max_over_u_tau", 1.0)
        raise Exception("Either Re or Re_tau has to be given as a parameter.")

    @classmethod
    def FromVelocityField(cls, velocity_field: VectorField, **params: Any):
        vh = velocity_field.hat()
        vh.set_name("velocity_hat")
        return cls(vh, **params)

    @classmethod
    def FromDomain(cls, domain: PhysicalDomain, **params: Any):
        device = params.pop("device", "cuda")
        vel = VectorField([PhysicalField.Zeros(domain, device=device) for _ in range(3)], "velocity")
        return cls.FromVelocityField(vel, **params)

    def get_Re_tau(self) -> float:
        return self.Re_tau

    def get_u_max_over_u_tau(self) -> float:
        return self.u_max_over_u_tau

    def get_physical_domain(self) -> PhysicalDomain:
        return self.physical_domain

    def get_rk_parameters(self):
        """navier_stokes.py:898-918."""
        return ([29 / 96, -3 / 40, 1 / 6], [37 / 160, 5 / 24, 1 / 6], [8 / 15, 5 / 12, 3 / 4], [0, -17 / 60, -5 / 12])

    # ------------------------------------------------------------------
    def plan(self) -> StepPlan:
        if self._plan is None:
            g = self.physical_domain.grid_math
            self._plan = StepPlan(g, self.Re_tau, self.get_dt(), self.kind, self.linearise_flag, self.u_max_over_u_tau)
        return self._plan

    def _state(self, vel_hat_data: Optional[torch.Tensor]) -> torch.Tensor:
        if vel_hat_data is None:
            vel_hat_data = self.get_latest_field("velocity_hat").get_data()
        return vel_hat_data.to(device="cuda", dtype=torch.complex128).contiguous()

    def perform_time_step(
        self, vel_hat_data: Optional[torch.Tensor] = None, dPdx: Optional[float] = None, time_step: Optional[int] = None
    ) -> Tuple[torch.Tensor, float]:
        """navier_stokes.py:1778-1813: one RK3/CN step; without explicit data
        the latest field is advanced and appended to the history."""
        explicit = vel_hat_data is not None
        state = self._state(vel_hat_data).clone()
        dp = self.dPdx if dPdx is None else dPdx
        self.plan().step(state, dp, 1)
        if not explicit:
            vh = VectorField.FromData(FourierField, self.physical_domain, state, "velocity_hat")
            vh.set_time_step(self.get_latest_field("velocity_hat").get_time_step() + 1)
            self.append_field("velocity_hat", vh, in_place=False)
        return state, dp

    def number_of_time_steps(self) -> int:
        return max(0, int(self.end_time / self.get_dt()))

    def solve_scan(self) -> Tuple[torch.Tensor, List[float], int]:
        """navier_stokes.py:1826-2013: n = int(T/dt) steps split into
        outer x inner (median factor); returns the outer-step trajectory
        (with the initial state prepended) when write_intermediate_output or
        write_entire_output is set, else the final state."""
        n = self.number_of_time_steps()
        u = self._state(None).clone()
        u0 = u.clone()
        dp = self.dPdx
        if self.write_entire_output:
            inner, outer = 1, n
        else:
            inner = median_factor(n) if n > 0 else 1
            outer = n // inner if n > 0 else 0
        self.number_of_inner_steps, self.number_of_outer_steps, self.number_of_time_steps_ = inner, outer, n
        traj = [u0]
        plan = self.plan()
        for _ in range(outer):
            plan.step(u, dp, inner)
            if self.write_intermediate_output or self.write_entire_output:
                traj.append(u.clone())
        final = VectorField.FromData(FourierField, self.physical_domain, u.clone(), "velocity_hat")
        final.set_time_step(outer)
        self.append_field("velocity_hat", final, in_place=False)
        if self.write_intermediate_output or self.write_entire_output:
            return torch.stack(traj), [dp] * outer, n
        return u.unsqueeze(0), [dp], n

    def energy(self, vel_hat_data: torch.Tensor) -> float:
        """Energy of a spectral state on the GPU (field.py:607-611)."""
        return self.plan().energy(self._state(vel_hat_data))
