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
from .fixed_parameters import NavierStokesVelVortFixedParameters
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
        # navier_stokes.py:335: only the perturbation equation may set it
        self.enforce_base = False
        # the nonlinear_update_fn plug-in (navier_stokes.py:324-332): by default the
        # library's fused term; a callable (vel_hat, time_step) -> (h_v, h_g, vort,
        # conv_ns) assigned here replaces it in every substep
        self._library_nonlinear_update_fn = self._library_nonlinear
        self.nonlinear_update_fn: Callable = self._library_nonlinear_update_fn
        self.nse_fixed_parameters = NavierStokesVelVortFixedParameters(
            physical_domain=pd,
            poisson_mat=None,
            rk_mats_rhs=None,
            rk_mats_lhs_inv=None,
            rk_rhs_inhom=None,
            rk_mats_lhs_inv_inhom=None,
            rk_mats_rhs_ns=None,
            rk_mats_lhs_inv_ns=None,
            Re_tau=self.Re_tau,
            max_cfl=self.max_cfl,
            u_max_over_u_tau=self.u_max_over_u_tau,
            number_of_rk_steps=3,
        )

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

    def get_fixed_params(self) -> NavierStokesVelVortFixedParameters:
        return self.nse_fixed_parameters

    def get_source_term(self, _time_step: int) -> Optional[torch.Tensor]:
        """Source hook (navier_stokes.py:1216-1217): None, or a complex128
        (3, Nx_modes, Ny, Nz_modes) tensor added to the step's right-hand
        side the way perform_runge_kutta_step adds it."""
        return None

    def _library_nonlinear(self, vel_hat: torch.Tensor, _time_step: int):
        """The library's fused term as a nonlinear_update_fn: (h_v, h_g,
        vort, conv_ns) of vel_hat (navier_stokes.py:48-73, 236-261)."""
        h_v, h_g, conv = self.plan().nonlinear(self._state(vel_hat))
        vort = self.get_domain().curl(self._state(vel_hat))
        return h_v, h_g, vort, conv

    def _uses_library_nonlinear(self) -> bool:
        return self.nonlinear_update_fn is self._library_nonlinear_update_fn

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

    def plan_for_step(self, _time_step: int) -> StepPlan:
        """The plan that advances step time_step (the perturbation equation
        switches plans with linearise_switch / coupling_term_switch)."""
        return self.plan()

    def _advance(self, state: torch.Tensor, dPdx: float, time_step: int, nsteps: int) -> None:
        """nsteps steps from time_step on, in place: one library call for the
        plain step; step by step when a source term, enforce_base, a
        nonlinear_update_fn plug-in or a per-step plan switch is in use."""
        plain = self._uses_library_nonlinear() and not self.enforce_base and not self._has_step_hooks()
        if plain and self._plans_uniform(time_step, nsteps):
            self.plan_for_step(time_step).step(state, dPdx, nsteps)
            return
        for t in range(time_step, time_step + nsteps):
            self._one_step(state, dPdx, t)

    def _has_step_hooks(self) -> bool:
        """get_source_term overridden (in a subclass or on the instance)."""
        return getattr(self.get_source_term, "__func__", None) is not NavierStokesVelVort.get_source_term

    def _plans_uniform(self, time_step: int, nsteps: int) -> bool:
        return True

    def _one_step(self, state: torch.Tensor, dPdx: float, time_step: int, fwd_hat: Optional[torch.Tensor] = None):
        plan = self.plan_for_step(time_step)
        src = self.get_source_term(time_step)
        if src is not None:
            src = self._state(src)
        if self._uses_library_nonlinear():
            if src is None and not self.enforce_base:
                if fwd_hat is None:
                    plan.step(state, dPdx, 1)
                else:
                    plan.dual_step(state, fwd_hat)
            else:
                plan.step_source(state, dPdx, src, self.enforce_base, fwd_hat)
            return
        # the plug-in: terms from the user's function in every substep
        for s in range(3):
            h_v, h_g, _vort, conv = self.nonlinear_update_fn(state, time_step)
            plan.substep_terms(
                state, s, self._state(h_v), self._state(h_g), self._state(conv), dPdx, src, self.enforce_base
            )

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
        t = self.time_step if time_step is None else int(time_step)
        self._advance(state, dp, t, 1)
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
        for o in range(outer):
            self._advance(u, dp, o * inner, inner)
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
