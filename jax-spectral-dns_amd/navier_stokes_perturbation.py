"""NavierStokesVelVortPerturbation (reference navier_stokes_perturbation.py:341-580):
perturbation about the laminar base flow U = u_max (1 - y^2) with the
``linearise`` switch; the base flow enters the HIP kernels as the profiles
U(y), U'(y).  dPdx = 0 (navier_stokes_perturbation.py:503)."""

from __future__ import annotations

from typing import Any

from . import _lib
from .navier_stokes import NavierStokesVelVort
from .field import FourierField, PhysicalField, VectorField
from .plan import StepPlan


class NavierStokesVelVortPerturbation(NavierStokesVelVort):
    name = "Navier Stokes equation (velocity-vorticity formulation) for perturbations on top of a base flow."
    kind = _lib.KIND_PERTURBATION

    def __init__(self, velocity_field: VectorField, **params: Any):
        super().__init__(velocity_field, **params)
        if params.get("linearise_switch") is not None or params.get("coupling_term_switch") is not None:
            raise NotImplementedError("time-dependent linearise / coupling switches are not part of the MI355X step")
        if "velocity_base_hat" in params:
            raise NotImplementedError("the MI355X step uses the laminar base flow u_max (1 - y^2)")
        self.linearise_flag = bool(params.get("linearise", False))
        # coupling_term = False adds d_z hat(U' v) to h_g (navier_stokes_perturbation.py:214-224)
        self.coupling_term_flag = bool(params.get("coupling_term", True))
        self.dPdx = 0.0
        pd = self.physical_domain
        um = self.u_max_over_u_tau
        ub = PhysicalField.FromFunc(pd, lambda X: um * (1 - X[1] ** 2) + 0.0 * X[0] * X[2], "velocity_x_base")
        zero = PhysicalField.FromFunc(pd, lambda X: 0.0 * X[0] * X[1] * X[2], "velocity_base")
        base = VectorField([ub.hat(), zero.hat(), zero.hat()], "velocity_base_hat")
        self.add_field("velocity_base_hat", base)

    def linearise(self, _t: int = 0) -> bool:
        return self.linearise_flag

    def coupling_term(self, _t: int = 0) -> bool:
        return self.coupling_term_flag

    def plan(self) -> StepPlan:
        if self._plan is None:
            g = self.physical_domain.grid_math
            self._plan = StepPlan(
                g, self.Re_tau, self.get_dt(), self.kind, self.linearise_flag, self.u_max_over_u_tau,
                coupling_term=self.coupling_term_flag,
            )
        return self._plan
