"""NavierStokesVelVortPerturbation (reference navier_stokes_perturbation.py:341-580):
perturbation about a base flow with the ``linearise`` / ``coupling_term``
switches, their time-dependent forms ``linearise_switch`` /
``coupling_term_switch`` (:385-416) and ``enforce_base`` (:418).  dPdx = 0
(navier_stokes_perturbation.py:503).

The base flow is the reference's ``velocity_base_hat`` (:351-380): by default
the laminar profile u_max (1 - y^2); any base flow given as a streamwise
profile U(y) (the (0,0) mode of the x component, nothing else) is accepted,
and it enters the HIP kernels as the profiles U(y), U'(y) = D U.  A base flow
with other components or x / z dependence is rejected: the product kernels
evaluate the base-flow products as y-only factors.
"""

from __future__ import annotations

from typing import Any, Dict, Optional, Tuple

import numpy as np
import torch

from . import _lib
from .field import FourierField, PhysicalField, VectorField
from .navier_stokes import NavierStokesVelVort
from .plan import StepPlan


def base_profile_from_hat(grid, base_hat: torch.Tensor, tol: float = 1e-13) -> np.ndarray:
    """U(y) of a spectral base flow (3, Nx_modes, Ny, Nz_modes) that is a
    streamwise profile: the physical value of a lone (0,0) mode is
    Re(c) * prod(L/2pi) / sqrt(Px Pz) (the inverse of field_hat's scaling)."""
    b = base_hat.detach().to("cpu", torch.complex128)
    if tuple(b.shape) != (3, grid.nxm, grid.ny, grid.nzm):
        raise ValueError("velocity_base_hat has shape %s" % (tuple(b.shape),))
    line = b[0, 0, :, 0].clone()
    scale = float(line.abs().max()) if line.numel() else 0.0
    rest = b.clone()
    rest[0, 0, :, 0] = 0
    if float(rest.abs().max()) > tol * max(scale, 1e-300) or float(line.imag.abs().max()) > tol * max(scale, 1e-300):
        raise NotImplementedError(
            "the MI355X step takes base flows U(y) in x only (the (0,0) mode of the x component)"
        )
    return line.real.numpy() * grid.scaling / np.sqrt(grid.px * grid.pz)


class NavierStokesVelVortPerturbation(NavierStokesVelVort):
    name = "Navier Stokes equation (velocity-vorticity formulation) for perturbations on top of a base flow."
    kind = _lib.KIND_PERTURBATION
    max_cfl = 0.5
    max_dt = 2e-2

    def __init__(self, velocity_field: VectorField, **params: Any):
        super().__init__(velocity_field, **params)
        pd = self.physical_domain
        um = self.u_max_over_u_tau
        if params.get("velocity_base_hat") is not None:
            base = params["velocity_base_hat"]
            data = base.get_data() if hasattr(base, "get_data") else torch.as_tensor(base)
            self.base_profile = base_profile_from_hat(pd.grid_math, data)
            if not hasattr(base, "get_data"):
                base = VectorField.FromData(FourierField, pd, data, "velocity_base_hat")
        else:
            ub = PhysicalField.FromFunc(pd, lambda X: um * (1 - X[1] ** 2) + 0.0 * X[0] * X[2], "velocity_x_base")
            zero = PhysicalField.FromFunc(pd, lambda X: 0.0 * X[0] * X[1] * X[2], "velocity_base")
            base = VectorField([ub.hat(), zero.hat(), zero.hat()], "velocity_base_hat")
            self.base_profile = um * (1 - pd.grid_math.y**2)
        self.add_field("velocity_base_hat", base)
        # linearise / coupling_term and their time-dependent switches (:385-416):
        # linearised while t < n * linearise_switch, coupling term off while
        # t < n * coupling_term_switch, n = len(arange(0, end_time, dt))
        n = len(np.arange(0, abs(self.end_time), abs(self.get_dt()))) if self.end_time else 0
        self.linearise_switch: Optional[float] = params.get("linearise_switch")
        self.coupling_term_switch: Optional[float] = params.get("coupling_term_switch")
        if self.linearise_switch is not None:
            assert params.get("linearise") is None, "need to either pass linearise or linearise_switch."
            sw = float(self.linearise_switch)
            self.linearise = lambda t=0: t < n * sw
        else:
            lin = bool(params.get("linearise", False) or False)
            self.linearise = lambda _t=0: lin
        if self.coupling_term_switch is not None:
            assert params.get("coupling_term") is None, "need to either pass coupling_term or coupling_term_switch."
            csw = float(self.coupling_term_switch)
            self.coupling_term = lambda t=0: not (t < n * csw)
        else:
            cpl = params.get("coupling_term", True)
            cpl = True if cpl is None else bool(cpl)
            self.coupling_term = lambda _t=0: cpl
        self.linearise_flag = bool(self.linearise(0))
        self.coupling_term_flag = bool(self.coupling_term(0))
        self.enforce_base = bool(params.get("enforce_base", False))
        self.dPdx = 0.0
        self._plans: Dict[Tuple[bool, bool], StepPlan] = {}
        self._ops = None

    def _make_plan(self, linearise: bool, coupling_term: bool) -> StepPlan:
        g = self.physical_domain.grid_math
        p = StepPlan(
            g, self.Re_tau, self.get_dt(), self.kind, linearise, self.u_max_over_u_tau,
            ops=self._ops, coupling_term=coupling_term, base_profile=self.base_profile,
        )
        self._ops = p.ops  # the operator tables are shared by every plan of the equation
        return p

    def plan_for_step(self, time_step: int) -> StepPlan:
        key = (bool(self.linearise(time_step)), bool(self.coupling_term(time_step)))
        if key not in self._plans:
            self._plans[key] = self._make_plan(*key)
        return self._plans[key]

    def plan(self) -> StepPlan:
        return self.plan_for_step(0)

    def _plans_uniform(self, time_step: int, nsteps: int) -> bool:
        if self.linearise_switch is None and self.coupling_term_switch is None:
            return True
        k0 = (self.linearise(time_step), self.coupling_term(time_step))
        return all((self.linearise(t), self.coupling_term(t)) == k0 for t in range(time_step, time_step + nsteps))
