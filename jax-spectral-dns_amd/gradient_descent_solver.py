"""Steepest adaptive descent for the linear / nonlinear optimal-perturbation
problem on top of the dual (adjoint) pipeline: the caller that turns
gradient evaluations into optimisation throughput.

Restates SteepestAdaptiveDescentSolver (gradient_descent_solver.py:340-487):
each iteration moves the current guess along the energy-constrained
projected gradient (get_projected_grad, navier_stokes_perturbation_dual.py:
1122-1167), renormalises it to the initial energy, and accepts the step if
the objective (the gain, get_objective_fun) increased, growing the step size
by 1.5; otherwise it shrinks the step size by 5 and retries from the last
accepted guess with its gradient, at most ``max_number_of_sub_iterations``
times (gradient_descent_solver.py:85-90, 377-487).  The time step is kept
fixed (the reference re-derives it from the CFL number of the latest
field, Equation.find_suitable_dt, which is outside the step path).
"""

from __future__ import annotations

from typing import Any, List, Optional

import torch

from .field import FourierField, VectorField
from .navier_stokes_perturbation import NavierStokesVelVortPerturbation
from .navier_stokes_perturbation_dual import NavierStokesVelVortPerturbationDual


class SteepestAdaptiveDescentSolver:
    def __init__(self, nse: NavierStokesVelVortPerturbation, **params: Any):
        self.nse = nse
        self.step_size = float(params.get("step_size", 1e-2))
        self.number_of_steps = int(params.get("max_iter", 20))
        self.max_number_of_sub_iterations = int(params.get("max_number_of_sub_iterations", 10))
        self.re = nse.get_Re_tau() * nse.get_u_max_over_u_tau()
        self.dt = nse.get_dt()
        self.end_time = nse.end_time
        self.linearise = nse.linearise()
        self.physical_domain = nse.get_physical_domain()
        u0 = nse.get_initial_field("velocity_hat").get_data()
        self.current_guess = u0.clone()
        self.dual = NavierStokesVelVortPerturbationDual.FromNavierStokesVelVortPerturbation(nse)
        self.e_0 = self.dual.forward_equation.plan().energy(self.current_guess)
        self.value = self.dual.get_objective_fun()
        self.old_value = self.value
        self.grad, _ = self.dual.get_projected_grad(self.step_size)
        self.old_dual = self.dual
        self.values: List[float] = [self.value]
        self.i = 0

    def increase_step_size(self, tau: float = 1.5) -> None:
        self.step_size *= tau

    def decrease_step_size(self, tau: float = 5.0) -> None:
        self.step_size /= tau

    def normalize_field(self, v: torch.Tensor) -> torch.Tensor:
        """Rescale to the initial energy (gradient_descent_solver.py:307-316)."""
        e = self.dual.forward_equation.plan().energy(v)
        return v * (self.e_0 / e) ** 0.5

    def _equation(self, u0: torch.Tensor) -> NavierStokesVelVortPerturbation:
        vh = VectorField.FromData(FourierField, self.physical_domain, u0, "velocity_hat")
        return NavierStokesVelVortPerturbation(
            vh, Re=self.re, dt=self.dt, end_time=self.end_time, linearise=self.linearise
        )

    def update(self) -> None:
        j = 0
        while True:
            new = self.normalize_field(self.current_guess + self.step_size * self.grad)
            dual = NavierStokesVelVortPerturbationDual.FromNavierStokesVelVortPerturbation(self._equation(new))
            value = dual.get_objective_fun()
            if value - self.old_value > 0.0:
                self.increase_step_size()
                self.grad, _ = dual.get_projected_grad(self.step_size)
                self.current_guess = new
                self.dual = self.old_dual = dual
                self.value = self.old_value = value
                break
            self.decrease_step_size()
            self.grad, _ = self.old_dual.get_projected_grad(self.step_size)
            j += 1
            if j > self.max_number_of_sub_iterations:
                break
        self.values.append(self.value)
        self.i += 1

    def optimise(self, iterations: Optional[int] = None) -> List[float]:
        for _ in range(self.number_of_steps if iterations is None else iterations):
            self.update()
        return self.values
