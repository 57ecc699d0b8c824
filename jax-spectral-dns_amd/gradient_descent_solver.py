"""Steepest adaptive descent and conjugate-gradient descent for the linear /
nonlinear optimal-perturbation problem on top of the dual (adjoint)
pipeline: the callers that turn gradient evaluations into optimisation
throughput.

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


class ConjugateGradientDescentSolver:
    """Restates ConjugateGradientDescentSolver (gradient_descent_solver.py:490-860)
    on its default path (use_linesearch = False).

    Each iteration (update, gradient_descent_solver.py:710-842):

    * runs the forward problem from the current guess and takes its gain;
    * from the second iteration on, moves along the projected conjugate
      direction (get_projected_cg_grad with the current beta and the previous
      direction) and also forms the steepest projected gradient; the first
      iteration uses the steepest one with step_size_factor =
      max_step_size_initial_factor and caps the step at that factor times
      the largest admissible step (get_max_step_size_ls);
    * renormalises the guess to the initial energy;
    * grows the step size by 1.5 (at most max_step_size) if the gain rose
      since the previous iteration, else halves it (at least min_step_size)
      and resets beta to 0;
    * otherwise updates beta by Polak-Ribiere on the steepest gradients,
      beta = <g, g - g_old> / <g_old, g_old>, clipped at 0 (update_beta,
      :847-860).  The reference's jnp.dot of the flattened spectra is a plain
      (unconjugated) sum; for the Hermitian spectra of real fields it is real
      up to rounding, and beta keeps its real part here so the direction
      stays the spectrum of a real field.

    Iteration stops after max_iterations or when the relative gain change
    falls below value_change_threshold (is_done, :97-109)."""

    def __init__(self, nse: NavierStokesVelVortPerturbation, **params: Any):
        self.nse = nse
        self.step_size = float(params.get("step_size", 1e-2))
        self.max_step_size = float(params.get("max_step_size", 1e-1))
        self.min_step_size = float(params.get("min_step_size", 1e-4))
        self.max_step_size_initial_factor = float(params.get("max_step_size_initial_factor", 1.25e-1))
        self.number_of_steps = int(params.get("max_iterations", params.get("max_iter", 20)))
        self.value_change_threshold = float(params.get("value_change_threshold", 1e-8))
        self.re = nse.get_Re_tau() * nse.get_u_max_over_u_tau()
        self.dt = nse.get_dt()
        self.end_time = nse.end_time
        self.linearise = nse.linearise()
        self.physical_domain = nse.get_physical_domain()
        self.current_guess = nse.get_initial_field("velocity_hat").get_data().clone()
        dual = NavierStokesVelVortPerturbationDual.FromNavierStokesVelVortPerturbation(nse)
        self.plan = dual.forward_equation.plan()
        self.e_0 = self.plan.energy(self.current_guess)
        self.beta = 0.0
        self.value: Optional[float] = None
        self.old_value: Optional[float] = None
        self.old_grad: Optional[torch.Tensor] = None
        self.old_steepest_grad: Optional[torch.Tensor] = None
        self.values: List[float] = []
        self.betas: List[float] = []
        self.i = 0

    def normalize_field(self, v: torch.Tensor) -> torch.Tensor:
        """Rescale to the initial energy (gradient_descent_solver.py:307-316)."""
        return v * (self.e_0 / self.plan.energy(v)) ** 0.5

    def _dual(self, u0: torch.Tensor) -> NavierStokesVelVortPerturbationDual:
        vh = VectorField.FromData(FourierField, self.physical_domain, u0, "velocity_hat")
        nse = NavierStokesVelVortPerturbation(vh, Re=self.re, dt=self.dt, end_time=self.end_time, linearise=self.linearise)
        return NavierStokesVelVortPerturbationDual.FromNavierStokesVelVortPerturbation(nse)

    def update_beta(self, last_iteration_successful: bool, steepest: torch.Tensor) -> None:
        if last_iteration_successful and self.old_steepest_grad is not None:
            g, g0 = steepest.flatten(), self.old_steepest_grad.flatten()
            beta = (torch.sum(g * (g - g0)) / torch.sum(g0 * g0)).real.item()
            self.beta = max(beta, 0.0)
        else:
            self.beta = 0.0

    def update(self) -> None:
        dual = self._dual(self.current_guess)
        gain = dual.get_objective_fun()
        gain_change = None if self.value is None else gain - self.value
        if self.old_grad is not None:
            grad, _ = dual.get_projected_cg_grad(self.step_size, self.beta, self.old_grad)
            steepest, _ = dual.get_projected_grad(self.step_size)
        else:
            grad, _ = dual.get_projected_grad(self.step_size, step_size_factor=self.max_step_size_initial_factor)
            steepest = grad
            self.step_size = min(self.step_size, dual.get_max_step_size() * self.max_step_size_initial_factor)
        self.current_guess = self.normalize_field(self.current_guess + self.step_size * grad)
        if gain_change is not None:
            if gain_change > 0.0:
                self.step_size = min(self.max_step_size, self.step_size * 1.5)
                self.update_beta(True, steepest)
            else:
                self.step_size = max(self.step_size / 2.0, self.min_step_size)
                self.update_beta(False, steepest)
        self.old_grad = grad
        self.old_steepest_grad = steepest
        self.old_value = self.value
        self.value = gain
        self.values.append(gain)
        self.betas.append(self.beta)
        self.i += 1

    def is_done(self) -> bool:
        if self.i >= self.number_of_steps:
            return True
        if self.value is not None and self.old_value is not None:
            return abs((self.value - self.old_value) / self.value) < self.value_change_threshold
        return False

    def optimise(self, iterations: Optional[int] = None) -> List[float]:
        n = self.number_of_steps if iterations is None else iterations
        for _ in range(n):
            self.update()
            if self.is_done():
                break
        return self.values
