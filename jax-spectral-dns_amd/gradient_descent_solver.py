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

import numpy as np
import torch

from .field import FourierField, VectorField
from .navier_stokes_perturbation import NavierStokesVelVortPerturbation
from .navier_stokes_perturbation_dual import NavierStokesVelVortPerturbationDual
from .plan import StepPlan


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
        self.dual_params = {k: params[k] for k in ("optimisation_mode", "checkpointing", "checkpoint_every") if k in params}
        self.dual = NavierStokesVelVortPerturbationDual.FromNavierStokesVelVortPerturbation(nse, **self.dual_params)
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
        """The forward equation started from u0: the same equation object
        every iteration (its plans and operator tables, and the dual plans
        cached on it, are built once), with its initial field replaced."""
        vh = VectorField.FromData(FourierField, self.physical_domain, u0, "velocity_hat")
        self.nse.set_initial_field("velocity_hat", vh)
        return self.nse

    def update(self) -> None:
        j = 0
        while True:
            new = self.normalize_field(self.current_guess + self.step_size * self.grad)
            dual = NavierStokesVelVortPerturbationDual.FromNavierStokesVelVortPerturbation(
                self._equation(new), **self.dual_params
            )
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


class MultiStartSteepestAdaptiveDescentSolver:
    """SteepestAdaptiveDescentSolver (gradient_descent_solver.py:340-487) on
    several starts at once, with every objective and gradient evaluation of
    all starts batched: one batched forward of the linearised perturbation
    step (jsd_batch_step over the starts' plans) gives every candidate's
    gain, and one batched dual (the linearised dual plans of the same
    starts) gives every v(0) for the projected gradients
    (get_projected_grad, navier_stokes_perturbation_dual.py:1122-1167).

    Each start keeps its own guess, step size and sub-iteration count and
    follows the single-start rules: a candidate is accepted when its gain
    rose (the step size grows by 1.5 and the gradient is taken at the new
    guess), else the step size shrinks by 5 and the direction is re-projected
    from the last accepted guess, at most max_number_of_sub_iterations times
    per iteration.  The starts advance in lockstep rounds; a start whose
    iteration is finished waits (its candidate is its current guess).  The
    starts may differ in grid, Re, box and initial field; they share the
    number of time steps.  Linearised problems only (the batched dual is the
    linearised one)."""

    def __init__(self, nses, **params: Any):
        from .batch import BatchPlan
        from .navier_stokes_perturbation_dual import projected_gradient

        self._projected = projected_gradient
        self.nses = list(nses)
        if not self.nses:
            raise ValueError("no starts")
        if not all(bool(n.linearise(0)) for n in self.nses):
            raise ValueError("the batched dual is the linearised one: every start must be linearised")
        counts = {n.number_of_time_steps() for n in self.nses}
        if len(counts) != 1:
            raise ValueError("the starts must share the number of time steps")
        self.nsteps = counts.pop()
        self.step_size = [float(params.get("step_size", 1e-2))] * len(self.nses)
        self.number_of_steps = int(params.get("max_iter", 20))
        self.max_number_of_sub_iterations = int(params.get("max_number_of_sub_iterations", 10))
        self.fplans = [n.plan() for n in self.nses]
        self.dplans = [NavierStokesVelVortPerturbationDual(n).plan_for_step(0) for n in self.nses]
        dev = torch.device("cuda")
        self.u = [torch.zeros(p.shape, dtype=torch.complex128, device=dev) for p in self.fplans]
        self.v = [torch.zeros_like(x) for x in self.u]
        self.fwd = BatchPlan(self.fplans, self.u)
        self.dual = BatchPlan(self.dplans, self.v)
        self.current_guess = [n._state(n.get_initial_field("velocity_hat").get_data()).clone() for n in self.nses]
        self.e_0 = [p.energy(g) for p, g in zip(self.fplans, self.current_guess)]
        self.value = self._gains(self.current_guess)
        self.u0 = [g.clone() for g in self.current_guess]
        self.v0 = self._duals()
        self.grad = [self._projected(p, a, b, s) for p, a, b, s in zip(self.fplans, self.u0, self.v0, self.step_size)]
        self.values: List[List[float]] = [list(self.value)]
        self.i = 0
        self.evaluations = 1

    def _gains(self, guesses: List[torch.Tensor]) -> List[float]:
        """Batched forward: J_r = E(u_r(T)) / E(u_r(0)); leaves u_r(T) in self.u."""
        for u, g in zip(self.u, guesses):
            u.copy_(g)
        self.fwd.step(self.nsteps)
        return [p.energy(u) / p.energy(g) for p, u, g in zip(self.fplans, self.u, guesses)]

    def _duals(self) -> List[torch.Tensor]:
        """Batched dual from v(T) = -u(T) of the last forward: v_r(0)."""
        for v, u in zip(self.v, self.u):
            torch.neg(u, out=v)
        self.dual.step(self.nsteps)
        return [v.clone() for v in self.v]

    def normalize_field(self, r: int, v: torch.Tensor) -> torch.Tensor:
        return v * (self.e_0[r] / self.fplans[r].energy(v)) ** 0.5

    def update(self) -> None:
        R = len(self.nses)
        searching = [True] * R
        sub = [0] * R
        while any(searching):
            cand = [
                self.normalize_field(r, self.current_guess[r] + self.step_size[r] * self.grad[r])
                if searching[r] else self.current_guess[r]
                for r in range(R)
            ]
            gains = self._gains(cand)
            v0s = self._duals()
            self.evaluations += 1
            for r in range(R):
                if not searching[r]:
                    continue
                if gains[r] - self.value[r] > 0.0:
                    self.step_size[r] *= 1.5
                    self.current_guess[r] = cand[r]
                    self.value[r] = gains[r]
                    self.u0[r], self.v0[r] = cand[r], v0s[r]
                    self.grad[r] = self._projected(self.fplans[r], self.u0[r], self.v0[r], self.step_size[r])
                    searching[r] = False
                    continue
                self.step_size[r] /= 5.0
                self.grad[r] = self._projected(self.fplans[r], self.u0[r], self.v0[r], self.step_size[r])
                sub[r] += 1
                if sub[r] > self.max_number_of_sub_iterations:
                    searching[r] = False
        self.values.append(list(self.value))
        self.i += 1

    def optimise(self, iterations: Optional[int] = None) -> List[List[float]]:
        for _ in range(self.number_of_steps if iterations is None else iterations):
            self.update()
        return self.values

    def close(self) -> None:
        self.fwd.close()
        self.dual.close()


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
        self.dual_params = {k: params[k] for k in ("optimisation_mode", "checkpointing", "checkpoint_every") if k in params}
        dual = NavierStokesVelVortPerturbationDual.FromNavierStokesVelVortPerturbation(nse, **self.dual_params)
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
        """The dual of the forward equation started from u0; the equation
        object, its plans and the dual plans are reused across iterations."""
        vh = VectorField.FromData(FourierField, self.physical_domain, u0, "velocity_hat")
        self.nse.set_initial_field("velocity_hat", vh)
        return NavierStokesVelVortPerturbationDual.FromNavierStokesVelVortPerturbation(self.nse, **self.dual_params)

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


def _rdot(a: torch.Tensor, b: torch.Tensor) -> float:
    """Real inner product of two spectra viewed as real vectors."""
    return float(torch.sum(a.real * b.real + a.imag * b.imag))


class OptimiserWrapper:
    """Restates OptimiserWrapper (gradient_descent_solver.py:862-934), the
    driver that hands the gain and its dual gradient to an external
    optimiser: jaxopt.NonlinearCG with a zoom line search (optimiser.py:231-242,
    use_optax = False).  jaxopt is not available here, so the same method is
    written out from Nocedal & Wright, Numerical Optimization (2006):

    * the objective is the reference's, f = -gain (run_adjoint,
      minimise = False).  Its gradient is the exact one: the discrete adjoint
      of the step chained with the reference's gradient parameterisation
      (optimiser.gain_and_grad_parameters; parameters omega_y, v, u_00,
      w_00 through vort_yvel_to_vel, navier_stokes.py:1063-1203).  The
      parameterisation keeps every iterate solenoidal and no-slip, and the
      gain is scale-invariant in the parameters.  The reference's wrapper
      passes the velocity spectrum itself as parameters, with
      -get_grad = -(u0 - v0)/E0 of the continuous-adjoint (dual) solution.
      Two things go wrong when that pair is used as written:
        - that vector is not the Euclidean gradient a Wolfe line search
          needs (it is an ascent direction in the energy metric), so no step
          along it satisfies sufficient decrease;
        - the exact gradient over the raw spectrum leaves the solenoidal
          subspace (the step sees only v and omega_y) and inflates the gain
          without bound;
    * nonlinear conjugate gradients, Polak-Ribiere with beta clipped at 0
      (PR+), restarting along the steepest direction when the direction is
      not a descent direction.  By default the directions are
      preconditioned by the inverse energy metric of the parameters
      (Q^-1 g, Q the quadratic form of the initial energy in the
      parameters, applied by an inner conjugate-gradient solve): the gain
      is a Rayleigh quotient whose denominator is that form, and without it
      the iteration crawls.  On the reference's y-profile case (Re 3000,
      alpha = 1, T = 15, Ny = 90, starting from the 20-mode stability
      optimum at gain 1.242) it reaches 20.24 after one iteration and the
      optimum 20.366 after four, against 19.02 after twenty without it (the
      reference publishes 1.24 -> 19.29 in twenty).  Run on, both find a
      larger gain that the Orr-Sommerfeld analysis does not have: an
      unstable discrete mode of the collocation step at high Chebyshev
      degrees (the y products are not dealiased), which the oracle's
      restatement of the reference step reproduces;
    * the step along d satisfies the strong Wolfe conditions (c1 = 1e-4,
      c2 = 0.1, the curvature constant Nocedal & Wright give for nonlinear
      CG) by bracketing and zoom (Algorithms 3.5 and 3.6) with
      safeguarded cubic interpolation, at most maxls = 15 evaluations,
      starting from the previous accepted step ("current" initialisation);
      if no point in the budget satisfies both conditions, the lowest
      sufficient-decrease point is taken, and none at all means no move and
      a restart along -g.

    After each iteration current_guess is N(x) and value its gain
    (update, :936-944)."""

    c1 = 1e-4
    c2 = 0.1

    def __init__(self, nse: NavierStokesVelVortPerturbation, **params: Any):
        self.nse = nse
        self.number_of_steps = int(params.get("max_iterations", params.get("max_iter", 20)))
        self.maxls = int(params.get("maxls", 15))
        self.c2 = float(params.get("c2", self.c2))
        self.step_size = float(params.get("step_size", 1e-2))
        self.re = nse.get_Re_tau() * nse.get_u_max_over_u_tau()
        self.dt = nse.get_dt()
        self.end_time = nse.end_time
        self.linearise = nse.linearise()
        self.physical_domain = nse.get_physical_domain()
        u_init = nse.get_initial_field("velocity_hat").get_data()
        fplan = nse.plan()
        # the discrete adjoint: linearised perturbation plans always carry it,
        # nonlinear ones are rebuilt with adjoint = True
        self.plan = fplan if self.linearise else StepPlan(
            fplan.grid, fplan.re_tau, fplan.dt, fplan.kind, False, nse.u_max_over_u_tau, ops=fplan.ops, adjoint=True
        )
        self.nsteps = nse.number_of_time_steps()
        self.checkpoint_every = int(params.get("checkpoint_every", 50))
        self.e_0 = self.plan.energy(u_init)
        g = self.plan.grid
        dev = u_init.device
        kx = torch.as_tensor(g.kx, device=dev)[:, None, None]
        kz = torch.as_tensor(g.kz, device=dev)[None, None, :]
        # (omega_y, v, u_00, w_00) of the initial field
        init = [1j * kz * u_init[0] - 1j * kx * u_init[2], u_init[1], u_init[0, 0, :, 0], u_init[2, 0, :, 0]]
        self._shapes = [t.shape for t in init]
        # optional restriction to some (kx, kz) pencils (stored indices): omega_y
        # and v elsewhere stay zero, the mean-flow profiles stay free only if
        # (0, 0) is listed -- the reference's y-profile optimisation frees the
        # v profile of the alpha = 1 pencil and u_00 (examples.py:1634-1760)
        modes = params.get("modes")
        self._mask: Optional[torch.Tensor] = None
        if modes is not None:
            m3 = torch.zeros(u_init.shape[1:], dtype=torch.bool, device=dev)
            for i, j in modes:
                m3[i, :, j] = True
            m00 = torch.full((u_init.shape[2],), bool(m3[0, 0, 0]), dtype=torch.bool, device=dev)
            self._mask = self._pack([m3, m3, m00, m00]).to(torch.complex128)
        self.x = self._masked(self._pack(init))
        from .parameterisation import VelocityParameterisation

        self._param = VelocityParameterisation(g, dev)
        self.evaluations = 0
        # preconditioning by the energy metric of the parameters (below)
        self.precondition = bool(params.get("precondition", True))
        self.pc_iterations = int(params.get("pc_iterations", 40))
        self.f, self.g = self.value_and_grad(self.x)
        self.y = self._precondition(self.g)
        self.d = -self.y
        if self.precondition:
            # the first trial step moves the parameters by a tenth of their size
            self.step_size = 0.1 * float(torch.linalg.norm(self.x)) / max(float(torch.linalg.norm(self.d)), 1e-300)
        self.current_guess = self.velocity(self.x)
        self.value = -self.f
        self.values: List[float] = [self.value]
        self.steps: List[float] = []
        self.wolfe: List[bool] = []
        self.i = 0

    def _pack(self, ts) -> torch.Tensor:
        return torch.cat([t.reshape(-1) for t in ts])

    def _masked(self, x: torch.Tensor) -> torch.Tensor:
        return x if self._mask is None else x * self._mask

    def _unpack(self, x: torch.Tensor):
        out, o = [], 0
        for sh in self._shapes:
            n = int(np.prod(sh))
            out.append(x[o : o + n].reshape(sh))
            o += n
        return out

    def normalize_field(self, v: torch.Tensor) -> torch.Tensor:
        return v * (self.e_0 / self.plan.energy(v)) ** 0.5

    def velocity(self, x: torch.Tensor) -> torch.Tensor:
        """The initial velocity spectrum of parameters x at the initial energy
        (parameters_to_run_input and normalize_current_guess)."""
        with torch.no_grad():
            return self.normalize_field(self._param(*self._unpack(x)))

    def _energy_gradient(self, x: torch.Tensor) -> torch.Tensor:
        """Gradient in the parameters of E(u(x)), the initial energy.  E is a
        quadratic form x^T Q x of the parameters (u is linear in them), so
        Q g = grad E(g) / 2."""
        from .vjp import energy

        leaves = [torch.view_as_real(p.contiguous().clone()).requires_grad_(True) for p in self._unpack(x)]
        u = self._param(*[torch.view_as_complex(t) for t in leaves])
        e = energy(torch.view_as_real(u), self.plan)
        grads = torch.autograd.grad(e, leaves)
        return self._masked(self._pack([torch.view_as_complex(t.contiguous()) for t in grads]))

    def _precondition(self, g: torch.Tensor) -> torch.Tensor:
        """y = Q^-1 g by conjugate gradients on Q (the energy metric of the
        parameters; without preconditioning y = g).  The gain is a Rayleigh
        quotient N(x) / x^T Q x: in the metric of Q its level sets are round
        in the denominator, which the parameters' own Euclidean metric (v
        enters u through D, so high Chebyshev degrees weigh ~Ny^2 more)
        is far from.  g lies in the range of Q (both are built from the
        parameterisation's transpose), so the singular directions of Q
        (parameters the velocity does not see) do not matter."""
        if not self.precondition:
            return g
        y = torch.zeros_like(g)
        r = g.clone()
        p = r.clone()
        rr = _rdot(r, r)
        rr0 = rr
        for _ in range(self.pc_iterations):
            qp = 0.5 * self._energy_gradient(p)
            pq = _rdot(p, qp)
            if pq <= 0.0:
                break
            a = rr / pq
            y = y + a * p
            r = r - a * qp
            rr_new = _rdot(r, r)
            if rr_new <= 1e-20 * rr0:
                break
            p = r + (rr_new / rr) * p
            rr = rr_new
        return y

    def value_and_grad(self, x: torch.Tensor):
        """run_adjoint (gradient_descent_solver.py:880-906): -gain of the
        parameters' initial field and its exact gradient."""
        from .optimiser import gain_and_grad_parameters

        gain, grads = gain_and_grad_parameters(self.plan, self._unpack(x), self.nsteps, self.checkpoint_every)
        self.evaluations += 1
        return -gain, -self._masked(self._pack(grads))

    def _line_search(self, d: torch.Tensor, dphi0: float):
        """Strong Wolfe step along d from (self.x, self.f); returns
        (alpha, f, g, wolfe) or None."""
        f0 = self.f
        evals = [0]
        best = [None]

        def phi(a: float):
            f, g = self.value_and_grad(self.x + a * d)
            evals[0] += 1
            if f <= f0 + self.c1 * a * dphi0 and (best[0] is None or f < best[0][1]):
                best[0] = (a, f, g, False)
            return f, g, _rdot(g, d)

        def zoom(lo, f_lo, dp_lo, hi, f_hi, dp_hi):
            while evals[0] < self.maxls:
                # cubic interpolation of phi on [lo, hi], kept inside the
                # middle 80 % of the interval, else bisection
                d1 = dp_lo + dp_hi - 3.0 * (f_lo - f_hi) / (lo - hi)
                rad = d1 * d1 - dp_lo * dp_hi
                a = 0.5 * (lo + hi)
                if rad >= 0.0:
                    d2 = (1.0 if hi > lo else -1.0) * rad**0.5
                    den = dp_hi - dp_lo + 2.0 * d2
                    if den != 0.0:
                        c = hi - (hi - lo) * (dp_hi + d2 - d1) / den
                        lo_b, hi_b = min(lo, hi), max(lo, hi)
                        w = hi_b - lo_b
                        if lo_b + 0.1 * w <= c <= hi_b - 0.1 * w:
                            a = c
                f, g, dp = phi(a)
                if f > f0 + self.c1 * a * dphi0 or f >= f_lo:
                    hi, f_hi, dp_hi = a, f, dp
                else:
                    if abs(dp) <= -self.c2 * dphi0:
                        return a, f, g, True
                    if dp * (hi - lo) >= 0.0:
                        hi, f_hi, dp_hi = lo, f_lo, dp_lo
                    lo, f_lo, dp_lo = a, f, dp
            return best[0]

        a_prev, f_prev, dp_prev = 0.0, f0, dphi0
        a = self.step_size
        while evals[0] < self.maxls:
            f, g, dp = phi(a)
            if f > f0 + self.c1 * a * dphi0 or (a_prev > 0.0 and f >= f_prev):
                return zoom(a_prev, f_prev, dp_prev, a, f, dp)
            if abs(dp) <= -self.c2 * dphi0:
                return a, f, g, True
            if dp >= 0.0:
                return zoom(a, f, dp, a_prev, f_prev, dp_prev)
            a_prev, f_prev, dp_prev = a, f, dp
            a *= 2.0
        return best[0]

    def update(self) -> None:
        d = self.d
        dphi0 = _rdot(self.g, d)
        if dphi0 >= 0.0:  # not a descent direction: restart
            d = -self.y
            dphi0 = _rdot(self.g, d)
        res = self._line_search(d, dphi0)
        if res is None:
            self.d = -self.y
            self.wolfe.append(False)
        else:
            a, f, g, wolfe = res
            y = self._precondition(g)
            # preconditioned Polak-Ribiere, clipped at 0 (PR+)
            beta = max(_rdot(g, y - self.y) / _rdot(self.g, self.y), 0.0)
            self.x = self.x + a * d
            self.f, self.step_size = f, a
            self.d = -y + beta * d
            self.g, self.y = g, y
            self.steps.append(a)
            self.wolfe.append(wolfe)
        self.current_guess = self.velocity(self.x)
        self.value = -self.f
        self.values.append(self.value)
        self.i += 1

    def optimise(self, iterations: Optional[int] = None) -> List[float]:
        for _ in range(self.number_of_steps if iterations is None else iterations):
            self.update()
        return self.values
