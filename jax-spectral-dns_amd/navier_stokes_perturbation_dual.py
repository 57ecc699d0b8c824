"""NavierStokesVelVortPerturbationDual: the continuous adjoint of the
perturbation step (reference navier_stokes_perturbation_dual.py:376-1212).

The dual is the same RK3/CN stepper with Re -> -Re and dt -> -dt, reusing the
forward operator tables (dt L is unchanged; navier_stokes_perturbation_dual.py:
470-471, :542-589), with the adjoint nonlinear term linearised about the
forward state history[-1-t] (navier_stokes_perturbation_dual.py:591-602).
The forward trajectory is kept on the GPU either whole or, with
checkpointing, every ``inner`` steps and recomputed chunk by chunk before
each adjoint sweep (navier_stokes_perturbation_dual.py:690-757, 990-1035).
The gain gradient is (u0 - v0) / E0 (navier_stokes_perturbation_dual.py:
1106-1120)."""

from __future__ import annotations

import math
import os
from typing import Any, Dict, List, Optional, Tuple

import numpy as np
import torch

from . import _lib
from .equation import write_trajectory
from .field import FourierField, VectorField
from .navier_stokes import median_factor
from .navier_stokes_perturbation import NavierStokesVelVortPerturbation
from .plan import StepPlan


OPTIMISATION_MODES = ("gain", "gain_3d", "dissipation", "gain_no_00")


class NavierStokesVelVortPerturbationDual:
    name = "Dual Navier Stokes equation (velocity-vorticity formulation) for perturbations on top of a base flow."

    def __init__(self, forward_equation: NavierStokesVelVortPerturbation, **params: Any):
        self.forward_equation = forward_equation
        mode = params.get("optimisation_mode", "gain")
        if mode not in OPTIMISATION_MODES:
            raise Exception("unknown optimisation mode " + str(mode) + ". Valid choices are " + str(OPTIMISATION_MODES))
        self.optimisation_mode = mode
        fwd = forward_equation
        self.Re_tau = -fwd.get_Re_tau()
        self.dt = -fwd.get_dt()
        self.end_time = -fwd.end_time
        n = fwd.number_of_time_steps()
        # the dual's switches mirror the forward ones in dual time t (step
        # n - 1 - t of the forward run; navier_stokes_perturbation_dual.py:436-466)
        n_sw = len(np.arange(0, abs(self.end_time), abs(self.dt))) if self.end_time else 0
        if getattr(fwd, "linearise_switch", None) is not None:
            sw = float(fwd.linearise_switch)
            self.linearise = lambda t=0: t + 1 >= n_sw * (1.0 - sw)
        else:
            lin = bool(fwd.linearise(0))
            self.linearise = lambda _t=0: lin
        if getattr(fwd, "coupling_term_switch", None) is not None:
            csw = float(fwd.coupling_term_switch)
            self.coupling_term = lambda t=0: not (t + 1 >= n_sw * csw)
        else:
            # coupling_term = False: h_v gains (d_x^2 + d_z^2) hat(U' v_x)
            # (navier_stokes_perturbation_dual.py:269-278)
            cpl = bool(params.get("coupling_term", fwd.coupling_term(0)))
            self.coupling_term = lambda _t=0: cpl
        self.linearise_flag = bool(self.linearise(0))
        calc = abs(self.end_time / self.dt) if self.dt else 0.0
        for sh in fwd.physical_domain.get_shape():
            calc *= sh
        # checkpointing threshold of the reference (navier_stokes_perturbation_dual.py:420-434)
        self.checkpointing = params.get("checkpointing", calc > params.get("checkpointing_threshold", 1.4e8))
        self.number_of_time_steps = n
        self.number_of_inner_steps = params.get("checkpoint_every", median_factor(n) if n > 0 else 1)
        if n % self.number_of_inner_steps != 0:
            raise ValueError("checkpoint interval must divide the number of steps")
        self.number_of_outer_steps = n // self.number_of_inner_steps
        self._plans: Dict[Tuple[bool, bool], StepPlan] = {}
        self.velocity_u_hat_0: Optional[torch.Tensor] = None
        self.u_T: Optional[torch.Tensor] = None
        self.v_0: Optional[torch.Tensor] = None
        self._checkpoints: List[torch.Tensor] = []
        self._history: List[torch.Tensor] = []
        self._dissipation_sum: Optional[torch.Tensor] = None
        self.gain: Optional[float] = None
        # set by the optimisation drivers (gradient_descent_solver.py:547-629)
        self.write_trajectory = False

    @classmethod
    def FromNavierStokesVelVortPerturbation(cls, nse: NavierStokesVelVortPerturbation, **params: Any):
        return cls(nse, **params)

    def plan_for_step(self, t: int) -> StepPlan:
        """Dual plan of dual step t: Re -> -Re and dt -> -dt on the forward
        operator tables (navier_stokes_perturbation_dual.py:470-471)."""
        key = (bool(self.linearise(t)), bool(self.coupling_term(t)))
        if key not in self._plans:
            # kept on the forward equation, so the duals an optimisation
            # driver builds from it every iteration share the plans
            cache = self.forward_equation.__dict__.setdefault("_dual_plan_cache", {})
            if key in cache:
                self._plans[key] = cache[key]
                return cache[key]
            fplan = self.forward_equation.plan()
            self._plans[key] = cache[key] = StepPlan(
                fplan.grid,
                fplan.re_tau,
                fplan.dt,
                _lib.KIND_DUAL,
                key[0],
                self.forward_equation.u_max_over_u_tau,
                ops=fplan.ops,
                explicit_dt=self.dt,
                plan_re=self.Re_tau,
                coupling_term=key[1],
                base_profile=fplan.base_profile,
            )
        return self._plans[key]

    def plan(self) -> StepPlan:
        return self.plan_for_step(0)

    def get_Re_tau(self) -> float:
        return self.Re_tau

    def _needs_history(self) -> bool:
        """The forward states enter the dual steps when any dual step is
        nonlinear, and the dissipation source at every step."""
        n = self.number_of_time_steps
        return self.optimisation_mode == "dissipation" or not all(self.linearise(t) for t in range(n))

    # forward (navier_stokes_perturbation_dual.py:850-939)
    def run_forward_calculation(self) -> None:
        fwd = self.forward_equation
        u = fwd._state(fwd.get_initial_field("velocity_hat").get_data()).clone()
        self.velocity_u_hat_0 = u.clone()
        n, inner = self.number_of_time_steps, self.number_of_inner_steps
        need_history = self._needs_history()
        self._history, self._checkpoints = [], []
        diss = self.optimisation_mode == "dissipation"
        if diss:
            # sum over the n + 1 forward states of the trajectory
            # (get_dissipation_average, navier_stokes_perturbation_dual.py:1087-1104)
            g = fwd.physical_domain.grid_math
            self._dissipation_sum = torch.zeros((g.nxm, g.ny, g.nzm), dtype=torch.complex128, device=u.device)
            fwd.plan().dissipation_accumulate(u, self._dissipation_sum)
        if need_history and not self.checkpointing:
            self._history.append(u.clone())
            for k in range(n):
                self._advance_forward(u, k, 1)
                self._history.append(u.clone())
                if diss:
                    fwd.plan().dissipation_accumulate(u, self._dissipation_sum)
        else:
            self._checkpoints.append(u.clone())
            step = 1 if diss else inner
            for k in range(0, n, step):
                self._advance_forward(u, k, step)
                if diss:
                    fwd.plan().dissipation_accumulate(u, self._dissipation_sum)
                if need_history and (k + step) % inner == 0:
                    self._checkpoints.append(u.clone())
        self.u_T = u
        if os.environ.get("JAX_SPECTRAL_DNS_FIELD_DIR") is not None and self.write_trajectory:
            # the forward states this run keeps: every step without
            # checkpointing, else the initial and final states and the
            # checkpoints (navier_stokes_perturbation_dual.py:877-889)
            states = self._history or (self._checkpoints if len(self._checkpoints) > 1 else [self.velocity_u_hat_0, u])
            write_trajectory(torch.stack(states))

    def _advance_forward(self, u: torch.Tensor, k0: int, nsteps: int) -> None:
        """Forward steps k0 .. k0 + nsteps - 1 in place (with the forward
        equation's switches, source hook and enforce_base)."""
        self.forward_equation._advance(u, 0.0, k0, nsteps)

    def get_source_term(self, t: int, fwd_state: Optional[torch.Tensor]) -> Optional[torch.Tensor]:
        """navier_stokes_perturbation_dual.py:962-988: none for the gain
        objectives; -laplacian(u(t)) / |Re_tau T| of the forward state for
        the dissipation objective."""
        if self.optimisation_mode != "dissipation":
            return None
        assert fwd_state is not None
        scale = -1.0 / abs(self.Re_tau * self.end_time)
        return self.plan().laplacian(fwd_state, scale)

    def _dual_step(self, v: torch.Tensor, t: int, fwd_state: Optional[torch.Tensor]) -> None:
        """Dual step t (0: from the final time) in place."""
        plan = self.plan_for_step(t)
        src = self.get_source_term(t, fwd_state)
        lin = self.linearise(t)
        fwd_arg = None if lin else fwd_state
        if src is None and not self.forward_equation.enforce_base:
            plan.dual_step(v, fwd_arg)
        else:
            plan.step_source(v, 0.0, src, self.forward_equation.enforce_base, fwd_arg)

    def _forward_state(self, t: int, chunk: List[torch.Tensor], start: int) -> Optional[torch.Tensor]:
        if not self._needs_history():
            return None
        if not self.checkpointing:
            return self._history[-1 - t]
        return chunk[-1 - (t - start)]

    def _dual_initial_condition(self) -> torch.Tensor:
        """navier_stokes_perturbation_dual.py:903-935: -u(T) for the gain,
        without its kx = 0 part (gain_3d) or its (0,0) pencil (gain_no_00),
        zero for the dissipation."""
        assert self.u_T is not None
        v = -self.u_T.clone()
        if self.optimisation_mode == "gain_3d":
            v[:, 0] = 0
        elif self.optimisation_mode == "gain_no_00":
            v[:, 0, :, 0] = 0
        elif self.optimisation_mode == "dissipation":
            v.zero_()
        return v

    # backward (navier_stokes_perturbation_dual.py:690-839, 1037-1056)
    def run_backward_calculation(self) -> torch.Tensor:
        if self.u_T is None:
            self.run_forward_calculation()
        assert self.u_T is not None
        v = self._dual_initial_condition()
        n, inner = self.number_of_time_steps, self.number_of_inner_steps
        if not self._needs_history() or not self.checkpointing:
            for t in range(n):
                self._dual_step(v, t, self._forward_state(t, [], 0))
        else:
            for o in range(self.number_of_outer_steps):
                u = self._checkpoints[-1 - (o + 1)].clone()
                k0 = n - (o + 1) * inner  # forward step index of the chunk's start
                chunk = [u.clone()]
                for j in range(inner):
                    self._advance_forward(u, k0 + j, 1)
                    chunk.append(u.clone())
                for t in range(o * inner, (o + 1) * inner):
                    self._dual_step(v, t, self._forward_state(t, chunk, o * inner))
        self.v_0 = v
        return v

    def get_gain(self) -> float:
        if self.u_T is None:
            self.run_forward_calculation()
        p = self.forward_equation.plan()
        self.gain = p.energy(self.u_T) / p.energy(self.velocity_u_hat_0)
        return self.gain

    def get_gain_3d(self) -> float:
        """E(u(T) - its kx = 0 part) / E(u(0)) (navier_stokes_perturbation_dual.py:1065-1077)."""
        if self.u_T is None:
            self.run_forward_calculation()
        p = self.forward_equation.plan()
        u3 = self.u_T.clone()
        u3[:, 0] = 0
        self.gain = p.energy(u3) / p.energy(self.velocity_u_hat_0)
        return self.gain

    def get_gain_no_00(self) -> float:
        """E(u(T) - its (0,0) pencil) / E(u(0)) (navier_stokes_perturbation_dual.py:1079-1089)."""
        if self.u_T is None:
            self.run_forward_calculation()
        p = self.forward_equation.plan()
        u3 = self.u_T.clone()
        u3[:, 0, :, 0] = 0
        self.gain = p.energy(u3) / p.energy(self.velocity_u_hat_0)
        return self.gain

    def get_dissipation_average(self) -> float:
        """Volume integral of no_hat(sum_states sum_ij (d_j u_i)^2) / |Re_tau (n + 1)|
        (navier_stokes_perturbation_dual.py:1087-1104; the squares are taken
        of the spectral coefficients, as there)."""
        if self.u_T is None:
            self.run_forward_calculation()
        assert self._dissipation_sum is not None
        n_states = self.number_of_time_steps + 1
        return self.plan().volume_integral(self._dissipation_sum) / abs(self.Re_tau * n_states)

    def get_objective_fun(self) -> float:
        """navier_stokes_perturbation_dual.py:941-957."""
        if self.optimisation_mode == "gain":
            return self.get_gain()
        if self.optimisation_mode == "gain_3d":
            return self.get_gain_3d()
        if self.optimisation_mode == "gain_no_00":
            return self.get_gain_no_00()
        self.gain = self.get_gain()
        return self.get_dissipation_average()

    def get_objective_fun_name(self) -> str:
        return self.optimisation_mode

    def get_grad(self, u_hat_0: Optional[torch.Tensor] = None, v_hat_0: Optional[torch.Tensor] = None) -> torch.Tensor:
        """(u0 - v0) / E0 (navier_stokes_perturbation_dual.py:1106-1120)."""
        if u_hat_0 is None and v_hat_0 is None:
            self.run_backward_calculation()
        u0 = self.velocity_u_hat_0 if u_hat_0 is None else u_hat_0
        v0 = self.v_0 if v_hat_0 is None else v_hat_0
        self.get_gain()
        e0 = self.forward_equation.plan().energy(u0)
        return (u0 - v0) / e0

    def get_projected_grad(self, step_size: float, step_size_factor: Optional[float] = None) -> Tuple[torch.Tensor, bool]:
        """Energy-constrained update direction (navier_stokes_perturbation_dual.py:1122-1167)."""
        if self.v_0 is None:
            self.run_backward_calculation()
        return projected_gradient(self.forward_equation.plan(), self.velocity_u_hat_0, self.v_0, step_size,
                                  step_size_factor), True


    def get_projected_cg_grad(
        self,
        step_size: float,
        beta: complex,
        old_grad: torch.Tensor,
        step_size_factor: Optional[float] = None,
    ) -> Tuple[torch.Tensor, bool]:
        """Conjugate-gradient direction under the energy constraint
        (navier_stokes_perturbation_dual.py:1169-1201): the projected gradient
        with the adjoint field v_0 replaced by v_0 - beta * old_grad."""
        if self.v_0 is None:
            self.run_backward_calculation()
        f = 0.9 if step_size_factor is None else step_size_factor
        p = self.forward_equation.plan()
        u0 = self.velocity_u_hat_0
        v0 = self.v_0 - beta * old_grad
        e0 = p.energy(u0)
        e0_adj = p.energy(v0)
        c0 = 2 * 0.5 * (p.energy(u0 + v0) - e0 - e0_adj)  # 2 <u0, v0>_E
        max_step = f * (e0_adj / e0 - c0**2 / (4 * e0**2)) ** (-1 / 2)
        step_size = min(step_size, max_step)
        A = (step_size * c0 - 2 * e0) / (2 * step_size * e0)
        lam = A + math.sqrt(A**2 - (step_size * e0_adj - c0) / (step_size * e0))
        return lam * u0 + (-1.0 * self.v_0 + beta * old_grad), True

    def get_max_step_size(self, beta: complex = 0.0, old_grad: Optional[torch.Tensor] = None) -> float:
        """Largest step the energy constraint admits
        (gradient_descent_solver.py:510-520, ConjugateGradientDescentSolver.get_max_step_size_ls)."""
        if self.v_0 is None:
            self.run_backward_calculation()
        p = self.forward_equation.plan()
        u0 = self.velocity_u_hat_0
        v0 = self.v_0 if old_grad is None else self.v_0 - beta * old_grad
        e0 = p.energy(u0)
        e0_adj = p.energy(v0)
        c0 = 2 * 0.5 * (p.energy(u0 + v0) - e0 - e0_adj)
        return (e0_adj / e0 - c0**2 / (4 * e0**2)) ** (-1 / 2)


def projected_gradient(
    plan: StepPlan, u0: torch.Tensor, v0: torch.Tensor, step_size: float, step_size_factor: Optional[float] = None
) -> torch.Tensor:
    """lam u0 - v0 with lam such that a step of step_size (capped at
    step_size_factor times the largest admissible one) keeps the energy of
    u0 (get_projected_grad_from_u_and_v, navier_stokes_perturbation_dual.py:1122-1167)."""
    f = 0.9 if step_size_factor is None else step_size_factor
    e0 = plan.energy(u0)
    e0_adj = plan.energy(v0)
    c0 = 2 * 0.5 * (plan.energy(u0 + v0) - e0 - e0_adj)  # 2 <u0, v0>_E
    max_step = f * (e0_adj / e0 - c0**2 / (4 * e0**2)) ** (-1 / 2)
    step_size = min(step_size, max_step)
    A = (step_size * c0 - 2 * e0) / (2 * step_size * e0)
    lam = A + math.sqrt(A**2 - (step_size * e0_adj - c0) / (step_size * e0))
    return lam * u0 - v0


def perform_step_navier_stokes_perturbation_dual(nse: NavierStokesVelVortPerturbation, step_size: float = 1e-2):
    """navier_stokes_perturbation_dual.py:1204-1212."""
    dual = NavierStokesVelVortPerturbationDual.FromNavierStokesVelVortPerturbation(nse)
    return dual.get_gain(), dual.get_projected_grad(step_size)[0]
