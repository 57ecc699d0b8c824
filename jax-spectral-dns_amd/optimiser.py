"""Gain and gradient with respect to the reference's optimisation
parameters (omega_y, v, u_00, w_00), the quantity jax.value_and_grad returns
in optimiser.py:115-121 for the linear optimal-perturbation problem.

u_0 = vort_yvel_to_vel(params) (``parameterisation``), u_T = n steps of the
plan (``vjp.run``: hand-written discrete adjoint, checkpoint + recompute),
J = E(u_T) / E(u_0); torch.autograd chains the adjoint of the steps with
the parameterisation.
"""

from __future__ import annotations

from typing import List, Sequence, Tuple

import torch

from .parameterisation import VelocityParameterisation
from .plan import StepPlan
from .vjp import energy, run


def gain_and_grad_parameters(
    plan: StepPlan,
    params: Sequence[torch.Tensor],
    nsteps: int,
    checkpoint_every: int = 50,
    dPdx: float = 0.0,
) -> Tuple[float, List[torch.Tensor]]:
    """params = (vort (Nx,Ny,Nz), vel_y (Nx,Ny,Nz), vel_x_00 (Ny), vel_z_00 (Ny)),
    complex128 on the GPU.  Returns J and dJ/dparam for each (real-pair
    convention d/dRe + i d/dIm)."""
    dev = params[0].device
    param = VelocityParameterisation(plan.grid, dev)
    leaves = [torch.view_as_real(p.contiguous().clone()).requires_grad_(True) for p in params]
    u0 = param(*[torch.view_as_complex(x) for x in leaves])
    x = torch.view_as_real(u0)
    J = energy(run(x, plan, nsteps, checkpoint_every, dPdx), plan) / energy(x, plan)
    grads = torch.autograd.grad(J, leaves)
    return float(J.detach()), [torch.view_as_complex(g.contiguous()) for g in grads]
