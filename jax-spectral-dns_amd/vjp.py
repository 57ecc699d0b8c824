"""Discrete adjoint of the step behind ``torch.autograd`` (the build's
counterpart of reverse-mode AD through the reference's scan,
jax.value_and_grad in optimiser.py:115-121): the linearised perturbation
step, and the NS / nonlinear perturbation step linearised about its forward
state.

``run(u0, plan, nsteps, checkpoint_every)`` advances a complex128 state
(given as its float64 real-pair view, shape (3, Nx, Ny, Nz, 2)) through
nsteps RK3/CN steps on the GPU; its backward applies the hand-written
transposed kernels (``jsd_step_vjp``).  The forward keeps a checkpoint every
``checkpoint_every`` steps and the backward walks the chunks in reverse,
recomputing a chunk's step-start states from its checkpoint when the
adjoint needs them (nonlinear steps: ``jsd_step_vjp_state`` then recomputes
the three substep states of each step); the linearised step is linear, so
its transpose needs no forward state and nothing is stored or recomputed.
``checkpoint_every=1`` keeps every step-start state (one state of 3 x Nx x Ny
x Nz complex128 per step, 48.6 MB at 96x129x80, so a 288 GB GPU holds
thousands) and the backward recomputes no chunk; the substep states inside a
step are still recomputed by ``jsd_step_vjp_state``.  ``checkpoint_every=0``
records every step (``jsd_step_record``: its start state, two substep states
and the kept transforms, 0.47 GB a step at 96x129x80) and the backward
recomputes nothing (``jsd_step_vjp_recorded``).
``energy(u, plan)`` is the energy functional with its gradient from
``jsd_energy_grad``; ``gain_and_grad`` evaluates J = E(u_T)/E(u_0) and
dJ/du_0.
"""

from __future__ import annotations

import ctypes
from typing import Tuple

import torch

from . import _lib
from .plan import StepPlan, _dev, _stream


def _vjp(plan: StepPlan, g: torch.Tensor, nsteps: int) -> None:
    _lib.check(_lib.load().jsd_step_vjp(plan._h, _dev(g), int(nsteps), _stream()))


def _energy_grad(plan: StepPlan, u: torch.Tensor, scale: float) -> torch.Tensor:
    out = torch.empty_like(u)
    _lib.check(_lib.load().jsd_energy_grad(plan._h, _dev(u), _dev(out), float(scale), _stream()))
    return out


class _Run(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u0r: torch.Tensor, plan: StepPlan, nsteps: int, checkpoint_every: int, dPdx: float):  # type: ignore[override]
        u = torch.view_as_complex(u0r.contiguous()).clone()
        ctx.recs = None
        if checkpoint_every == 0 and plan.adjoint_needs_state:
            # every step recorded: its start state and its record (substep
            # states and kept transforms); the backward recomputes nothing
            n = plan.record_elems()
            recs = []
            for _ in range(nsteps):
                st = u.clone()
                rec = torch.empty(n, dtype=torch.complex128, device=u.device)
                plan.step_record(u, rec, dPdx)
                recs.append((st, rec))
            ctx.plan = plan
            ctx.recs = recs
            ctx.dPdx = dPdx
            return torch.view_as_real(u).clone()
        k = max(1, min(checkpoint_every, max(1, nsteps)))
        chunks = []
        done = 0
        while done < nsteps:
            c = min(k, nsteps - done)
            chunks.append((done, c, u.clone() if plan.adjoint_needs_state else None))
            plan.step(u, dPdx, c)
            done += c
        ctx.plan = plan
        ctx.chunks = chunks
        ctx.dPdx = dPdx
        return torch.view_as_real(u).clone()

    @staticmethod
    def backward(ctx, gr: torch.Tensor):  # type: ignore[override]
        plan: StepPlan = ctx.plan
        g = torch.view_as_complex(gr.contiguous()).clone()
        if ctx.recs is not None:
            for st, rec in reversed(ctx.recs):
                plan.vjp_recorded(st, rec, g, ctx.dPdx)
            ctx.recs = None  # the records are consumed
            return torch.view_as_real(g), None, None, None, None
        for start, count, state in reversed(ctx.chunks):
            if state is None:
                _vjp(plan, g, count)
                continue
            # recompute the chunk's step-start states from its checkpoint, then
            # walk its steps backwards
            states = [state]
            x = state.clone()
            for _ in range(count - 1):
                plan.step(x, ctx.dPdx, 1)
                states.append(x.clone())
            for st in reversed(states):
                plan.vjp_state(st, g, ctx.dPdx)
        return torch.view_as_real(g), None, None, None, None


class _Energy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, ur: torch.Tensor, plan: StepPlan):  # type: ignore[override]
        u = torch.view_as_complex(ur.contiguous())
        ctx.plan = plan
        ctx.save_for_backward(ur)
        return torch.tensor(plan.energy(u), dtype=torch.float64, device=ur.device)

    @staticmethod
    def backward(ctx, ge: torch.Tensor):  # type: ignore[override]
        (ur,) = ctx.saved_tensors
        u = torch.view_as_complex(ur.contiguous())
        g = _energy_grad(ctx.plan, u, float(ge))
        return torch.view_as_real(g), None


def run(u0r: torch.Tensor, plan: StepPlan, nsteps: int, checkpoint_every: int = 50, dPdx: float = 0.0) -> torch.Tensor:
    linear = plan.kind == _lib.KIND_PERTURBATION and plan.linearise
    if not linear and not (plan.kind in (_lib.KIND_NS, _lib.KIND_PERTURBATION) and plan.has_adjoint):
        raise _lib.JsdError(
            "the discrete adjoint needs a linearised perturbation plan, or an NS / perturbation plan built with adjoint=True"
        )
    return _Run.apply(u0r, plan, nsteps, checkpoint_every, float(dPdx))


def energy(ur: torch.Tensor, plan: StepPlan) -> torch.Tensor:
    return _Energy.apply(ur, plan)


def gain_and_grad(
    u0: torch.Tensor, plan: StepPlan, nsteps: int, checkpoint_every: int = 50, dPdx: float = 0.0
) -> Tuple[float, torch.Tensor]:
    """J = E(u_T) / E(u_0) and dJ/du_0 (complex128, real-pair convention:
    d/dRe + i d/dIm)."""
    x = torch.view_as_real(u0.contiguous()).clone().requires_grad_(True)
    J = energy(run(x, plan, nsteps, checkpoint_every, dPdx), plan) / energy(x, plan)
    (gx,) = torch.autograd.grad(J, x)
    return float(J.detach()), torch.view_as_complex(gx.contiguous())
