"""Batches of independent runs on one GPU.

``BatchPlan`` binds a list of ``StepPlan`` objects (linearised perturbation
or dual plans of any shape, Re, dt and box) to their device state buffers and
advances all of them with one launch per stage (``jsd_batch_step``): the
stage kernels walk the concatenated per-run grids with an autotuned
persistent grid and one RK3 step is replayed from a captured HIP graph.  With
adjoint buffers it also runs the batched discrete adjoint
(``jsd_batch_step_vjp``).  Nonlinear perturbation plans form a batch whose
runs step concurrently, one stream each (``jsd_batch_step``), with the
adjoint taken about each run's forward state (``jsd_batch_step_vjp_state``).
The reference advances such runs one at a time
(a loop of perform_time_step, navier_stokes.py:1778-1813, per run);
``batch_gain_and_grad`` is the batched form of ``vjp.gain_and_grad`` for the
gradient evaluations of an optimisation driver.
"""

from __future__ import annotations

import ctypes
from typing import List, Optional, Sequence, Tuple

import torch

from . import _lib
from .plan import StepPlan, _dev, _stream

STAGES = ("ygemm", "nl_lin", "solve", "solve_T", "hv_T", "nl_lin_T", "acc")


def seeded_linearised_runs(n: int = 64, seed: int = 0) -> List[Tuple[Tuple[int, int, int], float, float, float]]:
    """The batch workload of uneven linearised runs, drawn from one seed:
    (requested shape, Re, Lx, Lz) with Re ~ U[500, 6000], alpha ~ U[0.5, 2],
    beta ~ U[0, 2], Ny in {65, 97, 129}, (Nx, Nz) in {(16, 16), (32, 16),
    (48, 32)}, Lx = 2 pi / alpha and Lz = 2 pi / beta (1e-3 for
    beta < 1e-3, the reference's quasi-2D box, examples.py:1303)."""
    import numpy as np

    rng = np.random.default_rng(seed)
    shapes = [(16, 16), (32, 16), (48, 32)]
    out = []
    for _ in range(n):
        re = rng.uniform(500.0, 6000.0)
        alpha = rng.uniform(0.5, 2.0)
        beta = rng.uniform(0.0, 2.0)
        ny = int(rng.choice([65, 97, 129]))
        nx, nz = shapes[int(rng.integers(0, 3))]
        lz = 2 * np.pi / beta if beta >= 1e-3 else 1e-3
        out.append(((nx, ny, nz), float(re), float(2 * np.pi / alpha), float(lz)))
    return out


def check_distinct(plans: Sequence[object], states: Sequence[torch.Tensor],
                   adjoints: Optional[Sequence[torch.Tensor]] = None) -> None:
    """Every run of a batch writes its own plan's workspaces inside each stage
    launch, so two runs on one plan (or one state / adjoint buffer) would race
    on them: reject repeated plans and buffers."""
    if len({id(p) for p in plans}) != len(plans):
        raise _lib.JsdError("a plan appears twice in the batch: each run needs its own StepPlan")
    bufs = list(states) + (list(adjoints) if adjoints is not None else [])
    ptrs = [b.data_ptr() for b in bufs]
    if len(set(ptrs)) != len(ptrs):
        raise _lib.JsdError("a state or adjoint buffer appears twice in the batch")


class BatchPlan:
    def __init__(
        self,
        plans: Sequence[StepPlan],
        states: Sequence[torch.Tensor],
        adjoints: Optional[Sequence[torch.Tensor]] = None,
    ):
        if len(plans) == 0 or len(plans) != len(states):
            raise _lib.JsdError("one state per plan")
        if adjoints is not None and len(adjoints) != len(plans):
            raise _lib.JsdError("one adjoint buffer per plan")
        check_distinct(plans, states, adjoints)
        for p, s in zip(plans, states):
            p._check_state(s)
        self.plans = list(plans)
        self.states = list(states)
        self.adjoints = list(adjoints) if adjoints is not None else None
        n = len(plans)
        vp = ctypes.c_void_p
        hp = (vp * n)(*[p._h for p in plans])
        sp = (vp * n)(*[_dev(s) for s in states])
        ap = (vp * n)(*[_dev(a) for a in self.adjoints]) if self.adjoints is not None else None
        h = vp()
        _lib.check(_lib.load().jsd_batch_create(hp, sp, ap, n, _stream(), ctypes.byref(h)))
        self._h = h

    def step(self, nsteps: int = 1) -> None:
        _lib.check(_lib.load().jsd_batch_step(self._h, int(nsteps), _stream()))

    def vjp(self, nsteps: int = 1) -> None:
        _lib.check(_lib.load().jsd_batch_step_vjp(self._h, int(nsteps), _stream()))

    def vjp_state(self, states: Sequence[torch.Tensor]) -> None:
        """One adjoint step of every run about its forward state at the start
        of the step (nonlinear runs; linearised batches ignore the states)."""
        if len(states) != len(self.plans):
            raise _lib.JsdError("one forward state per run")
        for p, st in zip(self.plans, states):
            p._check_state(st)
        vp = ctypes.c_void_p
        arr = (vp * len(states))(*[_dev(st) for st in states])
        _lib.check(_lib.load().jsd_batch_step_vjp_state(self._h, arr, _stream()))

    def info(self) -> dict:
        g = (ctypes.c_int * len(STAGES))()
        b = (ctypes.c_int * len(STAGES))()
        t = (ctypes.c_int * len(STAGES))()
        w = (ctypes.c_int * len(STAGES))()
        _lib.check(_lib.load().jsd_batch_info(self._h, g, b))
        _lib.check(_lib.load().jsd_batch_tiles(self._h, t, w))
        return {s: {"grid": g[i], "blocks": b[i], "tile": t[i], "waves": w[i]} for i, s in enumerate(STAGES)}

    def close(self) -> None:
        if getattr(self, "_h", None):
            _lib.load().jsd_batch_destroy(self._h)
            self._h = None

    def __del__(self) -> None:
        try:
            self.close()
        except Exception:
            pass


def _energy_grad(plan: StepPlan, u: torch.Tensor, scale: float) -> torch.Tensor:
    out = torch.empty_like(u)
    _lib.check(_lib.load().jsd_energy_grad(plan._h, _dev(u), _dev(out), float(scale), _stream()))
    return out


class BatchGradient:
    """J_r = E(u_T)/E(u_0) and dJ_r/du_0 for every run of a batch.

    Linearised perturbation plans: one batched forward of nsteps, the
    energies, one batched adjoint of nsteps (the linearised step's transpose
    needs no forward state, so nothing is stored or recomputed).

    Nonlinear perturbation plans (created with adjoint = True): the batched
    forward keeps a checkpoint of every run every ``checkpoint_every``
    steps; the adjoint walks the chunks backwards, recomputes each chunk's
    step-start states of all runs with batched steps and takes one batched
    adjoint step per forward step about them (the reference's checkpoint /
    recompute scheme, navier_stokes_perturbation_dual.py:730-757, 990-1035,
    applied to the discrete adjoint)."""

    def __init__(self, plans: Sequence[StepPlan], checkpoint_every: int = 100):
        self.plans = list(plans)
        dev = torch.device("cuda")
        self.u = [torch.zeros(p.shape, dtype=torch.complex128, device=dev) for p in self.plans]
        self.g = [torch.zeros_like(u) for u in self.u]
        self.batch = BatchPlan(self.plans, self.u, self.g)
        self.nonlinear = any(p.adjoint_needs_state for p in self.plans)
        self.checkpoint_every = max(1, int(checkpoint_every))

    def _nonlinear_backward(self, u0s: Sequence[torch.Tensor], nsteps: int, ckpts: List[List[torch.Tensor]]) -> None:
        k = self.checkpoint_every
        starts = list(range(0, nsteps, k))
        for ci in reversed(range(len(starts))):
            c = min(k, nsteps - starts[ci])
            for u, ck in zip(self.u, ckpts[ci]):
                u.copy_(ck)
            hist = [[u.clone()] for u in self.u]
            for _ in range(c - 1):
                self.batch.step(1)
                for h, u in zip(hist, self.u):
                    h.append(u.clone())
            for j in reversed(range(c)):
                self.batch.vjp_state([h[j] for h in hist])

    def __call__(self, u0s: Sequence[torch.Tensor], nsteps: int) -> Tuple[List[float], List[torch.Tensor]]:
        for u, u0 in zip(self.u, u0s):
            u.copy_(u0)
        ckpts: List[List[torch.Tensor]] = []
        if self.nonlinear:
            done = 0
            while done < nsteps:
                ckpts.append([u.clone() for u in self.u])
                c = min(self.checkpoint_every, nsteps - done)
                self.batch.step(c)
                done += c
        else:
            self.batch.step(nsteps)
        e0 = [p.energy(u0) for p, u0 in zip(self.plans, u0s)]
        eT = [p.energy(u) for p, u in zip(self.plans, self.u)]
        J = [a / b for a, b in zip(eT, e0)]
        for p, u, g, E0 in zip(self.plans, self.u, self.g, e0):
            g.copy_(_energy_grad(p, u, 1.0 / E0))
        if self.nonlinear:
            self._nonlinear_backward(u0s, nsteps, ckpts)
        else:
            self.batch.vjp(nsteps)
        grads = []
        for p, u0, g, j, E0 in zip(self.plans, u0s, self.g, J, e0):
            grads.append(g - _energy_grad(p, u0, j / E0))
        return J, grads

    def close(self) -> None:
        self.batch.close()


def batch_gain_and_grad(
    plans: Sequence[StepPlan], u0s: Sequence[torch.Tensor], nsteps: int, checkpoint_every: int = 100
) -> Tuple[List[float], List[torch.Tensor]]:
    bg = BatchGradient(plans, checkpoint_every)
    try:
        return bg(u0s, nsteps)
    finally:
        bg.close()
