"""Batches of independent linearised runs (jsd_batch_*): one launch per stage
for runs of uneven shape, Re, dt and box, replayed from a captured HIP graph.

* every replay equals stepping each plan on its own, with eager work and
  synchronisations between the replays (the same kernel bodies; inlined into
  the batched kernel the compiler may contract differently, so agreement is
  to rounding, <= 1e-13);
* one run of the batch against the oracle;
* the batched adjoint equals each plan's own adjoint, and the batched gain
  gradient equals the single-run one;
* plans the batch cannot hold are rejected.
"""

import numpy as np
import pytest
import torch

from jax_spectral_dns_amd import KIND_NS, KIND_PERTURBATION, _lib
from jax_spectral_dns_amd import domain_math as DM
from jax_spectral_dns_amd.batch import BatchPlan, batch_gain_and_grad
from jax_spectral_dns_amd.plan import StepPlan
from jax_spectral_dns_amd.vjp import gain_and_grad
from oracle import jsd_oracle as O
from tests.helpers import perturbation_field, rel

pytestmark = pytest.mark.gpu
DEV = "cuda"

# (shape, Re, alpha, beta, dt): uneven shapes and parameters
RUNS = [
    ((16, 33, 16), 2000.0, 1.3, 0.7, 0.01),
    ((32, 65, 16), 4000.0, 0.8, 1.9, 0.01),
    ((24, 65, 4), 3000.0, 1.0, 1.0, 0.01),
    ((16, 97, 16), 800.0, 1.7, 0.4, 0.005),
]


def _plans():
    plans, u0 = [], []
    for k, (shape, re, a, b, dt) in enumerate(RUNS):
        g = DM.make_grid(shape, 2 * np.pi / a, 2 * np.pi / b)
        plans.append(StepPlan(g, re, dt, KIND_PERTURBATION, linearise=True))
        u0.append(perturbation_field(g, seed=40 + k, energy=1e-6).to(DEV))
    return plans, u0


def test_batch_step_equals_individual_steps():
    plans, u0 = _plans()
    ind = [u.clone() for u in u0]
    bat = [u.clone() for u in u0]
    batch = BatchPlan(plans, bat)
    info = batch.info()
    print(info)
    for st in ("ygemm", "nl_lin", "solve"):
        assert 1 <= info[st]["grid"] <= info[st]["blocks"]
    # creation (autotuning) leaves the states unchanged
    for a, b in zip(bat, u0):
        assert torch.equal(a, b)
    for n in range(3):
        for p, u in zip(plans, ind):
            p.step(u, 0.0, 1)
        torch.cuda.synchronize()
        batch.step(1)
        torch.cuda.synchronize()
        for r, (a, b) in enumerate(zip(bat, ind)):
            err = rel(a, b)
            print("replay", n, "run", r, "rel", err)
            assert err <= 1e-13
    batch.close()


def test_batch_run_against_oracle():
    plans, u0 = _plans()
    bat = [u.clone() for u in u0]
    batch = BatchPlan(plans, bat)
    batch.step(2)
    shape, re, a, b, dt = RUNS[1]
    d = O.make_domain(shape, 2 * np.pi / a, 2 * np.pi / b)
    sp = O.make_perturbation(d, re, dt, True)
    ref = u0[1].cpu()
    for n in range(2):
        ref = O.rk_step(sp, ref, n)
    err = rel(bat[1], ref)
    print("run 1 vs oracle after 2 steps:", err)
    assert err < 2e-11
    batch.close()


def test_batch_vjp_equals_individual():
    plans, u0 = _plans()
    rng = np.random.default_rng(5)
    g0 = []
    for p in plans:
        x = torch.from_numpy(rng.standard_normal(p.shape) + 1j * rng.standard_normal(p.shape)).to(DEV)
        x[:, :, 0] = 0
        x[:, :, -1] = 0
        g0.append(x)
    states = [u.clone() for u in u0]
    adj = [x.clone() for x in g0]
    ind = [x.clone() for x in g0]
    batch = BatchPlan(plans, states, adj)
    for a, b in zip(adj, g0):
        assert torch.equal(a, b)
    for n in range(2):
        for p, x in zip(plans, ind):
            p.vjp(x, 1)
        torch.cuda.synchronize()
        batch.vjp(1)
        torch.cuda.synchronize()
        for r, (a, b) in enumerate(zip(adj, ind)):
            err = rel(a, b)
            print("adjoint replay", n, "run", r, "rel", err)
            assert err <= 1e-13
    batch.close()


def test_batch_gradient_equals_single_run_gradient():
    plans, u0 = _plans()
    J, G = batch_gain_and_grad(plans, u0, 20)
    for r, p in enumerate(plans):
        j1, g1 = gain_and_grad(u0[r], p, 20, checkpoint_every=7)
        print("run", r, "J", J[r], j1, "grad rel", rel(G[r], g1))
        assert abs(J[r] - j1) <= 1e-13 * abs(j1)
        assert rel(G[r], g1) < 1e-12


def test_batch_rejects_unsupported_plans():
    g = DM.make_grid((16, 33, 16), 2 * np.pi, 2 * np.pi)
    ns = StepPlan(g, 2000.0, 0.01, KIND_NS)
    u = torch.zeros(ns.shape, dtype=torch.complex128, device=DEV)
    with pytest.raises(_lib.JsdError):
        BatchPlan([ns], [u])
    nonlin = StepPlan(g, 2000.0, 0.01, KIND_PERTURBATION, linearise=False)
    with pytest.raises(_lib.JsdError):
        BatchPlan([nonlin], [u])


def test_library_rejects_repeated_plans_in_a_batch():
    """jsd_batch_create itself (bypassing the Python check) rejects a plan or
    a state buffer that appears twice: the runs would race on one plan's
    workspaces."""
    import ctypes

    g = DM.make_grid((16, 33, 16), 2 * np.pi, 2 * np.pi)
    p = StepPlan(g, 2000.0, 0.01, KIND_PERTURBATION, linearise=True)
    q = StepPlan(g, 2000.0, 0.01, KIND_PERTURBATION, linearise=True)
    u = torch.zeros(p.shape, dtype=torch.complex128, device=DEV)
    w = torch.zeros_like(u)
    lib = _lib.load()
    vp = ctypes.c_void_p
    st = vp(torch.cuda.current_stream().cuda_stream)
    for plans, states in (([p, p], [u, w]), ([p, q], [u, u])):
        hp = (vp * 2)(*[x._h for x in plans])
        sp = (vp * 2)(*[vp(s.data_ptr()) for s in states])
        h = vp()
        rc = lib.jsd_batch_create(hp, sp, None, 2, st, ctypes.byref(h))
        assert rc == -1 and not h.value


def test_cfg3_seeded_64_run_batch():
    """The batch workload itself: the 64 runs drawn from seed 0 (Ny up to 129,
    (Nx, Nz) up to (48, 32)), 20 batched steps against the same runs stepped
    one at a time (<= 1e-12), and one Ny = 129, 48 x 32 run of the draw
    against the oracle for 2 steps (<= 2e-11)."""
    from jax_spectral_dns_amd import initial
    from jax_spectral_dns_amd.batch import seeded_linearised_runs

    runs = seeded_linearised_runs(64, 0)
    assert {r[0][1] for r in runs} == {65, 97, 129}
    assert {(r[0][0], r[0][2]) for r in runs} == {(16, 16), (32, 16), (48, 32)}
    plans, u0 = [], []
    for r, (shape, re, lx, lz) in enumerate(runs):
        g = DM.make_grid(shape, lx, lz)
        plans.append(StepPlan(g, re, 0.01, KIND_PERTURBATION, linearise=True))
        plans[-1].ops = None
        u0.append(torch.from_numpy(initial.perturbation(g, seed=r, energy=1e-6)).to(DEV))
    ind = [u.clone() for u in u0]
    for p, u in zip(plans, ind):
        p.step(u, 0.0, 20)
    bat = [u.clone() for u in u0]
    batch = BatchPlan(plans, bat)
    batch.step(20)
    torch.cuda.synchronize()
    worst = max(rel(a, b) for a, b in zip(bat, ind))
    print("64-run batch vs one at a time after 20 steps: %.3e" % worst)
    assert worst <= 1e-12
    batch.close()
    k = next(i for i, r in enumerate(runs) if r[0] == (48, 129, 32))
    shape, re, lx, lz = runs[k]
    d = O.make_domain(shape, lx, lz)
    sp = O.make_perturbation(d, re, 0.01, True)
    ref = u0[k].cpu()
    for n in range(2):
        ref = O.rk_step(sp, ref, n)
    b2 = [u.clone() for u in u0]
    batch = BatchPlan(plans, b2)
    batch.step(2)
    err = rel(b2[k], ref)
    print("run %d %s vs oracle after 2 steps: %.3e" % (k, shape, err))
    assert err < 2e-11
    batch.close()
