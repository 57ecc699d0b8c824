"""Batches of independent runs (jsd_batch_*).  Linearised runs: one launch per
stage for runs of uneven shape, Re, dt and box, replayed from a captured HIP
graph.  Nonlinear runs: concurrent streams, checked bit for bit against
single runs and their gradient against the oracle.

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


@pytest.mark.parametrize("tile", ["4", "8", "8p"])
def test_batch_solve_tile_shapes(monkeypatch, tile):
    """The tile shapes the autotuner chooses between for the batched solve
    (4 pencils on 4 waves, one MFMA column tile per wave; 8 pencils on 4
    waves, two tiles per wave; 8 pencils on 8 waves, one parity per wave),
    forced with JSD_BATCH_TILE, step every run as its own plan does, to
    rounding; the batch reports the shape it runs."""
    monkeypatch.setenv("JSD_BATCH_TILE", tile)
    plans, u0 = _plans()
    ind = [u.clone() for u in u0]
    bat = [u.clone() for u in u0]
    batch = BatchPlan(plans, bat)
    info = batch.info()
    assert info["solve"]["tile"] == (4 if tile == "4" else 8)
    assert info["solve"]["waves"] == (8 if tile == "8p" else 4)
    assert all(v["tile"] == 0 for k, v in info.items() if k != "solve")
    for p, u in zip(plans, ind):
        p.step(u, 0.0, 3)
    batch.step(3)
    torch.cuda.synchronize()
    for a, b in zip(bat, ind):
        assert rel(a, b) <= 1e-13
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
    # linearised and nonlinear runs do not share a batch
    nonlin = StepPlan(g, 2000.0, 0.01, KIND_PERTURBATION, linearise=False)
    lin = StepPlan(g, 2000.0, 0.01, KIND_PERTURBATION, linearise=True)
    with pytest.raises(_lib.JsdError):
        BatchPlan([nonlin, lin], [u, torch.zeros_like(u)])
    # the nonlinear adjoint needs plans built with adjoint = True
    with pytest.raises(_lib.JsdError):
        BatchPlan([nonlin], [u], [torch.zeros_like(u)])


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


def test_seeded_64_uneven_run_batch():
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


def _nonlinear_batch_reduced(n, adjoint=True):
    """Reduced copies of the nonlinear gradient workload: 24 x 65 x 16
    requested modes, Re 3000, Lx 2 pi, Lz pi, dt 0.005, initial energies
    log-spaced in [1e-6, 1e-2] (seeded fields), one plan per run."""
    from jax_spectral_dns_amd import initial

    g = DM.make_grid((24, 65, 16), 2 * np.pi, np.pi)
    plans, u0 = [], []
    for r, e0 in enumerate(np.logspace(-6, -2, n)):
        plans.append(StepPlan(g, 3000.0, 0.005, KIND_PERTURBATION, linearise=False, adjoint=adjoint))
        u0.append(torch.from_numpy(initial.perturbation(g, seed=70 + r, energy=float(e0))).to(DEV))
    return g, plans, u0


def test_nonlinear_batch_steps_equal_individual_steps():
    """A batch of nonlinear perturbation runs steps them concurrently, one
    stream each; every run is bit-identical to stepping its plan alone."""
    _, plans, u0 = _nonlinear_batch_reduced(8, adjoint=False)
    ind = [u.clone() for u in u0]
    for p, u in zip(plans, ind):
        p.step(u, 0.0, 6)
    bat = [u.clone() for u in u0]
    batch = BatchPlan(plans, bat)
    assert all(v["grid"] == 0 for v in batch.info().values())
    batch.step(2)
    batch.step(4)
    torch.cuda.synchronize()
    for a, b in zip(bat, ind):
        assert torch.equal(a, b)
    batch.close()


def test_nonlinear_batch_gradient_matches_single_runs_and_oracle():
    """The batched gain gradient of 8 nonlinear runs (checkpoint every 4 of 10
    steps, recompute) equals each run's own discrete adjoint, and run 0 and
    run 7 match autograd of the oracle (the stand-in for jax.grad) <= 1e-9."""
    g, plans, u0 = _nonlinear_batch_reduced(8)
    n = 10
    J, G = batch_gain_and_grad(plans, u0, n, checkpoint_every=4)
    for r in range(8):
        j1, g1 = gain_and_grad(u0[r], plans[r], n, checkpoint_every=4)
        assert J[r] == j1
        assert torch.equal(G[r], g1)
    d = O.make_domain((24, 65, 16), 2 * np.pi, np.pi)
    sp = O.make_perturbation(d, 3000.0, 0.005, False)
    for r in (0, 7):
        x = torch.view_as_real(u0[r].cpu()).clone().requires_grad_(True)
        u = torch.view_as_complex(x)
        Jr = O.energy(d, O.run(sp, u, n)) / O.energy(d, u)
        Jr.backward()
        err = rel(G[r], torch.view_as_complex(x.grad))
        print("run %d (E0 %.1e): J %.15f vs %.15f, grad rel %.3e" % (r, float(O.energy(d, u0[r].cpu())), J[r], float(Jr), err))
        assert abs(J[r] - float(Jr)) <= 1e-11 * abs(float(Jr))
        assert err < 1e-9


def test_nonlinear_gradient_full_size_against_finite_differences():
    """One run of the nonlinear gradient workload at full size (96 x 129 x 80,
    Re 3000, dt 0.005, E0 1e-2), 20 steps with a checkpoint every 10: the
    directional derivative along two seeded directions against centred
    differences (h = 1e-4 relative; the O(h^2) truncation is ~1e-8)."""
    from jax_spectral_dns_amd import initial

    g = DM.make_grid((96, 129, 80), 2 * np.pi, np.pi)
    plan = StepPlan(g, 3000.0, 0.005, KIND_PERTURBATION, linearise=False, adjoint=True)
    plan.ops = None
    u0 = torch.from_numpy(initial.perturbation(g, seed=3, energy=1e-2)).to(DEV)
    n = 20
    J, grad = batch_gain_and_grad([plan], [u0], n, checkpoint_every=10)
    J, grad = J[0], grad[0]

    def gain(u):
        v = u.clone()
        plan.step(v, 0.0, n)
        return plan.energy(v) / plan.energy(u)

    assert abs(gain(u0) - J) <= 1e-13 * J
    for k in range(2):
        rng = np.random.default_rng(80 + k)
        dirn = torch.from_numpy(initial.perturbation(g, seed=90 + k, energy=1e-2)).to(DEV)
        h = 1e-4
        fd = (gain(u0 + h * dirn) - gain(u0 - h * dirn)) / (2 * h)
        ad = float(torch.sum(grad.real * dirn.real + grad.imag * dirn.imag))
        print("dir %d fd %.10e ad %.10e rel %.2e" % (k, fd, ad, abs(fd - ad) / abs(ad)))
        assert abs(fd - ad) < 1e-6 * abs(ad)
