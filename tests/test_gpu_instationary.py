"""Perturbation steps about the time-dependent base flow with
coupling_term = False (navier_stokes_perturbation_instationary.py:29-108)
against the oracle, stage and step."""

import numpy as np
import pytest
import torch

from jax_spectral_dns_amd import KIND_PERTURBATION
from jax_spectral_dns_amd.navier_stokes_perturbation_instationary import instationary_base_profile
from jax_spectral_dns_amd.plan import StepPlan
from oracle import jsd_oracle as O
from tests.helpers import pair, perturbation_field, rel

pytestmark = pytest.mark.gpu
DEV = "cuda"


@pytest.mark.parametrize("accelerate", [True, False])
def test_instationary_nonlinear_term(accelerate):
    g, d = pair((24, 33, 16), 2 * np.pi, np.pi)
    re, dt, t = 500.0, 0.01, 7
    plan = StepPlan(g, re, dt, KIND_PERTURBATION, linearise=False, coupling_term=False)
    U = instationary_base_profile(g.y, t * dt, re, 1.0, accelerate, 100)
    plan.set_base_flow(U)
    u = perturbation_field(g, seed=41, energy=1e-2)
    hv, hg, conv = plan.nonlinear(u.to(DEV))
    sp = O.make_perturbation_instationary(d, re, dt, 1.0, accelerate, 100)
    rv, rg, _, rc = sp.nonlinear(u, t)
    assert rel(hv, rv) < 1e-12
    assert rel(hg, rg) < 1e-12
    assert rel(conv, rc) < 1e-12


def test_instationary_steps():
    g, d = pair((24, 33, 16), 2 * np.pi, np.pi)
    re, dt = 500.0, 0.01
    plan = StepPlan(g, re, dt, KIND_PERTURBATION, linearise=False, coupling_term=False)
    sp = O.make_perturbation_instationary(d, re, dt, 1.0, True, 100)
    u = perturbation_field(g, seed=42, energy=1e-2)
    ref, out = u.clone(), u.to(DEV).clone()
    for k, n in enumerate(range(3, 7)):
        ref = O.rk_step(sp, ref, n)
        plan.set_base_flow(instationary_base_profile(g.y, n * dt, re, 1.0, True, 100))
        plan.step(out, 0.0, 1)
        err = rel(out, ref)
        print("step", n, "rel", err)
        assert err < 1e-11 * (k + 1)


def test_coupling_correction_changes_only_h_g():
    """coupling_term = False differs from the default only in h_g, by
    d_z hat(U' v) (navier_stokes_perturbation.py:214-224)."""
    g, d = pair((24, 33, 16), 2 * np.pi, np.pi)
    u = perturbation_field(g, seed=43, energy=1e-2).to(DEV)
    a = StepPlan(g, 1000.0, 0.01, KIND_PERTURBATION).nonlinear(u)
    b = StepPlan(g, 1000.0, 0.01, KIND_PERTURBATION, coupling_term=False).nonlinear(u)
    assert torch.equal(a[0], b[0]) and torch.equal(a[2], b[2])
    base = O.base_flow_hat(d)
    comb = O.field_hat(d, O.diff_phys_y(d, O.field_no_hat(d, base[0])) * O.field_no_hat(d, u[1].cpu()))
    assert rel(b[1] - a[1], O.diff(d, comb, 2)) < 1e-12


def test_equation_class_solve_scan():
    """NavierStokesVelVortPerturbationInstationary.solve_scan: four steps, each
    with the base flow of its time step, against the oracle."""
    from jax_spectral_dns_amd.domain import PhysicalDomain
    from jax_spectral_dns_amd.field import FourierField, VectorField
    from jax_spectral_dns_amd.navier_stokes_perturbation_instationary import NavierStokesVelVortPerturbationInstationary

    shape, lx, lz = (24, 33, 16), 2 * np.pi, np.pi
    pd = PhysicalDomain.create(shape, (True, False, True), scale_factors=(lx, 1.0, lz))
    g, d = pair(shape, lx, lz)
    u = perturbation_field(g, seed=44, energy=1e-2)
    vh = VectorField.FromData(FourierField, pd, u.to(DEV), "velocity_hat")
    nse = NavierStokesVelVortPerturbationInstationary(vh, Re=500.0, dt=0.01, end_time=0.04, kappa=1.0, accelerate=False)
    out, _, n = nse.solve_scan()
    assert n == 4
    sp = O.make_perturbation_instationary(d, 500.0, 0.01, 1.0, False, 100)
    ref = u.clone()
    for k in range(n):
        ref = O.rk_step(sp, ref, k)
    assert rel(out[-1], ref) < 4e-11


@pytest.mark.parametrize("linearise", [False, True])
def test_dual_term_without_coupling(linearise):
    """Dual term with coupling_term = False about a time-dependent profile:
    h_v gains (d_x^2 + d_z^2) hat(U' v_x); conv_ns stays the uncorrected
    one (navier_stokes_perturbation_dual.py:269-291)."""
    from jax_spectral_dns_amd import KIND_DUAL

    g, d = pair((24, 33, 16), 2 * np.pi, np.pi)
    re, dt, t = 500.0, 0.01, 5
    fwd = StepPlan(g, re, dt, KIND_PERTURBATION, linearise=linearise)
    plan = StepPlan(g, re, dt, KIND_DUAL, linearise=linearise, ops=fwd.ops, explicit_dt=-dt, plan_re=-re,
                    coupling_term=False)
    U = instationary_base_profile(g.y, t * dt, re, 1.0, True, 100)
    plan.set_base_flow(U)
    v = perturbation_field(g, seed=63, energy=1e-2)
    uf = perturbation_field(g, seed=64, energy=1e-2)
    hv, hg, conv = plan.nonlinear(v.to(DEV), None if linearise else uf.to(DEV))
    X0, X1, X2 = np.meshgrid(d.x, d.y, d.z, indexing="ij")
    zero = O.field_hat(d, torch.from_numpy(0.0 * X0))
    base = torch.stack([O.field_hat(d, torch.from_numpy(U[None, :, None] + 0.0 * X0 * X2)), zero, zero])
    rv, rg, _, rc = O.nl_dual_skew(d, v, base, uf, linearise, coupling_term=False)
    assert rel(hv, rv) < 1e-12
    assert rel(hg, rg) < 1e-12
    assert rel(conv, rc) < 1e-12


@pytest.mark.parametrize("checkpoint_every", [None, 2])
def test_instationary_dual_run(checkpoint_every):
    """NavierStokesVelVortPerturbationInstationaryDual: forward run with a
    per-step base flow, then the dual run with the base flow of forward step
    n - t and the forward history (with or without checkpointed
    recomputation), against the oracle's forward and dual steps."""
    from jax_spectral_dns_amd.domain import PhysicalDomain
    from jax_spectral_dns_amd.field import FourierField, VectorField
    from jax_spectral_dns_amd.navier_stokes_perturbation_instationary import NavierStokesVelVortPerturbationInstationary
    from jax_spectral_dns_amd.navier_stokes_perturbation_instationary_dual import (
        NavierStokesVelVortPerturbationInstationaryDual,
    )

    shape, lx, lz = (24, 33, 16), 2 * np.pi, np.pi
    re, dt, n = 500.0, 0.01, 4
    pd = PhysicalDomain.create(shape, (True, False, True), scale_factors=(lx, 1.0, lz))
    g, d = pair(shape, lx, lz)
    u0 = perturbation_field(g, seed=65, energy=1e-2)
    vh = VectorField.FromData(FourierField, pd, u0.to(DEV), "velocity_hat")
    nse = NavierStokesVelVortPerturbationInstationary(vh, Re=re, dt=dt, end_time=n * dt, kappa=1.0, accelerate=True)
    kw = {"checkpointing": False} if checkpoint_every is None else {"checkpointing": True, "checkpoint_every": checkpoint_every}
    dual = NavierStokesVelVortPerturbationInstationaryDual.FromNavierStokesVelVortPerturbation(nse, **kw)
    dual.run_forward_calculation()
    v0 = dual.run_backward_calculation()
    sp = O.make_perturbation_instationary(d, re, dt, 1.0, True, 100)
    hist = [u0.clone()]
    for k in range(n):
        hist.append(O.rk_step(sp, hist[-1], k))
    assert rel(dual.u_T, hist[-1]) < 1e-11
    sd = O.make_dual_instationary(d, re, dt, n, lambda t: hist[n - t], 1.0, True, 100)
    v = -hist[-1]
    for t in range(n):
        v = O.rk_step(sd, v, t)
    err = rel(v0, v)
    print("instationary dual rel %.3e" % err)
    assert err < 1e-10


def test_coupling_off_rejections():
    """coupling_term = False is refused where it has no meaning or no
    implementation: on NS plans, in batches, and in the discrete adjoint.
    All three are host-side checks made before any kernel runs."""
    from jax_spectral_dns_amd import KIND_DUAL, KIND_NS
    from jax_spectral_dns_amd._lib import JsdError
    from jax_spectral_dns_amd.batch import BatchPlan

    g, _ = pair((24, 33, 16), 2 * np.pi, np.pi)
    with pytest.raises(JsdError):
        StepPlan(g, 500.0, 0.01, KIND_NS, coupling_term=False)
    fwd = StepPlan(g, 500.0, 0.01, KIND_PERTURBATION, linearise=True)
    dual = StepPlan(g, 500.0, 0.01, KIND_DUAL, linearise=True, ops=fwd.ops, explicit_dt=-0.01, plan_re=-500.0,
                    coupling_term=False)
    state = torch.zeros(dual.shape, dtype=torch.complex128, device=DEV)
    with pytest.raises(JsdError):
        BatchPlan([dual], [state])
    pert = StepPlan(g, 500.0, 0.01, KIND_PERTURBATION, linearise=True, coupling_term=False)
    with pytest.raises(JsdError):
        pert.vjp(torch.zeros(pert.shape, dtype=torch.complex128, device=DEV), 1)


def test_device_profile_table_matches_host_profiles_and_captures():
    """The device table of profiles (jsd_set_base_flow_device: device-to-device
    copies, no wait) steps exactly as the host profiles of each step do, and
    a graph holding three (set profile, step) pairs replays bit for bit,
    with eager work between the replays."""
    from jax_spectral_dns_amd.domain import PhysicalDomain
    from jax_spectral_dns_amd.field import FourierField, VectorField
    from jax_spectral_dns_amd.navier_stokes_perturbation_instationary import NavierStokesVelVortPerturbationInstationary

    shape, lx, lz = (24, 33, 16), 2 * np.pi, np.pi
    pd = PhysicalDomain.create(shape, (True, False, True), scale_factors=(lx, 1.0, lz))
    g, _ = pair(shape, lx, lz)
    u0 = perturbation_field(g, seed=45, energy=1e-2).to(DEV)
    vh = VectorField.FromData(FourierField, pd, u0.clone(), "velocity_hat")
    nse = NavierStokesVelVortPerturbationInstationary(vh, Re=500.0, dt=0.01, end_time=0.06, kappa=1.0)
    table = nse.base_flow_table()
    assert table.shape == (7, 2, g.ny)
    plan = nse.plan()
    host = u0.clone()
    for k in range(6):
        plan.set_base_flow(nse.vel_base_profile(k))
        plan.step(host, 0.0, 1)
    out, _, n = nse.solve_scan()
    assert n == 6 and torch.equal(out[-1], host)
    # capture steps 0..2 through the table, replay against eager steps
    ref = u0.clone()
    for k in range(3):
        plan.set_base_flow(table[k, 0], table[k, 1])
        plan.step(ref, 0.0, 1)
    state = u0.clone()
    for k in range(3):  # warm-up before capture
        plan.set_base_flow(table[k, 0], table[k, 1])
        plan.step(state, 0.0, 1)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        for k in range(3):
            plan.set_base_flow(table[k, 0], table[k, 1])
            plan.step(state, 0.0, 1)
    for r in range(3):
        state.copy_(u0)
        plan.set_base_flow(table[5, 0], table[5, 1])  # a different profile set eagerly in between
        graph.replay()
        torch.cuda.synchronize()
        assert np.isfinite(torch.randn(1 << 16, device=DEV).sum().item())
        assert torch.equal(state, ref), r
